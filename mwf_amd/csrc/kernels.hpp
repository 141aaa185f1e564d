// kernels.hpp -- the per-step kernels of the Model Waleffe Flow timestepper.
//
// One step (program/velocity.f90:21-41) is three launches:
//   k_xinv : state c  -> U   mpt2spec + laminar + d/dx, padded inverse x-DFT
//   k_z    : U        -> NL  z c2r of u, du/dx, du/dz; udotgradu; z r2c
//   k_xfwd : NL, c    -> c   forward x-DFT, spec2mpt, AB2 + implicit update
//
// Device layouts (complex values interleaved):
//   c        double2 [NN][K][ldM]      m in the reference order (m<0 at M+m)
//   onl      cplx<R> [NN][K][ldM]      AB2 history in the precision of the
//                                      transforms that produce it
//   U        cplx<R> [P][2*KK][NN]     x-physical, z-spectral; lines 0..KK-1
//                                      are u (spec slot order), KK..2KK-1 du/dx
//   NL       cplx<R> [P][KK][NN]       x-physical, z-spectral nonlinear term
//
// Padded transforms.  The x transform of length P = 3*MM over modes
// |m| <= MM-1 is split with x = 3a + b into three length-MM transforms of the
// folded spectra Y_b[r] = X[r] w^{rb} + X[r-MM] w^{(r-MM)b}, w = exp(2 pi i/P)
// (inverse), and the forward one into three length-MM transforms of the
// decimated rows followed by X[m] = sum_b w^{-mb} F_b[m mod MM].  The z
// transforms (length Q = 3*NN, real data) fold the same way; two real rows
// share one complex transform.
#pragma once
#include "common.hpp"
#include "fft.hpp"
#include "galerkin_gen.hpp"

constexpr int XT = 512;     // threads per block of the x kernels
constexpr int LOG_XTILE = 11;
constexpr int XTILE = 1 << LOG_XTILE; // NTx * MM: x-kernel tile (complex values per line)
constexpr int ZT = 512;     // threads per block of the z kernels (= XB * NN)
constexpr int ZPB = 2;      // component pairs per z batch

// x wavenumber tables of program/variables.f90:516-523 for storage index mi.
__device__ __forceinline__ void xwave(const Geo& g, int mi, double& am1, double& am2)
{
    const int ma = mi < g.MM ? mi : g.M - mi; // |m|
    const double a = g.alpha * ma;
    am1 = mi < g.MM ? a : -a;
    am2 = -(a * g.alpha * ma);
}

__device__ __forceinline__ void zwave(const Geo& g, int n, double& an1, double& an2)
{
    const double c = g.gamma * n;
    an1 = c;
    an2 = -(c * g.gamma * n);
}

// Phase b of the folded spectrum of the padded inverse x transform,
// Y_b[r] = X[r] w^{rb} + X[r-MM] w^{(r-MM)b} with w1 = w^{rb}, w2 = w^{(MM-r)b}
// (conjugated), in this operation order and without contraction: every
// kernel that builds a U line (the fused step, its inverse-only form, k_xinv)
// then produces the same bits, whatever each compilation would fuse.
__device__ __forceinline__ double2 xfold_b(double2 X1, double2 X2, double2 w1, double2 w2)
{
#pragma clang fp contract(off)
    return double2{(X1.x * w1.x - X1.y * w1.y) + (X2.x * w2.x + X2.y * w2.y),
                   (X1.x * w1.y + X1.y * w1.x) + (X2.y * w2.x - X2.x * w2.y)};
}

// var_mpt2spec (program/variables.f90:390-434) for y-mode k at one (m, n):
// from psi_k (slot k) and phi_k (slot K0-1+k, k >= 1) to u_k, v_k, w_k.
// At (m,n) = (0,0) the mean-mode rows apply (:418-433) and, when lam is set,
// the laminar profile of vel_addlam (program/velocity.f90:76-82).
__device__ __forceinline__ void mpt2spec_k(const Geo& g, int k, int mi, int n, double2 psi, double2 phi,
                                           bool lam, double2& u, double2& v, double2& w)
{
#pragma clang fp contract(off)
    double am1, am2, an1, an2;
    xwave(g, mi, am1, am2);
    zwave(g, n, an1, an2);
    const double ak1 = ak1_of(k, g.beta);
    if (k == 0) phi = double2{0.0, 0.0};
    u = double2{-(-(psi.y * an1) + phi.y * ak1 * am1), (-(psi.x * an1) + phi.x * ak1 * am1)};
    v = double2{-(phi.x * (am2 + an2)), -(phi.y * (am2 + an2))};
    w = double2{-(psi.y * am1 + phi.y * ak1 * an1), (psi.x * am1 + phi.x * ak1 * an1)};
    if (mi == 0 && n == 0) {
        if (k == 0) {
            u = double2{0.0, 0.0};
            w = double2{0.0, 0.0};
        } else {
            u = psi;
            w = phi;
            if (lam && k == 1) {
                u.x = u.x + g.cth;
                w.x = w.x + g.sth;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// x inverse: block (n-tile, y-mode k).  Writes U for the k-group lines and,
// when epart is set, the block's share of the perturbation energy (Parseval
// of the spec fields without the laminar profile, program/velocity.f90:115-162).
// ---------------------------------------------------------------------------
template <typename R, int LOGMM, bool FLUX>
__global__ __launch_bounds__(XT) void k_xinv(Batch bt, const double2* __restrict__ c, cplx_t<R>* __restrict__ U,
                                             double* __restrict__ epart, const cplx_t<R>* __restrict__ twMM,
                                             const double2* __restrict__ twP, int addlam)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.z];
    c += blockIdx.z * bt.c_stride;
    if (U) U += blockIdx.z * bt.u_stride;
    if (epart) epart += blockIdx.z * bt.ep_stride;
    constexpr int MM = 1 << LOGMM;
    constexpr int LP = lpad(MM);
    constexpr int IPT = XTILE / XT; // items per thread
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);
    __shared__ double red[XT];

    const int NT = g.NTx;
    const int n0 = xcd_tile(blockIdx.x, gridDim.x) * NT;
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    const int nitems = NT * MM;
    const int nlines3 = k == 0 ? 2 : 3;
    const int slot_psi = k, slot_phi = g.K0 - 1 + k;
    const int spec_slot[3] = {k, g.K0 - 1 + k, 2 * g.K0 - 1 + k};

    // spec values at m = r (h = 0) and m = r - MM (h = 1): u, v, w
    double2 sp[IPT][2][3];
    double esum = 0.0;
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int it = tid + q * XT;
        if (it >= nitems) continue;
        const int t = it >> LOGMM, r = it & (MM - 1);
        const int n = n0 + t;
        const double2* cn = c + (size_t)n * g.K * g.ldM;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int mi = h == 0 ? r : MM - 1 + r;
            if (h == 1 && r == 0) {
                sp[q][h][0] = sp[q][h][1] = sp[q][h][2] = double2{0.0, 0.0};
                continue;
            }
            const double2 psi = cn[(size_t)slot_psi * g.ldM + mi];
            const double2 phi = k > 0 ? cn[(size_t)slot_phi * g.ldM + mi] : double2{0.0, 0.0};
            double2 u, v, w;
            mpt2spec_k(g, k, mi, n, psi, phi, false, u, v, w);
            if (epart) {
                // y weights of epos: 1 for the k = 0 mean of u and w, 1/2 otherwise
                const double yw = k == 0 ? 1.0 : 0.5;
                if (n > 0) {
                    esum += 2.0 * yw * (u.x * u.x + u.y * u.y + w.x * w.x + w.y * w.y);
                    if (k > 0) esum += 2.0 * yw * (v.x * v.x + v.y * v.y);
                } else if (h == 0) {
                    // n = 0: the c2r keeps the Hermitian part of the column
                    if (r == 0) {
                        esum += yw * (u.x * u.x + w.x * w.x);
                    } else {
                        const int mneg = g.M - r;
                        const double2 psn = cn[(size_t)slot_psi * g.ldM + mneg];
                        const double2 phn = k > 0 ? cn[(size_t)slot_phi * g.ldM + mneg] : double2{0.0, 0.0};
                        double2 un, vn, wn;
                        mpt2spec_k(g, k, mneg, n, psn, phn, false, un, vn, wn);
                        const double hu_r = 0.5 * (u.x + un.x), hu_i = 0.5 * (u.y - un.y);
                        const double hv_r = 0.5 * (v.x + vn.x), hv_i = 0.5 * (v.y - vn.y);
                        const double hw_r = 0.5 * (w.x + wn.x), hw_i = 0.5 * (w.y - wn.y);
                        esum += 2.0 * yw * (hu_r * hu_r + hu_i * hu_i + hw_r * hw_r + hw_i * hw_i);
                        if (k > 0) esum += 2.0 * yw * (hv_r * hv_r + hv_i * hv_i);
                    }
                }
            }
            if (addlam) mpt2spec_k(g, k, mi, n, psi, phi, true, u, v, w);
            sp[q][h][0] = u;
            sp[q][h][1] = v;
            sp[q][h][2] = w;
        }
    }

    if (epart) {
        red[tid] = esum;
        __syncthreads();
        for (int s = XT / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) epart[blockIdx.y * gridDim.x + (n0 >> g.lgNT)] = red[0];
    }
    if (!U) return;

    // advective form: u, v, w and their x-derivatives; flux form: u, v, w only
#pragma unroll
    for (int li = 0; li < (FLUX ? 3 : 6); li++) {
        const int comp = li % 3, deriv = li / 3;
        if (nlines3 == 2 && comp == 1) continue;
        const int line = spec_slot[comp] + deriv * g.KK;
        // fold into Y_b[r], b = 0..2
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int it = tid + q * XT;
            if (it >= nitems) continue;
            const int t = it >> LOGMM, r = it & (MM - 1);
            double2 X1 = sp[q][0][comp], X2 = sp[q][1][comp];
            if (deriv) { // i alpha m (program/variables.f90:606-607)
                double am1, am2;
                xwave(g, r, am1, am2);
                X1 = double2{-(X1.y * am1), X1.x * am1};
                xwave(g, MM - 1 + r, am1, am2);
                X2 = double2{-(X2.y * am1), X2.x * am1};
            }
            CT* bt = buf + (size_t)t * 3 * LP;
            bt[pidx(r)] = CT{(R)(X1.x + X2.x), (R)(X1.y + X2.y)};
#pragma unroll
            for (int b = 1; b < 3; b++) {
                const double2 w1 = twP[r * b];
                const double2 w2 = twP[(MM - r) * b]; // conj: exp(2 pi i (r-MM) b / P)
                const double2 y = xfold_b(X1, X2, w1, w2);
                bt[b * LP + pidx(r)] = CT{(R)y.x, (R)y.y};
            }
        }
        __syncthreads();
        ip_fft<CT, LOGMM, +1, XT, 3 * (XTILE >> LOGMM)>(buf, 3 * NT, twMM);
        // write U[x][line][n0 + t]
        CT* Ul = U + (size_t)line * g.NN + n0;
        const int ustride = g.UL * g.NN;
#pragma unroll
        for (int q = 0; q < 3 * IPT; q++) {
            const int it = tid + q * XT;
            if (it >= 3 * nitems) continue;
            const int x = it >> g.lgNT, t = it & (NT - 1);
            const int a = x / 3, b = x - 3 * a;
            Ul[x * ustride + t] = buf[((size_t)t * 3 + b) * LP + pidx(a)];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// z fused: block of XB = ZT/NN x-rows; thread (row xl, a) owns the physical
// points z = 3a + b, b = 0..2 of its row.  The 3*KK input rows (u, du/dx,
// du/dz) are taken in batches of ZP pairs.  A fold task (row, pair, r) with
// r in [0, NN/2) holds X[r] and X[NN-r] of both rows of the pair and writes
// the folded spectra at r and NN-r (r = 0 covers 0 and NN/2); its four values
// for the next batch are loaded from HBM while the current batch is
// transformed and consumed.
// ---------------------------------------------------------------------------
template <typename R> constexpr int zpairs() { return sizeof(R) == 8 ? 2 : 4; }

// Order of the 2*NP real rows that the z kernel transforms, two per complex
// sequence (pair pr = rows 2pr, 2pr+1).  A row is a base value (kind 0, spec
// slot s), an x-derivative (kind 1, U line KK+s), a z-derivative (kind 2,
// i*gamma*n times base line s) or zero padding (kind 3).  The w rows come
// first, so that every z-derivative can be folded in as it arrives (dz<s>
// needs only w); base u/v rows are paired with their own z-derivative and
// loaded once; the x-derivatives (which need only u) come last.  The
// z-derivatives of w (a second read of the w lines) start at pair D0, any
// position from NW to NW + NS; the kernels put them at the first pair of the
// batch after the w rows (zorder_d0), so that the second read follows the
// first by one batch and still hits L2.
template <int K0, int D0_> struct ZOrder {
    static_assert(K0 % 2 == 0, "w rows are paired among themselves");
    static constexpr int KK = 3 * K0 - 1;
    static constexpr int NW = K0 / 2;          // pairs of w rows (base, then dz)
    static constexpr int NS = 2 * K0 - 1;      // (base, dz) pairs of u and v
    static constexpr int NX = (KK + 1) / 2;    // pairs of x-derivatives
    static constexpr int NP = 2 * NW + NS + NX;
    static constexpr int W0 = 2 * K0 - 1;      // spec slot of w_0
    static constexpr int D0 = D0_;                // first dz-w pair
    static_assert(D0 >= NW && D0 <= NW + NS, "dz-w pairs between the w rows and the dx rows");
    __host__ __device__ static constexpr bool dzw(int pr) { return pr >= D0 && pr < D0 + NW; }
    // (base, dz) u/v pairs fill [NW, 2 NW + NS) around the dz-w pairs
    __host__ __device__ static constexpr bool shared(int pr) { return pr >= NW && pr < 2 * NW + NS && !dzw(pr); }
    __host__ __device__ static constexpr int kind(int q)
    {
        const int pr = q >> 1, h = q & 1;
        if (pr < NW) return 0;
        if (dzw(pr)) return 2;
        if (shared(pr)) return h == 0 ? 0 : 2;
        if (pr < NP && 2 * (pr - 2 * NW - NS) + h < KK) return 1;
        return 3;
    }
    __host__ __device__ static constexpr int slot(int q)
    {
        const int pr = q >> 1, h = q & 1;
        if (pr < NW) return W0 + 2 * pr + h;
        if (dzw(pr)) return W0 + 2 * (pr - D0) + h;
        if (shared(pr)) return pr < D0 ? pr - NW : pr - 2 * NW;
        return 2 * (pr - 2 * NW - NS) + h;
    }
};

// Every row appears once; each dz row after all w base rows, each dx row after
// all u base rows (the order in which the z kernel consumes them).
// first pair of the batch after the one holding the last w row (batches of ZP pairs)
template <int K0, int ZP> constexpr int zorder_d0()
{
    constexpr int NW = K0 / 2, NS = 2 * K0 - 1;
    constexpr int d = (NW + ZP - 1) / ZP * ZP;
    return d < NW + NS ? d : NW + NS;
}

template <int K0, int D0> constexpr bool zorder_ok()
{
    using ZO = ZOrder<K0, D0>;
    constexpr int KK = ZO::KK;
    int cnt[3][KK] = {};
    int last_w = -1, last_u = -1, first_dz = 1 << 30, first_dx = 1 << 30;
    for (int q = 0; q < 2 * ZO::NP; q++) {
        const int k = ZO::kind(q), sl = ZO::slot(q);
        if (k == 3) continue;
        if (k < 0 || k > 3 || sl < 0 || sl >= KK) return false;
        cnt[k][sl]++;
        if (k == 0 && sl >= ZO::W0) last_w = q;
        if (k == 0 && sl < K0) last_u = q;
        if (k == 2 && q < first_dz) first_dz = q;
        if (k == 1 && q < first_dx) first_dx = q;
    }
    for (int k = 0; k < 3; k++)
        for (int sl = 0; sl < KK; sl++)
            if (cnt[k][sl] != 1) return false;
    for (int pr = 0; pr < ZO::NP; pr++)
        if (ZO::shared(pr) && (ZO::kind(2 * pr) != 0 || ZO::kind(2 * pr + 1) != 2 ||
                               ZO::slot(2 * pr) != ZO::slot(2 * pr + 1)))
            return false;
    return last_w < first_dz && last_u < first_dx;
}
static_assert(zorder_ok<4, 2>() && zorder_ok<4, 4>() && zorder_ok<4, 9>() && zorder_ok<6, 3>() &&
                  zorder_ok<6, 4>() && zorder_ok<6, 14>(),
              "z row order");

// i gamma n v (program/variables.f90:608-609)
template <typename R>
__device__ __forceinline__ cplx_t<R> zderiv(const Geo& g, cplx_t<R> v, int n)
{
    const R an1 = (R)(g.gamma * n);
    return cplx_t<R>{-(v.y * an1), v.x * an1};
}

template <typename R, int K0>
__device__ __forceinline__ cplx_t<R> zrow(const Geo& g, const cplx_t<R>* __restrict__ U, int x, int kind, int s, int n)
{
    constexpr int KK = 3 * K0 - 1;
    if (kind == 3) return cplx_t<R>{(R)0, (R)0};
    const cplx_t<R> v = U[uidx(g, x, kind == 1 ? KK + s : s, n)];
    return kind == 2 ? zderiv<R>(g, v, n) : v;
}

// Radix-3 fold of one phase b of the 3/2-padded z inverse (c2r of length
// Q = 3 NN, program/transform.fftw5.f90:83): physical z = 3 r + b is the NN-point
// inverse FFT of Y_b[s] = X[s] w^{sb} + conj(X[NN-s]) w^{-(NN-s)b}, w = exp(2 pi i/Q).
// Two real rows share one complex sequence Z = Y(row 1) + i Y(row 2).  Fold task
// fr < NN/2 writes Z[fr] and Z[NN-fr] from a = X[fr], m = X[NN-fr] of both rows
// (fr = 0: the DC bin with its imaginary part dropped, and the self-mirrored
// NN/2 bin, Y = 2 Re(X w^{(NN/2) b})).
// wr = w^{fr b} and wm = w^{(NN-fr) b} (fr = 0: wm = w^{(NN/2) b}); unused for b = 0.
template <typename R, int LS = 0>
__device__ __forceinline__ void zfold_bw(cplx_t<R>* __restrict__ seq, int b, int fr, int NN, cplx_t<R> a1,
                                         cplx_t<R> b1, cplx_t<R> a2, cplx_t<R> b2, cplx_t<R> wr, cplx_t<R> wm)
{
    using CT = cplx_t<R>;
    const int H = NN / 2;
    if (fr == 0) {
        seq[lslot<LS>(0)] = CT{a1.x, a2.x};
        if (b == 0) {
            seq[lslot<LS>(H)] = CT{(R)2 * b1.x, (R)2 * b2.x};
        } else {
            const CT p1 = cmul(b1, wm), p2 = cmul(b2, wm);
            seq[lslot<LS>(H)] = CT{(R)2 * p1.x, (R)2 * p2.x};
        }
        return;
    }
    const int fm = NN - fr;
    CT P1a = a1, P2a = b1, P1b = a2, P2b = b2; // X[s] w^{sb}, X[NN-s] w^{(NN-s)b}
    if (b > 0) {
        const CT w1 = wr, w2 = wm;
        P1a = cmul(a1, w1);
        P2a = cmul(b1, w2);
        P1b = cmul(a2, w1);
        P2b = cmul(b2, w2);
    }
    // row 1: Y[fr] = P1a + conj(P2a), Y[fm] = P2a + conj(P1a); row 2 likewise
    const CT y1r = CT{P1a.x + P2a.x, P1a.y - P2a.y}, y1m = CT{P2a.x + P1a.x, P2a.y - P1a.y};
    const CT y2r = CT{P1b.x + P2b.x, P1b.y - P2b.y}, y2m = CT{P2b.x + P1b.x, P2b.y - P1b.y};
    seq[lslot<LS>(fr)] = CT{y1r.x - y2r.y, y1r.y + y2r.x};
    seq[lslot<LS>(fm)] = CT{y1m.x - y2m.y, y1m.y + y2m.x};
}

template <typename R>
__device__ __forceinline__ void zfold_b(cplx_t<R>* __restrict__ seq, int b, int fr, int NN, cplx_t<R> a1,
                                        cplx_t<R> b1, cplx_t<R> a2, cplx_t<R> b2,
                                        const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    CT wr{(R)1, (R)0}, wm{(R)1, (R)0};
    if (b > 0) {
        wr = twQ[fr * b];
        wm = twQ[(fr == 0 ? NN / 2 : NN - fr) * b];
    }
    zfold_bw<R>(seq, b, fr, NN, a1, b1, a2, b2, wr, wm);
}

// w^k = exp(2 pi i k/Q), Q = 3 NN, for 0 <= k < 2 NN from a table of the
// first NN entries: w^{NN + j} = exp(2 pi i/3) w^j.
template <typename R>
__device__ __forceinline__ cplx_t<R> twq2(const cplx_t<R>* __restrict__ tw, int NN, int k)
{
    using CT = cplx_t<R>;
    if (k < NN) return tw[k];
    const CT w = tw[k - NN];
    const R c = (R)-0.5, sn = (R)0.86602540378443864676; // exp(2 pi i/3)
    return CT{c * w.x - sn * w.y, c * w.y + sn * w.x};
}

// two transform buffers, then the pass twiddle table of the NN-point
// transforms (ptw_size entries) and the first 2 NN entries of
// exp(2 pi i j/Q) (fold and r2c unpacking use indices below 2 NN).  Twiddles
// read from LDS instead of global memory keep their latency off every pass.
template <typename R, int K0, int LOGNN> constexpr size_t z_lds_bytes()
{
    constexpr int NN = 1 << LOGNN, XB = ZT / NN, ZP = zpairs<R>();
    return sizeof(cplx_t<R>) * (2 * (size_t)XB * ZP * 3 * lpad(NN) + ptw_size<NN>() + 2 * (size_t)NN);
}

template <typename R, int K0, int LOGNN>
__global__ __launch_bounds__(ZT) void k_z(Batch bt, const cplx_t<R>* __restrict__ U, cplx_t<R>* __restrict__ NL,
                                          const cplx_t<R>* __restrict__ twNN, const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.z];
    U += blockIdx.z * bt.u_stride;
    NL += blockIdx.z * bt.nl_stride;
    using G = Gal<K0>;
    constexpr int KK = 3 * K0 - 1;
    constexpr int NN = 1 << LOGNN;
    constexpr int XB = ZT / NN;
    using ZO = ZOrder<K0, zorder_d0<K0, zpairs<R>()>()>;
    static_assert(zorder_ok<K0, ZO::D0>(), "z row order");
    constexpr int ZP = zpairs<R>();                 // pairs per batch
    constexpr int NBI = (ZO::NP + ZP - 1) / ZP;
    constexpr int NPO = (KK + 1) / 2, NBO = (NPO + ZP - 1) / ZP;
    constexpr int SEQ = XB * ZP * 3;
    constexpr int LP = lpad(NN);
    constexpr int H = NN / 2;
    constexpr int FT = XB * ZP * H / ZT;            // fold tasks per thread
    constexpr int GL = group_lanes<NN>();           // lanes per transform group
    // FP32: one group of GL lanes per sequence (below); FP64: block-wide
    // ping-pong passes, whose butterflies need fewer live registers next to
    // the 2 * 3 * KK double point values
    constexpr bool GRP = sizeof(R) == 4;
    constexpr bool INB = fft_in_b<LOGNN>();
    static_assert(XB * ZP * H % ZT == 0, "fold tasks must tile the block");
    extern __shared__ unsigned char smem_raw[];
    CT* A = reinterpret_cast<CT*>(smem_raw);        // [SEQ][LP]
    CT* B = A + SEQ * LP;                           // [SEQ][LP]
    CT* twNl = B + SEQ * LP;                        // [ptw_size<NN>()] pass twiddles
    CT* twQl = twNl + ptw_size<NN>();               // [2 NN]
    // Group form: each sequence of a batch is transformed by its own group of
    // GL lanes (group_fft: 8 points per lane, wave-local exchanges), so the
    // waves run their transforms without block barriers and overlap each
    // other's LDS and VALU work; batches alternate between A and B.  Ping-pong
    // form: folds go to A, results come back in B when the pass count is odd.
    // With one x-row per block (XB = 1) the sequences of the valid pairs of a
    // batch come first, so a batch whose last pairs are padding transforms
    // only nv pairs (3 nv sequences).
    auto nseq = [](int nv) { return XB == 1 ? 3 * nv : SEQ; };
    auto seq_fft = [&](CT* buf, auto sign, int S) {
        const int gl = threadIdx.x % GL;
        for (int s = threadIdx.x / GL; s < S; s += ZT / GL) {
            CT* sr = buf + s * LP;
            CT v[NN / GL];
#pragma unroll
            for (int q = 0; q < NN / GL; q++) v[q] = sr[pidx(gl + GL * q)];
            group_fft<CT, LOGNN, decltype(sign)::value, true>(v, sr, gl, twNl);
#pragma unroll
            for (int q = 0; q < NN / GL; q++) sr[pidx(gl + GL * q)] = v[q];
        }
    };

    const int tid = threadIdx.x;
    const int xl = tid >> LOGNN, r = tid & (NN - 1);
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * XB;
    const int x = x0 + xl;
    const bool valid = x < g.P;
    const bool wide = g.NTs * (int)sizeof(CT) >= 64;

    // fold task f of this thread: tau = tid + ZT*f -> (row fx, pair fp, index fr)
    CT pf[FT][4];
    auto prefetch = [&](int batch) {
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int tau = tid + ZT * f;
            const int fx = tau / (ZP * H), fp = (tau / H) % ZP, fr = tau % H;
            const int xg = x0 + fx;
            const int pr = batch * ZP + fp;
            const int r2 = fr == 0 ? H : NN - fr;
            if (xg < g.P) {
                const int ka = ZO::kind(2 * pr), sa = ZO::slot(2 * pr);
                pf[f][0] = zrow<R, K0>(g, U, xg, ka, sa, fr);
                pf[f][1] = zrow<R, K0>(g, U, xg, ka, sa, r2);
                if (ZO::shared(pr)) {
                    pf[f][2] = zderiv<R>(g, pf[f][0], fr);
                    pf[f][3] = zderiv<R>(g, pf[f][1], r2);
                } else {
                    const int kb = ZO::kind(2 * pr + 1), sb = ZO::slot(2 * pr + 1);
                    pf[f][2] = zrow<R, K0>(g, U, xg, kb, sb, fr);
                    pf[f][3] = zrow<R, K0>(g, U, xg, kb, sb, r2);
                }
            } else {
                pf[f][0] = pf[f][1] = pf[f][2] = pf[f][3] = CT{(R)0, (R)0};
            }
        }
    };

    R vb[3][KK];
    R an[3][KK];
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int s = 0; s < KK; s++) an[b][s] = (R)0;

    prefetch(0);
    ptw_fill<CT, LOGNN>(twNl, twNN, tid, ZT);
    for (int i = tid; i < 2 * NN; i += ZT) twQl[i] = twQ[i];
    __syncthreads();
    static_for<0, NBI>([&](auto bi) {
        constexpr int BN = decltype(bi)::value;
        constexpr int NV = ZO::NP - BN * ZP < ZP ? ZO::NP - BN * ZP : ZP; // valid pairs
        // fold: Y_b[s] = X[s] w^{sb} + conj(X[NN-s]) w^{-(NN-s)b}, w = exp(2 pi i/Q),
        // two real rows per complex sequence Z = Y(row 1) + i Y(row 2)
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int tau = tid + ZT * f;
            const int fx = tau / (ZP * H), fp = (tau / H) % ZP, fr = tau % H;
            if (XB == 1 && fp >= NV) continue;
            CT* seq = (GRP && (BN & 1) ? B : A) + ((fx * ZP + fp) * 3) * LP;
#pragma unroll
            for (int b = 0; b < 3; b++)
                zfold_b<R>(seq + b * LP, b, fr, NN, pf[f][0], pf[f][1], pf[f][2], pf[f][3], twQl);
        }
        if constexpr (BN + 1 < NBI) prefetch(BN + 1);
        __syncthreads();
        CT* F;
        if constexpr (GRP) {
            F = BN & 1 ? B : A;
            seq_fft(F, std::integral_constant<int, +1>{}, nseq(NV));
            __syncthreads();
        } else {
            pp_fft<CT, LOGNN, +1, ZT, true>(A, B, nseq(NV), twNl);
            F = INB ? B : A;
        }
        // consume: thread (xl, a = r)
        static_for<0, ZP>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const CT zv = F[((xl * ZP + p) * 3 + b) * LP + pidx(r)];
                static_for<0, 2>([&](auto hi) {
                    constexpr int q = BN * 2 * ZP + 2 * p + decltype(hi)::value;
                    constexpr int kd = ZO::kind(q), s = ZO::slot(q);
                    const R val = decltype(hi)::value == 0 ? zv.x : zv.y;
                    if constexpr (kd == 0) vb[b][s] = val;
                    else if constexpr (kd == 1) G::template dx<s>(vb[b], val, an[b]);
                    else if constexpr (kd == 2) G::template dz<s>(vb[b], val, an[b]);
                });
            }
        });
        if constexpr (!GRP && !INB) __syncthreads();
    });
#pragma unroll
    for (int b = 0; b < 3; b++) G::yterms(vb[b], an[b], (R)g.beta);

    // forward z transform of the KK products, two per complex sequence
    const R scale = (R)g.scale;
    static_for<0, NBO>([&](auto ri) {
        constexpr int RD = decltype(ri)::value;
        constexpr int NV = NPO - RD * ZP < ZP ? NPO - RD * ZP : ZP; // valid output pairs
        CT* F = GRP && ((NBI + RD) & 1) ? B : A;
        static_for<0, ZP>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int c1 = 2 * (RD * ZP + p), c2 = c1 + 1;
            if constexpr (c1 < KK) {
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    R v2 = (R)0;
                    if constexpr (c2 < KK) v2 = an[b][c2];
                    F[((xl * ZP + p) * 3 + b) * LP + pidx(r)] = CT{an[b][c1], v2};
                }
            }
        });
        __syncthreads();
        if constexpr (GRP) {
            seq_fft(F, std::integral_constant<int, -1>{}, nseq(NV));
            __syncthreads();
        } else {
            pp_fft<CT, LOGNN, -1, ZT, true>(A, B, nseq(NV), twNl);
            F = INB ? B : A;
        }
        static_for<0, ZP>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int c1 = 2 * (RD * ZP + p), c2 = c1 + 1;
            if constexpr (c1 < KK) {
                CT X1{(R)0, (R)0}, X2{(R)0, (R)0};
                const int rm = (NN - r) & (NN - 1);
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    const CT zn = F[((xl * ZP + p) * 3 + b) * LP + pidx(r)];
                    const CT zm = F[((xl * ZP + p) * 3 + b) * LP + pidx(rm)];
                    const CT F1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                    const CT F2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                    if (b == 0) {
                        X1 = F1;
                        X2 = F2;
                    } else {
                        const CT w = cconj(twQl[r * b]);
                        X1 = cadd(X1, cmul(F1, w));
                        X2 = cadd(X2, cmul(F2, w));
                    }
                }
                if (valid) {
                    st_piece(&NL[nidx(g, x, c1, r)], cscale(X1, scale), wide);
                    if constexpr (c2 < KK) st_piece(&NL[nidx(g, x, c2, r)], cscale(X2, scale), wide);
                }
            }
        });
        if constexpr (!GRP && !INB) __syncthreads();
    });
}

// ---------------------------------------------------------------------------
// z fused with two points per thread.  A block of Z2T = 768 threads covers
// XB = 512/NN x-rows, 1536 physical points; thread tid owns the flat points
// f = tid and tid + Z2T, f = (xl*3 + b)*NN + a (row xl, z = 3a + b), so each
// thread keeps 2 (not 3) points of base values and accumulators and the
// block runs 3 waves per SIMD instead of 2.  Batches hold Z2P = 3 pairs
// (9 sequences per x-row, 576 radix-8 butterflies per pass: most threads
// busy).  The fold, transforms, products and r2c unpacking are those of k_z;
// the passes ping-pong between two buffers.  The fold and unpack twiddles
// w^k, k < 2 NN, come from an LDS copy of the first NN entries of the Q
// table (twq2), which keeps the block within 160 KB at NN = 512 in FP64.
// ---------------------------------------------------------------------------
constexpr int Z2T = 768;
constexpr int Z2P = 3;

template <typename R, int LOGNN> constexpr size_t z2_lds_bytes()
{
    constexpr int NN = 1 << LOGNN, XB = 512 / NN;
    constexpr size_t seq = (size_t)XB * Z2P * 3 * lpad(NN);
    return sizeof(cplx_t<R>) * (2 * seq + ptw_size<NN>() + (size_t)NN);
}

template <typename R, int K0, int LOGNN>
__global__ __launch_bounds__(Z2T) void k_z2(Batch bt, const cplx_t<R>* __restrict__ U, cplx_t<R>* __restrict__ NL,
                                            const cplx_t<R>* __restrict__ twNN, const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.z];
    U += blockIdx.z * bt.u_stride;
    NL += blockIdx.z * bt.nl_stride;
    using G = Gal<K0>;
    constexpr int KK = 3 * K0 - 1;
    constexpr int NN = 1 << LOGNN;
    constexpr int XB = 512 / NN;
    using ZO = ZOrder<K0, zorder_d0<K0, Z2P>()>;
    static_assert(zorder_ok<K0, ZO::D0>(), "z row order");
    constexpr int ZP = Z2P;
    constexpr int NBI = (ZO::NP + ZP - 1) / ZP;
    constexpr int NPO = (KK + 1) / 2, NBO = (NPO + ZP - 1) / ZP;
    constexpr int SEQ = XB * ZP * 3;
    constexpr int LP = lpad(NN);
    constexpr int H = NN / 2;
    constexpr int PT = 3 * XB * NN / Z2T;      // points per thread
    constexpr int UT = XB * ZP * NN / Z2T;     // r2c unpack tasks per thread
    constexpr bool INB = fft_in_b<LOGNN>();
    static_assert(XB * ZP * H == Z2T, "one fold task per thread");
    static_assert(PT * Z2T == 3 * XB * NN && UT * Z2T == XB * ZP * NN, "points and tasks tile the block");
    extern __shared__ unsigned char smem_raw[];
    CT* A = reinterpret_cast<CT*>(smem_raw);   // [SEQ][LP]
    CT* B = A + SEQ * LP;                      // [SEQ][LP]
    CT* twNl = B + SEQ * LP;                   // [ptw_size<NN>()] pass twiddles
    CT* twQl = twNl + ptw_size<NN>();          // [NN]: w^k, k < NN

    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * XB;
    const bool wide = g.NTs * (int)sizeof(CT) >= 64;
    // Thread-derived indices are recomputed per batch from an opaque copy of
    // the thread index, so that the compiler does not hoist every batch's
    // addresses into registers next to the point state.
    auto opaque_tid = [] {
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        return t;
    };
    // owned point j: LDS offset of (xl, pair 0, b, a); pair p adds 3 p LP
    auto poff = [](int t, int j) {
        const int f = t + Z2T * j;
        const int xl = f / (3 * NN), b = (f / NN) % 3, a = f % NN;
        return (xl * ZP * 3 + b) * LP + pidx(a);
    };

    // fold task of thread t: (row fx, pair fp, index fr)
    CT pf[4];
    auto prefetch = [&](int t, int batch) {
        const int fx = t / (ZP * H), fp = (t / H) % ZP, fr = t % H;
        const int xg = x0 + fx;
        const int pr = batch * ZP + fp;
        const int r2 = fr == 0 ? H : NN - fr;
        if (xg < g.P) {
            const int ka = ZO::kind(2 * pr), sa = ZO::slot(2 * pr);
            pf[0] = zrow<R, K0>(g, U, xg, ka, sa, fr);
            pf[1] = zrow<R, K0>(g, U, xg, ka, sa, r2);
            if (ZO::shared(pr)) {
                pf[2] = zderiv<R>(g, pf[0], fr);
                pf[3] = zderiv<R>(g, pf[1], r2);
            } else {
                const int kb = ZO::kind(2 * pr + 1), sb = ZO::slot(2 * pr + 1);
                pf[2] = zrow<R, K0>(g, U, xg, kb, sb, fr);
                pf[3] = zrow<R, K0>(g, U, xg, kb, sb, r2);
            }
        } else {
            pf[0] = pf[1] = pf[2] = pf[3] = CT{(R)0, (R)0};
        }
    };
    // with one x-row per block only the 3 nv sequences of the valid pairs
    // of a batch are transformed (the others are padding)
    auto transform = [&](auto sign, int nv) {
        pp_fft<CT, LOGNN, decltype(sign)::value, Z2T, true>(A, B, XB == 1 ? 3 * nv : SEQ, twNl);
    };

    R vb[PT][KK];
    R an[PT][KK];
#pragma unroll
    for (int j = 0; j < PT; j++)
#pragma unroll
        for (int s = 0; s < KK; s++) an[j][s] = (R)0;

    prefetch(threadIdx.x, 0);
    ptw_fill<CT, LOGNN>(twNl, twNN, threadIdx.x, Z2T);
    for (int i = threadIdx.x; i < NN; i += Z2T) twQl[i] = twQ[i];
    __syncthreads();
    CT* const F = INB ? B : A; // where a transform's result lies
    static_for<0, NBI>([&](auto bi) {
        constexpr int BN = decltype(bi)::value;
        constexpr int NV = ZO::NP - BN * ZP < ZP ? ZO::NP - BN * ZP : ZP; // valid pairs
        const int tid = opaque_tid();
        if (XB > 1 || (tid / H) % ZP < NV) {
            const int fx = tid / (ZP * H), fp = (tid / H) % ZP, fr = tid % H;
            const int fm = fr == 0 ? H : NN - fr;
            CT* seq = A + ((fx * ZP + fp) * 3) * LP;
            const CT one{(R)1, (R)0};
            zfold_bw<R>(seq, 0, fr, NN, pf[0], pf[1], pf[2], pf[3], one, one);
            zfold_bw<R>(seq + LP, 1, fr, NN, pf[0], pf[1], pf[2], pf[3], twQl[fr], twQl[fm]);
            zfold_bw<R>(seq + 2 * LP, 2, fr, NN, pf[0], pf[1], pf[2], pf[3], twQl[2 * fr],
                        twq2<R>(twQl, NN, 2 * fm));
        }
        if constexpr (BN + 1 < NBI) prefetch(tid, BN + 1);
        __syncthreads();
        transform(std::integral_constant<int, +1>{}, NV);
        static_for<0, ZP>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
#pragma unroll
            for (int j = 0; j < PT; j++) {
                const CT zv = F[poff(tid, j) + p * 3 * LP];
                static_for<0, 2>([&](auto hi) {
                    constexpr int q = BN * 2 * ZP + 2 * p + decltype(hi)::value;
                    constexpr int kd = ZO::kind(q), s = ZO::slot(q);
                    const R val = decltype(hi)::value == 0 ? zv.x : zv.y;
                    if constexpr (kd == 0) vb[j][s] = val;
                    else if constexpr (kd == 1) G::template dx<s>(vb[j], val, an[j]);
                    else if constexpr (kd == 2) G::template dz<s>(vb[j], val, an[j]);
                });
            }
        });
        if constexpr (!INB) __syncthreads();
    });
#pragma unroll
    for (int j = 0; j < PT; j++) G::yterms(vb[j], an[j], (R)g.beta);

    // forward z transform of the KK products, two per complex sequence
    const R scale = (R)g.scale;
    static_for<0, NBO>([&](auto ri) {
        constexpr int RD = decltype(ri)::value;
        constexpr int NV = NPO - RD * ZP < ZP ? NPO - RD * ZP : ZP; // valid output pairs
        const int tid = opaque_tid();
        static_for<0, ZP>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int c1 = 2 * (RD * ZP + p), c2 = c1 + 1;
            if constexpr (c1 < KK) {
#pragma unroll
                for (int j = 0; j < PT; j++) {
                    R v2 = (R)0;
                    if constexpr (c2 < KK) v2 = an[j][c2];
                    A[poff(tid, j) + p * 3 * LP] = CT{an[j][c1], v2};
                }
            }
        });
        __syncthreads();
        transform(std::integral_constant<int, -1>{}, NV);
#pragma unroll
        for (int u = 0; u < UT; u++) {
            const int t = tid + Z2T * u;
            const int r = t % NN, p = (t / NN) % ZP, xl = t / (NN * ZP);
            const int c1 = 2 * (RD * ZP + p), c2 = c1 + 1;
            if (c1 < KK) {
                const CT* fs = F + (xl * ZP + p) * 3 * LP;
                const int rm = (NN - r) & (NN - 1);
                CT X1{(R)0, (R)0}, X2{(R)0, (R)0};
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    const CT zn = fs[b * LP + pidx(r)];
                    const CT zm = fs[b * LP + pidx(rm)];
                    const CT F1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                    const CT F2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                    if (b == 0) {
                        X1 = F1;
                        X2 = F2;
                    } else {
                        const CT w = cconj(twq2<R>(twQl, NN, r * b));
                        X1 = cadd(X1, cmul(F1, w));
                        X2 = cadd(X2, cmul(F2, w));
                    }
                }
                const int x = x0 + xl;
                if (x < g.P) {
                    st_piece(&NL[nidx(g, x, c1, r)], cscale(X1, scale), wide);
                    if (c2 < KK) st_piece(&NL[nidx(g, x, c2, r)], cscale(X2, scale), wide);
                }
            }
        }
        if constexpr (!INB) __syncthreads();
    });
}

// ---------------------------------------------------------------------------
// z pass with the three fold phases in sequence (b = 0, 1, 2), for rows whose
// per-point state does not fit the registers of one block when all three
// phases are live (NN = 1024, and the 17-slot M6 basis).  A block of ZB
// threads holds XB = ZB/NN x-rows, one point (x, z = 3r + b) per thread in a
// phase: vb[KK] base values and an[KK] product accumulators.  Each phase
// re-reads the U row (from L2), runs the inverse transforms of its own fold,
// the products and the forward transforms, and adds its share
// F_b[m] w^{-mb} of the r2c output to NL; the third phase scales by 1/(PQ).
// NL therefore holds partial sums between phases; each element is read and
// written only by the thread that owns it, in phase order, which makes the
// result identical to the three-phase sum of k_z.
//
// PH3: the three phases of an x-row run in three blocks instead, each writing
// its unscaled share to its own NL partial (bt.nl_part apart); the x kernel
// sums ((P0 + P1) + P2) and scales by 1/(PQ) as it loads NL.  Block B maps to
// XCD B % 8 and, there, to phase (B / 8) % 3 of row (B / 24): the three
// phases of a row are dispatched back to back on one XCD and read the row's U
// lines through its L2 once, with no read-modify-write of NL; each XCD takes a
// contiguous run of rows, so rows sharing NL lines meet in one L2.
// ---------------------------------------------------------------------------
constexpr int ZB = 1024;
template <typename R> constexpr int zb_seq() { return 8; } // sequences per batch
// FP32: twiddles staged in LDS (4 % faster at M6 2048^2); FP64 keeps them in
// global memory, where the LDS they would take costs batch size and registers
template <typename R> constexpr bool zb_twl() { return sizeof(R) == 4; }
template <int NP, int SMAX> constexpr int zb_split() { return (NP + SMAX - 1) / SMAX; }
template <int NP, int SMAX> constexpr int zb_per() { return (NP + zb_split<NP, SMAX>() - 1) / zb_split<NP, SMAX>(); }
// transform buffer, then the twiddle copies (NN of exp(2 pi i j/NN), the first
// 2 NN of exp(2 pi i j/Q)) for NN up to ZB
template <typename R, int LOGNN> constexpr size_t zb_lds_bytes()
{
    return sizeof(cplx_t<R>) * ((size_t)zb_seq<R>() * ZB + ptw_size<(1 << LOGNN)>() + (zb_twl<R>() ? 2 * ((size_t)1 << LOGNN) : 0));
}

template <typename R, int K0, int LOGNN, bool PH3 = false>
__global__ __launch_bounds__(ZB) void k_zb(Batch bt, const cplx_t<R>* __restrict__ U, cplx_t<R>* __restrict__ NL,
                                           const cplx_t<R>* __restrict__ twNN, const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.z];
    U += blockIdx.z * bt.u_stride;
    NL += blockIdx.z * bt.nl_stride;
    using G = Gal<K0>;
    constexpr int KK = 3 * K0 - 1;
    constexpr int NN = 1 << LOGNN;
    static_assert(NN <= ZB, "one point per thread");
    constexpr int XB = ZB / NN;
    constexpr int LP = lpad(NN);
    constexpr int H = NN / 2;
    constexpr int SMAX = zb_seq<R>() / XB;                    // pairs per batch that fit
    constexpr int NP0 = ZOrder<K0, K0 / 2>::NP;               // pair count (any order)
    using ZO = ZOrder<K0, zorder_d0<K0, zb_per<NP0, SMAX>()>()>;
    static_assert(zorder_ok<K0, ZO::D0>(), "z row order");
    constexpr int ZP = zb_per<ZO::NP, SMAX>();                // inverse pairs per batch
    constexpr int NBI = zb_split<ZO::NP, SMAX>();
    constexpr int NPO = (KK + 1) / 2;
    constexpr int ZO_ = zb_per<NPO, SMAX>();                  // forward pairs per batch
    constexpr int NBO = zb_split<NPO, SMAX>();
    extern __shared__ unsigned char smem_raw[];
    CT* A = reinterpret_cast<CT*>(smem_raw);                  // [XB*ZP][LP]
    CT* twNl = A + zb_seq<R>() * ZB;                          // [ptw_size<NN>()] pass twiddles
    CT* twQl = twNl + ptw_size<NN>();                         // [2 NN] (when zb_twl)
    ptw_fill<CT, LOGNN>(twNl, twNN, threadIdx.x, ZB);
    if constexpr (zb_twl<R>())
        for (int i = threadIdx.x; i < 2 * NN; i += ZB) twQl[i] = twQ[i];
    __syncthreads();
    const CT* tN = twNl;
    const CT* tQ = zb_twl<R>() ? twQl : twQ;

    int xt = 0, b_first = 0, b_end = 3;
    if constexpr (PH3) {
        const int B = blockIdx.x, idx = B >> 3, nb = gridDim.x / 3;
        b_first = idx % 3;
        b_end = b_first + 1;
        xt = (B & 7) * (nb >> 3) + idx / 3;
        NL += b_first * bt.nl_part;
    } else {
        xt = xcd_tile(blockIdx.x, gridDim.x);
    }
    const int x0 = xt * XB;
    const R scale = (R)g.scale;

#pragma unroll 1
    for (int b = b_first; b < b_end; b++) {
        R vb[KK], an[KK];
#pragma unroll
        for (int s = 0; s < KK; s++) an[s] = (R)0;

        static_for<0, NBI>([&](auto bi) {
            constexpr int BN = decltype(bi)::value;
            // opaque copy of the thread index: every index below is recomputed
            // per batch instead of being hoisted (and kept live) across batches
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            const int xl = tid >> LOGNN, r = tid & (NN - 1);
            // fold tasks tau = (row fx, pair fp, index fr)
#pragma unroll 1
            for (int tau = tid; tau < XB * ZP * H; tau += ZB) {
                const int fx = tau / (ZP * H), fp = (tau / H) % ZP, fr = tau % H;
                const int xg = x0 + fx;
                const int pr = BN * ZP + fp;
                const int r2 = fr == 0 ? H : NN - fr;
                CT a1{(R)0, (R)0}, b1{(R)0, (R)0}, a2{(R)0, (R)0}, b2{(R)0, (R)0};
                if (xg < g.P && pr < ZO::NP) {
                    const int ka = ZO::kind(2 * pr), sa = ZO::slot(2 * pr);
                    a1 = zrow<R, K0>(g, U, xg, ka, sa, fr);
                    b1 = zrow<R, K0>(g, U, xg, ka, sa, r2);
                    if (ZO::shared(pr)) {
                        a2 = zderiv<R>(g, a1, fr);
                        b2 = zderiv<R>(g, b1, r2);
                    } else {
                        const int kb = ZO::kind(2 * pr + 1), sb = ZO::slot(2 * pr + 1);
                        a2 = zrow<R, K0>(g, U, xg, kb, sb, fr);
                        b2 = zrow<R, K0>(g, U, xg, kb, sb, r2);
                    }
                }
                zfold_b<R>(A + (fx * ZP + fp) * LP, b, fr, NN, a1, b1, a2, b2, tQ);
            }
            __syncthreads();
            ip_fft<CT, LOGNN, +1, ZB, XB * ZP, true>(A, XB * ZP, tN, tid);
            static_for<0, ZP>([&](auto pi) {
                constexpr int p = decltype(pi)::value;
                const CT zv = A[(xl * ZP + p) * LP + pidx(r)];
                static_for<0, 2>([&](auto hi) {
                    constexpr int q = BN * 2 * ZP + 2 * p + decltype(hi)::value;
                    constexpr int kd = ZO::kind(q), sl = ZO::slot(q);
                    const R val = decltype(hi)::value == 0 ? zv.x : zv.y;
                    if constexpr (kd == 0) vb[sl] = val;
                    else if constexpr (kd == 1) G::template dx<sl>(vb, val, an);
                    else if constexpr (kd == 2) G::template dz<sl>(vb, val, an);
                });
            });
            __syncthreads();
        });
        G::yterms(vb, an, (R)g.beta);

        // forward z transform of this phase's products, two per sequence
        static_for<0, NBO>([&](auto ri) {
            constexpr int RD = decltype(ri)::value;
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            const int xl = tid >> LOGNN, r = tid & (NN - 1);
            const int rm = (NN - r) & (NN - 1);
            const int x = x0 + xl;
            const bool valid = x < g.P;
            // FP32: the partial sums of the earlier phases are loaded before the
            // transform, so that their latency hides behind it (FP64 has no
            // registers to spare for them and loads them after)
            constexpr bool PRE = sizeof(R) == 4 && !PH3;
            CT o1[ZO_], o2[ZO_];
            if (PRE && b > 0 && valid) {
                static_for<0, ZO_>([&](auto pi) {
                    constexpr int p = decltype(pi)::value;
                    constexpr int c1 = 2 * (RD * ZO_ + p), c2 = c1 + 1;
                    if constexpr (c1 < KK) o1[p] = NL[nidx(g, x, c1, r)];
                    if constexpr (c2 < KK) o2[p] = NL[nidx(g, x, c2, r)];
                });
            }
            static_for<0, ZO_>([&](auto pi) {
                constexpr int p = decltype(pi)::value;
                constexpr int c1 = 2 * (RD * ZO_ + p), c2 = c1 + 1;
                if constexpr (c1 < KK) {
                    R v2 = (R)0;
                    if constexpr (c2 < KK) v2 = an[c2];
                    A[(xl * ZO_ + p) * LP + pidx(r)] = CT{an[c1], v2};
                }
            });
            __syncthreads();
            ip_fft<CT, LOGNN, -1, ZB, XB * ZO_, true>(A, XB * ZO_, tN, tid);
            static_for<0, ZO_>([&](auto pi) {
                constexpr int p = decltype(pi)::value;
                constexpr int c1 = 2 * (RD * ZO_ + p), c2 = c1 + 1;
                if constexpr (c1 < KK) {
                    const CT zn = A[(xl * ZO_ + p) * LP + pidx(r)];
                    const CT zm = A[(xl * ZO_ + p) * LP + pidx(rm)];
                    CT X1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                    CT X2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                    if (b > 0) {
                        const CT w = cconj(tQ[r * b]);
                        X1 = cmul(X1, w);
                        X2 = cmul(X2, w);
                    }
                    if (valid && PH3) {
                        NL[nidx(g, x, c1, r)] = X1;
                        if constexpr (c2 < KK) NL[nidx(g, x, c2, r)] = X2;
                    } else if (valid) {
                        if (b > 0) X1 = cadd(PRE ? o1[p] : NL[nidx(g, x, c1, r)], X1);
                        NL[nidx(g, x, c1, r)] = b == 2 ? cscale(X1, scale) : X1;
                        if constexpr (c2 < KK) {
                            if (b > 0) X2 = cadd(PRE ? o2[p] : NL[nidx(g, x, c2, r)], X2);
                            NL[nidx(g, x, c2, r)] = b == 2 ? cscale(X2, scale) : X2;
                        }
                    }
                }
            });
            __syncthreads();
        });
    }
}

// z inverse only, to the reference phys layout (vel_mpt2phys); blockIdx.y
// selects the member, whose phys array follows the previous one's.
template <int LOGNN> constexpr int zphys_threads() { return (1 << LOGNN) > ZT ? (1 << LOGNN) : ZT; }
template <typename R, int K0, int LOGNN>
__global__ __launch_bounds__(zphys_threads<LOGNN>()) void k_zphys(Batch bt, const cplx_t<R>* __restrict__ U, double* __restrict__ phys,
                                              const cplx_t<R>* __restrict__ twNN, const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.y];
    U += blockIdx.y * bt.u_stride;
    phys += blockIdx.y * ((size_t)g.KK * g.Q * g.P);
    constexpr int KK = 3 * K0 - 1;
    constexpr int NN = 1 << LOGNN;
    constexpr int TZ = zphys_threads<LOGNN>();
    constexpr int XB = TZ / NN;
    constexpr int NBATCH = (KK + 2 * ZPB - 1) / (2 * ZPB);
    constexpr int SEQ = XB * ZPB * 3;
    constexpr int LP = lpad(NN);
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);
    const int tid = threadIdx.x;
    const int xl = tid >> LOGNN, r = tid & (NN - 1);
    const int x = blockIdx.x * XB + xl;
    const bool valid = x < g.P;
    const int xr = valid ? x : 0;
    const int Q = 3 * NN;
    for (int B = 0; B < NBATCH; B++) {
        for (int p = 0; p < ZPB; p++) {
            const int g1 = B * 2 * ZPB + 2 * p, g2 = g1 + 1;
            CT Z[3];
            if (valid && g1 < KK) {
                const CT a1 = U[uidx(g, xr, g1, r)];
                const CT a2 = g2 < KK ? U[uidx(g, xr, g2, r)] : CT{(R)0, (R)0};
                if (r == 0) {
                    Z[0] = Z[1] = Z[2] = CT{a1.x, a2.x};
                } else {
                    const CT b1 = U[uidx(g, xr, g1, NN - r)];
                    const CT b2 = g2 < KK ? U[uidx(g, xr, g2, NN - r)] : CT{(R)0, (R)0};
                    Z[0] = CT{(a1.x + b1.x) - (a2.y - b2.y), (a1.y - b1.y) + (a2.x + b2.x)};
                    for (int b = 1; b < 3; b++) {
                        const CT w1 = twQ[r * b];
                        const CT w2 = twQ[(NN - r) * b];
                        const CT y1 = cadd(cmul(a1, w1), cconj(cmul(b1, w2)));
                        const CT y2 = cadd(cmul(a2, w1), cconj(cmul(b2, w2)));
                        Z[b] = CT{y1.x - y2.y, y1.y + y2.x};
                    }
                }
            } else {
                Z[0] = Z[1] = Z[2] = CT{(R)0, (R)0};
            }
            for (int b = 0; b < 3; b++) buf[((xl * ZPB + p) * 3 + b) * LP + pidx(r)] = Z[b];
        }
        __syncthreads();
        ip_fft<CT, LOGNN, +1, TZ, SEQ>(buf, SEQ, twNN);
        for (int p = 0; p < ZPB; p++) {
            for (int b = 0; b < 3; b++) {
                const CT zv = buf[((xl * ZPB + p) * 3 + b) * LP + pidx(r)];
                const int g1 = B * 2 * ZPB + 2 * p;
                const int z = 3 * r + b;
                if (valid && g1 < KK) phys[((size_t)x * Q + z) * KK + g1] = (double)zv.x;
                if (valid && g1 + 1 < KK) phys[((size_t)x * Q + z) * KK + g1 + 1] = (double)zv.y;
            }
        }
        __syncthreads();
    }
}


// var_spec2mpt (program/variables.f90:440-480), split by the order in which
// the x-forward kernel produces u_k, v_k, w_k.
__device__ __forceinline__ double2 s2m_vterm(double2 v, double am2, double an2)
{
#pragma clang fp contract(off)
    return double2{v.x * (am2 + an2), v.y * (am2 + an2)};
}

__device__ __forceinline__ void s2m_finish(int k, int mi, int n, double2 u, double2 w, double2 vh, double am1,
                                           double an1, double ak1, double2& psi, double2& phi)
{
#pragma clang fp contract(off)
    psi = double2{(w.y * am1 - u.y * an1), -(w.x * am1 - u.x * an1)};
    if (k > 0) phi = double2{-(u.y * am1 * ak1 + w.y * an1 * ak1) + vh.x, (u.x * am1 * ak1 + w.x * an1 * ak1) + vh.y};
    if (mi == 0 && n == 0) {
        if (k == 0) {
            psi = double2{0.0, 0.0};
        } else {
            psi = double2{-u.x, -u.y};
            phi = double2{-w.x, -w.y};
        }
    }
}

// var_LHSRHS (program/variables.f90:565-597) for psi_k (which = 0) or
// phi_k (which = 1) at one (m, n).
__device__ __forceinline__ void lhsrhs_k(const Geo& g, int k, int which, bool mean, double am2, double an2,
                                         double ak2, double& lhs, double& rhs)
{
#pragma clang fp contract(off)
    const double dt = g.dt;
    const double lap = (am2 + ak2) + an2;
    const double hlap = am2 + an2;
    const double ilap16 = g.hypo != 0.0 ? g.hypo / lap : 0.0; // lap = 0 only where tvl is unused
    const double tvl = (g.dragall || k == 0) ? 1.0 - dt * ((lap / g.re + ilap16) - g.drag)
                                             : 1.0 - dt * (lap / g.re + ilap16);
    if (mean) {
        if (which == 0 && k == 0) {
            lhs = 1.0;
            rhs = 0.0;
        } else {
            lhs = tvl;
            rhs = 1.0;
        }
    } else if (which == 0) {
        lhs = tvl * hlap;
        rhs = hlap;
    } else {
        lhs = g.dragall ? tvl * hlap * lap + dt * (g.vdrag - g.drag) * hlap * hlap : tvl * hlap * lap;
        rhs = hlap * lap;
    }
}

// The AB2 history (the previous step's nonlinear term) is stored in the
// precision of the transforms that produced it; the AB2 combination is always
// evaluated in FP64.  In the mixed form this halves the history's bytes and
// loses nothing: the stored values carry FP32 transform rounding already.
__device__ __forceinline__ double2 hist_load(const double2* p) { return *p; }
__device__ __forceinline__ double2 hist_load(const float2* p)
{
    const float2 v = *p;
    return double2{(double)v.x, (double)v.y};
}
__device__ __forceinline__ void hist_store(double2* p, double2 v, bool stream = false)
{
    if (stream) st_stream(p, v);
    else *p = v;
}
__device__ __forceinline__ void hist_store(float2* p, double2 v, bool stream = false)
{
    const float2 f = float2{(float)v.x, (float)v.y};
    if (stream) st_stream(p, f);
    else *p = f;
}

// vel_TS update (program/velocity.f90:28-35) in the reference's order
__device__ __forceinline__ double2 ab2_update(double2 cv, double2 nl, double2 ov, double lhs, double rhs, double dt)
{
#pragma clang fp contract(off)
    double cr = rhs * cv.x;
    double ci = rhs * cv.y;
    cr = cr + dt / 2.0 * (3.0 * nl.x - ov.x);
    ci = ci + dt / 2.0 * (3.0 * nl.y - ov.y);
    return double2{cr / lhs, ci / lhs};
}

// ---------------------------------------------------------------------------
// x forward + spec2mpt + update: block (n-tile, y-mode k).
// mode 0: AB2/implicit update of c and the history onl (cplx<R>; vel_TS);
//         nst = first step.
// mode 1: write the nonlinear term into nlout only (double2; vel_nonlinear).
// ---------------------------------------------------------------------------
template <typename R, int LOGMM, bool FLUX>
__global__ __launch_bounds__(XT) void k_xfwd(Batch bt, const cplx_t<R>* __restrict__ NL, double2* __restrict__ c,
                                             cplx_t<R>* __restrict__ onl, double2* __restrict__ nlout,
                                             const cplx_t<R>* __restrict__ twMM, const double2* __restrict__ twP,
                                             int mode, int nst)
{
    using CT = cplx_t<R>;
    const Geo g = bt.geo[blockIdx.z];
    NL += blockIdx.z * bt.nl_stride;
    c += blockIdx.z * bt.c_stride;
    onl += blockIdx.z * bt.c_stride;
    if (nlout) nlout += blockIdx.z * bt.c_stride;
    constexpr int MM = 1 << LOGMM;
    constexpr int LP = lpad(MM);
    constexpr int IPT = XTILE / XT;
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);

    const int NT = g.NTx;
    const int n0 = xcd_tile(blockIdx.x, gridDim.x) * NT;
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    const int nitems = NT * MM;
    const double ak1 = ak1_of(k, g.beta);

    // psi[] holds u_k until w_k arrives, phi[] holds v_k (am2+an2) until then;
    // xd[] holds i alpha m B of the flux form until its A line arrives
    double2 psi[IPT][2], phi[IPT][2];
    double2 xd[FLUX ? IPT : 1][2];

    for (int li = 0; li < (FLUX ? 6 : 3); li++) {
        const int comp = FLUX ? li >> 1 : li, part = FLUX ? li & 1 : 0;
        if (k == 0 && comp == 1) continue;
        const int slot = comp == 0 ? k : (comp == 1 ? g.K0 - 1 + k : 2 * g.K0 - 1 + k);
        const CT* pan = nl_panel(g, NL, n0, slot + part * g.KK);
        // load NL rows, decimated by b = x mod 3
#pragma unroll
        for (int q = 0; q < 3 * IPT; q++) {
            const int it = tid + q * XT;
            if (it >= 3 * nitems) continue;
            const int x = it >> g.lgNT, t = it & (NT - 1);
            const int a = x / 3, b = x - 3 * a;
            const int po = pan_at(g, it);
            CT v = pan[po];
            if (bt.nl_part) { // the z pass's three fold-phase partials: ((P0 + P1) + P2) / (PQ)
                const CT v1 = pan[po + bt.nl_part], v2 = pan[po + 2 * bt.nl_part];
                v = cscale(cadd(cadd(v, v1), v2), (R)g.scale);
            }
            buf[((size_t)t * 3 + b) * LP + pidx(a)] = v;
        }
        __syncthreads();
        ip_fft<CT, LOGMM, -1, XT, 3 * (XTILE >> LOGMM)>(buf, 3 * NT, twMM);
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int it = tid + q * XT;
            if (it >= nitems) continue;
            const int t = it >> LOGMM, r = it & (MM - 1);
            const int n = n0 + t;
            const CT* bt = buf + (size_t)t * 3 * LP;
            double2 Xp, Xn;
            {
                const CT F0 = bt[pidx(r)];
                Xp = double2{(double)F0.x, (double)F0.y};
                Xn = Xp;
#pragma unroll
                for (int b = 1; b < 3; b++) {
                    const CT Fb = bt[b * LP + pidx(r)];
                    const double fr = (double)Fb.x, fi = (double)Fb.y;
                    const double2 w1 = twP[r * b];        // m = r:      times conj(w1)
                    const double2 w2 = twP[(MM - r) * b]; // m = r - MM: times w2
                    Xp.x += fr * w1.x + fi * w1.y;
                    Xp.y += fi * w1.x - fr * w1.y;
                    Xn.x += fr * w2.x - fi * w2.y;
                    Xn.y += fi * w2.x + fr * w2.y;
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double2 X0 = h == 0 ? Xp : Xn;
                if (h == 1 && r == 0) continue;
                const int mi = h == 0 ? r : MM - 1 + r;
                double am1, am2, an1, an2;
                xwave(g, mi, am1, am2);
                zwave(g, n, an1, an2);
                double2 X = X0;
                if constexpr (FLUX) { // N = i alpha m B + A
                    if (part == 0) {
                        xd[q][h] = double2{-(X.y * am1), X.x * am1};
                        continue;
                    }
                    X = double2{xd[q][h].x + X.x, xd[q][h].y + X.y};
                }
                // var_spec2mpt (program/variables.f90:440-480)
                if (comp == 0) {
                    psi[q][h] = X;
                } else if (comp == 1) {
                    phi[q][h] = s2m_vterm(X, am2, an2);
                } else {
                    const double2 u = psi[q][h], vh = phi[q][h];
                    s2m_finish(k, mi, n, u, X, vh, am1, an1, ak1, psi[q][h], phi[q][h]);
                }
            }
        }
        __syncthreads();
    }

    // update (program/velocity.f90:24-40) with var_LHSRHS (program/variables.f90
    // :565-597) evaluated in place
    const double dt = g.dt;
    const double ak2 = ak2_of(k, g.beta);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int it = tid + q * XT;
        if (it >= nitems) continue;
        const int t = it >> LOGMM, r = it & (MM - 1);
        const int n = n0 + t;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (h == 1 && r == 0) continue;
            const int mi = h == 0 ? r : MM - 1 + r;
            double am1, am2, an1, an2;
            xwave(g, mi, am1, am2);
            zwave(g, n, an1, an2);
            const bool mean = (mi == 0 && n == 0);
#pragma unroll
            for (int which = 0; which < 2; which++) {
                if (which == 1 && k == 0) continue;
                const int slot = which == 0 ? k : g.K0 - 1 + k;
                const size_t idx = ((size_t)n * g.K + slot) * g.ldM + mi;
                const double2 nl = which == 0 ? psi[q][h] : phi[q][h];
                if (mode == 1) {
                    nlout[idx] = nl;
                    continue;
                }
                double lhs, rhs;
                lhsrhs_k(g, k, which, mean, am2, an2, ak2, lhs, rhs);
                const double2 cv = c[idx];
                const double2 ov = nst ? nl : hist_load(onl + idx);
                double2 cn = ab2_update(cv, nl, ov, lhs, rhs, dt);
                if (mean && which == 0 && k == 0) cn.x = 0.0;
                c[idx] = cn;
                hist_store(onl + idx, nl);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused x kernel of one step: forward x transform of NL, spec2mpt, AB2 and
// implicit update of c and onl, then the inverse x transform of the updated
// state for the next step's z kernel (mpt2spec + laminar + d/dx), so c is read
// once per step.  Block (n-tile, y-mode k), XS = 3 * NT * G threads: each
// group of G = min(64, MM/8) lanes owns one of the 3*NT sequences (t, b), 8
// points per lane (16 at MM = 1024), and transforms it with group_fft (no
// block barrier inside a transform); items (t, r), up to XTILE/XS per thread,
// hold psi_k / phi_k at m = r and m = r - MM.  768 threads at every MM (at
// MM = 1024, twice the 384 lanes of its six sequences).
// ---------------------------------------------------------------------------
// Threads per block: one lane group per sequence, 3 XTILE/MM sequences; at MM
// = 1024 (6 sequences) twice that, so that a thread holds 3 items instead of
// 6 (the extra waves take items but no sequence).
template <int LOGMM> constexpr int xs_mult() { return LOGMM >= 10 ? 2 : 1; }
template <int LOGMM> constexpr int xs_of() { return xs_mult<LOGMM>() * 3 * (XTILE >> LOGMM) * group_lanes<(1 << LOGMM)>(); }

// FP32 k_xstep streams the next NL panel straight into an LDS staging area
// (global_load_lds, no registers) while the current line is transformed: the
// staging copy of one panel (3 XTILE values, 48 KB) fits beside the transform
// buffer in FP32, not in FP64.
template <typename R> constexpr bool xstep_async() { return sizeof(R) == 4; }

// Sequence (t, b) of the x kernels' transform buffer holds element a at slot
// pidx(a) ^ xsw(t, b).  The decimating transposes between panels ([x][t],
// x = 3a + b) and sequences have up to 16 lanes of one bank phase on one a;
// this per-sequence XOR sends them to distinct banks (conflict-free at MM <=
// 256, at most 2-way at 512 and 1024, from 6- to 16-way unswizzled), while
// accesses within one sequence keep their conflict-free patterns.
template <typename R> __device__ __forceinline__ int xsw(int t, int b)
{
    if constexpr (sizeof(R) == 4) return ((t * 9) ^ (b * 3)) & 15; // 8-byte slots, 16 per bank phase
    else return ((t * 5) ^ b) & 7;                                 // 16-byte slots, 8 per bank phase
}
template <typename R> constexpr size_t xstep_stage_bytes() { return xstep_async<R>() ? 3 * (size_t)XTILE * sizeof(cplx_t<R>) : 0; }

// Wave-cooperative copy of a block's NL stripe (nch 16-byte chunks, chunk c
// holding stripe elements c E .. c E + E - 1, E = 16 / sizeof(CT), which lie
// in one x-row's piece) from global memory into LDS, asynchronously: lane l of
// a wave moves chunk c0 + l to dst + 16 (c0 + l), c0 a multiple of 64 taken by
// the wave.  nch is a multiple of 64.  The data is in LDS once the wave's
// vector memory count has drained and a barrier has passed.
template <int T, typename CT>
__device__ __forceinline__ void panel_to_lds(const Geo& g, const CT* __restrict__ src, CT* dst, int nch, int tid)
{
    constexpr int E = 16 / (int)sizeof(CT);
    const int wave = tid >> 6, lane = tid & 63;
    char* l = reinterpret_cast<char*>(dst);
    for (int c0 = wave * 64; c0 < nch; c0 += T)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const char*>(src + pan_at(g, (c0 + lane) * E)),
                                         (__attribute__((address_space(3))) void*)(l + (size_t)c0 * 16), 16, 0, 0);
}
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <typename R, int LOGMM, bool INV_ONLY, bool FLUX>
__global__ __launch_bounds__(xs_of<LOGMM>()) void k_xstep(Batch bt, const cplx_t<R>* __restrict__ NL, double2* __restrict__ c,
                                              cplx_t<R>* __restrict__ onl, cplx_t<R>* __restrict__ U,
                                              const cplx_t<R>* __restrict__ twMM, const double2* __restrict__ twP,
                                              int nst)
{
    using CT = cplx_t<R>;
    constexpr int MM = 1 << LOGMM;
    constexpr int G = group_lanes<MM>();
    constexpr int PL = MM / G;
    constexpr int XS = xs_of<LOGMM>();
    constexpr int IPT = (XTILE + XS - 1) / XS; // items per thread
    constexpr int LPT = 3 * XTILE / XS;        // rows per thread in the cooperative load/store
    // After the update each thread holds the new state of its items in
    // psi/phi; the inverse half rebuilds its 6 lines from those registers
    // instead of re-reading c (k_xstep 10 % faster in FP64, 18 % in FP32 at
    // MM = 512).  At MM = 1024 (6 items per thread) the registers spill and
    // the kernel re-reads c instead.
    constexpr bool XREG = !INV_ONLY && LOGMM < 10;
    extern __shared__ unsigned char smem_raw[];
    CT* reg = reinterpret_cast<CT*>(smem_raw); // [XS/G sequences][MM], swizzled
    CT* twMl = reg + 3 * XTILE;                // pass twiddles of the MM-point transforms (seven-entry)
    static_assert(ptw_size<MM, 1, 7>() % 2 == 0, "twPl must stay 16-byte aligned");
    double2* twPl = reinterpret_cast<double2*>(twMl + ptw_size<MM, 1, 7>()); // exp(2 pi i j/P), j <= 2 MM
    constexpr bool ASYNC = xstep_async<R>() && !INV_ONLY;
    // FP32 at MM <= 512: the forward transforms' first pass reads a whole
    // tile's sequences straight from the staged panel (element a of sequence
    // (t, b) at panel row 3a + b, column t), with no copy into the transform
    // layout and no barrier between (k_xstep 0.130 -> 0.123 ms at M4 1024^2);
    // at MM = 1024 that form spills 32 VGPRs and is 30 % slower
    constexpr bool SREAD = ASYNC && LOGMM < 10;
    CT* stage = reinterpret_cast<CT*>(twPl + 2 * MM + 1);               // [3 XTILE] next NL panel

    const Geo g = bt.geo[blockIdx.z];
    NL += blockIdx.z * bt.nl_stride;
    c += blockIdx.z * bt.c_stride;
    onl += blockIdx.z * bt.c_stride;
    U += blockIdx.z * bt.u_stride;

    const int NT = g.NTx;
    const int n0 = xcd_tile(blockIdx.x, gridDim.x) * NT;
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    const int nitems = NT * MM;
    const int nrows = 3 * nitems; // (x, t) pairs of one line
    const int s = tid / G, gl = tid % G;
    const bool seq_on = s < 3 * NT;
    CT* myreg = reg + s * MM;
    const int sxo = xsw<R>(s / 3, s % 3); // swizzle of this thread's group sequence
    // Decimating transposes between a panel ([x][t], i = x NT + t, x = 3a + b)
    // and the sequences (t, b).  When the panel is a whole tile (NT = NTC, the
    // usual case) item q of thread tid, i = tid + XS q, keeps t and b and has
    // a = a0 + DA q (DA = XS / (3 NTC), a power of two above a0), so its slot
    // is one per-thread base and a constant (swz); otherwise every item's
    // indices are computed.
    constexpr int NTC = XTILE >> LOGMM, LOGNTC = LOG_XTILE - LOGMM;
    constexpr int DA = XS / (3 * NTC);
    static_assert((1 << LOGNTC) == NTC && XS == 3 * NTC * DA && (DA & (DA - 1)) == 0, "whole-tile transposes");
    const bool whole = NT == NTC;
    auto tile_item = [&](int tidv, int& t0, int& xq, CT*& rb, int& pa) {
        t0 = tidv & (NTC - 1);
        xq = tidv >> LOGNTC;
        const int a0 = xq / 3, b0 = xq - 3 * a0;
        rb = reg + (t0 * 3 + b0) * MM;
        pa = pidx(a0) ^ xsw<R>(t0, b0);
    };
    const double ak1 = ak1_of(k, g.beta);
    const double ak2 = ak2_of(k, g.beta);

    // psi[] holds u_k until w_k arrives, phi[] holds v_k (am2+an2); after the
    // update both hold the new state.  xd[] holds i alpha m B of the flux
    // form until its A line arrives.
    double2 psi[IPT][2], phi[IPT][2];
    double2 xd[FLUX ? IPT : 1][2];

    ptw_fill<CT, LOGMM, 1, 7>(twMl, twMM, tid, XS);
    for (int i = tid; i <= 2 * MM; i += XS) twPl[i] = twP[i];
    // the tables are read by other threads; the inverse-only form reads twPl
    // before its first transform barrier
    __syncthreads();

    if constexpr (!INV_ONLY) {
        // ---------------- forward x transform + spec2mpt ----------------
        // NL lines in order: comp u, v, w (flux form: B then A of each);
        // y-mode 0 has no v
        constexpr int NLI = FLUX ? 6 : 3;
        auto line_on = [&](int li) { return !(k == 0 && (FLUX ? li >> 1 : li) == 1); };
        auto panel_of = [&](int li) {
            const int comp = FLUX ? li >> 1 : li, part = FLUX ? li & 1 : 0;
            const int slot = comp == 0 ? k : (comp == 1 ? g.K0 - 1 + k : 2 * g.K0 - 1 + k);
            return nl_panel(g, NL, n0, slot + part * g.KK);
        };
        const int nch = nrows * (int)sizeof(CT) / 16; // 16-byte chunks of one stripe
        if constexpr (ASYNC) panel_to_lds<XS>(g, panel_of(0), stage, nch, tid);
#pragma unroll 1
        for (int li = 0; li < NLI; li++) {
            const int comp = FLUX ? li >> 1 : li, part = FLUX ? li & 1 : 0;
            if (!line_on(li)) continue;
            // opaque copy of the thread index: keeps item arithmetic inside the loop
            int tidv = tid;
            asm volatile("" : "+v"(tidv));
            if constexpr (ASYNC) {
                // this line's panel has landed in the staging area.  A whole
                // tile's transforms read it there (MM <= 512); otherwise it is
                // laid out for them first, and the next line's panel starts then
                wait_vmem();
                __syncthreads();
                if (!(SREAD && whole)) {
                    if (whole) {
                        int t0, xq, pa;
                        CT* rb;
                        tile_item(tidv, t0, xq, rb, pa);
                        static_for<0, LPT>([&](auto qi) {
                            constexpr int q = decltype(qi)::value;
                            swz<q * DA>(rb, pa) = stage[tidv + XS * q];
                        });
                    } else {
#pragma unroll
                        for (int q = 0; q < LPT; q++) {
                            const int i = tidv + XS * q;
                            if (i >= nrows) continue;
                            const int x = i >> g.lgNT, t = i & (NT - 1);
                            const int a = x / 3, b = x - 3 * a;
                            reg[(t * 3 + b) * MM + (pidx(a) ^ xsw<R>(t, b))] = stage[i];
                        }
                    }
                    __syncthreads();
                    int ln = li + 1;
                    while (ln < NLI && !line_on(ln)) ln++;
                    if (ln < NLI) panel_to_lds<XS>(g, panel_of(ln), stage, nch, tid);
                }
            } else {
                const CT* pan = panel_of(li);
                // NL element, or the sum ((P0 + P1) + P2) / (PQ) of the three
                // fold-phase partials of the phase-split z kernel
                const size_t nlp = bt.nl_part;
                auto nl_at = [&](const CT* p) {
                    CT v = ld_stream(p);
                    if (nlp) v = cscale(cadd(cadd(v, ld_stream(p + nlp)), ld_stream(p + 2 * nlp)), (R)g.scale);
                    return v;
                };
                if (whole) {
                    int t0, xq, pa;
                    CT* rb;
                    tile_item(tidv, t0, xq, rb, pa);
                    // item q is element tidv + XS q of the stripe: XS/NTx
                    // x-rows further
                    const CT* pb = pan + pan_at(g, tidv);
                    const int pstep = (XS >> LOGNTC) << g.lgNS;
                    static_for<0, LPT>([&](auto qi) {
                        constexpr int q = decltype(qi)::value;
                        swz<q * DA>(rb, pa) = nl_at(pb + q * pstep);
                    });
                } else {
#pragma unroll
                    for (int q = 0; q < LPT; q++) {
                        const int i = tidv + XS * q;
                        if (i >= nrows) continue;
                        const int x = i >> g.lgNT, t = i & (NT - 1);
                        const int a = x / 3, b = x - 3 * a;
                        reg[(t * 3 + b) * MM + (pidx(a) ^ xsw<R>(t, b))] = nl_at(pan + pan_at(g, i));
                    }
                }
                __syncthreads();
            }
            if (seq_on) {
                CT v[PL];
                const int pg = pidx(gl) ^ sxo;
                if (SREAD && whole) {
                    // element a of sequence (t, b) = (s/3, s%3) is panel row
                    // x = 3a + b, column t of the staged panel
                    const CT* sb = stage + (3 * gl + s % 3) * NTC + s / 3;
                    static_for<0, PL>([&](auto qi) { v[qi] = sb[3 * G * NTC * decltype(qi)::value]; });
                } else {
                    static_for<0, PL>([&](auto qi) { v[qi] = swz<decltype(qi)::value * G>(myreg, pg); });
                }
                group_fft<CT, LOGMM, -1, true, 7>(v, myreg, gl, twMl, sxo);
                static_for<0, PL>([&](auto qi) { swz<decltype(qi)::value * G>(myreg, pg) = v[qi]; });
            }
            __syncthreads();
            if constexpr (SREAD) {
                if (whole) { // the staged panel has been read: start the next line's
                    int ln = li + 1;
                    while (ln < NLI && !line_on(ln)) ln++;
                    if (ln < NLI) panel_to_lds<XS>(g, panel_of(ln), stage, nch, tid);
                }
            }
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const int it = tidv + XS * q;
                if (it >= nitems) continue;
                const int t = it >> LOGMM, r = it & (MM - 1);
                const int n = n0 + t;
                const CT* bt3 = reg + t * 3 * MM;
                const CT F0 = bt3[pidx(r) ^ xsw<R>(t, 0)];
                double2 Xp = double2{(double)F0.x, (double)F0.y}, Xn = Xp;
#pragma unroll
                for (int b = 1; b < 3; b++) {
                    const CT Fb = bt3[b * MM + (pidx(r) ^ xsw<R>(t, b))];
                    const double fr = (double)Fb.x, fi = (double)Fb.y;
                    const double2 w1 = twPl[r * b];        // m = r:      times conj(w1)
                    const double2 w2 = twPl[(MM - r) * b]; // m = r - MM: times w2
                    Xp.x += fr * w1.x + fi * w1.y;
                    Xp.y += fi * w1.x - fr * w1.y;
                    Xn.x += fr * w2.x - fi * w2.y;
                    Xn.y += fi * w2.x + fr * w2.y;
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (h == 1 && r == 0) continue;
                    double2 X = h == 0 ? Xp : Xn;
                    const int mi = h == 0 ? r : MM - 1 + r;
                    double am1, am2, an1, an2;
                    xwave(g, mi, am1, am2);
                    zwave(g, n, an1, an2);
                    if constexpr (FLUX) { // N = i alpha m B + A
                        if (part == 0) {
                            xd[q][h] = double2{-(X.y * am1), X.x * am1};
                            continue;
                        }
                        X = double2{xd[q][h].x + X.x, xd[q][h].y + X.y};
                    }
                    if (comp == 0) {
                        psi[q][h] = X;
                    } else if (comp == 1) {
                        phi[q][h] = s2m_vterm(X, am2, an2);
                    } else {
                        const double2 u = psi[q][h], vh = phi[q][h];
                        s2m_finish(k, mi, n, u, X, vh, am1, an1, ak1, psi[q][h], phi[q][h]);
                    }
                }
            }
            __syncthreads();
        }

        // ---------------- update (program/velocity.f90:24-40) ----------------
        const double dt = g.dt;
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int it = tid + XS * q;
            if (it >= nitems) continue;
            const int t = it >> LOGMM, r = it & (MM - 1);
            const int n = n0 + t;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (h == 1 && r == 0) continue;
                const int mi = h == 0 ? r : MM - 1 + r;
                double am1, am2, an1, an2;
                xwave(g, mi, am1, am2);
                zwave(g, n, an1, an2);
                const bool mean = (mi == 0 && n == 0);
#pragma unroll
                for (int which = 0; which < 2; which++) {
                    if (which == 1 && k == 0) continue;
                    const int slot = which == 0 ? k : g.K0 - 1 + k;
                    const size_t idx = ((size_t)n * g.K + slot) * g.ldM + mi;
                    const double2 nl = which == 0 ? psi[q][h] : phi[q][h];
                    double lhs, rhs;
                    lhsrhs_k(g, k, which, mean, am2, an2, ak2, lhs, rhs);
                    const double2 cv = c[idx];
                    const double2 ov = nst ? nl : hist_load(onl + idx);
                    double2 cn = ab2_update(cv, nl, ov, lhs, rhs, dt);
                    if (mean && which == 0 && k == 0) cn.x = 0.0;
                    c[idx] = cn;
                    hist_store(onl + idx, nl, true);
                    // psi/phi now hold the new state, which the inverse half uses
                    if constexpr (XREG) {
                        if (which == 0) psi[q][h] = cn;
                        else phi[q][h] = cn;
                    }
                }
            }
        }
    }

    const bool wide = g.NTx * (int)sizeof(CT) >= 64;
    // ---------------- inverse x transform of the new state ----------------
#pragma unroll 1
    for (int li = 0; li < (FLUX ? 3 : 6); li++) {
        const int comp = li % 3, deriv = li / 3;
        if (k == 0 && comp == 1) continue;
        // spec slot of u_k, v_k, w_k: k, K0-1+k, 2K0-1+k
        const int line = (comp == 0 ? k : (comp == 1 ? g.K0 - 1 + k : 2 * g.K0 - 1 + k)) + deriv * g.KK;
        int tidv = tid;
        asm volatile("" : "+v"(tidv));
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int it = tidv + XS * q;
            if (it >= nitems) continue;
            const int t = it >> LOGMM, r = it & (MM - 1);
            const int n = n0 + t;
            const double2* cn = c + (size_t)n * g.K * g.ldM;
            double2 X[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (h == 1 && r == 0) {
                    X[h] = double2{0.0, 0.0};
                    continue;
                }
                const int mi = h == 0 ? r : MM - 1 + r;
                // the state this thread has just written: from the registers
                // after an update (XREG), else from c (an L2 hit)
                double2 ps, ph = double2{0.0, 0.0};
                if constexpr (XREG) {
                    ps = psi[q][h];
                    if (k > 0) ph = phi[q][h];
                } else {
                    ps = cn[(size_t)k * g.ldM + mi];
                    if (k > 0) ph = cn[(size_t)(g.K0 - 1 + k) * g.ldM + mi];
                }
                double2 u, v, w;
                mpt2spec_k(g, k, mi, n, ps, ph, true, u, v, w);
                double2 val = comp == 0 ? u : (comp == 1 ? v : w);
                if (deriv) { // i alpha m (program/variables.f90:606-607)
                    double am1, am2;
                    xwave(g, mi, am1, am2);
                    val = double2{-(val.y * am1), val.x * am1};
                }
                X[h] = val;
            }
            CT* bt3 = reg + t * 3 * MM;
            bt3[pidx(r) ^ xsw<R>(t, 0)] = CT{(R)(X[0].x + X[1].x), (R)(X[0].y + X[1].y)};
#pragma unroll
            for (int b = 1; b < 3; b++) {
                const double2 w1 = twPl[r * b];
                const double2 w2 = twPl[(MM - r) * b]; // conj: exp(2 pi i (r-MM) b / P)
                const double2 y = xfold_b(X[0], X[1], w1, w2);
                bt3[b * MM + (pidx(r) ^ xsw<R>(t, b))] = CT{(R)y.x, (R)y.y};
            }
        }
        __syncthreads();
        if (seq_on) {
            CT v[PL];
            const int pg = pidx(gl) ^ sxo;
            static_for<0, PL>([&](auto qi) { v[qi] = swz<decltype(qi)::value * G>(myreg, pg); });
            group_fft<CT, LOGMM, +1, true, 7>(v, myreg, gl, twMl, sxo);
            static_for<0, PL>([&](auto qi) { swz<decltype(qi)::value * G>(myreg, pg) = v[qi]; });
        }
        __syncthreads();
        CT* Ul = U + (size_t)line * g.NN + n0;
        const int ustride = g.UL * g.NN;
        if (whole) {
            int t0, xq, pa;
            CT* rb;
            tile_item(tidv, t0, xq, rb, pa);
            CT* ub = Ul + (size_t)xq * ustride + t0;
            const size_t ustep = (size_t)(3 * DA) * ustride; // x advances 3 DA per item
            static_for<0, LPT>([&](auto qi) {
                constexpr int q = decltype(qi)::value;
                st_piece(ub + q * ustep, swz<q * DA>(rb, pa), wide);
            });
        } else {
#pragma unroll
            for (int q = 0; q < LPT; q++) {
                const int i = tidv + XS * q;
                if (i >= nrows) continue;
                const int x = i >> g.lgNT, t = i & (NT - 1);
                const int a = x / 3, b = x - 3 * a;
                st_piece(Ul + x * ustride + t, reg[(t * 3 + b) * MM + (pidx(a) ^ xsw<R>(t, b))], wide);
            }
        }
        __syncthreads();
    }
}
