// fft.hpp -- batched Stockham FFTs of power-of-two length, in LDS, for gfx950.
//
// The 3/2-padded transforms of the reference (lengths P = 3*MM, Q = 3*NN,
// program/transform.fftw5.f90:33-59) are split by the kernels into a radix-3
// fold done in registers and three power-of-two transforms of length MM (or
// NN) done here.  A block transforms S sequences stored back to back in LDS.
// Each pass is a Stockham radix-R step: every butterfly reads R points L/R
// apart, applies the twiddles of the current sub-transform, does a radix-R
// DFT in registers and writes the R results Ns apart, so the output comes out
// in natural order.
//
// LDS layout: sequence s starts at s*L and element i sits at the swizzled
// slot pidx(i) = i ^ ((i >> 3) & 15), a permutation within each aligned block
// of 128.  Every read and write pattern of these passes then hits distinct
// 16-byte bank groups within each 8- or 16-lane group of a b128 access (and
// each 16- or 32-lane group of a b64 access), so the passes run free of bank
// conflicts without padding.
//
// Transforms are unnormalised: sign +1 computes sum_j x_j exp(+2 pi i jk/L)
// (FFTW's backward), sign -1 exp(-2 pi i jk/L).  Twiddles come from a table
// tw[j] = exp(+2 pi i j / L) rounded from long double on the host; a radix-8
// butterfly loads w, w^2, w^4 and forms w^3, w^5, w^6, w^7 by products.
#pragma once
#include "common.hpp"

__host__ __device__ constexpr int lpad(int L) { return L; }
__host__ __device__ constexpr int pidx(int i) { return i ^ ((i >> 3) & 15); }

// The swizzle is XOR-linear over disjoint bit fields: for a lane index a and a
// constant c with (a & c) == 0, pidx(a + c) = pidx(a) ^ pidx(c), and since
// pidx changes only bits 0..3, pidx(a) ^ pidx(c) = (pidx(a) ^ lo(c)) + hi(c)
// with lo(c) = pidx(c) & 15 and hi(c) = c & ~15.  Every butterfly address is
// such a sum (j < L/R below r L/R; o = (j - k) R + k around r Ns), so a pass
// computes pidx of its lane part once, and each of its R accesses costs at
// most one XOR with a small constant, the rest being the instruction's
// immediate offset.  swz<c>(base, pa) is base[pidx(a + c)] given pa = pidx(a).
template <int C, typename CT> __device__ __forceinline__ CT& swz(CT* base, int pa)
{
    constexpr int LO = pidx(C) & 15, HI = C & ~15;
    return (base + (pa ^ LO))[HI];
}

// Padded layout (shift S > 0): element i at i + (i >> S), one spare slot per
// 2^S, with sequences lspan<L, S>() slots apart.  The map is additive over
// disjoint bit fields (no carries), so every butterfly address is the lane
// part's slot plus a constant: no per-access index arithmetic at all.  A CPU
// model of the lane-group patterns puts S = 3 at 1.02-way bank conflicts on
// average for 16-byte values and S = 4 at 1.12-way for 8-byte ones.  S = 0
// is the XOR swizzle above.
template <int S> __host__ __device__ constexpr int lslot(int i)
{
    if constexpr (S == 0) return pidx(i);
    else return i + (i >> S);
}
template <int C, int S, typename CT> __device__ __forceinline__ CT& lsw(CT* base, int pa)
{
    if constexpr (S == 0) return swz<C>(base, pa);
    else return base[pa + lslot<S>(C)];
}
template <int L, int S> constexpr int lspan() { return S == 0 ? L : L + (L >> S); }

template <int SIGN, typename CT> __device__ __forceinline__ void dft2(CT* x)
{
    CT a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <int SIGN, typename CT> __device__ __forceinline__ void dft4(CT& x0, CT& x1, CT& x2, CT& x3)
{
    CT t0 = cadd(x0, x2), t1 = csub(x0, x2);
    CT t2 = cadd(x1, x3), t3 = cmuli<SIGN>(csub(x1, x3));
    x0 = cadd(t0, t2);
    x2 = csub(t0, t2);
    x1 = cadd(t1, t3);
    x3 = csub(t1, t3);
}

template <int SIGN, typename CT> __device__ __forceinline__ void dft4(CT* x) { dft4<SIGN>(x[0], x[1], x[2], x[3]); }

template <int SIGN, typename CT> __device__ __forceinline__ void dft8(CT* x)
{
    using R = decltype(x[0].x);
    const R h = (R)0.70710678118654752440;
    CT e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    CT o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    // o_s *= exp(SIGN * 2 pi i s / 8)
    CT w1 = CT{(o1.x - (R)SIGN * o1.y) * h, (o1.y + (R)SIGN * o1.x) * h};
    CT w2 = cmuli<SIGN>(o2);
    CT w3 = CT{(-o3.x - (R)SIGN * o3.y) * h, (-o3.y + (R)SIGN * o3.x) * h};
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cadd(e1, w1);
    x[5] = csub(e1, w1);
    x[2] = cadd(e2, w2);
    x[6] = csub(e2, w2);
    x[3] = cadd(e3, w3);
    x[7] = csub(e3, w3);
}

template <int R, int SIGN, typename CT> __device__ __forceinline__ void dftR(CT* x)
{
    if constexpr (R == 2) dft2<SIGN>(x);
    else if constexpr (R == 4) dft4<SIGN>(x);
    else dft8<SIGN>(x);
}

// x[r] *= w^r, w = tw[e] (conjugated for SIGN < 0)
template <int R, int SIGN, typename CT>
__device__ __forceinline__ void twiddle(CT* x, const CT* __restrict__ tw, int e)
{
    auto ld = [&](int i) {
        CT w = tw[i];
        if (SIGN < 0) w = cconj(w);
        return w;
    };
    if constexpr (R == 2) {
        x[1] = cmul(x[1], ld(e));
    } else if constexpr (R == 4) {
        const CT w1 = ld(e), w2 = ld(2 * e);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], cmul(w1, w2));
    } else {
        const CT w1 = ld(e), w2 = ld(2 * e), w4 = ld(4 * e);
        const CT w3 = cmul(w1, w2);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
        x[4] = cmul(x[4], w4);
        x[5] = cmul(x[5], cmul(w1, w4));
        x[6] = cmul(x[6], cmul(w2, w4));
        x[7] = cmul(x[7], cmul(w3, w4));
    }
}

template <int L, int Ns> constexpr int radix_at() { return (L / Ns) >= 8 ? 8 : (L / Ns); }

// ---------------------------------------------------------------------------
// Pass twiddle tables.  In the pass with sub-transform length Ns, butterfly
// j uses w^e, w^{2e}, w^{4e} with e = (j mod Ns) L/(Ns R).  Read from the
// natural table, consecutive lanes of the Ns = 8 pass are 8 entries (128 B)
// apart and all hit one bank group.  The pass table stores, for every pass
// with Ns > 1 and every k < Ns, the triple (w^e, w^{2e}, w^{4e}) at
// ptw_off<L, Ns>() + 3k (radix 2: w^e alone at ptw_off + k): lanes 48 (16) B
// apart cover all eight 16-byte bank groups, the loads of one butterfly share
// one address, and the values are those of the natural table (the third
// entry of a radix-4 triple is filler).
//
// Width PW = 7 stores all seven powers w^e .. w^{7e} of a radix-8 butterfly
// instead (lanes 7 entries apart: 7 is odd, so still every bank group): the
// butterfly then loads four more values and forms no products of twiddles,
// 16 fewer VALU operations.  The lane-group transforms use it; the block-wide
// passes keep the triples, whose tables are smaller.
// ---------------------------------------------------------------------------
template <int R, int PW = 3> constexpr int ptw_width() { return R == 2 ? 1 : (R == 8 && PW == 7 ? 7 : 3); }
template <int L, int Ns = 1, int PW = 3> constexpr int ptw_size()
{
    if constexpr (Ns >= L) return 0;
    else
        return (Ns > 1 ? ptw_width<radix_at<L, Ns>(), PW>() * Ns : 0) + ptw_size<L, Ns * radix_at<L, Ns>(), PW>();
}
template <int L, int NsT, int Ns = 1, int PW = 3> constexpr int ptw_off()
{
    if constexpr (Ns >= NsT) return 0;
    else
        return (Ns > 1 ? ptw_width<radix_at<L, Ns>(), PW>() * Ns : 0) + ptw_off<L, NsT, Ns * radix_at<L, Ns>(), PW>();
}

// dst[i] for i = t, t + T, ... < ptw_size<L, 1, PW>() from the natural table tw (L entries).
template <typename CT, int LOG, int Ns = 1, int PW = 3>
__device__ __forceinline__ void ptw_fill(CT* __restrict__ dst, const CT* __restrict__ tw, int t, int T)
{
    constexpr int L = 1 << LOG;
    if constexpr (Ns < L) {
        constexpr int R = radix_at<L, Ns>();
        if constexpr (Ns > 1) {
            constexpr int off = ptw_off<L, Ns, 1, PW>(), W = ptw_width<R, PW>();
            for (int i = t; i < W * Ns; i += T) {
                const int k = i / W, m = i % W;
                const int e = k * (L / (Ns * R));
                dst[off + i] = tw[(W == 7 ? e * (m + 1) : e << m) & (L - 1)];
            }
        }
        ptw_fill<CT, LOG, Ns * R, PW>(dst, tw, t, T);
    }
}

// x[r] *= w^r from the pass table entries at p (conjugated for SIGN < 0)
template <int R, int SIGN, int PW = 3, typename CT>
__device__ __forceinline__ void twiddle_p(CT* x, const CT* __restrict__ p)
{
    auto ld = [&](int i) {
        CT w = p[i];
        if (SIGN < 0) w = cconj(w);
        return w;
    };
    if constexpr (R == 2) {
        x[1] = cmul(x[1], ld(0));
    } else if constexpr (R == 8 && PW == 7) {
#pragma unroll
        for (int r = 1; r < 8; r++) x[r] = cmul(x[r], ld(r - 1));
    } else if constexpr (R == 4) {
        const CT w1 = ld(0), w2 = ld(1);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], cmul(w1, w2));
    } else {
        const CT w1 = ld(0), w2 = ld(1), w4 = ld(2);
        const CT w3 = cmul(w1, w2);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
        x[4] = cmul(x[4], w4);
        x[5] = cmul(x[5], cmul(w1, w4));
        x[6] = cmul(x[6], cmul(w2, w4));
        x[7] = cmul(x[7], cmul(w3, w4));
    }
}

// Twiddles of butterfly j in the pass (L, R, Ns): from the natural table, or
// with PTW from the pass table.
template <int L, int R, int Ns, int SIGN, bool PTW, int PW = 3, typename CT>
__device__ __forceinline__ void twiddle_j(CT* x, const CT* __restrict__ tw, int j)
{
    if constexpr (PTW)
        twiddle_p<R, SIGN, PW>(x, tw + ptw_off<L, Ns, 1, PW>() + ptw_width<R, PW>() * (j & (Ns - 1)));
    else twiddle<R, SIGN>(x, tw, (j & (Ns - 1)) * (L / (Ns * R)));
}

// One radix-R butterfly j of sequence base (load, twiddle, DFT); outputs stay in x.
template <typename CT, int L, int R, int Ns, int SIGN, bool PTW = false>
__device__ __forceinline__ void bfly_load(const CT* __restrict__ sb, int j, CT* x, const CT* __restrict__ tw)
{
    constexpr int NB = L / R;
    const int pj = pidx(j);
    static_for<0, R>([&](auto ri) { x[ri] = swz<decltype(ri)::value * NB>(sb, pj); });
    if constexpr (Ns > 1) twiddle_j<L, R, Ns, SIGN, PTW>(x, tw, j);
    dftR<R, SIGN>(x);
}

template <typename CT, int R, int Ns>
__device__ __forceinline__ void bfly_store(CT* __restrict__ db, int j, const CT* x)
{
    const int k = j & (Ns - 1);
    const int po = pidx((j - k) * R + k);
    static_for<0, R>([&](auto ri) { swz<decltype(ri)::value * Ns>(db, po) = x[ri]; });
}

template <int L, int Ns = 1> constexpr int npasses()
{
    if constexpr (Ns >= L) return 0;
    else return 1 + npasses<L, Ns * radix_at<L, Ns>()>();
}

// Ping-pong FFT over S sequences: pass 0 reads A and writes B, pass 1 reads B
// and writes A, ...  Ends with a barrier.  The result is in B when the pass
// count is odd (fft_in_b<LOG>()), else in A.
template <typename CT, int LOG, int SIGN, int T, bool PTW = false, int Ns = 1, int PASS = 0>
__device__ __forceinline__ void pp_fft(CT* A, CT* B, int S, const CT* __restrict__ tw, int tid = threadIdx.x)
{
    constexpr int L = 1 << LOG, LP = lpad(L);
    if constexpr (Ns < L) {
        constexpr int R = radix_at<L, Ns>();
        constexpr int NB = L / R;
        const CT* src = (PASS & 1) ? B : A;
        CT* dst = (PASS & 1) ? A : B;
        const int total = S * NB;
        for (int i = tid; i < total; i += T) {
            const int s = i / NB, j = i % NB;
            CT x[R];
            bfly_load<CT, L, R, Ns, SIGN, PTW>(src + s * LP, j, x, tw);
            bfly_store<CT, R, Ns>(dst + s * LP, j, x);
        }
        __syncthreads();
        pp_fft<CT, LOG, SIGN, T, PTW, Ns * R, PASS + 1>(A, B, S, tw, tid);
    }
}
template <int LOG> constexpr bool fft_in_b() { return (npasses<(1 << LOG)>() & 1) != 0; }

// In-place FFT (all butterflies of a pass are loaded before any is stored);
// for kernels whose LDS cannot hold two buffers.  SMAX bounds S.
template <typename CT, int LOG, int SIGN, int T, int SMAX, bool PTW = false, int Ns = 1>
__device__ __forceinline__ void ip_fft(CT* buf, int S, const CT* __restrict__ tw, int tid = threadIdx.x)
{
    constexpr int L = 1 << LOG, LP = lpad(L);
    if constexpr (Ns < L) {
        constexpr int R = radix_at<L, Ns>();
        constexpr int NB = L / R;
        constexpr int MAXB = (SMAX * NB + T - 1) / T;
        const int total = S * NB;
        CT x[MAXB][R];
#pragma unroll
        for (int q = 0; q < MAXB; q++) {
            const int i = tid + q * T;
            if (i < total) bfly_load<CT, L, R, Ns, SIGN, PTW>(buf + (i / NB) * LP, i % NB, x[q], tw);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXB; q++) {
            const int i = tid + q * T;
            if (i < total) bfly_store<CT, R, Ns>(buf + (i / NB) * LP, i % NB, x[q]);
        }
        __syncthreads();
        ip_fft<CT, LOG, SIGN, T, SMAX, PTW, Ns * R>(buf, S, tw, tid);
    }
}

// ---------------------------------------------------------------------------
// Group FFT: one sequence of length L held by a group of G = min(64, L/8)
// lanes of one wave, lane gl holding elements gl + G*q, q < L/G, in v[] on
// entry and on exit (natural order).  Passes exchange data through the
// group's own LDS region (L values, swizzled), so no block barrier is
// needed: the lanes of a wave issue their LDS operations in program order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int L> constexpr int group_lanes() { return L / 8 < 64 ? L / 8 : 64; }

// xo: a constant XORed into every slot of the group's region (a
// per-sequence swizzle of the caller's layout; any xo < 16 keeps the access
// patterns of the passes free of bank conflicts).  LS: the region's layout
// (lslot; 0 = XOR swizzle, xo must be 0 otherwise).
template <typename CT, int LOG, int SIGN, bool PTW = false, int PW = 3, int Ns = 1, int PASS = 0, int LS = 0>
__device__ __forceinline__ void group_fft(CT* v, CT* __restrict__ reg, int gl, const CT* __restrict__ tw, int xo = 0)
{
    constexpr int L = 1 << LOG;
    constexpr int G = group_lanes<L>();
    if constexpr (Ns < L) {
        constexpr int R = radix_at<L, Ns>();
        constexpr int NBL = L / (R * G); // butterflies per lane
        constexpr bool LAST = Ns * R == L;
        CT x[NBL][R];
#pragma unroll
        for (int i = 0; i < NBL; i++) {
            const int j = gl + G * i;
            if constexpr (PASS == 0) {
#pragma unroll
                for (int r = 0; r < R; r++) x[i][r] = v[i + r * NBL];
            } else {
                const int pj = lslot<LS>(j) ^ xo;
                static_for<0, R>([&](auto ri) { x[i][ri] = lsw<decltype(ri)::value * (L / R), LS>(reg, pj); });
            }
        }
        if constexpr (PASS > 0) wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NBL; i++) {
            const int j = gl + G * i;
            if constexpr (Ns > 1) twiddle_j<L, R, Ns, SIGN, PTW, PW>(x[i], tw, j);
            dftR<R, SIGN>(x[i]);
        }
        if constexpr (LAST) {
            // element j + r*Ns = gl + G*(i + r*NBL)
#pragma unroll
            for (int i = 0; i < NBL; i++)
#pragma unroll
                for (int r = 0; r < R; r++) v[i + r * NBL] = x[i][r];
        } else {
#pragma unroll
            for (int i = 0; i < NBL; i++) {
                const int j = gl + G * i;
                const int k = j & (Ns - 1);
                const int po = lslot<LS>((j - k) * R + k) ^ xo;
                static_for<0, R>([&](auto ri) { lsw<decltype(ri)::value * Ns, LS>(reg, po) = x[i][ri]; });
            }
            wave_lds_sync();
            group_fft<CT, LOG, SIGN, PTW, PW, Ns * R, PASS + 1, LS>(v, reg, gl, tw, xo);
        }
    }
}
