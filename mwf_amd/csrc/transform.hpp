// transform.hpp -- the reference's transform module on the device: the
// 3/2-dealiased 2-D transforms between type(spec) and type(phys)
// (tra_spec2phys, program/transform.fftw5.f90:64-87; tra_phys2spec, :122-145)
// as standalone passes, for the kept host code and utilities that call them
// directly.  The time step never uses these: its transforms are fused into
// the step kernels.
//
// Device layouts:
//   spec   double2 [NN][KK][ldM]   m in the reference order (m<0 at M+m)
//   U      cplx<R> [P][2*KK][NN]   lines 0..KK-1 used (the step's U buffer)
//   NL     cplx<R> panels, nidx()  (the step's NL buffer)
//   phys   double  [P][Q][KK]      type(phys), element (c, z, x) at c + KK*(z + Q*x)
#pragma once
#include "kernels.hpp"

// spec -> U, x inverse of one spec slot per block: block (n-tile, slot).
// The padded inverse x transform (c2c, +1, unnormalised, length P = 3 MM,
// transform.fftw5.f90:71-77) as three folded MM-point transforms.
template <typename R, int LOGMM>
__global__ __launch_bounds__(XT) void k_sxinv(Geo g, const double2* __restrict__ s, cplx_t<R>* __restrict__ U,
                                              const cplx_t<R>* __restrict__ twMM, const double2* __restrict__ twP)
{
    using CT = cplx_t<R>;
    constexpr int MM = 1 << LOGMM;
    constexpr int LP = lpad(MM);
    constexpr int IPT = XTILE / XT;
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);
    const int NT = g.NTx;
    const int n0 = blockIdx.x * NT;
    const int slot = blockIdx.y;
    const int tid = threadIdx.x;
    const int nitems = NT * MM;
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int it = tid + q * XT;
        if (it >= nitems) continue;
        const int t = it >> LOGMM, r = it & (MM - 1);
        const double2* sn = s + ((size_t)(n0 + t) * g.KK + slot) * g.ldM;
        const double2 X1 = sn[r];                                        // m = r
        const double2 X2 = r ? sn[MM - 1 + r] : double2{0.0, 0.0};       // m = r - MM
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
    CT* Ul = U + (size_t)slot * g.NN + n0;
    const size_t ustride = (size_t)g.UL * g.NN;
#pragma unroll
    for (int q = 0; q < 3 * IPT; q++) {
        const int it = tid + q * XT;
        if (it >= 3 * nitems) continue;
        const int x = it >> g.lgNT, t = it & (NT - 1);
        const int a = x / 3, b = x - 3 * a;
        Ul[x * ustride + t] = buf[((size_t)t * 3 + b) * LP + pidx(a)];
    }
}

// phys -> NL, z forward of one block of x-rows: scale 1/(PQ) (transform.fftw5
// .f90:128-130), r2c of length Q = 3 NN (:131), keep n < NN.  Two real
// components share one complex transform; the three phases z = 3r + b are
// NN-point transforms combined as X[n] = sum_b F_b[n] w^{-nb}, w = exp(2 pi i/Q).
template <int LOGNN> constexpr int zfp_threads() { return (1 << LOGNN) > ZT ? (1 << LOGNN) : ZT; }
template <typename R, int LOGNN> constexpr size_t zfp_lds()
{
    return (size_t)(zfp_threads<LOGNN>() >> LOGNN) * ZPB * 3 * lpad(1 << LOGNN) * sizeof(cplx_t<R>);
}

template <typename R, int LOGNN>
__global__ __launch_bounds__(zfp_threads<LOGNN>()) void k_zfwd_phys(Geo g, const double* __restrict__ phys,
                                                                    cplx_t<R>* __restrict__ NL,
                                                                    const cplx_t<R>* __restrict__ twNN,
                                                                    const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    constexpr int NN = 1 << LOGNN;
    constexpr int TZ = zfp_threads<LOGNN>();
    constexpr int XB = TZ / NN;
    constexpr int SEQ = XB * ZPB * 3;
    constexpr int LP = lpad(NN);
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);
    const int tid = threadIdx.x;
    const int xl = tid >> LOGNN, r = tid & (NN - 1);
    const int x = blockIdx.x * XB + xl;
    const bool valid = x < g.P;
    const int KK = g.KK, Q = g.Q;
    const R scale = (R)g.scale;
    const int nbatch = (KK + 2 * ZPB - 1) / (2 * ZPB);
    for (int B = 0; B < nbatch; B++) {
        for (int p = 0; p < ZPB; p++) {
            const int c1 = B * 2 * ZPB + 2 * p, c2 = c1 + 1;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                CT z{(R)0, (R)0};
                if (valid && c1 < KK) {
                    const double* v = phys + ((size_t)x * Q + 3 * r + b) * KK;
                    z = CT{(R)v[c1], c2 < KK ? (R)v[c2] : (R)0};
                }
                buf[((xl * ZPB + p) * 3 + b) * LP + pidx(r)] = z;
            }
        }
        __syncthreads();
        ip_fft<CT, LOGNN, -1, TZ, SEQ>(buf, SEQ, twNN);
        for (int p = 0; p < ZPB; p++) {
            const int c1 = B * 2 * ZPB + 2 * p, c2 = c1 + 1;
            const int rm = (NN - r) & (NN - 1);
            CT X1{(R)0, (R)0}, X2{(R)0, (R)0};
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const CT zn = buf[((xl * ZPB + p) * 3 + b) * LP + pidx(r)];
                const CT zm = buf[((xl * ZPB + p) * 3 + b) * LP + pidx(rm)];
                // F1, F2: transforms of the two real components (Hermitian split)
                const CT F1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                const CT F2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                if (b == 0) {
                    X1 = F1;
                    X2 = F2;
                } else {
                    const CT w = cconj(twQ[r * b]);
                    X1 = cadd(X1, cmul(F1, w));
                    X2 = cadd(X2, cmul(F2, w));
                }
            }
            if (valid && c1 < KK) {
                NL[nidx(g, x, c1, r)] = cscale(X1, scale);
                if (c2 < KK) NL[nidx(g, x, c2, r)] = cscale(X2, scale);
            }
        }
        __syncthreads();
    }
}

// NL -> spec, x forward of one slot per block: block (n-tile, slot).  The
// forward x transform (c2c, -1, length P, transform.fftw5.f90:137) as three
// MM-point transforms of the decimated rows, X[m] = sum_b w^{-mb} F_b[m mod MM],
// truncated to |m| < MM.
template <typename R, int LOGMM>
__global__ __launch_bounds__(XT) void k_sxfwd(Geo g, const cplx_t<R>* __restrict__ NL, double2* __restrict__ s,
                                              const cplx_t<R>* __restrict__ twMM, const double2* __restrict__ twP)
{
    using CT = cplx_t<R>;
    constexpr int MM = 1 << LOGMM;
    constexpr int LP = lpad(MM);
    constexpr int IPT = XTILE / XT;
    extern __shared__ unsigned char smem_raw[];
    CT* buf = reinterpret_cast<CT*>(smem_raw);
    const int NT = g.NTx;
    const int n0 = blockIdx.x * NT;
    const int slot = blockIdx.y;
    const int tid = threadIdx.x;
    const int nitems = NT * MM;
    const CT* pan = nl_panel(g, NL, n0, slot);
#pragma unroll
    for (int q = 0; q < 3 * IPT; q++) {
        const int it = tid + q * XT;
        if (it >= 3 * nitems) continue;
        const int x = it >> g.lgNT, t = it & (NT - 1);
        const int a = x / 3, b = x - 3 * a;
        buf[((size_t)t * 3 + b) * LP + pidx(a)] = pan[pan_at(g, it)];
    }
    __syncthreads();
    ip_fft<CT, LOGMM, -1, XT, 3 * (XTILE >> LOGMM)>(buf, 3 * NT, twMM);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int it = tid + q * XT;
        if (it >= nitems) continue;
        const int t = it >> LOGMM, r = it & (MM - 1);
        const CT* bt = buf + (size_t)t * 3 * LP;
        const CT F0 = bt[pidx(r)];
        double2 Xp = double2{(double)F0.x, (double)F0.y}, Xn = Xp;
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
        double2* sn = s + ((size_t)(n0 + t) * g.KK + slot) * g.ldM;
        sn[r] = Xp;
        if (r) sn[MM - 1 + r] = Xn;
    }
}
