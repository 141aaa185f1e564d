// fft.hpp -- batched Stockham FFTs of power-of-two length, in LDS, for gfx950.
//
// The 3/2-padded transforms of the reference (lengths P = 3*MM, Q = 3*NN,
// program/transform.fftw5.f90:33-59) are split by the kernels into a radix-3
// fold done in registers and three power-of-two transforms of length MM (or
// NN) done here.  A block transforms S sequences stored back to back in LDS
// (sequence s at buf[s*L]).  Each pass is a Stockham radix-R step:
// every butterfly reads R points L/R apart, applies the twiddles of the
// current sub-transform, does a radix-R DFT in registers and writes the R
// results Ns apart, so the output comes out in natural order.  Passes run in
// place: a block first loads all its butterflies into registers, then
// synchronises, then writes.
//
// Transforms are unnormalised: sign +1 computes sum_j x_j exp(+2 pi i jk/L)
// (FFTW's backward), sign -1 exp(-2 pi i jk/L).  Twiddles come from a table
// tw[j] = exp(+2 pi i j / L) rounded from long double on the host.
#pragma once
#include "common.hpp"

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

// One in-place Stockham pass over S sequences of length L.
template <typename CT, int L, int R, int Ns, int SIGN, int T, int MAXB>
__device__ __forceinline__ void stockham_pass(CT* __restrict__ buf, int S, const CT* __restrict__ tw)
{
    constexpr int NB = L / R;
    const int total = S * NB;
    CT x[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int i = threadIdx.x + q * T;
        if (i < total) {
            const int s = i / NB, j = i % NB;
            const CT* b = buf + s * L;
#pragma unroll
            for (int r = 0; r < R; r++) x[q][r] = b[j + r * NB];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int i = threadIdx.x + q * T;
        if (i < total) {
            const int s = i / NB, j = i % NB;
            const int k = j & (Ns - 1);
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    CT w = ld_tw(tw, r * k * (L / (Ns * R)));
                    if (SIGN < 0) w = cconj(w);
                    x[q][r] = cmul(x[q][r], w);
                }
            }
            dftR<R, SIGN>(x[q]);
            CT* b = buf + s * L;
            const int o = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) b[o + r * Ns] = x[q][r];
        }
    }
    __syncthreads();
}

// All passes: radix 8 while at least 8 remain, then 4 or 2.
template <typename CT, int LOG, int SIGN, int T, int SMAX, int Ns = 1>
__device__ __forceinline__ void lds_fft(CT* __restrict__ buf, int S, const CT* __restrict__ tw)
{
    constexpr int L = 1 << LOG;
    if constexpr (Ns < L) {
        constexpr int rem = L / Ns;
        constexpr int R = rem >= 8 ? 8 : rem;
        constexpr int MAXB = (SMAX * (L / R) + T - 1) / T;
        stockham_pass<CT, L, R, Ns, SIGN, T, MAXB>(buf, S, tw);
        lds_fft<CT, LOG, SIGN, T, SMAX, Ns * R>(buf, S, tw);
    }
}
