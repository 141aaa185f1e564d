// common.hpp -- complex arithmetic, geometry and compile-time helpers for the
// gfx950 kernels of the Model Waleffe Flow timestepper.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

template <typename R> struct cplx_of;
template <> struct cplx_of<double> { using type = double2; };
template <> struct cplx_of<float> { using type = float2; };
template <typename R> using cplx_t = typename cplx_of<R>::type;

// Streaming (non-temporal) vector load/store of one complex value: data
// that is touched once per step and should not displace reused lines in L2.
template <typename CT> __device__ __forceinline__ void st_stream(CT* p, CT v)
{
    using E = decltype(v.x);
    typedef E V2 __attribute__((ext_vector_type(2)));
    const V2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}
// Store of one element of an NTx-long piece: streamed when the piece spans at
// least 64 bytes; shorter pieces are left to merge in L2 with their neighbours.
template <typename CT> __device__ __forceinline__ void st_piece(CT* p, CT v, bool wide)
{
    if (wide) st_stream(p, v);
    else *p = v;
}
template <typename CT> __device__ __forceinline__ CT ld_stream(const CT* p)
{
    using E = decltype(p->x);
    typedef E V2 __attribute__((ext_vector_type(2)));
    const V2 w = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return CT{w.x, w.y};
}

template <typename CT> __device__ __forceinline__ CT cadd(CT a, CT b) { return CT{a.x + b.x, a.y + b.y}; }
template <typename CT> __device__ __forceinline__ CT csub(CT a, CT b) { return CT{a.x - b.x, a.y - b.y}; }
template <typename CT> __device__ __forceinline__ CT cmul(CT a, CT b)
{
    return CT{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename CT> __device__ __forceinline__ CT cconj(CT a) { return CT{a.x, -a.y}; }
template <typename CT> __device__ __forceinline__ CT cscale(CT a, decltype(a.x) s) { return CT{a.x * s, a.y * s}; }
// multiply by SIGN * i
template <int SIGN, typename CT> __device__ __forceinline__ CT cmuli(CT a)
{
    return SIGN > 0 ? CT{-a.y, a.x} : CT{a.y, -a.x};
}

template <typename CT> __device__ __forceinline__ CT ld_tw(const CT* __restrict__ tw, int e)
{
    return tw[e];
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(std::forward<F>(f));
    }
}

// Geometry and physical parameters of one model instance.
struct Geo {
    int K0, K, KK;     // y-modes, mpt slots (2K0-1), spec slots (3K0-1)
    int MM, M, ldM;    // x: modes |m| < MM, M = 2MM-1 stored, ldM = row pitch
    int NN, P, Q;      // z modes, padded grid P = 3MM (x), Q = 3NN (z)
    int NTx;           // z-modes per x-kernel block (power of two)
    int lgNT;          // log2(NTx)
    int NTs;           // z-modes per NL piece (a multiple of NTx, power of two)
    int lgNS;          // log2(NTs)
    int XB;            // x-rows per z-kernel block
    int UL, NLL;       // lines per x-row of U and NL: 2 KK and KK (advective form), KK and 2 KK (flux form)
    double alpha, gamma, re, dt, drag, vdrag, hypo, cth, sth, beta;
    double scale;      // 1/(P*Q), transform.fftw5.f90:242
    int dragall;
};

// Member addressing of one launch: blockIdx.z selects the member; its
// geometry and the offsets of its buffers follow from this.
struct Batch {
    const Geo* geo;      // [members] geometry and physical parameters
    size_t c_stride;     // double2 elements between members' c / onl
    size_t u_stride;     // complex elements between members' U
    size_t nl_stride;    // complex elements between members' NL
    int ep_stride;       // energy partials per member
    size_t nl_part = 0;  // complex elements between the three fold-phase partials of NL (0: NL is whole)
};

// Workgroups of a launch go round-robin to the 8 XCDs (block b on XCD b % 8),
// each with its own L2.  xcd_tile(b, n) is a bijection of [0, n) that gives
// every XCD a contiguous run of tiles, so that tiles sharing 128-byte lines
// (neighbouring n-tiles or x-rows) meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int per = n >> 3;
    if (b >= per * 8) return b;
    return (b & 7) * per + (b >> 3);
}

// U (the x-transformed lines, UL per x row) is stored U[x][line][n]: the z
// kernel reads whole rows.  NL (the z-forward products, NLL lines) is stored
// in panels NL[tile][line][x][s], n = tile*NTs + s: a z-kernel thread writes
// NTs-long pieces, and an x-kernel block, which takes NTx <= NTs columns,
// reads its NTx-wide stripe of a panel.  With NTs = NTx the stripe is the
// whole panel.  Where NTx values are under 128 bytes (FP64 at MM >= 512, FP32
// at MM >= 256), NTs = 2 NTx: the writers store twice as long runs (whole L2
// lines at FP64 MM = 512), and the two blocks that read the halves of each
// piece run side by side on one XCD, so the lines are fetched once.
// U's writers store NTx- or NN-long pieces; partial-line writes cost no extra
// HBM reads, partial-line reads would.
__device__ __forceinline__ size_t uidx(const Geo& g, int x, int line, int n)
{
    return ((size_t)x * g.UL + line) * g.NN + n;
}
// NL stripe of an x-kernel block (columns n0 .. n0 + NTx - 1, slot): element
// i = x * NTx + t at offset pan_at(g, i) from the returned base.
template <typename CT> __device__ __forceinline__ const CT* nl_panel(const Geo& g, const CT* NL, int n0, int slot)
{
    return NL + ((size_t)(n0 >> g.lgNS) * g.NLL + slot) * g.P * g.NTs + (n0 & (g.NTs - 1));
}
__device__ __forceinline__ int pan_at(const Geo& g, int i) { return ((i >> g.lgNT) << g.lgNS) + (i & (g.NTx - 1)); }
__device__ __forceinline__ size_t nidx(const Geo& g, int x, int slot, int n)
{
    const int tile = n >> g.lgNS, t = n & (g.NTs - 1);
    return (((size_t)tile * g.NLL + slot) * g.P + x) * g.NTs + t;
}

// wall-normal derivative table ad_k1, ad_k2 (program/variables.f90:528-531)
__device__ __forceinline__ double ak1_of(int k, double beta) { return ((k & 1) ? -1.0 : 1.0) * k * beta; }
__device__ __forceinline__ double ak2_of(int k, double beta) { return -k * k * beta * beta; }
