// diag.hpp -- small kernels of the host-side entry points: the energy
// reduction, the on-device diagnostics of the kept host I/O and the box
// utilities.  Included by mwf.hip only.
#pragma once
#include "kernels.hpp"

// Sum of each member's energy partials in a fixed order with compensation;
// one thread per member.
__global__ void k_ereduce(const double* __restrict__ part, int n, int nmem, double* __restrict__ out)
{
    const int mbr = blockIdx.x * blockDim.x + threadIdx.x;
    if (mbr < nmem) {
        part += (size_t)mbr * n;
        double s = 0.0, comp = 0.0;
        for (int i = 0; i < n; i++) {
            const double y = part[i] - comp;
            const double t = s + y;
            comp = (t - s) - y;
            s = t;
        }
        out[mbr] = s;
    }
}

// ---------------------------------------------------------------------------
// On-device diagnostics of the kept host I/O, so that the 3/2-grid field
// never has to leave the GPU for them.
// ---------------------------------------------------------------------------

// io_save_Uspec (program/io.f90:345-395): d = |spec(k, m, n)| of
// var_mpt2spec(vel_c) (no laminar profile); n_max[n] = max over (k, m),
// m_max[|m|] = max over (k, n).  Block n, threads over m.  The m maxima go
// through an atomic max on the bit pattern, which orders non-negative doubles
// like their values, so the result does not depend on the order of updates.
__global__ __launch_bounds__(256) void k_specmax(Geo g, const double2* __restrict__ c,
                                                 double* __restrict__ n_max,
                                                 unsigned long long* __restrict__ m_max)
{
    const int n = blockIdx.x;
    __shared__ double red[256];
    double dn = 0.0;
    for (int mi = threadIdx.x; mi < g.M; mi += blockDim.x) {
        const double2* cn = c + (size_t)n * g.K * g.ldM;
        double dm = 0.0;
        for (int k = 0; k < g.K0; k++) {
            const double2 ps = cn[(size_t)k * g.ldM + mi];
            const double2 ph = k > 0 ? cn[(size_t)(g.K0 - 1 + k) * g.ldM + mi] : double2{0.0, 0.0};
            double2 u, v, w;
            mpt2spec_k(g, k, mi, n, ps, ph, false, u, v, w);
            dm = fmax(dm, sqrt(u.x * u.x + u.y * u.y));
            if (k > 0) dm = fmax(dm, sqrt(v.x * v.x + v.y * v.y));
            dm = fmax(dm, sqrt(w.x * w.x + w.y * w.y));
        }
        dn = fmax(dn, dm);
        const int ma = mi < g.MM ? mi : g.M - mi;
        atomicMax(&m_max[ma], (unsigned long long)__double_as_longlong(dm));
    }
    red[threadIdx.x] = dn;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) n_max[n] = red[0];
}

// vel_history (program/velocity.f90:164-180, called from io.f90:472-483):
// the u, v, w profiles at wall-normal position y of the physical field at x
// index 0 and nh z-samples ji = (j*Q)/nh + 1.  hb holds the basis at y:
// hb[k] = h_k(y), k < K0 (1, sin, cos, ...), hb[K0 - 1 + k] = g_k(y), k >= 1.
// ans is (4, nh) column-major: z, u, v, w.  The sums run in modes.M4.f90:83-108
// order (velU, velV, velW).
__global__ void k_velhist(const double* __restrict__ phys, int Q, int KK, int K0, int nh, double lz,
                          const double* __restrict__ hb, double* __restrict__ ans)
{
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nh) return;
    const int ji = (int)(((long long)j * Q) / nh) + 1;
    const double* V = phys + (size_t)ji * KK;
    double u = V[0], v = 0.0, w = V[2 * K0 - 1];
    for (int k = 1; k < K0; k++) u = u + hb[k] * V[k];
    for (int k = 1; k < K0; k++) v = k == 1 ? hb[K0 - 1 + k] * V[K0 - 1 + k] : v + hb[K0 - 1 + k] * V[K0 - 1 + k];
    for (int k = 1; k < K0; k++) w = w + hb[k] * V[2 * K0 - 1 + k];
    ans[4 * j + 0] = (double)ji * lz / Q;
    ans[4 * j + 1] = u;
    ans[4 * j + 2] = v;
    ans[4 * j + 3] = w;
}

// eposC (program/velocity.f90:128-142) of one point's KK wall-normal values:
// the y-means of u^2, v^2, w^2, i.e. nluw(u,u)(0) = u0 u0 + (u1 u1)/2 + ...,
// nlvv(v)(0) = v1**2/2 + ... (modes.M4.f90:38-81), in the reference's order.
__device__ __forceinline__ void eposc_of(const double* __restrict__ V, int K0, double& eu, double& ev, double& ew)
{
#pragma clang fp contract(off)
    eu = V[0] * V[0];
    ev = 0.0;
    ew = V[2 * K0 - 1] * V[2 * K0 - 1];
    for (int k = 1; k < K0; k++) eu = eu + (V[k] * V[k]) / 2;
    for (int k = 1; k < K0; k++) {
        const double t = V[K0 - 1 + k] * V[K0 - 1 + k] / 2;
        ev = k == 1 ? t : ev + t;
    }
    for (int k = 1; k < K0; k++) ew = ew + (V[2 * K0 - 1 + k] * V[2 * K0 - 1 + k]) / 2;
}

// turb_sample (program/turb.f90:37-47): eposC of the physical field at
// x = m*dm, z = n*dn.  out is (3, SN, SM) column-major.
__global__ void k_eposc(const double* __restrict__ phys, int Q, int KK, int K0, int dm, int dn, int SM, int SN,
                        double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= SM * SN) return;
    const int m = i / SN, n = i - m * SN;
    double eu, ev, ew;
    eposc_of(phys + ((size_t)m * dm * Q + (size_t)n * dn) * KK, K0, eu, ev, ew);
    out[3 * i + 0] = eu;
    out[3 * i + 1] = ev;
    out[3 * i + 2] = ew;
}

// Turbulent points of io_writeVTK_txz (program/io.f90:550-617, driven by
// utils/turbfrac.f90:14-38): on the grid iz = 0, zs, ... < Q, ix = 0, xs, ...
// < P, the points with epos > ct (epos = (eu + ev) + ew, program/velocity.f90
// :115-126).  blockIdx.y selects the member (phys arrays back to back);
// count[member] accumulates with integer atomics, so the count does not
// depend on the order of the updates.
__global__ void k_turbcount(const double* __restrict__ phys, const Geo* __restrict__ geos, int xs, int zs,
                            double ct, unsigned long long* __restrict__ count)
{
#pragma clang fp contract(off)
    const Geo& g = geos[blockIdx.y];
    phys += blockIdx.y * ((size_t)g.KK * g.Q * g.P);
    const int nx = (g.P + xs - 1) / xs, nz = (g.Q + zs - 1) / zs;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int hit = 0;
    if (i < nx * nz) {
        const int ix = (i / nz) * xs, iz = (i % nz) * zs;
        double eu, ev, ew;
        eposc_of(phys + ((size_t)ix * g.Q + iz) * g.KK, g.K0, eu, ev, ew);
        hit = ((eu + ev) + ew) > ct;
    }
    // one atomic per wave
    const unsigned long long wave = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && wave) atomicAdd(&count[blockIdx.y], (unsigned long long)__popcll(wave));
}

// One sample of the ensemble's on-device tracking, for the nmem members of
// a class: energies e[] (vel_energy) and, when count is set, turbulent point
// counts (k_turbcount, then cleared for the next sample) as fractions
// count / den.  The sample goes to row *fill of the trace buffers (cap rows
// of nmem), the first step with energy <= thr to relam[] (step = sample index
// since tracking began times every); *fill and *total then advance.  The host
// drains the buffers before they fill.  One block.
__global__ void k_track_push(int nmem, const double* __restrict__ e, unsigned long long* __restrict__ count,
                             double den, double thr, long long every, double* __restrict__ etrace,
                             double* __restrict__ ftrace, int* __restrict__ fill, long long* __restrict__ total,
                             long long* __restrict__ relam, int cap)
{
    const int row = *fill;
    const long long step = *total * every;
    for (int i = threadIdx.x; i < nmem; i += blockDim.x) {
        const double ei = e[i];
        if (row < cap) {
            etrace[(size_t)row * nmem + i] = ei;
            if (count) ftrace[(size_t)row * nmem + i] = (double)count[i] / den;
        }
        if (count) count[i] = 0ull;
        if (relam[i] < 0 && ei <= thr) relam[i] = step;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *fill = row + 1;
        *total = *total + 1;
    }
}

// ---------------------------------------------------------------------------
// State utilities for large-box runs (utils/doubleX.f90, utils/quadX.f90,
// utils/interpstate.f90).
// ---------------------------------------------------------------------------

// var_doubleX / var_quadX (program/variables.f90:147-185): the state of a box
// `factor` times longer in x holding the same field, mode m -> factor*m.
// Gather form over the destination: signed m_d = factor*m_s where divisible,
// zero elsewhere (modes with |factor*m| > MM-1 are dropped).  One thread per
// (n, slot, m_d), padding slots of ldM included.
__global__ void k_extend_x(Geo g, const double2* __restrict__ src, double2* __restrict__ dst, int factor)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)g.NN * g.K * g.ldM;
    if (i >= total) return;
    const int mi = (int)(i % g.ldM);
    const size_t row = i / g.ldM; // n * K + slot
    double2 v = double2{0.0, 0.0};
    if (mi < g.M) {
        const int md = mi < g.MM ? mi : mi - g.M;
        if (md % factor == 0) {
            const int ms = md / factor;
            v = src[row * g.ldM + (ms >= 0 ? ms : g.M + ms)];
        }
    }
    dst[i] = v;
}

// utils/interpstate.f90:30-31: u := (1 - d) u1 + d u2, in that order
__global__ void k_blend(size_t n, const double2* __restrict__ a, const double2* __restrict__ b, double d,
                        double2* __restrict__ out)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 x = a[i], y = b[i];
    out[i] = double2{(1.0 - d) * x.x + d * y.x, (1.0 - d) * x.y + d * y.y};
}
