// ic.cpp -- seeded Gaussian-spectrum initial condition (host only).
//
// The reference has no IC generator (its README names a randIC utility that
// is not in the tree, and program/variables.f90:233-286 draw unseeded
// uniforms), so the synthetic inputs are defined here:
//  1. SplitMix64(seed), uniforms u = (x >> 11) 2^-53, Box-Muller pairs.
//  2. For n = 0..NN-1, m = 0..M-1 (reference storage order, m < 0 at M+m),
//     spec slot c = 0..KK-1: s = A(m,n) (g1 + i g2)/sqrt(2) with
//     A = scl exp(-(kx^2 + kz^2)/k0^2), kx = alpha m, kz = gamma n.
//  3. Per wavenumber, orthogonal projection of (u_k, v_k, w_k) onto the
//     divergence-free directions of the toroidal/poloidal form that
//     var_mpt2spec spans (program/variables.f90:390-416):
//       psi: (-i kz, 0, i kx)              psi_k = a^H s / kh^2
//       phi: (i kx D_k, kh^2, i kz D_k)    phi_k = b^H s / (kh^2 (D_k^2 + kh^2))
//     with kh^2 = kx^2 + kz^2 and D_k = ad_k1(k) = (-1)^k k beta.
//  4. Zero deviation at (m,n) = (0,0): the mean flow is the laminar profile.
//  5. n = 0 column made Hermitian: c(M-m, 0) = conj(c(m, 0)).
//  6. Scaled so that the perturbation energy of vel_energy (the mean square
//     velocity, program/velocity.f90:115-162) equals scl^2: scl is the rms
//     velocity of the perturbation and A(m,n) only shapes its spectrum.
//     (Unscaled, every mode at amplitude ~scl puts 10^2-10^4 times the
//     laminar energy into the box and the step diverges within tens of steps.)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mwf.h"

namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
} // namespace

extern "C" int mwf_ic_gaussian(const mwf_params* p, uint64_t seed, double scl, double k0, double* re, double* im)
{
    if (!p || !re || !im || p->k0 < 2 || p->mm < 2 || p->nn < 1 || !(k0 > 0)) return MWF_E_ARG;
    const int K0 = p->k0, K = 2 * K0 - 1, KK = 3 * K0 - 1, MM = p->mm, M = 2 * MM - 1, NN = p->nn;
    const double beta = 3.1415926535897931 / 2.0, twopi = 6.283185307179586476925286766559;
    SplitMix64 rng{seed};
    std::vector<double> sr(KK), si(KK);
    double esum = 0.0; // Parseval energy of the projected field
    std::memset(re, 0, sizeof(double) * (size_t)K * M * NN);
    std::memset(im, 0, sizeof(double) * (size_t)K * M * NN);
    for (int n = 0; n < NN; n++)
        for (int mi = 0; mi < M; mi++) {
            const int m = mi < MM ? mi : mi - M;
            const double kx = p->alpha * m, kz = p->gamma * n;
            const double amp = scl * std::exp(-(kx * kx + kz * kz) / (k0 * k0));
            for (int c = 0; c < KK; c++) {
                const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
                const double rad = std::sqrt(-2.0 * std::log(u1));
                sr[c] = amp * rad * std::cos(twopi * u2) / std::sqrt(2.0);
                si[c] = amp * rad * std::sin(twopi * u2) / std::sqrt(2.0);
            }
            const double kh2 = kx * kx + kz * kz;
            if (kh2 == 0.0) continue; // (0,0): zero deviation
            for (int k = 0; k < K0; k++) {
                const int cu = k, cv = K0 - 1 + k, cw = 2 * K0 - 1 + k;
                // a^H s = i (kz u - kx w)
                const double ar = kz * sr[cu] - kx * sr[cw], ai = kz * si[cu] - kx * si[cw];
                const double psr = -ai / kh2, psi = ar / kh2;
                const size_t ip = (size_t)k + (size_t)K * ((size_t)mi + (size_t)M * n);
                re[ip] = psr;
                im[ip] = psi;
                // energy weights: 2 for n > 0 (the -n half), m > 0 only at n = 0
                // (the column is made Hermitian below); y weight 1 for k = 0
                const double wn = n > 0 ? 2.0 : (m > 0 ? 2.0 : (m == 0 ? 1.0 : 0.0));
                const double yw = k == 0 ? 1.0 : 0.5;
                // u = -i kz psi, w = i kx psi (+ phi parts below)
                double ur = kz * psi, ui = -kz * psr, wr = -kx * psi, wi = kx * psr, vr = 0.0, vi = 0.0;
                if (k == 0) {
                    esum += wn * yw * (ur * ur + ui * ui + wr * wr + wi * wi);
                    continue;
                }
                const double D = ((k & 1) ? -1.0 : 1.0) * k * beta;
                // b^H s = -i kx D u + kh2 v - i kz D w
                const double tr = kx * D * si[cu] + kh2 * sr[cv] + kz * D * si[cw];
                const double ti = -kx * D * sr[cu] + kh2 * si[cv] - kz * D * sr[cw];
                const double den = kh2 * (D * D + kh2);
                const size_t iph = (size_t)(K0 - 1 + k) + (size_t)K * ((size_t)mi + (size_t)M * n);
                re[iph] = tr / den;
                im[iph] = ti / den;
                const double fr = tr / den, fi = ti / den;
                ur += -kx * D * fi;
                ui += kx * D * fr;
                wr += -kz * D * fi;
                wi += kz * D * fr;
                vr = kh2 * fr;
                vi = kh2 * fi;
                esum += wn * yw * (ur * ur + ui * ui + vr * vr + vi * vi + wr * wr + wi * wi);
            }
        }
    if (esum > 0.0) {
        const double f = scl / std::sqrt(esum);
        for (size_t i = 0; i < (size_t)K * M * NN; i++) {
            re[i] *= f;
            im[i] *= f;
        }
    }
    // Hermitian n = 0 column
    for (int m = 1; m < MM; m++)
        for (int k = 0; k < K; k++) {
            const size_t ip = (size_t)k + (size_t)K * (size_t)m;
            const size_t in = (size_t)k + (size_t)K * (size_t)(M - m);
            re[in] = re[ip];
            im[in] = -im[ip];
        }
    return 0;
}
