// plans.hpp -- the launchers of the per-step kernels, one table per size
// class.  The kernels are compiled in separate translation units, one per
// precision (x kernels, plan_x.hip) and per precision and basis (z kernels,
// plan_z.hip), which build in parallel; the host side (mwf.hip) only sees
// these function pointers.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

struct Plan {
    hipError_t (*xinv)(const Batch&, int, const Geo&, hipStream_t, const double2*, void*, double*, const void*,
                       const double2*, int) = nullptr;
    // x forward + update: (NL, c, history onl (precision R), nonlinear-term
    // output nlout (double2, mode 1 only), twiddles, mode, nst)
    hipError_t (*xfwd)(const Batch&, int, const Geo&, hipStream_t, const void*, double2*, void*, double2*,
                       const void*, const double2*, int, int) = nullptr;
    hipError_t (*zf)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*, const void*) = nullptr;
    // z inverse to the phys layout, members back to back
    hipError_t (*zphys)(const Batch&, int, const Geo&, hipStream_t, const void*, double*, const void*,
                        const void*) = nullptr;
    // fused forward + update + next inverse
    hipError_t (*xstep)(const Batch&, int, const Geo&, hipStream_t, const void*, double2*, void*, void*,
                        const void*, const double2*, int, int) = nullptr;
    // alternative z kernels the autotuner may pick: k_zb where k_z also runs,
    // and the two-points-per-thread k_z2
    hipError_t (*zf_alt)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*,
                         const void*) = nullptr;
    hipError_t (*zf_alt2)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*,
                          const void*) = nullptr;
    // k_zb with the fold phases of a row in three blocks and three NL partials
    // (bt.nl_part), summed by the split x kernels
    hipError_t (*zf_ph3)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*,
                         const void*) = nullptr;
    // flux form (zflux.hpp): z kernel and the x kernels that read its B/A lines
    // and write U without x-derivative lines
    hipError_t (*zc)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*, const void*) = nullptr;
    hipError_t (*zc_alt)(const Batch&, int, const Geo&, hipStream_t, const void*, void*, const void*,
                         const void*) = nullptr; // the flux z kernel at 2 waves per SIMD instead of 4
    hipError_t (*xinv_f)(const Batch&, int, const Geo&, hipStream_t, const double2*, void*, double*, const void*,
                         const double2*, int) = nullptr;
    hipError_t (*xfwd_f)(const Batch&, int, const Geo&, hipStream_t, const void*, double2*, void*, double2*,
                         const void*, const double2*, int, int) = nullptr;
    hipError_t (*xstep_f)(const Batch&, int, const Geo&, hipStream_t, const void*, double2*, void*, void*,
                          const void*, const double2*, int, int) = nullptr;
    // standalone transforms of the transform module (transform.hpp):
    // spec -> U lines, phys -> NL, NL -> spec
    hipError_t (*sxinv)(const Geo&, hipStream_t, const double2*, void*, const void*, const double2*) = nullptr;
    hipError_t (*zfwdp)(const Geo&, hipStream_t, const double*, void*, const void*, const void*) = nullptr;
    hipError_t (*sxfwd)(const Geo&, hipStream_t, const void*, double2*, const void*, const double2*) = nullptr;
};

// x kernels depend on the precision and MM only, z kernels on the precision,
// K0 and NN only.  pick_* fill the plan's launchers for log2(MM) / log2(NN);
// prepare_* raise the dynamic LDS limit of the kernels they pick.
template <typename R> void pick_x(int lm, Plan& p);
template <typename R> void prepare_x(int lm);
template <typename R, int K0> void pick_z(int ln, Plan& p);
template <typename R, int K0> void prepare_z(int ln);

template <typename F> static inline void set_lds(F* f, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}
