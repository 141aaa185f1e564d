// plan_x.hip -- launchers of the x kernels for one precision, MWF_R (double
// or float, set on the compile line).  See plans.hpp.
#include "kernels.hpp"
#include "plans.hpp"
#include "transform.hpp"

#ifndef MWF_R
#error "compile with -DMWF_R=double or -DMWF_R=float"
#endif

template <typename R, int LOGMM> static size_t x_lds() { return 3 * (size_t)(XTILE >> LOGMM) * lpad(1 << LOGMM) * sizeof(cplx_t<R>); }
template <typename R, int LOGMM, bool FLUX>
static hipError_t launch_xinv(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const double2* c, void* U,
                              double* ep, const void* twMM, const double2* twP, int addlam)
{
    using CT = cplx_t<R>;
    dim3 grid(g.NN / g.NTx, g.K0, nmem);
    hipLaunchKernelGGL((k_xinv<R, LOGMM, FLUX>), grid, dim3(XT), (x_lds<R, LOGMM>()), s, bt, c, (CT*)U, ep, (const CT*)twMM,
                       twP, addlam);
    return hipGetLastError();
}

template <typename R, int LOGMM, bool FLUX>
static hipError_t launch_xfwd(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* NL, double2* c,
                              void* onl, double2* nlout, const void* twMM, const double2* twP, int mode, int nst)
{
    using CT = cplx_t<R>;
    dim3 grid(g.NN / g.NTx, g.K0, nmem);
    hipLaunchKernelGGL((k_xfwd<R, LOGMM, FLUX>), grid, dim3(XT), (x_lds<R, LOGMM>()), s, bt, (const CT*)NL, c, (CT*)onl,
                       nlout,
                       (const CT*)twMM, twP, mode, nst);
    return hipGetLastError();
}

// transform buffers, then the pass twiddles of the MM-point transforms and
// 2 MM + 1 fold twiddles
template <typename R, int LOGMM> static size_t xstep_lds()
{
    constexpr size_t MM = (size_t)1 << LOGMM;
    return (3 * (size_t)XTILE + ptw_size<(1 << LOGMM), 1, 7>()) * sizeof(cplx_t<R>) +
           (2 * MM + 1) * sizeof(double2) + xstep_stage_bytes<R>();
}

template <typename R, int LOGMM, bool FLUX>
static hipError_t launch_xstep(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* NL, double2* c,
                               void* onl, void* U, const void* twMM, const double2* twP, int nst, int inv_only)
{
    using CT = cplx_t<R>;
    dim3 grid(g.NN / g.NTx, g.K0, nmem);
    if (inv_only)
        hipLaunchKernelGGL((k_xstep<R, LOGMM, true, FLUX>), grid, dim3(xs_of<LOGMM>()), (xstep_lds<R, LOGMM>()), s, bt, (const CT*)NL,
                           c, (CT*)onl, (CT*)U, (const CT*)twMM, twP, nst);
    else
        hipLaunchKernelGGL((k_xstep<R, LOGMM, false, FLUX>), grid, dim3(xs_of<LOGMM>()), (xstep_lds<R, LOGMM>()), s, bt, (const CT*)NL,
                           c, (CT*)onl, (CT*)U, (const CT*)twMM, twP, nst);
    return hipGetLastError();
}

template <typename R, int LOGMM>
static hipError_t launch_sxinv(const Geo& g, hipStream_t s, const double2* spec, void* U, const void* twMM,
                               const double2* twP)
{
    using CT = cplx_t<R>;
    hipLaunchKernelGGL((k_sxinv<R, LOGMM>), dim3(g.NN / g.NTx, g.KK), dim3(XT), (x_lds<R, LOGMM>()), s, g, spec,
                       (CT*)U, (const CT*)twMM, twP);
    return hipGetLastError();
}

template <typename R, int LOGMM>
static hipError_t launch_sxfwd(const Geo& g, hipStream_t s, const void* NL, double2* spec, const void* twMM,
                               const double2* twP)
{
    using CT = cplx_t<R>;
    hipLaunchKernelGGL((k_sxfwd<R, LOGMM>), dim3(g.NN / g.NTx, g.KK), dim3(XT), (x_lds<R, LOGMM>()), s, g,
                       (const CT*)NL, spec, (const CT*)twMM, twP);
    return hipGetLastError();
}

// x kernels depend on the precision and MM only
template <typename R> void pick_x(int lm, Plan& p)
{
    switch (lm) {
#define C_(L)                                   \
    case L:                                     \
        p.xinv = launch_xinv<R, L, false>;      \
        p.xfwd = launch_xfwd<R, L, false>;      \
        p.xstep = launch_xstep<R, L, false>;    \
        p.xinv_f = launch_xinv<R, L, true>;     \
        p.xfwd_f = launch_xfwd<R, L, true>;     \
        p.xstep_f = launch_xstep<R, L, true>;   \
        p.sxinv = launch_sxinv<R, L>;           \
        p.sxfwd = launch_sxfwd<R, L>;           \
        break;
        C_(4) C_(5) C_(6) C_(7) C_(8) C_(9)
#undef C_
    case 10:
        p.xinv = launch_xinv<R, 10, false>;
        p.xfwd = launch_xfwd<R, 10, false>;
        p.xinv_f = launch_xinv<R, 10, true>;
        p.xfwd_f = launch_xfwd<R, 10, true>;
        p.sxinv = launch_sxinv<R, 10>;
        p.sxfwd = launch_sxfwd<R, 10>;
        // fused kernel at MM = 1024: a plan candidate next to the split kernels
        p.xstep = launch_xstep<R, 10, false>;
        p.xstep_f = launch_xstep<R, 10, true>;
        break;
    default: break;
    }
}

template <typename R> void prepare_x(int lm)
{
    switch (lm) {
#define C_(L)                                         \
    case L:                                           \
        set_lds(k_xinv<R, L, false>, x_lds<R, L>());            \
        set_lds(k_xfwd<R, L, false>, x_lds<R, L>());            \
        set_lds(k_xstep<R, L, false, false>, xstep_lds<R, L>()); \
        set_lds(k_xstep<R, L, true, false>, xstep_lds<R, L>());  \
        set_lds(k_xinv<R, L, true>, x_lds<R, L>());             \
        set_lds(k_xfwd<R, L, true>, x_lds<R, L>());             \
        set_lds(k_xstep<R, L, false, true>, xstep_lds<R, L>());  \
        set_lds(k_xstep<R, L, true, true>, xstep_lds<R, L>());   \
        set_lds(k_sxinv<R, L>, x_lds<R, L>());         \
        set_lds(k_sxfwd<R, L>, x_lds<R, L>());         \
        break;
        C_(4) C_(5) C_(6) C_(7) C_(8) C_(9)
#undef C_
    case 10:
        set_lds(k_xinv<R, 10, false>, x_lds<R, 10>());
        set_lds(k_xfwd<R, 10, false>, x_lds<R, 10>());
        set_lds(k_xinv<R, 10, true>, x_lds<R, 10>());
        set_lds(k_xfwd<R, 10, true>, x_lds<R, 10>());
        set_lds(k_sxinv<R, 10>, x_lds<R, 10>());
        set_lds(k_sxfwd<R, 10>, x_lds<R, 10>());
        set_lds(k_xstep<R, 10, false, false>, xstep_lds<R, 10>());
        set_lds(k_xstep<R, 10, true, false>, xstep_lds<R, 10>());
        set_lds(k_xstep<R, 10, false, true>, xstep_lds<R, 10>());
        set_lds(k_xstep<R, 10, true, true>, xstep_lds<R, 10>());
        break;
    default: break;
    }
}

template void pick_x<MWF_R>(int, Plan&);
template void prepare_x<MWF_R>(int);
