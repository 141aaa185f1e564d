// plan_z.hip -- launchers of the z kernels for one precision MWF_R and
// wall-normal basis MWF_K0 (4 or 6), both set on the compile line.  See
// plans.hpp.
#include "kernels.hpp"
#include "plans.hpp"
#include "transform.hpp"
#include "zflux.hpp"

#if !defined(MWF_R) || !defined(MWF_K0)
#error "compile with -DMWF_R=double|float -DMWF_K0=4|6"
#endif

template <typename R, int LOGNN> static size_t zphys_lds()
{
    return (size_t)(zphys_threads<LOGNN>() >> LOGNN) * ZPB * 3 * lpad(1 << LOGNN) * sizeof(cplx_t<R>);
}
template <typename R, int K0, int LOGNN>
static hipError_t launch_z(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* U, void* NL,
                           const void* twNN, const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = ZT >> LOGNN;
    dim3 grid((g.P + XB - 1) / XB, 1, nmem);
    hipLaunchKernelGGL((k_z<R, K0, LOGNN>), grid, dim3(ZT), (z_lds_bytes<R, K0, LOGNN>()), s, bt, (const CT*)U,
                       (CT*)NL, (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int K0, int LOGNN>
static hipError_t launch_z2(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* U, void* NL,
                            const void* twNN, const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = 512 >> LOGNN;
    dim3 grid((g.P + XB - 1) / XB, 1, nmem);
    hipLaunchKernelGGL((k_z2<R, K0, LOGNN>), grid, dim3(Z2T), (z2_lds_bytes<R, LOGNN>()), s, bt,
                       (const CT*)U, (CT*)NL, (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int K0, int LOGNN, bool PH3 = false>
static hipError_t launch_zb(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* U, void* NL,
                            const void* twNN, const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = ZB >> LOGNN;
    const int nb = (g.P + XB - 1) / XB;
    // PH3: three blocks per row block; the XCD mapping needs whole rounds of 8
    if (PH3 && (nb % 8 != 0 || g.P % XB != 0 || bt.nl_part == 0)) return hipErrorInvalidValue;
    dim3 grid(PH3 ? 3 * nb : nb, 1, nmem);
    hipLaunchKernelGGL((k_zb<R, K0, LOGNN, PH3>), grid, dim3(ZB), (zb_lds_bytes<R, LOGNN>()), s, bt, (const CT*)U,
                       (CT*)NL, (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int K0, int LOGNN>
static hipError_t launch_zphys(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* U, double* phys,
                               const void* twNN, const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int TZ = zphys_threads<LOGNN>();
    constexpr int XB = TZ >> LOGNN;
    dim3 grid((g.P + XB - 1) / XB, nmem);
    hipLaunchKernelGGL((k_zphys<R, K0, LOGNN>), grid, dim3(TZ), (zphys_lds<R, LOGNN>()), s, bt, (const CT*)U, phys,
                       (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int LOGNN>
static hipError_t launch_zfwdp(const Geo& g, hipStream_t s, const double* phys, void* NL, const void* twNN,
                               const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = zfp_threads<LOGNN>() >> LOGNN;
    hipLaunchKernelGGL((k_zfwd_phys<R, LOGNN>), dim3((g.P + XB - 1) / XB), dim3(zfp_threads<LOGNN>()),
                       (zfp_lds<R, LOGNN>()), s, g, phys, (CT*)NL, (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

// flux-form z kernel: two pairs per batch
constexpr int ZC_ZP = 2;
// W: waves per SIMD the kernel is compiled for: 4 (two 512-thread blocks per
// CU, 128 VGPRs, transforms in place) or 2 (one block, 256 VGPRs, ping-pong
// transforms)
template <typename R, int K0, int LOGNN, int W>
static hipError_t launch_zc(const Batch& bt, int nmem, const Geo& g, hipStream_t s, const void* U, void* NL,
                            const void* twNN, const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = zc_threads<LOGNN>() >> LOGNN;
    dim3 grid((g.P + XB - 1) / XB, 1, nmem);
    hipLaunchKernelGGL((k_zc<R, K0, LOGNN, ZC_ZP, W, W == 2>), grid, dim3(zc_threads<LOGNN>()),
                       (zc_lds_bytes<R, LOGNN, ZC_ZP, W == 2>()), s, bt, (const CT*)U, (CT*)NL, (const CT*)twNN,
                       (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int K0> void pick_z(int ln, Plan& p)
{
    switch (ln) {
#define C_(L)                                   \
    case L:                                     \
        p.zf = launch_z<R, K0, L>;              \
        p.zf_alt2 = launch_z2<R, K0, L>;        \
        p.zphys = launch_zphys<R, K0, L>;       \
        p.zfwdp = launch_zfwdp<R, L>;           \
        p.zc = launch_zc<R, K0, L, 4>;          \
        p.zc_alt = launch_zc<R, K0, L, 2>;      \
        break;
        C_(4) C_(5) C_(6) C_(7) C_(8) C_(9)
#undef C_
    default: break;
    }
    switch (ln) { // phase-sequential alternatives for the autotuner
    case 8: p.zf_alt = launch_zb<R, K0, 8>; break;
    case 9: p.zf_alt = launch_zb<R, K0, 9>; break;
    case 10: // one x-row per 1024-thread block: the three fold phases in sequence
        p.zf = launch_zb<R, K0, 10>;
        p.zf_ph3 = launch_zb<R, K0, 10, true>;
        p.zphys = launch_zphys<R, K0, 10>;
        p.zfwdp = launch_zfwdp<R, 10>;
        // the flux z kernel holds all three phases of a 3072-point row at
        // 1024 threads (4 waves per SIMD) in FP32 only
        if constexpr (std::is_same<R, float>::value) p.zc = launch_zc<R, K0, 10, 4>;
        break;
    default: break;
    }
}

template <typename R, int K0> void prepare_z(int ln)
{
    switch (ln) {
#define C_(L)                                                     \
    case L:                                                       \
        set_lds(k_z<R, K0, L>, z_lds_bytes<R, K0, L>());          \
        set_lds(k_z2<R, K0, L>, z2_lds_bytes<R, L>());            \
        set_lds(k_zphys<R, K0, L>, zphys_lds<R, L>());            \
        set_lds(k_zfwd_phys<R, L>, zfp_lds<R, L>());              \
        set_lds(k_zc<R, K0, L, ZC_ZP, 4, false>, zc_lds_bytes<R, L, ZC_ZP, false>()); \
        set_lds(k_zc<R, K0, L, ZC_ZP, 2, true>, zc_lds_bytes<R, L, ZC_ZP, true>());   \
        break;
        C_(4) C_(5) C_(6) C_(7) C_(8) C_(9)
#undef C_
    default: break;
    }
    switch (ln) {
    case 8: set_lds(k_zb<R, K0, 8>, zb_lds_bytes<R, 8>()); break;
    case 9: set_lds(k_zb<R, K0, 9>, zb_lds_bytes<R, 9>()); break;
    case 10:
        set_lds(k_zb<R, K0, 10>, zb_lds_bytes<R, 10>());
        set_lds(k_zb<R, K0, 10, true>, zb_lds_bytes<R, 10>());
        set_lds(k_zphys<R, K0, 10>, zphys_lds<R, 10>());
        set_lds(k_zfwd_phys<R, 10>, zfp_lds<R, 10>());
        if constexpr (std::is_same<R, float>::value)
            set_lds(k_zc<R, K0, 10, ZC_ZP, 4, false>, zc_lds_bytes<R, 10, ZC_ZP, false>());
        break;
    default: break;
    }
}

template void pick_z<MWF_R, MWF_K0>(int, Plan&);
template void prepare_z<MWF_R, MWF_K0>(int);
