// mwf.hip -- C-ABI of libmwf.so: handle, device buffers, plans and launches.
//
// One handle is one model instance (the module state of program/velocity.f90
// :15-17 plus the plans of program/transform.fftw5.f90:33-59).  All device
// work goes to the handle's own stream at default priority.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mwf.h"
#include "kernels.hpp"

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(MWF_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char* mwf_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// launchers and dispatch
// ---------------------------------------------------------------------------
using xinv_fn = hipError_t (*)(const Geo&, hipStream_t, const double2*, void*, double*, const void*,
                               const double2*, int);
using xfwd_fn = hipError_t (*)(const Geo&, hipStream_t, const void*, double2*, double2*, const void*,
                               const double2*, int, int);
using z_fn = hipError_t (*)(const Geo&, hipStream_t, const void*, void*, const void*, const void*);
using zphys_fn = hipError_t (*)(const Geo&, hipStream_t, const void*, double*, const void*, const void*);

template <typename F> static void set_lds(F* f, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

template <typename R, int LOGMM>
static hipError_t launch_xinv(const Geo& g, hipStream_t s, const double2* c, void* U, double* ep,
                              const void* twMM, const double2* twP, int addlam)
{
    using CT = cplx_t<R>;
    const size_t lds = 3 * (size_t)(XTILE >> LOGMM) * lpad(1 << LOGMM) * sizeof(CT);
    set_lds(k_xinv<R, LOGMM>, lds);
    dim3 grid(g.NN / g.NTx, g.K0);
    hipLaunchKernelGGL((k_xinv<R, LOGMM>), grid, dim3(XT), lds, s, g, c, (CT*)U, ep, (const CT*)twMM, twP,
                       addlam);
    return hipGetLastError();
}

template <typename R, int LOGMM>
static hipError_t launch_xfwd(const Geo& g, hipStream_t s, const void* NL, double2* c, double2* onl,
                              const void* twMM, const double2* twP, int mode, int nst)
{
    using CT = cplx_t<R>;
    const size_t lds = 3 * (size_t)(XTILE >> LOGMM) * lpad(1 << LOGMM) * sizeof(CT);
    set_lds(k_xfwd<R, LOGMM>, lds);
    dim3 grid(g.NN / g.NTx, g.K0);
    hipLaunchKernelGGL((k_xfwd<R, LOGMM>), grid, dim3(XT), lds, s, g, (const CT*)NL, c, onl, (const CT*)twMM,
                       twP, mode, nst);
    return hipGetLastError();
}

template <typename R, int K0, int LOGNN>
static hipError_t launch_z(const Geo& g, hipStream_t s, const void* U, void* NL, const void* twNN,
                           const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = ZT >> LOGNN;
    const size_t lds = z_lds_bytes<R, K0, LOGNN>();
    set_lds(k_z<R, K0, LOGNN>, lds);
    dim3 grid((g.P + XB - 1) / XB);
    hipLaunchKernelGGL((k_z<R, K0, LOGNN>), grid, dim3(ZT), lds, s, g, (const CT*)U, (CT*)NL,
                       (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

template <typename R, int K0, int LOGNN>
static hipError_t launch_zphys(const Geo& g, hipStream_t s, const void* U, double* phys, const void* twNN,
                               const void* twQ)
{
    using CT = cplx_t<R>;
    constexpr int XB = ZT >> LOGNN;
    const size_t lds = (size_t)XB * ZPB * 3 * lpad(1 << LOGNN) * sizeof(CT);
    set_lds(k_zphys<R, K0, LOGNN>, lds);
    dim3 grid((g.P + XB - 1) / XB);
    hipLaunchKernelGGL((k_zphys<R, K0, LOGNN>), grid, dim3(ZT), lds, s, g, (const CT*)U, phys,
                       (const CT*)twNN, (const CT*)twQ);
    return hipGetLastError();
}

#define LOG_CASES(M, ...) \
    M(4, __VA_ARGS__) M(5, __VA_ARGS__) M(6, __VA_ARGS__) M(7, __VA_ARGS__) M(8, __VA_ARGS__) M(9, __VA_ARGS__) M(10, __VA_ARGS__)
#define LOGZ_CASES(M, ...) \
    M(4, __VA_ARGS__) M(5, __VA_ARGS__) M(6, __VA_ARGS__) M(7, __VA_ARGS__) M(8, __VA_ARGS__) M(9, __VA_ARGS__)

template <typename R> static xinv_fn pick_xinv(int lg)
{
#define C_(L, R_) \
    case L: return launch_xinv<R_, L>;
    switch (lg) { LOG_CASES(C_, R) }
#undef C_
    return nullptr;
}

template <typename R> static xfwd_fn pick_xfwd(int lg)
{
#define C_(L, R_) \
    case L: return launch_xfwd<R_, L>;
    switch (lg) { LOG_CASES(C_, R) }
#undef C_
    return nullptr;
}

template <typename R, int K0> static z_fn pick_z(int lg)
{
#define C_(L, R_, K_) \
    case L: return launch_z<R_, K_, L>;
    switch (lg) { LOGZ_CASES(C_, R, K0) }
#undef C_
    return nullptr;
}

template <typename R, int K0> static zphys_fn pick_zphys(int lg)
{
#define C_(L, R_, K_) \
    case L: return launch_zphys<R_, K_, L>;
    switch (lg) { LOGZ_CASES(C_, R, K0) }
#undef C_
    return nullptr;
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
enum { KID_XINV = 0, KID_Z = 1, KID_XFWD = 2, KID_ENERGY = 3, NKID = 4 };

struct mwf_handle {
    mwf_params p{};
    Geo g{};
    hipStream_t st = nullptr;
    double2* c = nullptr;   // state vel_c
    double2* onl = nullptr; // AB2 history vel_onl
    double2* cs = nullptr;  // scratch mpt (nonlinear output, mpt2phys input)
    void* U = nullptr;
    void* NL = nullptr;
    double* phys = nullptr;
    size_t phys_elems = 0;
    double2 *twMMd = nullptr, *twNNd = nullptr, *twPd = nullptr, *twQd = nullptr;
    float2 *twMMf = nullptr, *twNNf = nullptr, *twQf = nullptr;
    double *epart = nullptr, *edev = nullptr;
    int nepart = 0;
    double2* hstage = nullptr; // pinned host staging, one device mpt
    size_t mpt_elems = 0;      // NN * K * ldM
    int nst = 1;
    xinv_fn xinv = nullptr;
    xfwd_fn xfwd = nullptr;
    z_fn zf = nullptr;
    zphys_fn zphys = nullptr;
    // profiling
    struct Ev {
        int id;
        hipEvent_t a, b;
    };
    int prof = 0;
    std::vector<Ev> pending;
    std::vector<hipEvent_t> pool;
    double ms[NKID] = {0, 0, 0, 0};
    int64_t cnt[NKID] = {0, 0, 0, 0};
};

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

template <typename CT> static std::vector<CT> twiddles(int n)
{
    std::vector<CT> t(n);
    const long double twopi = 6.283185307179586476925286766559L;
    for (int j = 0; j < n; j++) {
        long double a = twopi * (long double)j / (long double)n;
        t[j].x = (decltype(t[j].x))cosl(a);
        t[j].y = (decltype(t[j].y))sinl(a);
    }
    return t;
}

template <typename T> static int dev_upload(T** d, const std::vector<T>& h)
{
    HIPCHK(hipMalloc((void**)d, sizeof(T) * h.size()));
    HIPCHK(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

static hipEvent_t ev_get(mwf_handle* h)
{
    if (!h->pool.empty()) {
        hipEvent_t e = h->pool.back();
        h->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

static int prof_drain(mwf_handle* h)
{
    for (auto& e : h->pending) {
        HIPCHK(hipEventSynchronize(e.b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
        h->ms[e.id] += ms;
        h->cnt[e.id] += 1;
        h->pool.push_back(e.a);
        h->pool.push_back(e.b);
    }
    h->pending.clear();
    return 0;
}

// wraps one launch in profiling events when enabled
template <typename F> static int timed(mwf_handle* h, int id, F&& f)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (h->prof) {
        a = ev_get(h);
        b = ev_get(h);
        HIPCHK(hipEventRecord(a, h->st));
    }
    hipError_t e = f();
    if (e != hipSuccess) return fail(MWF_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (h->prof) {
        HIPCHK(hipEventRecord(b, h->st));
        h->pending.push_back({id, a, b});
        if (h->pending.size() > 4096) return prof_drain(h);
    }
    return 0;
}

static void destroy(mwf_handle* h)
{
    if (!h) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    prof_drain(h);
    for (auto e : h->pool) (void)hipEventDestroy(e);
    void* dptrs[] = {h->c, h->onl, h->cs, h->U, h->NL, h->phys, h->twMMd, h->twNNd, h->twPd, h->twQd,
                     h->twMMf, h->twNNf, h->twQf, h->epart, h->edev};
    for (void* p : dptrs)
        if (p) (void)hipFree(p);
    if (h->hstage) (void)hipHostFree(h->hstage);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

extern "C" int mwf_create(const mwf_params* p, mwf_handle** out)
{
    if (!p || !out) return fail(MWF_E_ARG, "null argument");
    *out = nullptr;
    if (p->k0 != 4 && p->k0 != 6) return fail(MWF_E_ARG, "k0 must be 4 (M4) or 6 (M6)");
    if (!is_pow2(p->mm) || p->mm < 16 || p->mm > 1024) return fail(MWF_E_ARG, "mm must be a power of two in [16,1024]");
    if (!is_pow2(p->nn) || p->nn < 16 || p->nn > 512) return fail(MWF_E_ARG, "nn must be a power of two in [16,512]");
    if (p->precision != MWF_FP64 && p->precision != MWF_MIXED) return fail(MWF_E_ARG, "precision must be MWF_FP64 or MWF_MIXED");
    if (p->reserved != 0) return fail(MWF_E_ARG, "reserved must be 0");
    if (!(p->dt > 0) || !(p->re > 0)) return fail(MWF_E_ARG, "dt and re must be positive");

    mwf_handle* h = new mwf_handle();
    h->p = *p;
    Geo& g = h->g;
    g.K0 = p->k0;
    g.K = 2 * p->k0 - 1;
    g.KK = 3 * p->k0 - 1;
    g.MM = p->mm;
    g.M = 2 * p->mm - 1;
    g.ldM = 2 * p->mm;
    g.NN = p->nn;
    g.P = 3 * p->mm;
    g.Q = 3 * p->nn;
    g.NTx = std::min(p->nn, XTILE / p->mm);
    g.XB = ZT / p->nn;
    g.alpha = p->alpha;
    g.gamma = p->gamma;
    g.re = p->re;
    g.dt = p->dt;
    g.drag = p->drag;
    g.vdrag = p->vdrag;
    g.hypo = p->hypo;
    g.cth = std::cos(p->theta);
    g.sth = std::sin(p->theta);
    g.beta = 3.1415926535897931 / 2.0; // d_beta, program/parameters.f90:60
    g.scale = 1.0 / (double)(g.P * g.Q);
    g.dragall = p->dragall ? 1 : 0;

    const int lm = ilog2(g.MM), ln = ilog2(g.NN);
    const bool mixed = p->precision == MWF_MIXED;
    if (mixed) {
        h->xinv = pick_xinv<float>(lm);
        h->xfwd = pick_xfwd<float>(lm);
        h->zf = g.K0 == 4 ? pick_z<float, 4>(ln) : pick_z<float, 6>(ln);
        h->zphys = g.K0 == 4 ? pick_zphys<float, 4>(ln) : pick_zphys<float, 6>(ln);
    } else {
        h->xinv = pick_xinv<double>(lm);
        h->xfwd = pick_xfwd<double>(lm);
        h->zf = g.K0 == 4 ? pick_z<double, 4>(ln) : pick_z<double, 6>(ln);
        h->zphys = g.K0 == 4 ? pick_zphys<double, 4>(ln) : pick_zphys<double, 6>(ln);
    }
    if (!h->xinv || !h->xfwd || !h->zf || !h->zphys) {
        destroy(h);
        return fail(MWF_E_ARG, "no kernel for this size");
    }

    int rc = 0;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == 0) rc = fail(MWF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(hipStreamCreate(&h->st), "hipStreamCreate");
    h->mpt_elems = (size_t)g.NN * g.K * g.ldM;
    const size_t csz = sizeof(double2) * h->mpt_elems;
    const size_t rsz = mixed ? sizeof(float2) : sizeof(double2);
    chk(hipMalloc((void**)&h->c, csz), "hipMalloc c");
    chk(hipMalloc((void**)&h->onl, csz), "hipMalloc onl");
    chk(hipMalloc((void**)&h->cs, csz), "hipMalloc scratch");
    chk(hipMalloc(&h->U, rsz * (size_t)g.P * 2 * g.KK * g.NN), "hipMalloc U");
    chk(hipMalloc(&h->NL, rsz * (size_t)g.P * g.KK * g.NN), "hipMalloc NL");
    h->nepart = (g.NN / g.NTx) * g.K0;
    chk(hipMalloc((void**)&h->epart, sizeof(double) * h->nepart), "hipMalloc epart");
    chk(hipMalloc((void**)&h->edev, sizeof(double)), "hipMalloc e");
    chk(hipHostMalloc((void**)&h->hstage, csz, hipHostMallocDefault), "hipHostMalloc");
    if (rc) {
        destroy(h);
        return rc;
    }
    if ((rc = dev_upload(&h->twMMd, twiddles<double2>(g.MM))) || (rc = dev_upload(&h->twNNd, twiddles<double2>(g.NN))) ||
        (rc = dev_upload(&h->twPd, twiddles<double2>(g.P))) || (rc = dev_upload(&h->twQd, twiddles<double2>(g.Q))) ||
        (rc = dev_upload(&h->twMMf, twiddles<float2>(g.MM))) || (rc = dev_upload(&h->twNNf, twiddles<float2>(g.NN))) ||
        (rc = dev_upload(&h->twQf, twiddles<float2>(g.Q)))) {
        destroy(h);
        return rc;
    }
    chk(hipMemsetAsync(h->c, 0, csz, h->st), "hipMemset c");
    chk(hipMemsetAsync(h->onl, 0, csz, h->st), "hipMemset onl");
    chk(hipStreamSynchronize(h->st), "hipStreamSynchronize");
    if (rc) {
        destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

extern "C" int mwf_destroy(mwf_handle* h)
{
    destroy(h);
    return 0;
}

static const void* twMM_of(const mwf_handle* h) { return h->p.precision == MWF_MIXED ? (const void*)h->twMMf : (const void*)h->twMMd; }
static const void* twNN_of(const mwf_handle* h) { return h->p.precision == MWF_MIXED ? (const void*)h->twNNf : (const void*)h->twNNd; }
static const void* twQ_of(const mwf_handle* h) { return h->p.precision == MWF_MIXED ? (const void*)h->twQf : (const void*)h->twQd; }

// reference mpt (re, im planes, (K, M, NN) column-major) <-> device rows
static void to_dev_layout(const mwf_handle* h, const double* re, const double* im, double2* d)
{
    const Geo& g = h->g;
    std::memset(d, 0, sizeof(double2) * h->mpt_elems);
    for (int n = 0; n < g.NN; n++)
        for (int m = 0; m < g.M; m++)
            for (int k = 0; k < g.K; k++) {
                const size_t i = (size_t)k + (size_t)g.K * ((size_t)m + (size_t)g.M * n);
                d[((size_t)n * g.K + k) * g.ldM + m] = double2{re[i], im[i]};
            }
}

static void from_dev_layout(const mwf_handle* h, const double2* d, double* re, double* im)
{
    const Geo& g = h->g;
    for (int n = 0; n < g.NN; n++)
        for (int m = 0; m < g.M; m++)
            for (int k = 0; k < g.K; k++) {
                const size_t i = (size_t)k + (size_t)g.K * ((size_t)m + (size_t)g.M * n);
                const double2 v = d[((size_t)n * g.K + k) * g.ldM + m];
                re[i] = v.x;
                im[i] = v.y;
            }
}

static int upload_mpt(mwf_handle* h, const double* re, const double* im, double2* dst)
{
    to_dev_layout(h, re, im, h->hstage);
    HIPCHK(hipMemcpyAsync(dst, h->hstage, sizeof(double2) * h->mpt_elems, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

static int download_mpt(mwf_handle* h, const double2* src, double* re, double* im)
{
    HIPCHK(hipMemcpyAsync(h->hstage, src, sizeof(double2) * h->mpt_elems, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    from_dev_layout(h, h->hstage, re, im);
    return 0;
}

extern "C" int mwf_set_state(mwf_handle* h, const double* re, const double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    int rc = upload_mpt(h, re, im, h->c);
    if (rc) return rc;
    h->nst = 1;
    return 0;
}

extern "C" int mwf_get_state(mwf_handle* h, double* re, double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    return download_mpt(h, h->c, re, im);
}

static int run_nonlinear(mwf_handle* h, double2* cin, double2* cout, double2* hist, int mode)
{
    const Geo& g = h->g;
    int rc;
    if ((rc = timed(h, KID_XINV, [&] { return h->xinv(g, h->st, cin, h->U, nullptr, twMM_of(h), h->twPd, 1); })))
        return rc;
    if ((rc = timed(h, KID_Z, [&] { return h->zf(g, h->st, h->U, h->NL, twNN_of(h), twQ_of(h)); }))) return rc;
    if ((rc = timed(h, KID_XFWD, [&] {
             return h->xfwd(g, h->st, h->NL, cout, hist, twMM_of(h), h->twPd, mode, h->nst);
         })))
        return rc;
    return 0;
}

extern "C" int mwf_step(mwf_handle* h, int32_t nsteps)
{
    if (!h) return fail(MWF_E_ARG, "null handle");
    if (nsteps < 0) return fail(MWF_E_ARG, "nsteps < 0");
    for (int s = 0; s < nsteps; s++) {
        int rc = run_nonlinear(h, h->c, h->c, h->onl, 0);
        if (rc) return rc;
        h->nst = 0;
    }
    return 0;
}

extern "C" int mwf_nonlinear(mwf_handle* h, double* nlre, double* nlim)
{
    if (!h || !nlre || !nlim) return fail(MWF_E_ARG, "null argument");
    HIPCHK(hipMemsetAsync(h->cs, 0, sizeof(double2) * h->mpt_elems, h->st));
    int rc = run_nonlinear(h, h->c, h->c, h->cs, 1);
    if (rc) return rc;
    return download_mpt(h, h->cs, nlre, nlim);
}

extern "C" int mwf_energy(mwf_handle* h, double* e)
{
    if (!h || !e) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->g;
    int rc = timed(h, KID_ENERGY, [&] {
        hipError_t r = h->xinv(g, h->st, h->c, nullptr, h->epart, twMM_of(h), h->twPd, 0);
        if (r != hipSuccess) return r;
        hipLaunchKernelGGL(k_ereduce, dim3(1), dim3(64), 0, h->st, h->epart, h->nepart, h->edev);
        return hipGetLastError();
    });
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(e, h->edev, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_mpt2phys(mwf_handle* h, const double* re, const double* im, double* phys)
{
    if (!h || !re || !im || !phys) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->g;
    const size_t np = (size_t)g.KK * g.Q * g.P;
    if (h->phys_elems < np) {
        if (h->phys) HIPCHK(hipFree(h->phys));
        h->phys = nullptr;
        HIPCHK(hipMalloc((void**)&h->phys, sizeof(double) * np));
        h->phys_elems = np;
    }
    int rc = upload_mpt(h, re, im, h->cs);
    if (rc) return rc;
    hipError_t e = h->xinv(g, h->st, h->cs, h->U, nullptr, twMM_of(h), h->twPd, 0);
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    e = h->zphys(g, h->st, h->U, h->phys, twNN_of(h), twQ_of(h));
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(phys, h->phys, sizeof(double) * np, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_sync(mwf_handle* h)
{
    if (!h) return fail(MWF_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_profile_enable(mwf_handle* h, int32_t on)
{
    if (!h) return fail(MWF_E_ARG, "null handle");
    int rc = prof_drain(h);
    if (rc) return rc;
    h->prof = on ? 1 : 0;
    for (int i = 0; i < NKID; i++) {
        h->ms[i] = 0;
        h->cnt[i] = 0;
    }
    return 0;
}

extern "C" int mwf_profile_read(mwf_handle* h, int32_t kernel, double* total_ms, int64_t* count)
{
    if (!h || !total_ms || !count) return fail(MWF_E_ARG, "null argument");
    if (kernel < 0 || kernel >= NKID) return fail(MWF_E_ARG, "kernel id out of range");
    int rc = prof_drain(h);
    if (rc) return rc;
    *total_ms = h->ms[kernel];
    *count = h->cnt[kernel];
    return 0;
}

extern "C" int mwf_kernel_bytes(mwf_handle* h, int32_t kernel, double* bytes)
{
    if (!h || !bytes) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->g;
    const double s = h->p.precision == MWF_MIXED ? 8.0 : 16.0; // complex intermediate
    const double state = 16.0 * g.K * g.M * g.NN;             // one complex FP64 mpt
    const double uline = s * (double)g.P * g.NN;                // one x-physical line set
    switch (kernel) {
    case KID_XINV: *bytes = state + 2.0 * g.KK * uline; break;           // read c, write u and du/dx
    case KID_Z: *bytes = 2.0 * g.KK * uline + g.KK * uline; break;        // read U, write NL
    case KID_XFWD: *bytes = g.KK * uline + 4.0 * state; break;           // read NL, c, onl; write c, onl
    case KID_ENERGY: *bytes = state; break;
    default: return fail(MWF_E_ARG, "kernel id out of range");
    }
    return 0;
}
