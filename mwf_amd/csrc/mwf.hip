// mwf.hip -- C-ABI of libmwf.so: handles, device buffers, plans and launches.
//
// A size class is a set of members that share (K0, MM, NN, precision): their
// states sit back to back in device memory and every kernel steps all of
// them in one launch, blockIdx.z selecting the member and its parameters.
// A single model (mwf_handle, the module state of program/velocity.f90:15-17
// plus the plans of program/transform.fftw5.f90:33-59) is one class with one
// member; an ensemble (mwf_ens) is any number of classes whose step sequence
// is captured once into a hipGraph and replayed.  All device work goes to the
// handle's own stream at default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mwf.h"
#include "diag.hpp"
#include "plans.hpp"

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                          \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) return fail(MWF_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define RCCHK(x)                 \
    do {                         \
        int rc_ = (x);           \
        if (rc_) return rc_;     \
    } while (0)

extern "C" const char* mwf_last_error(void) { return g_err.c_str(); }

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

static Plan plan_for(int prec, int k0, int mm, int nn)
{
    Plan p;
    const int lm = ilog2(mm), ln = ilog2(nn);
    if (prec == MWF_MIXED) {
        pick_x<float>(lm, p);
        if (k0 == 4) pick_z<float, 4>(ln, p);
        else pick_z<float, 6>(ln, p);
        prepare_x<float>(lm);
        if (k0 == 4) prepare_z<float, 4>(ln);
        else prepare_z<float, 6>(ln);
    } else {
        pick_x<double>(lm, p);
        if (k0 == 4) pick_z<double, 4>(ln, p);
        else pick_z<double, 6>(ln, p);
        prepare_x<double>(lm);
        if (k0 == 4) prepare_z<double, 4>(ln);
        else prepare_z<double, 6>(ln);
    }
    return p;
}

static int check_params(const mwf_params* p)
{
    if (!p) return fail(MWF_E_ARG, "null parameters");
    if (p->k0 != 4 && p->k0 != 6) return fail(MWF_E_ARG, "k0 must be 4 (M4) or 6 (M6)");
    if (!is_pow2(p->mm) || p->mm < 16 || p->mm > 1024) return fail(MWF_E_ARG, "mm must be a power of two in [16,1024]");
    if (!is_pow2(p->nn) || p->nn < 16 || p->nn > 1024) return fail(MWF_E_ARG, "nn must be a power of two in [16,1024]");
    if (p->precision != MWF_FP64 && p->precision != MWF_MIXED) return fail(MWF_E_ARG, "precision must be MWF_FP64 or MWF_MIXED");
    if (p->reserved != 0) return fail(MWF_E_ARG, "reserved must be 0");
    if (!(p->dt > 0) || !(p->re > 0)) return fail(MWF_E_ARG, "dt and re must be positive");
    return 0;
}

static Geo geo_of(const mwf_params* p)
{
    Geo g{};
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
    g.lgNT = ilog2(g.NTx);
    // NL pieces: two x blocks' columns where a block's are under 128 bytes
    // (FP64 at MM >= 512, FP32 at MM >= 256), so that the z kernels write
    // longer runs and the two blocks that read them run side by side on one
    // XCD.  FP64 M4 1024^2 (128-byte pieces): k_zc 0.196 -> 0.163 ms, k_xstep
    // unchanged; FP64 M6 2048^2 (64-byte): 182 -> 193 steps/s; mixed 4770 ->
    // 4868 (M4) and 568 -> 586 (M6).  Four blocks' columns per piece cost the
    // x kernels more than they save.
    const int cb = p->precision == MWF_FP64 ? 16 : 8; // bytes per complex value in NL
    g.NTs = g.NTx * cb < 128 ? std::min(p->nn, 2 * g.NTx) : g.NTx;
    g.lgNS = ilog2(g.NTs);
    g.XB = ZT / p->nn;
    g.UL = 2 * g.KK;
    g.NLL = g.KK;
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
    return g;
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

template <typename T> static int dev_upload(T** d, const T* h, size_t n)
{
    HIPCHK(hipMalloc((void**)d, sizeof(T) * n));
    HIPCHK(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------------------
// size class: members sharing (K0, MM, NN, precision)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Autotuner: the role of tra_precompute's FFTW_PATIENT planning
// (program/transform.fftw5.f90:33-59).  Mode 0 (default): the rule-based plan
// (k_z for NN <= 512, the fused x kernel wherever it exists).  Mode 1: each
// new size class times its candidate plans on its own buffers (zero state,
// best of 3 interleaved rounds of 5 steps) and keeps the fastest; the choice
// is cached per (precision, K0, MM, NN) for the process, so every handle and
// ensemble class of one size runs the same kernels.  Modes 2..9 force choice
// mode - 2 (where the alternative exists), for tests and plan comparisons.
// Choice bits: 1 = phase-sequential z kernel (k_zb) instead of k_z, 2 =
// x-inverse + x-forward instead of the fused x kernel, 4 = two-points-per-
// thread z kernel (k_z2; takes precedence over bit 1), 8 = flux form of the
// nonlinear term (k_zc and the flux x kernels; takes precedence over bits 1
// and 4, where it exists), 16 = with bit 8, the flux z kernel compiled for
// 2 waves per SIMD (no spills) instead of 4 (two blocks per CU), 32 = at
// NN = 1024 in the advective form, k_zb with the fold phases of a row in
// three blocks and three NL partials (with the split x kernels).
// ---------------------------------------------------------------------------
static std::atomic<int> g_tune{0};
static std::mutex g_tune_mu;
static std::map<std::array<int, 4>, int> g_tune_cache;

struct SizeClass {
    int nmem = 0;
    int prec = MWF_FP64;
    Geo g0{};                 // shared dimensions (member 0's parameters)
    std::vector<Geo> geos_h;  // per member
    Geo* geos = nullptr;      // device copy
    Batch bt{};
    Plan plan;
    size_t mpt_elems = 0;     // per member: NN * K * ldM
    double2 *c = nullptr, *cs = nullptr;
    void* onl = nullptr;      // AB2 history: double2 (FP64) or float2 (mixed) per element
    size_t hsz = 0;           // bytes of one history element
    void *U = nullptr, *NL = nullptr;
    double *epart = nullptr, *edev = nullptr;
    int nepart = 0;           // per member
    double2 *twMMd = nullptr, *twNNd = nullptr, *twPd = nullptr, *twQd = nullptr;
    float2 *twMMf = nullptr, *twNNf = nullptr, *twQf = nullptr;
    int nst = 1;
    bool u_valid = false; // U holds the x-inverse of the current c (with the laminar profile)
    Plan base;            // the rule-based plan; plan = base with the chosen alternatives
    int choice = 0;       // autotuner choice bits (see g_tune_cache)

    const void* twMM() const { return prec == MWF_MIXED ? (const void*)twMMf : (const void*)twMMd; }
    const void* twNN() const { return prec == MWF_MIXED ? (const void*)twNNf : (const void*)twNNd; }
    const void* twQ() const { return prec == MWF_MIXED ? (const void*)twQf : (const void*)twQd; }

    ~SizeClass()
    {
        void* d[] = {geos, c, onl, cs, U, NL, epart, edev, twMMd, twNNd, twPd, twQd, twMMf, twNNf, twQf};
        for (void* p : d)
            if (p) (void)hipFree(p);
    }

    int init(const std::vector<mwf_params>& ps, hipStream_t st)
    {
        nmem = (int)ps.size();
        prec = ps[0].precision;
        g0 = geo_of(&ps[0]);
        for (auto& p : ps) geos_h.push_back(geo_of(&p));
        plan = base = plan_for(prec, g0.K0, g0.MM, g0.NN);
        if (!plan.xinv || !plan.xfwd || !plan.zf || !plan.zphys || !plan.sxinv || !plan.sxfwd || !plan.zfwdp)
            return fail(MWF_E_ARG, "no kernel for this size");
        const size_t rsz = prec == MWF_MIXED ? sizeof(float2) : sizeof(double2);
        mpt_elems = (size_t)g0.NN * g0.K * g0.ldM;
        // room for either form: U and NL each hold up to 2 KK lines per x-row
        const size_t u_el = (size_t)g0.P * 2 * g0.KK * g0.NN, nl_el = (size_t)g0.P * 2 * g0.KK * g0.NN;
        nepart = (g0.NN / g0.NTx) * g0.K0;
        RCCHK(dev_upload(&geos, geos_h.data(), geos_h.size()));
        HIPCHK(hipMalloc((void**)&c, sizeof(double2) * mpt_elems * nmem));
        hsz = prec == MWF_MIXED ? sizeof(float2) : sizeof(double2);
        HIPCHK(hipMalloc(&onl, hsz * mpt_elems * nmem));
        HIPCHK(hipMalloc((void**)&cs, sizeof(double2) * mpt_elems));
        HIPCHK(hipMalloc(&U, rsz * u_el * nmem));
        // three NL partials where the phase-split z kernel may run
        HIPCHK(hipMalloc(&NL, rsz * nl_el * nmem * (base.zf_ph3 ? 3 : 1)));
        HIPCHK(hipMalloc((void**)&epart, sizeof(double) * nepart * nmem));
        HIPCHK(hipMalloc((void**)&edev, sizeof(double) * nmem));
        bt = Batch{geos, mpt_elems, u_el, nl_el, nepart};
        auto td = [&](double2** d, int n) { auto t = twiddles<double2>(n); return dev_upload(d, t.data(), t.size()); };
        auto tf = [&](float2** d, int n) { auto t = twiddles<float2>(n); return dev_upload(d, t.data(), t.size()); };
        RCCHK(td(&twMMd, g0.MM));
        RCCHK(td(&twNNd, g0.NN));
        RCCHK(td(&twPd, g0.P));
        RCCHK(td(&twQd, g0.Q));
        RCCHK(tf(&twMMf, g0.MM));
        RCCHK(tf(&twNNf, g0.NN));
        RCCHK(tf(&twQf, g0.Q));
        HIPCHK(hipMemsetAsync(c, 0, sizeof(double2) * mpt_elems * nmem, st));
        HIPCHK(hipMemsetAsync(onl, 0, hsz * mpt_elems * nmem, st));
        HIPCHK(hipStreamSynchronize(st));
        const int mode = g_tune.load();
        if (mode == 1) RCCHK(tune(st));
        else if (mode >= 2) RCCHK(apply(mode - 2));
        return 0;
    }

    // the plan of choice bits ch; the flux form also changes how many lines U
    // and NL hold per x-row, in every member's geometry
    int apply(int ch)
    {
        choice = ch;
        plan = base;
        const bool flux = (ch & 8) && base.zc;
        if (flux) {
            plan.zf = (ch & 16) && base.zc_alt ? base.zc_alt : base.zc;
            plan.xinv = base.xinv_f;
            plan.xfwd = base.xfwd_f;
            plan.xstep = base.xstep_f;
        } else if ((ch & 4) && base.zf_alt2) {
            plan.zf = base.zf_alt2;
        } else if ((ch & 1) && base.zf_alt) {
            plan.zf = base.zf_alt;
        }
        // phase-split z kernel: three NL partials, summed by the split x kernels
        const bool ph3 = !flux && (ch & 32) && base.zf_ph3;
        if (ph3) plan.zf = base.zf_ph3;
        // the fused x kernel sums the partials in FP64; FP32 stages whole NL
        // panels into LDS and takes them with the split x kernels
        if ((ch & 2) || (ph3 && prec == MWF_MIXED)) plan.xstep = nullptr;
        bt.nl_part = ph3 ? bt.nl_stride * nmem : 0;
        const int UL = flux ? g0.KK : 2 * g0.KK, NLL = flux ? 2 * g0.KK : g0.KK;
        if (g0.UL != UL) {
            g0.UL = UL;
            g0.NLL = NLL;
            for (auto& g : geos_h) {
                g.UL = UL;
                g.NLL = NLL;
            }
            HIPCHK(hipMemcpy(geos, geos_h.data(), sizeof(Geo) * geos_h.size(), hipMemcpyHostToDevice));
            u_valid = false;
        }
        return 0;
    }

    bool flux() const { return g0.UL == g0.KK; }

    int reset_state(hipStream_t st)
    {
        HIPCHK(hipMemsetAsync(c, 0, sizeof(double2) * mpt_elems * nmem, st));
        HIPCHK(hipMemsetAsync(onl, 0, hsz * mpt_elems * nmem, st));
        HIPCHK(hipStreamSynchronize(st));
        nst = 1;
        u_valid = false;
        return 0;
    }

    int tune(hipStream_t st)
    {
        const std::array<int, 4> key{prec, g0.K0, g0.MM, g0.NN};
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) return apply(it->second);
        std::vector<int> cand;
        std::vector<int> zc{0};
        if (base.zf_alt) zc.push_back(1);
        if (base.zf_alt2) zc.push_back(4);
        if (base.zc) zc.push_back(8);
        if (base.zc_alt) zc.push_back(8 | 16);
        if (base.zf_ph3) {
            cand.push_back(32 | 2);
            if (prec != MWF_MIXED && base.xstep) cand.push_back(32);
        }
        for (int z : zc) {
            if ((z & 8) ? (bool)base.xstep_f : (bool)base.xstep) cand.push_back(z);
            cand.push_back(z | 2);
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            return fail(MWF_E_HIP, "hipEventCreate");
        }
        auto run = [&](int, auto&& f) -> int {
            const hipError_t e = f();
            return e == hipSuccess ? 0 : fail(MWF_E_HIP, hipGetErrorString(e));
        };
        // every candidate runs 2 untimed steps (code load, first-step path),
        // then NR rounds, interleaved across candidates, of one untimed and NS
        // timed steps from a fresh zero state; a candidate's time is its best
        // round, which keeps a single slow round from deciding the choice.
        // Candidates a few percent apart (M4 1024^2 mixed: the one- and two-
        // buffer z kernels) need this many steps to be ordered reliably; it
        // costs about 0.2 s at M4 1024^2 and 1.5 s at M6 2048^2.
        constexpr int NR = 5, NS = 10;
        std::vector<float> tbest(cand.size(), -1.0f);
        int rc = 0;
        for (size_t i = 0; i < cand.size() && !rc; i++) {
            if ((rc = apply(cand[i]))) break;
            if ((rc = reset_state(st))) break;
            for (int w = 0; w < 2 && !rc; w++) rc = step(st, run);
        }
        for (int round = 0; round < NR && !rc; round++) {
            for (size_t i = 0; i < cand.size() && !rc; i++) {
                if ((rc = apply(cand[i]))) break;
                if ((rc = reset_state(st))) break;
                rc = step(st, run);
                if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = fail(MWF_E_HIP, "hipEventRecord");
                for (int w = 0; w < NS && !rc; w++) rc = step(st, run);
                if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = fail(MWF_E_HIP, "hipEventRecord");
                float ms = 0.0f;
                if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = fail(MWF_E_HIP, "hipEventElapsedTime");
                if (!rc && (tbest[i] < 0.0f || ms < tbest[i])) tbest[i] = ms;
            }
        }
        const int best = rc ? cand[0] : cand[std::min_element(tbest.begin(), tbest.end()) - tbest.begin()];
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        RCCHK(apply(best));
        RCCHK(reset_state(st));
        g_tune_cache[key] = best;
        return 0;
    }

    // one step of every member: vel_TS (program/velocity.f90:21-41).  With the
    // fused x kernel a step is k_z + k_xstep, k_xstep leaving U ready for the
    // next step.  When U is stale (first step, new state) it is rebuilt by the
    // inverse half of k_xstep itself, so a run restarted from (c, onl) takes
    // bit for bit the path of the run it continues.
    template <typename TIMED> int step(hipStream_t st, TIMED&& timed)
    {
        if (plan.xstep && !u_valid)
            RCCHK(timed(0, [&] { return plan.xstep(bt, nmem, g0, st, NL, c, onl, U, twMM(), twPd, 0, 1); }));
        else if (!plan.xstep)
            RCCHK(timed(0, [&] { return plan.xinv(bt, nmem, g0, st, c, U, nullptr, twMM(), twPd, 1); }));
        RCCHK(timed(1, [&] { return plan.zf(bt, nmem, g0, st, U, NL, twNN(), twQ()); }));
        if (plan.xstep) {
            RCCHK(timed(4, [&] { return plan.xstep(bt, nmem, g0, st, NL, c, onl, U, twMM(), twPd, nst, 0); }));
            u_valid = true;
        } else {
            RCCHK(timed(2, [&] { return plan.xfwd(bt, nmem, g0, st, NL, c, onl, nullptr, twMM(), twPd, 0, nst); }));
        }
        nst = 0;
        return 0;
    }

    // energies of all members into out[0..nmem) (default edev), with the
    // block partials in part (default epart, nepart * nmem doubles)
    int energy(hipStream_t st, double* part = nullptr, double* out = nullptr)
    {
        if (!part) part = epart;
        if (!out) out = edev;
        hipError_t e = plan.xinv(bt, nmem, g0, st, c, nullptr, part, twMM(), twPd, 0);
        if (e != hipSuccess) return fail(MWF_E_HIP, std::string("energy launch: ") + hipGetErrorString(e));
        hipLaunchKernelGGL(k_ereduce, dim3((nmem + 63) / 64), dim3(64), 0, st, part, nepart, nmem, out);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

// reference mpt / spec (re, im planes, (slots, M, NN) column-major) <-> device
// rows [NN][slots][ldM]
static void to_dev_rows(const Geo& g, int slots, size_t elems, const double* re, const double* im, double2* d)
{
    std::memset(d, 0, sizeof(double2) * elems);
    for (int n = 0; n < g.NN; n++)
        for (int m = 0; m < g.M; m++)
            for (int k = 0; k < slots; k++) {
                const size_t i = (size_t)k + (size_t)slots * ((size_t)m + (size_t)g.M * n);
                d[((size_t)n * slots + k) * g.ldM + m] = double2{re[i], im[i]};
            }
}

static void from_dev_rows(const Geo& g, int slots, const double2* d, double* re, double* im)
{
    for (int n = 0; n < g.NN; n++)
        for (int m = 0; m < g.M; m++)
            for (int k = 0; k < slots; k++) {
                const size_t i = (size_t)k + (size_t)slots * ((size_t)m + (size_t)g.M * n);
                const double2 v = d[((size_t)n * slots + k) * g.ldM + m];
                re[i] = v.x;
                im[i] = v.y;
            }
}

static void to_dev_layout(const Geo& g, size_t elems, const double* re, const double* im, double2* d)
{
    to_dev_rows(g, g.K, elems, re, im, d);
}

static void from_dev_layout(const Geo& g, const double2* d, double* re, double* im)
{
    from_dev_rows(g, g.K, d, re, im);
}

// ---------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------
enum { KID_XINV = 0, KID_Z = 1, KID_XFWD = 2, KID_ENERGY = 3, KID_XSTEP = 4, NKID = 5 };

struct Profiler {
    struct Ev {
        int id;
        hipEvent_t a, b;
    };
    int on = 0;
    std::vector<Ev> pending;
    std::vector<hipEvent_t> pool;
    double ms[NKID] = {0, 0, 0, 0, 0};
    int64_t cnt[NKID] = {0, 0, 0, 0, 0};

    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    int drain()
    {
        for (auto& e : pending) {
            HIPCHK(hipEventSynchronize(e.b));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e.a, e.b));
            ms[e.id] += t;
            cnt[e.id] += 1;
            pool.push_back(e.a);
            pool.push_back(e.b);
        }
        pending.clear();
        return 0;
    }
    template <typename F> int timed(hipStream_t st, int id, F&& f)
    {
        hipEvent_t a = nullptr, b = nullptr;
        if (on) {
            a = get();
            b = get();
            HIPCHK(hipEventRecord(a, st));
        }
        hipError_t e = f();
        if (e != hipSuccess) return fail(MWF_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        if (on) {
            HIPCHK(hipEventRecord(b, st));
            pending.push_back({id, a, b});
            if (pending.size() > 4096) return drain();
        }
        return 0;
    }
    void release()
    {
        (void)drain();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

// ---------------------------------------------------------------------------
// Scratch of the entry point that allocates nothing (mwf_energy_noalloc): the
// energy partials of one handle, in a file-scope device array that every GPU
// the library runs on holds until the process ends (64 KB).  In host code the
// name `scratch` is a host-side stand-in, not a device address: its device
// address is looked up once per device, with that device current, and only
// that address (g_scratch_dev) is passed to kernels.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_PARTS = 8192; // doubles
static __device__ __align__(256) double scratch[NOALLOC_PARTS];
constexpr int MAX_DEVICES = 64;
static double* g_scratch_dev[MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current device (looked up on first use)
static int scratch_setup(int* dev_out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(MWF_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MWF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *dev_out = dev;
    return 0;
}

// ---------------------------------------------------------------------------
// single model handle
// ---------------------------------------------------------------------------
struct mwf_handle {
    mwf_params p{};
    int dev = 0;               // the GPU the handle lives on
    hipStream_t st = nullptr;
    std::unique_ptr<SizeClass> cls;
    double* phys = nullptr;
    size_t phys_elems = 0;
    void* diag = nullptr;      // device scratch of the diagnostics
    size_t diag_bytes = 0;
    double2* hstage = nullptr; // pinned host staging, one device mpt
    double2* tspec = nullptr;  // device spec of the transform entry points
    double2* hspec = nullptr;  // pinned host staging of one spec
    Profiler prof;
};

static void destroy(mwf_handle* h)
{
    if (!h) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    h->prof.release();
    h->cls.reset();
    if (h->phys) (void)hipFree(h->phys);
    if (h->diag) (void)hipFree(h->diag);
    if (h->hstage) (void)hipHostFree(h->hstage);
    if (h->tspec) (void)hipFree(h->tspec);
    if (h->hspec) (void)hipHostFree(h->hspec);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

extern "C" int mwf_create(const mwf_params* p, mwf_handle** out)
{
    if (!out) return fail(MWF_E_ARG, "null argument");
    *out = nullptr;
    RCCHK(check_params(p));
    mwf_handle* h = new mwf_handle();
    h->p = *p;
    int rc = scratch_setup(&h->dev);
    if (!rc && hipStreamCreate(&h->st) != hipSuccess) rc = fail(MWF_E_HIP, "hipStreamCreate failed");
    if (!rc) {
        h->cls.reset(new SizeClass());
        rc = h->cls->init({*p}, h->st);
    }
    if (!rc && hipHostMalloc((void**)&h->hstage, sizeof(double2) * h->cls->mpt_elems, hipHostMallocDefault) != hipSuccess)
        rc = fail(MWF_E_NOMEM, "hipHostMalloc failed");
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

static int upload_mpt(mwf_handle* h, const double* re, const double* im, double2* dst)
{
    SizeClass& k = *h->cls;
    to_dev_layout(k.g0, k.mpt_elems, re, im, h->hstage);
    HIPCHK(hipMemcpyAsync(dst, h->hstage, sizeof(double2) * k.mpt_elems, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

static int download_mpt(mwf_handle* h, const double2* src, double* re, double* im)
{
    SizeClass& k = *h->cls;
    HIPCHK(hipMemcpyAsync(h->hstage, src, sizeof(double2) * k.mpt_elems, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    from_dev_layout(k.g0, h->hstage, re, im);
    return 0;
}

extern "C" int mwf_set_state(mwf_handle* h, const double* re, const double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    RCCHK(upload_mpt(h, re, im, h->cls->c));
    h->cls->nst = 1;
    h->cls->u_valid = false;
    return 0;
}

extern "C" int mwf_get_state(mwf_handle* h, double* re, double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    return download_mpt(h, h->cls->c, re, im);
}

extern "C" int mwf_put_state(mwf_handle* h, const double* re, const double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    RCCHK(upload_mpt(h, re, im, h->cls->c));
    h->cls->u_valid = false; // the AB2 history and nst are kept
    return 0;
}

extern "C" int mwf_step(mwf_handle* h, int32_t nsteps)
{
    if (!h) return fail(MWF_E_ARG, "null handle");
    if (nsteps < 0) return fail(MWF_E_ARG, "nsteps < 0");
    auto timed = [&](int id, auto&& f) { return h->prof.timed(h->st, id, f); };
    for (int s = 0; s < nsteps; s++) RCCHK(h->cls->step(h->st, timed));
    return 0;
}

extern "C" int mwf_nonlinear(mwf_handle* h, double* nlre, double* nlim)
{
    if (!h || !nlre || !nlim) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    HIPCHK(hipMemsetAsync(k.cs, 0, sizeof(double2) * k.mpt_elems, h->st));
    auto launch = [&](hipError_t e) { return e == hipSuccess ? 0 : fail(MWF_E_HIP, hipGetErrorString(e)); };
    // U is rebuilt here by k_xinv; the next step rebuilds it with its own
    // inverse, so that stepping does not depend on this call
    k.u_valid = false;
    RCCHK(launch(k.plan.xinv(k.bt, 1, k.g0, h->st, k.c, k.U, nullptr, k.twMM(), k.twPd, 1)));
    RCCHK(launch(k.plan.zf(k.bt, 1, k.g0, h->st, k.U, k.NL, k.twNN(), k.twQ())));
    // mode 1 writes the nonlinear term into nlout (the scratch mpt) only
    RCCHK(launch(k.plan.xfwd(k.bt, 1, k.g0, h->st, k.NL, k.c, k.onl, k.cs, k.twMM(), k.twPd, 1, 0)));
    return download_mpt(h, k.cs, nlre, nlim);
}

extern "C" int mwf_energy(mwf_handle* h, double* e)
{
    if (!h || !e) return fail(MWF_E_ARG, "null argument");
    RCCHK(h->prof.timed(h->st, KID_ENERGY, [&] {
        int rc = h->cls->energy(h->st);
        return rc ? hipErrorUnknown : hipSuccess;
    }));
    HIPCHK(hipMemcpyAsync(e, h->cls->edev, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_energy_noalloc(mwf_handle* h, double* e_dev)
{
    if (!h || !e_dev) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    if ((long long)k.nepart * k.nmem > NOALLOC_PARTS)
        return fail(MWF_E_ARG, "the handle's energy partials exceed the no-allocation scratch (" +
                                   std::to_string(NOALLOC_PARTS) + " doubles)");
    double* part = g_scratch_dev[h->dev];
    if (!part) return fail(MWF_E_STATE, "no-allocation scratch not set up on this device");
    return k.energy(h->st, part, e_dev);
}

static int ensure_phys(mwf_handle* h)
{
    const Geo& g = h->cls->g0;
    const size_t np = (size_t)g.KK * g.Q * g.P;
    if (h->phys_elems < np) {
        if (h->phys) HIPCHK(hipFree(h->phys));
        h->phys = nullptr;
        h->phys_elems = 0;
        HIPCHK(hipMalloc((void**)&h->phys, sizeof(double) * np));
        h->phys_elems = np;
    }
    return 0;
}

static int ensure_diag(mwf_handle* h, size_t bytes)
{
    if (h->diag_bytes < bytes) {
        if (h->diag) HIPCHK(hipFree(h->diag));
        h->diag = nullptr;
        h->diag_bytes = 0;
        HIPCHK(hipMalloc(&h->diag, bytes));
        h->diag_bytes = bytes;
    }
    return 0;
}

// Denominator of the turbulent fraction of io_writeVTK_txz (program/io.f90
// :564-565,615): (N_+1)(M_+1), M_ = 3M/xs, N_ = 3N/zs, which counts the
// periodic copies of the first row and column that the VTK file repeats.
static double turb_den(const Geo& g, int xs, int zs)
{
    return (double)(g.Q / zs + 1) * (double)(g.P / xs + 1);
}

// vel_mpt2phys of a device mpt into h->phys (program/velocity.f90:84-90):
// var_mpt2spec without the laminar profile, then tra_spec2phys.  U is
// overwritten, so the next step re-derives it from c.
static int phys_of(mwf_handle* h, const double2* src)
{
    SizeClass& k = *h->cls;
    const Geo& g = k.g0;
    RCCHK(ensure_phys(h));
    k.u_valid = false;
    hipError_t e = k.plan.xinv(k.bt, 1, g, h->st, src, k.U, nullptr, k.twMM(), k.twPd, 0);
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    e = k.plan.zphys(k.bt, 1, g, h->st, k.U, h->phys, k.twNN(), k.twQ());
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    return 0;
}

extern "C" int mwf_mpt2phys(mwf_handle* h, const double* re, const double* im, double* phys)
{
    if (!h || !re || !im || !phys) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    const Geo& g = k.g0;
    RCCHK(upload_mpt(h, re, im, k.cs));
    RCCHK(phys_of(h, k.cs));
    HIPCHK(hipMemcpyAsync(phys, h->phys, sizeof(double) * g.KK * g.Q * g.P, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

static int ensure_tspec(mwf_handle* h)
{
    const Geo& g = h->cls->g0;
    const size_t n = (size_t)g.NN * g.KK * g.ldM;
    if (!h->tspec) HIPCHK(hipMalloc((void**)&h->tspec, sizeof(double2) * n));
    if (!h->hspec && hipHostMalloc((void**)&h->hspec, sizeof(double2) * n, hipHostMallocDefault) != hipSuccess)
        return fail(MWF_E_NOMEM, "hipHostMalloc failed");
    return 0;
}

extern "C" int mwf_spec2phys(mwf_handle* h, const double* re, const double* im, double* phys)
{
    if (!h || !re || !im || !phys) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    const Geo& g = k.g0;
    RCCHK(ensure_tspec(h));
    RCCHK(ensure_phys(h));
    const size_t n = (size_t)g.NN * g.KK * g.ldM;
    to_dev_rows(g, g.KK, n, re, im, h->hspec);
    HIPCHK(hipMemcpyAsync(h->tspec, h->hspec, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
    k.u_valid = false; // U is overwritten
    hipError_t e = k.plan.sxinv(g, h->st, h->tspec, k.U, k.twMM(), k.twPd);
    if (e == hipSuccess) e = k.plan.zphys(k.bt, 1, g, h->st, k.U, h->phys, k.twNN(), k.twQ());
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(phys, h->phys, sizeof(double) * g.KK * g.Q * g.P, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_phys2spec(mwf_handle* h, const double* phys, double* re, double* im)
{
    if (!h || !phys || !re || !im) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    const Geo& g = k.g0;
    RCCHK(ensure_tspec(h));
    RCCHK(ensure_phys(h));
    const size_t n = (size_t)g.NN * g.KK * g.ldM;
    HIPCHK(hipMemcpyAsync(h->phys, phys, sizeof(double) * g.KK * g.Q * g.P, hipMemcpyHostToDevice, h->st));
    hipError_t e = k.plan.zfwdp(g, h->st, h->phys, k.NL, k.twNN(), k.twQ());
    if (e == hipSuccess) e = k.plan.sxfwd(g, h->st, k.NL, h->tspec, k.twMM(), k.twPd);
    if (e != hipSuccess) return fail(MWF_E_HIP, hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(h->hspec, h->tspec, sizeof(double2) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    from_dev_rows(g, g.KK, h->hspec, re, im);
    return 0;
}

// the history in the mixed form is FP32: widened on the way out, rounded on
// the way in (what the mixed step itself stores)
extern "C" int mwf_get_onl(mwf_handle* h, double* re, double* im, int32_t* first_step)
{
    if (!h || !re || !im || !first_step) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    *first_step = k.nst;
    if (k.hsz == sizeof(double2)) return download_mpt(h, static_cast<double2*>(k.onl), re, im);
    float2* f = reinterpret_cast<float2*>(h->hstage);
    HIPCHK(hipMemcpyAsync(f, k.onl, sizeof(float2) * k.mpt_elems, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t i = k.mpt_elems; i-- > 0;) h->hstage[i] = double2{(double)f[i].x, (double)f[i].y};
    from_dev_layout(k.g0, h->hstage, re, im);
    return 0;
}

extern "C" int mwf_set_onl(mwf_handle* h, const double* re, const double* im)
{
    if (!h || !re || !im) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    if (k.hsz == sizeof(double2)) {
        RCCHK(upload_mpt(h, re, im, static_cast<double2*>(k.onl)));
    } else {
        to_dev_layout(k.g0, k.mpt_elems, re, im, h->hstage);
        float2* f = reinterpret_cast<float2*>(h->hstage);
        for (size_t i = 0; i < k.mpt_elems; i++) f[i] = float2{(float)h->hstage[i].x, (float)h->hstage[i].y};
        HIPCHK(hipMemcpyAsync(k.onl, f, sizeof(float2) * k.mpt_elems, hipMemcpyHostToDevice, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
    }
    k.nst = 0;
    return 0;
}

extern "C" int mwf_spectrum_max(mwf_handle* h, double* m_max, double* n_max)
{
    if (!h || !m_max || !n_max) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->cls->g0;
    RCCHK(ensure_diag(h, sizeof(double) * (g.MM + g.NN)));
    unsigned long long* dm = reinterpret_cast<unsigned long long*>(h->diag);
    double* dn = reinterpret_cast<double*>(dm + g.MM);
    HIPCHK(hipMemsetAsync(dm, 0, sizeof(unsigned long long) * g.MM, h->st));
    hipLaunchKernelGGL(k_specmax, dim3(g.NN), dim3(256), 0, h->st, g, h->cls->c, dn, dm);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m_max, dm, sizeof(double) * g.MM, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(n_max, dn, sizeof(double) * g.NN, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_vel_history(mwf_handle* h, int32_t nh, double y, double* ans)
{
    if (!h || !ans) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->cls->g0;
    if (nh < 1 || nh > g.Q / 2) return fail(MWF_E_ARG, "nh must be in [1, Q/2]");
    // wall-normal bases at y (modes.M4.f90:83-108): h_k = sin/cos(k beta y) for
    // odd/even k, g_k = cos/sin(k beta y) for odd/even k, from the host libm
    std::vector<double> hb(2 * g.K0 - 1, 1.0);
    for (int k = 1; k < g.K0; k++) {
        const double a = k * g.beta * y;
        hb[k] = (k & 1) ? std::sin(a) : std::cos(a);
        hb[g.K0 - 1 + k] = (k & 1) ? std::cos(a) : std::sin(a);
    }
    RCCHK(phys_of(h, h->cls->c));
    RCCHK(ensure_diag(h, sizeof(double) * (hb.size() + 4 * (size_t)nh)));
    double* dhb = reinterpret_cast<double*>(h->diag);
    double* dans = dhb + hb.size();
    HIPCHK(hipMemcpyAsync(dhb, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, h->st));
    const double lz = 2.0 * M_PI / g.gamma;
    hipLaunchKernelGGL(k_velhist, dim3((nh + 63) / 64), dim3(64), 0, h->st, h->phys, g.Q, g.KK, g.K0, (int)nh, lz,
                       dhb, dans);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ans, dans, sizeof(double) * 4 * nh, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

extern "C" int mwf_turb_sample(mwf_handle* h, int32_t dm, int32_t dn, double* out)
{
    if (!h || !out) return fail(MWF_E_ARG, "null argument");
    const Geo& g = h->cls->g0;
    if (dm < 1 || dn < 1 || dm > g.P || dn > g.Q) return fail(MWF_E_ARG, "sampling steps must be in [1, P] x [1, Q]");
    const int SM = g.P / dm, SN = g.Q / dn;
    RCCHK(phys_of(h, h->cls->c));
    RCCHK(ensure_diag(h, sizeof(double) * 3 * (size_t)SM * SN));
    double* dout = reinterpret_cast<double*>(h->diag);
    const int npt = SM * SN;
    hipLaunchKernelGGL(k_eposc, dim3((npt + 255) / 256), dim3(256), 0, h->st, h->phys, g.Q, g.KK, g.K0, (int)dm,
                       (int)dn, SM, SN, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * 3 * npt, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

static bool same_size(const mwf_params& a, const mwf_params& b)
{
    return a.k0 == b.k0 && a.mm == b.mm && a.nn == b.nn;
}

extern "C" int mwf_extend_x(mwf_handle* src, int32_t factor, mwf_handle* dst)
{
    if (!src || !dst || src == dst) return fail(MWF_E_ARG, "two distinct handles required");
    if (factor < 2 || factor > src->p.mm || (factor & (factor - 1))) return fail(MWF_E_ARG, "factor must be a power of two in [2, mm]");
    if (!same_size(src->p, dst->p)) return fail(MWF_E_ARG, "src and dst differ in k0, mm or nn");
    const double want = src->p.alpha / factor;
    if (std::fabs(dst->p.alpha - want) > 1e-12 * want) return fail(MWF_E_ARG, "dst alpha must be src alpha / factor");
    SizeClass& s = *src->cls;
    SizeClass& d = *dst->cls;
    HIPCHK(hipStreamSynchronize(src->st));
    const size_t total = (size_t)s.g0.NN * s.g0.K * s.g0.ldM;
    hipLaunchKernelGGL(k_extend_x, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, dst->st, s.g0, s.c, d.c,
                       (int)factor);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(dst->st));
    d.nst = 1;
    d.u_valid = false;
    return 0;
}

extern "C" int mwf_blend(mwf_handle* a, mwf_handle* b, double d, mwf_handle* dst)
{
    if (!a || !b || !dst) return fail(MWF_E_ARG, "null handle");
    if (!same_size(a->p, b->p) || !same_size(a->p, dst->p)) return fail(MWF_E_ARG, "handles differ in k0, mm or nn");
    HIPCHK(hipStreamSynchronize(a->st));
    HIPCHK(hipStreamSynchronize(b->st));
    const size_t n = dst->cls->mpt_elems;
    hipLaunchKernelGGL(k_blend, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, dst->st, n, a->cls->c, b->cls->c, d,
                       dst->cls->c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(dst->st));
    dst->cls->nst = 1;
    dst->cls->u_valid = false;
    return 0;
}

extern "C" int mwf_autotune(int32_t mode)
{
    if (mode < 0 || mode > 65) return fail(MWF_E_ARG, "mode must be 0 (rules), 1 (timed) or 2..65 (forced plan)");
    g_tune.store(mode);
    return 0;
}

extern "C" int mwf_plan_info(mwf_handle* h, int32_t* choice)
{
    if (!h || !choice) return fail(MWF_E_ARG, "null argument");
    const SizeClass& k = *h->cls;
    *choice = ((k.base.zf_alt && k.plan.zf == k.base.zf_alt) ? 1 : 0) | (k.plan.xstep ? 0 : 2) |
              ((k.base.zf_alt2 && k.plan.zf == k.base.zf_alt2) ? 4 : 0) | (k.flux() ? 8 : 0) |
              ((k.base.zc_alt && k.plan.zf == k.base.zc_alt) ? 16 : 0) |
              ((k.base.zf_ph3 && k.plan.zf == k.base.zf_ph3) ? 32 : 0);
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
    RCCHK(h->prof.drain());
    h->prof.on = on ? 1 : 0;
    for (int i = 0; i < NKID; i++) {
        h->prof.ms[i] = 0;
        h->prof.cnt[i] = 0;
    }
    return 0;
}

extern "C" int mwf_profile_read(mwf_handle* h, int32_t kernel, double* total_ms, int64_t* count)
{
    if (!h || !total_ms || !count) return fail(MWF_E_ARG, "null argument");
    if (kernel < 0 || kernel >= NKID) return fail(MWF_E_ARG, "kernel id out of range");
    RCCHK(h->prof.drain());
    *total_ms = h->prof.ms[kernel];
    *count = h->prof.cnt[kernel];
    return 0;
}

static double kernel_bytes(const SizeClass& k, int kernel)
{
    const Geo& g = k.g0;
    const double s = k.prec == MWF_MIXED ? 8.0 : 16.0; // complex intermediate
    const double state = 16.0 * g.K * g.M * g.NN;      // one complex FP64 mpt
    const double hist = s * g.K * g.M * g.NN;          // the AB2 history (precision of the transforms)
    const double uline = s * (double)g.P * g.NN;         // one x-physical line set
    switch (kernel) {
    case KID_XINV: return state + g.UL * uline;               // read c, write U
    case KID_Z: return (double)(g.UL + g.NLL) * uline;        // read U, write NL
    case KID_XFWD: return g.NLL * uline + 2.0 * state + 2.0 * hist;          // read NL, c, onl; write c, onl
    case KID_ENERGY: return state;
    case KID_XSTEP: return (double)(g.NLL + g.UL) * uline + 2.0 * state + 2.0 * hist; // + write U
    default: return -1.0;
    }
}

extern "C" int mwf_kernel_bytes(mwf_handle* h, int32_t kernel, double* bytes)
{
    if (!h || !bytes) return fail(MWF_E_ARG, "null argument");
    const double b = kernel_bytes(*h->cls, kernel);
    if (b < 0) return fail(MWF_E_ARG, "kernel id out of range");
    *bytes = b;
    return 0;
}

extern "C" int mwf_turb_fraction(mwf_handle* h, int32_t xs, int32_t zs, double ct, double* frac)
{
    if (!h || !frac) return fail(MWF_E_ARG, "null argument");
    SizeClass& k = *h->cls;
    const Geo& g = k.g0;
    if (xs < 1 || zs < 1 || xs > g.P || zs > g.Q) return fail(MWF_E_ARG, "strides must be in [1, P] x [1, Q]");
    RCCHK(phys_of(h, k.c));
    RCCHK(ensure_diag(h, sizeof(unsigned long long)));
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(h->diag);
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), h->st));
    const int npts = ((g.P + xs - 1) / xs) * ((g.Q + zs - 1) / zs);
    hipLaunchKernelGGL(k_turbcount, dim3((npts + 255) / 256, 1), dim3(256), 0, h->st, h->phys, k.geos, (int)xs,
                       (int)zs, ct, cnt);
    HIPCHK(hipGetLastError());
    unsigned long long c = 0;
    HIPCHK(hipMemcpyAsync(&c, cnt, sizeof(c), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    *frac = (double)c / turb_den(g, xs, zs);
    return 0;
}

// ---------------------------------------------------------------------------
// ensemble
//
// Each size class steps on its own stream (default priority).  A chunk of
// steps forks from the ensemble stream to every class stream and joins back
// at its end, so under stream capture the classes become parallel branches
// of one graph.  Tracking samples (energies, turbulent fractions,
// relaminarisation steps) are taken on the device inside the chunk and kept
// in device trace buffers, which the host drains only when they fill or when
// it reads them: a tracked run has no host round trip per sample.
// ---------------------------------------------------------------------------
struct ClassTrack {
    double* etrace = nullptr;          // [cap][nmem] energies
    double* ftrace = nullptr;          // [cap][nmem] turbulent fractions
    int* fill = nullptr;               // rows written since the last drain
    long long* total = nullptr;        // samples since tracking began
    long long* relam = nullptr;        // [nmem] first step with E <= thr, or -1
    unsigned long long* count = nullptr; // [nmem] turbulent points of the sample
    void* Ud = nullptr;                // U of the turbulence sample (the step's U stays valid)
    double* physd = nullptr;           // [nmem] phys of the turbulence sample

    void release()
    {
        void* d[] = {etrace, ftrace, fill, total, relam, count, Ud, physd};
        for (void* p : d)
            if (p) (void)hipFree(p);
        *this = ClassTrack{};
    }
};

struct mwf_ens {
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<std::unique_ptr<SizeClass>> classes;
    std::vector<hipStream_t> cst;           // per class
    std::vector<hipEvent_t> cev;            // per class, joins
    std::vector<ClassTrack> trk;            // per class
    std::vector<std::pair<int, int>> where; // member -> (class, index in class)
    std::vector<mwf_params> params;
    double2* hstage = nullptr;
    size_t hstage_elems = 0;
    double* ebuf = nullptr; // host, one energy per member
    // step graphs by (steps, sample at the end), built on demand
    std::map<std::pair<int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    int nst_all = 1;
    std::vector<std::vector<double>> eclass; // host copy of each class's energies
    // on-device tracking
    static constexpr int CAP = 256; // samples held on the device between drains
    int track_every = 0;
    double track_thr = 0.0;
    int turb_on = 0, turb_xs = 1, turb_zs = 1;
    double turb_ct = 0.0;
    int pending = 0;                // samples in the device buffers (all classes alike)
    int64_t step_count = 0;
    std::vector<std::vector<double>> trace;  // [member][sample] energies
    std::vector<std::vector<double>> ftrace; // [member][sample] turbulent fractions
    std::vector<int64_t> relam;              // first step with E <= threshold, or -1
};

static void ens_drop_graphs(mwf_ens* e)
{
    for (auto& kv : e->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    e->graphs.clear();
}

static void ens_destroy(mwf_ens* e)
{
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (auto s : e->cst)
        if (s) (void)hipStreamSynchronize(s);
    ens_drop_graphs(e);
    for (auto& t : e->trk) t.release();
    e->classes.clear();
    for (auto s : e->cst)
        if (s) (void)hipStreamDestroy(s);
    for (auto v : e->cev)
        if (v) (void)hipEventDestroy(v);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->hstage) (void)hipHostFree(e->hstage);
    delete[] e->ebuf;
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

extern "C" int mwf_ens_create(const mwf_params* members, int32_t n, mwf_ens** out)
{
    if (!members || !out || n <= 0) return fail(MWF_E_ARG, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) RCCHK(check_params(&members[i]));
    mwf_ens* e = new mwf_ens();
    e->params.assign(members, members + n);
    e->where.resize(n);
    e->ebuf = new double[n];
    e->relam.assign(n, -1);
    e->trace.resize(n);
    e->ftrace.resize(n);
    int rc = 0;
    if (hipStreamCreate(&e->st) != hipSuccess || hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
        rc = fail(MWF_E_HIP, "hipStreamCreate / hipEventCreate failed");
    std::map<std::tuple<int, int, int, int>, std::vector<int>> groups;
    for (int i = 0; i < n; i++)
        groups[std::make_tuple(members[i].precision, members[i].k0, members[i].mm, members[i].nn)].push_back(i);
    size_t maxel = 0;
    for (auto& kv : groups) {
        if (rc) break;
        std::vector<mwf_params> ps;
        for (int i : kv.second) ps.push_back(members[i]);
        hipStream_t cs = nullptr;
        hipEvent_t ce = nullptr;
        if (hipStreamCreate(&cs) != hipSuccess) rc = fail(MWF_E_HIP, "hipStreamCreate failed");
        e->cst.push_back(cs);
        if (!rc && hipEventCreateWithFlags(&ce, hipEventDisableTiming) != hipSuccess)
            rc = fail(MWF_E_HIP, "hipEventCreate failed");
        e->cev.push_back(ce);
        e->trk.emplace_back();
        if (rc) break;
        std::unique_ptr<SizeClass> k(new SizeClass());
        rc = k->init(ps, cs);
        if (rc) break;
        for (size_t j = 0; j < kv.second.size(); j++) e->where[kv.second[j]] = {(int)e->classes.size(), (int)j};
        maxel = std::max(maxel, k->mpt_elems);
        e->classes.push_back(std::move(k));
    }
    if (!rc && hipHostMalloc((void**)&e->hstage, sizeof(double2) * maxel, hipHostMallocDefault) != hipSuccess)
        rc = fail(MWF_E_NOMEM, "hipHostMalloc failed");
    e->hstage_elems = maxel;
    if (rc) {
        ens_destroy(e);
        return rc;
    }
    *out = e;
    return 0;
}

extern "C" int mwf_ens_destroy(mwf_ens* e)
{
    ens_destroy(e);
    return 0;
}

extern "C" int mwf_ens_set_state(mwf_ens* e, int32_t member, const double* re, const double* im)
{
    if (!e || !re || !im || member < 0 || member >= (int)e->where.size()) return fail(MWF_E_ARG, "bad arguments");
    SizeClass& k = *e->classes[e->where[member].first];
    to_dev_layout(k.g0, k.mpt_elems, re, im, e->hstage);
    HIPCHK(hipMemcpyAsync(k.c + k.mpt_elems * e->where[member].second, e->hstage, sizeof(double2) * k.mpt_elems,
                          hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    // AB2 restarts with forward Euler for the whole ensemble (vel_precompute's nst)
    for (auto& c : e->classes) {
        c->nst = 1;
        c->u_valid = false;
    }
    e->nst_all = 1;
    return 0;
}

extern "C" int mwf_ens_get_state(mwf_ens* e, int32_t member, double* re, double* im)
{
    if (!e || !re || !im || member < 0 || member >= (int)e->where.size()) return fail(MWF_E_ARG, "bad arguments");
    SizeClass& k = *e->classes[e->where[member].first];
    HIPCHK(hipMemcpyAsync(e->hstage, k.c + k.mpt_elems * e->where[member].second, sizeof(double2) * k.mpt_elems,
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    from_dev_layout(k.g0, e->hstage, re, im);
    return 0;
}

static int ens_energies(mwf_ens* e)
{
    e->eclass.resize(e->classes.size());
    for (size_t c = 0; c < e->classes.size(); c++) {
        SizeClass& k = *e->classes[c];
        RCCHK(k.energy(e->st));
        e->eclass[c].resize(k.nmem);
        HIPCHK(hipMemcpyAsync(e->eclass[c].data(), k.edev, sizeof(double) * k.nmem, hipMemcpyDeviceToHost, e->st));
    }
    HIPCHK(hipStreamSynchronize(e->st));
    for (size_t i = 0; i < e->where.size(); i++) e->ebuf[i] = e->eclass[e->where[i].first][e->where[i].second];
    return 0;
}

extern "C" int mwf_ens_energy(mwf_ens* e, double* energies)
{
    if (!e || !energies) return fail(MWF_E_ARG, "null argument");
    RCCHK(ens_energies(e));
    std::memcpy(energies, e->ebuf, sizeof(double) * e->where.size());
    return 0;
}

// One tracking sample of class c on stream st: energies, optionally the
// turbulent points (vel_mpt2phys without the laminar profile into the class's
// own U and phys buffers, then k_turbcount), then k_track_push.
static int ens_sample(mwf_ens* e, size_t c, hipStream_t st)
{
    SizeClass& k = *e->classes[c];
    ClassTrack& t = e->trk[c];
    RCCHK(k.energy(st));
    if (e->turb_on) {
        hipError_t r = k.plan.xinv(k.bt, k.nmem, k.g0, st, k.c, t.Ud, nullptr, k.twMM(), k.twPd, 0);
        if (r == hipSuccess) r = k.plan.zphys(k.bt, k.nmem, k.g0, st, t.Ud, t.physd, k.twNN(), k.twQ());
        if (r != hipSuccess) return fail(MWF_E_HIP, std::string("turbulence sample: ") + hipGetErrorString(r));
        const int npts = ((k.g0.P + e->turb_xs - 1) / e->turb_xs) * ((k.g0.Q + e->turb_zs - 1) / e->turb_zs);
        hipLaunchKernelGGL(k_turbcount, dim3((npts + 255) / 256, k.nmem), dim3(256), 0, st, t.physd, k.geos,
                           e->turb_xs, e->turb_zs, e->turb_ct, t.count);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_track_push, dim3(1), dim3(256), 0, st, k.nmem, k.edev, e->turb_on ? t.count : nullptr,
                       turb_den(k.g0, e->turb_xs, e->turb_zs), e->track_thr, (long long)e->track_every, t.etrace,
                       t.ftrace, t.fill, t.total, t.relam, mwf_ens::CAP);
    HIPCHK(hipGetLastError());
    return 0;
}

// nsteps steps of every class (and a tracking sample after them when
// sample is set): forked from e->st onto the class streams, joined back.
static int ens_chunk(mwf_ens* e, int nsteps, bool sample)
{
    auto plain = [&](int, auto&& f) {
        hipError_t r = f();
        return r == hipSuccess ? 0 : fail(MWF_E_HIP, std::string("kernel launch: ") + hipGetErrorString(r));
    };
    HIPCHK(hipEventRecord(e->fork, e->st));
    for (size_t c = 0; c < e->classes.size(); c++) {
        hipStream_t cs = e->cst[c];
        HIPCHK(hipStreamWaitEvent(cs, e->fork, 0));
        for (int s = 0; s < nsteps; s++) RCCHK(e->classes[c]->step(cs, plain));
        if (sample) RCCHK(ens_sample(e, c, cs));
        HIPCHK(hipEventRecord(e->cev[c], cs));
        HIPCHK(hipStreamWaitEvent(e->st, e->cev[c], 0));
    }
    return 0;
}

static int ens_graph(mwf_ens* e, int steps, bool sample, hipGraphExec_t* exec)
{
    const auto key = std::make_pair(steps, (int)sample);
    auto it = e->graphs.find(key);
    if (it != e->graphs.end()) {
        *exec = it->second.second;
        return 0;
    }
    HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
    int rc = ens_chunk(e, steps, sample);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return fail(MWF_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipGraphExec_t x = nullptr;
    hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(MWF_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    e->graphs[key] = {g, x};
    *exec = x;
    return 0;
}

// Move the device trace buffers to the host vectors and empty them.
static int ens_drain(mwf_ens* e)
{
    if (e->track_every <= 0) return 0;
    HIPCHK(hipStreamSynchronize(e->st));
    std::vector<double> eb, fb;
    std::vector<long long> rl;
    for (size_t c = 0; c < e->classes.size(); c++) {
        SizeClass& k = *e->classes[c];
        ClassTrack& t = e->trk[c];
        int fill = 0;
        HIPCHK(hipMemcpy(&fill, t.fill, sizeof(int), hipMemcpyDeviceToHost));
        fill = std::min(fill, (int)mwf_ens::CAP);
        eb.resize((size_t)fill * k.nmem);
        rl.resize(k.nmem);
        if (fill) HIPCHK(hipMemcpy(eb.data(), t.etrace, sizeof(double) * eb.size(), hipMemcpyDeviceToHost));
        if (fill && e->turb_on) {
            fb.resize(eb.size());
            HIPCHK(hipMemcpy(fb.data(), t.ftrace, sizeof(double) * fb.size(), hipMemcpyDeviceToHost));
        }
        HIPCHK(hipMemcpy(rl.data(), t.relam, sizeof(long long) * k.nmem, hipMemcpyDeviceToHost));
        const int zero = 0;
        HIPCHK(hipMemcpy(t.fill, &zero, sizeof(int), hipMemcpyHostToDevice));
        for (size_t i = 0; i < e->where.size(); i++) {
            if (e->where[i].first != (int)c) continue;
            const int j = e->where[i].second;
            for (int r = 0; r < fill; r++) {
                e->trace[i].push_back(eb[(size_t)r * k.nmem + j]);
                if (e->turb_on) e->ftrace[i].push_back(fb[(size_t)r * k.nmem + j]);
            }
            e->relam[i] = rl[j];
        }
    }
    e->pending = 0;
    return 0;
}

// nsteps of every member; steps after the first are replayed from graphs.
// With tracking on, a chunk runs up to the next multiple of the tracking
// interval and ends with a sample, all inside the graph.
extern "C" int mwf_ens_step(mwf_ens* e, int32_t nsteps)
{
    if (!e) return fail(MWF_E_ARG, "null handle");
    if (nsteps < 0) return fail(MWF_E_ARG, "nsteps < 0");
    const int every = e->track_every;
    int done = 0;
    while (done < nsteps) {
        int chunk;
        bool sample = false;
        if (every > 0) {
            chunk = std::min(every - (int)(e->step_count % every), nsteps - done);
            sample = (e->step_count + chunk) % every == 0;
        } else {
            chunk = std::min(50, nsteps - done);
        }
        if (sample && e->pending + 1 > mwf_ens::CAP) RCCHK(ens_drain(e));
        if (e->nst_all) { // first step (forward Euler) eagerly, outside the graphs
            RCCHK(ens_chunk(e, 1, sample && chunk == 1));
            e->nst_all = 0;
            if (sample && chunk == 1) e->pending++;
            done++;
            e->step_count++;
            continue;
        }
        hipGraphExec_t x = nullptr;
        RCCHK(ens_graph(e, chunk, sample, &x));
        HIPCHK(hipGraphLaunch(x, e->st));
        if (sample) e->pending++;
        done += chunk;
        e->step_count += chunk;
    }
    return 0;
}

extern "C" int mwf_ens_sync(mwf_ens* e)
{
    if (!e) return fail(MWF_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

// (Re)allocate and reset every class's tracking buffers for the current
// settings; the first sample (the state as it is now) is taken at once.
static int ens_track_reset(mwf_ens* e)
{
    HIPCHK(hipStreamSynchronize(e->st));
    ens_drop_graphs(e); // the graphs hold the old sample kernels and buffers
    for (auto& v : e->trace) v.clear();
    for (auto& v : e->ftrace) v.clear();
    std::fill(e->relam.begin(), e->relam.end(), (int64_t)-1);
    e->pending = 0;
    e->step_count = 0;
    for (size_t c = 0; c < e->classes.size(); c++) {
        SizeClass& k = *e->classes[c];
        ClassTrack& t = e->trk[c];
        t.release();
        if (e->track_every <= 0) continue;
        const size_t n = k.nmem;
        HIPCHK(hipMalloc((void**)&t.etrace, sizeof(double) * n * mwf_ens::CAP));
        HIPCHK(hipMalloc((void**)&t.ftrace, sizeof(double) * n * mwf_ens::CAP));
        HIPCHK(hipMalloc((void**)&t.fill, sizeof(int)));
        HIPCHK(hipMalloc((void**)&t.total, sizeof(long long)));
        HIPCHK(hipMalloc((void**)&t.relam, sizeof(long long) * n));
        HIPCHK(hipMalloc((void**)&t.count, sizeof(unsigned long long) * n));
        HIPCHK(hipMemset(t.fill, 0, sizeof(int)));
        HIPCHK(hipMemset(t.total, 0, sizeof(long long)));
        HIPCHK(hipMemset(t.count, 0, sizeof(unsigned long long) * n));
        std::vector<long long> neg(n, -1);
        HIPCHK(hipMemcpy(t.relam, neg.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
        if (e->turb_on) {
            const size_t rsz = k.prec == MWF_MIXED ? sizeof(float2) : sizeof(double2);
            HIPCHK(hipMalloc(&t.Ud, rsz * (size_t)k.g0.P * 2 * k.g0.KK * k.g0.NN * n));
            HIPCHK(hipMalloc((void**)&t.physd, sizeof(double) * (size_t)k.g0.KK * k.g0.Q * k.g0.P * n));
        }
    }
    if (e->track_every <= 0) return 0;
    // the sample of the current state (step 0 of the tracked run)
    HIPCHK(hipEventRecord(e->fork, e->st));
    for (size_t c = 0; c < e->classes.size(); c++) {
        HIPCHK(hipStreamWaitEvent(e->cst[c], e->fork, 0));
        RCCHK(ens_sample(e, c, e->cst[c]));
        HIPCHK(hipEventRecord(e->cev[c], e->cst[c]));
        HIPCHK(hipStreamWaitEvent(e->st, e->cev[c], 0));
    }
    e->pending = 1;
    return 0;
}

extern "C" int mwf_ens_track(mwf_ens* e, int32_t every, double threshold)
{
    if (!e || every < 0) return fail(MWF_E_ARG, "bad arguments");
    e->track_every = every;
    e->track_thr = threshold;
    if (every == 0) e->turb_on = 0;
    return ens_track_reset(e);
}

extern "C" int mwf_ens_track_turb(mwf_ens* e, int32_t xs, int32_t zs, double ct)
{
    if (!e || xs < 1 || zs < 1) return fail(MWF_E_ARG, "bad arguments");
    if (e->track_every <= 0) return fail(MWF_E_STATE, "enable tracking with mwf_ens_track first");
    for (auto& k : e->classes)
        if (xs > k->g0.P || zs > k->g0.Q) return fail(MWF_E_ARG, "strides exceed a member's grid");
    e->turb_on = 1;
    e->turb_xs = xs;
    e->turb_zs = zs;
    e->turb_ct = ct;
    return ens_track_reset(e);
}

extern "C" int mwf_ens_trace(mwf_ens* e, int32_t member, double* energies, int32_t cap, int32_t* count,
                             int64_t* relam_step)
{
    if (!e || !count || !relam_step || member < 0 || member >= (int)e->where.size()) return fail(MWF_E_ARG, "bad arguments");
    RCCHK(ens_drain(e));
    const auto& t = e->trace[member];
    const int n = (int)std::min<size_t>(t.size(), cap > 0 ? (size_t)cap : 0);
    if (energies && n > 0) std::memcpy(energies, t.data(), sizeof(double) * n);
    *count = (int32_t)t.size();
    *relam_step = e->relam[member];
    return 0;
}

extern "C" int mwf_ens_trace_turb(mwf_ens* e, int32_t member, double* fractions, int32_t cap, int32_t* count)
{
    if (!e || !count || member < 0 || member >= (int)e->where.size()) return fail(MWF_E_ARG, "bad arguments");
    RCCHK(ens_drain(e));
    const auto& t = e->ftrace[member];
    const int n = (int)std::min<size_t>(t.size(), cap > 0 ? (size_t)cap : 0);
    if (fractions && n > 0) std::memcpy(fractions, t.data(), sizeof(double) * n);
    *count = (int32_t)t.size();
    return 0;
}

extern "C" int mwf_ens_classes(mwf_ens* e, int32_t* nclasses)
{
    if (!e || !nclasses) return fail(MWF_E_ARG, "null argument");
    *nclasses = (int32_t)e->classes.size();
    return 0;
}
