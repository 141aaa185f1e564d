/*
 * mwf_oracle.c -- CPU restatement of the Model Waleffe Flow per-step path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (see the header).
 *
 * Each routine below restates one reference routine, cited as file:line
 * relative to the reference root, in the same loop order and with the same
 * overwrite semantics.  Arithmetic is IEEE double, built with
 * -ffp-contract=off so no fused multiply-adds are formed.
 *
 * FFTW 3.3.8 (a third-party dependency of the reference, not present here) is
 * restated from its published definition ("What FFTW Really Computes",
 * FFTW 3 manual section 4.8): the transforms are unnormalised,
 *   sign -1: Y_k = sum_j X_j exp(-2 pi i j k / n),  sign +1: exp(+2 pi i j k / n),
 * c2r of length n reads the half spectrum X_0..X_{n/2}, extends it
 * Hermitian and ignores the imaginary parts of X_0 and X_{n/2}; r2c returns
 * Y_0..Y_{n/2} of the real input.  The DFT itself is a mixed-radix
 * Cooley-Tukey recursion with twiddles rounded from long double.
 *
 * Parity status of each function is pinned by tests/:
 *   Galerkin closures  -- compiled reference (M4) and exact y-quadrature (M4, M6)
 *   mpt2spec, spec2mpt, spec_grad, LHSRHS -- compiled reference
 *   DFT, spec2phys, phys2spec -- brute-force DFT, numpy.fft, round trips
 *   energy -- spectral Parseval identity
 *   vel_TS -- composition of the above; laminar fixed point; mean-mode decay
 *            closed form.  The chaotic multi-step trajectory of the reference
 *            itself is not pinned (the reference step needs FFTW to link).
 */
#include "mwf_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------------ */
/* DFT                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct fftplan {
    int n;
    cplx *tw; /* exp(+2 pi i j / n), j < n */
} fftplan;

static void plan_init(fftplan *pl, int n)
{
    pl->n = n;
    pl->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    const long double twopi = 6.283185307179586476925286766559L;
    for (int j = 0; j < n; j++) {
        long double a = twopi * (long double)j / (long double)n;
        pl->tw[j] = (double)cosl(a) + I * (double)sinl(a);
    }
}

static void plan_free(fftplan *pl) { free(pl->tw); pl->tw = NULL; }

static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    if (n % 7 == 0) return 7;
    return n;
}

static inline cplx twid(const fftplan *pl, long e, int sign)
{
    cplx w = pl->tw[e % pl->n];
    return sign > 0 ? w : conj(w);
}

/* out[k] = sum_j in[j*is] W^{sign*j*k}, W = exp(2 pi i / n); n divides pl->n */
static void fft_rec(const fftplan *pl, int n, int sign, const cplx *in, int is,
                    cplx *out, cplx *scratch)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int step = pl->n / n;
    int p = smallest_factor(n);
    if (p == n && n > 7) { /* prime length: direct sum */
        for (int k = 0; k < n; k++) {
            cplx acc = 0;
            for (int j = 0; j < n; j++)
                acc += in[(long)j * is] * twid(pl, (long)j * k % n * step, sign);
            scratch[k] = acc;
        }
        memcpy(out, scratch, sizeof(cplx) * (size_t)n);
        return;
    }
    const int m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(pl, m, sign, in + (long)r * is, is * p, out + (long)r * m, scratch);
    for (int k = 0; k < m; k++) {
        cplx t[7];
        for (int r = 0; r < p; r++)
            t[r] = out[r * m + k] * twid(pl, (long)r * k * step, sign);
        for (int s = 0; s < p; s++) {
            cplx acc = t[0];
            for (int r = 1; r < p; r++)
                acc += t[r] * twid(pl, (long)((r * s) % p) * (pl->n / p), sign);
            scratch[k + m * s] = acc;
        }
    }
    memcpy(out, scratch, sizeof(cplx) * (size_t)n);
}

static void fft_exec(const fftplan *pl, int sign, const cplx *in, cplx *out, cplx *scratch)
{
    fft_rec(pl, pl->n, sign, in, 1, out, scratch);
}

void orc_dft(int n, int sign, const double *in_reim, double *out_reim)
{
    fftplan pl;
    plan_init(&pl, n);
    cplx *in = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *out = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *sc = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; j++) in[j] = in_reim[2 * j] + I * in_reim[2 * j + 1];
    fft_exec(&pl, sign, in, out, sc);
    for (int j = 0; j < n; j++) {
        out_reim[2 * j] = creal(out[j]);
        out_reim[2 * j + 1] = cimag(out[j]);
    }
    free(in); free(out); free(sc);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------ */
/* Galerkin closures: projections of products of the wall-normal bases.     */
/* u/w family h_k (k = 0..K0-1): 1, sin(b y), cos(2 b y), sin(3 b y), ...   */
/* v family   g_k (k = 1..K0-1): cos(b y), sin(2 b y), cos(3 b y), ...      */
/* (program/modes.M4.f90:83-107, program/modes.M6.f90:81-111).              */
/* Projection onto h_k / g_k is the coefficient of that function in the     */
/* product-to-sum expansion, truncated at K0-1, which is what the closed    */
/* forms nluw, nlvv, nluv, nluyv, nlvvy in modes.M4.f90:38-81 and           */
/* modes.M6.f90:22-79 tabulate.  beta factors enter through d/dy.           */
/* ------------------------------------------------------------------------ */

enum { TC = 0, TS = 1 };
#define KMAX 8

typedef struct gterm { signed char i, j, l; double c; } gterm;
typedef struct gtab {
    int k0, ready;
    int n_uw, n_vv, n_uv, n_uyv, n_vvy;
    gterm uw[KMAX * KMAX * KMAX], vv[KMAX * KMAX * KMAX], uv[KMAX * KMAX * KMAX];
    gterm uyv[KMAX * KMAX * KMAX], vvy[KMAX * KMAX * KMAX]; /* coefficient / beta */
} gtab;

static gtab g_tabs[KMAX + 1];

static int utype(int k) { return (k % 2 == 0) ? TC : TS; }
static int vtype(int k) { return (k % 2 == 1) ? TC : TS; }

/* product (ta,fa)*(tb,fb) as up to two terms (type,freq,coef), freq >= 0 */
static int prod_terms(int ta, int fa, int tb, int fb, int *t, int *f, double *c)
{
    int n = 0;
    if (ta == TC && tb == TC) {
        t[n] = TC; f[n] = fa + fb; c[n++] = 0.5;
        t[n] = TC; f[n] = abs(fa - fb); c[n++] = 0.5;
    } else if (ta == TS && tb == TS) {
        t[n] = TC; f[n] = abs(fa - fb); c[n++] = 0.5;
        t[n] = TC; f[n] = fa + fb; c[n++] = -0.5;
    } else {
        int fs = (ta == TS) ? fa : fb; /* frequency of the sine factor */
        int fc = (ta == TS) ? fb : fa;
        t[n] = TS; f[n] = fs + fc; c[n++] = 0.5;
        int d = fs - fc; /* sin(a)cos(b) = (sin(a+b) + sin(a-b))/2 */
        if (d != 0) { t[n] = TS; f[n] = abs(d); c[n++] = d > 0 ? 0.5 : -0.5; }
    }
    return n;
}

/* coefficient of family function (fam 0 = u/w, 1 = v) index i in (ta,fa)*(tb,fb) */
static double proj(int fam, int i, int ta, int fa, int tb, int fb)
{
    int t[2], f[2];
    double c[2], s = 0.0;
    int n = prod_terms(ta, fa, tb, fb, t, f, c);
    int ti = fam == 0 ? utype(i) : vtype(i);
    for (int q = 0; q < n; q++)
        if (f[q] == i && t[q] == ti) s += c[q];
    return s;
}

static void gtab_build(int k0)
{
    gtab *g = &g_tabs[k0];
    if (g->ready) return;
    g->k0 = k0;
    g->n_uw = g->n_vv = g->n_uv = g->n_uyv = g->n_vvy = 0;
    for (int i = 0; i < k0; i++)
        for (int j = 0; j < k0; j++)
            for (int l = 0; l < k0; l++) {
                double c = proj(0, i, utype(j), j, utype(l), l);
                if (c != 0.0) g->uw[g->n_uw++] = (gterm){(signed char)i, (signed char)j, (signed char)l, c};
            }
    for (int i = 0; i < k0; i++)
        for (int j = 1; j < k0; j++)
            for (int l = 1; l < k0; l++) {
                double c = proj(0, i, vtype(j), j, vtype(l), l);
                if (c != 0.0) g->vv[g->n_vv++] = (gterm){(signed char)i, (signed char)j, (signed char)l, c};
            }
    for (int i = 1; i < k0; i++)
        for (int j = 0; j < k0; j++)
            for (int l = 1; l < k0; l++) {
                double c = proj(1, i, utype(j), j, vtype(l), l);
                if (c != 0.0) g->uv[g->n_uv++] = (gterm){(signed char)i, (signed char)j, (signed char)l, c};
            }
    /* v_l g_l * d/dy(u_j h_j):  d/dy cos(f) = -f sin(f), d/dy sin(f) = f cos(f) */
    for (int i = 0; i < k0; i++)
        for (int j = 1; j < k0; j++)
            for (int l = 1; l < k0; l++) {
                int td = utype(j) == TC ? TS : TC;
                double dsg = utype(j) == TC ? -(double)j : (double)j;
                double c = dsg * proj(0, i, vtype(l), l, td, j);
                if (c != 0.0) g->uyv[g->n_uyv++] = (gterm){(signed char)i, (signed char)j, (signed char)l, c};
            }
    /* v_j g_j * d/dy(v_l g_l) */
    for (int i = 1; i < k0; i++)
        for (int j = 1; j < k0; j++)
            for (int l = 1; l < k0; l++) {
                int td = vtype(l) == TC ? TS : TC;
                double dsg = vtype(l) == TC ? -(double)l : (double)l;
                double c = dsg * proj(1, i, vtype(j), j, td, l);
                if (c != 0.0) g->vvy[g->n_vvy++] = (gterm){(signed char)i, (signed char)j, (signed char)l, c};
            }
    g->ready = 1;
}

static const gtab *gt(int k0)
{
    if (k0 < 2 || k0 > KMAX) return NULL;
    if (!g_tabs[k0].ready) gtab_build(k0);
    return &g_tabs[k0];
}

/* nluw(u,w): modes.M4.f90:38-46, modes.M6.f90:22-31.  u,w,ans: K0 entries */
void orc_nluw(int k0, const double *u, const double *w, double *ans)
{
    const gtab *g = gt(k0);
    for (int i = 0; i < k0; i++) ans[i] = 0.0;
    for (int q = 0; q < g->n_uw; q++) ans[g->uw[q].i] += g->uw[q].c * (u[g->uw[q].j] * w[g->uw[q].l]);
}

/* nlvv(v): modes.M4.f90:74-81, modes.M6.f90:33-43.  v: K0-1 entries (modes 1..), ans: K0 */
void orc_nlvv(int k0, const double *v, double *ans)
{
    const gtab *g = gt(k0);
    for (int i = 0; i < k0; i++) ans[i] = 0.0;
    for (int q = 0; q < g->n_vv; q++)
        ans[g->vv[q].i] += g->vv[q].c * (v[g->vv[q].j - 1] * v[g->vv[q].l - 1]);
}

/* nluv(u,v): modes.M4.f90:58-64, modes.M6.f90:46-55.  ans: K0-1 entries (modes 1..) */
void orc_nluv(int k0, const double *u, const double *v, double *ans)
{
    const gtab *g = gt(k0);
    for (int i = 0; i < k0 - 1; i++) ans[i] = 0.0;
    for (int q = 0; q < g->n_uv; q++)
        ans[g->uv[q].i - 1] += g->uv[q].c * (u[g->uv[q].j] * v[g->uv[q].l - 1]);
}

/* nluyv(u,v) = P_u(v du/dy): modes.M4.f90:48-56, modes.M6.f90:57-67 */
void orc_nluyv(int k0, double beta, const double *u, const double *v, double *ans)
{
    const gtab *g = gt(k0);
    for (int i = 0; i < k0; i++) ans[i] = 0.0;
    for (int q = 0; q < g->n_uyv; q++)
        ans[g->uyv[q].i] += (beta * g->uyv[q].c) * (u[g->uyv[q].j] * v[g->uyv[q].l - 1]);
}

/* nlvvy(v) = P_v(v dv/dy): modes.M4.f90:66-72, modes.M6.f90:70-79 */
void orc_nlvvy(int k0, double beta, const double *v, double *ans)
{
    const gtab *g = gt(k0);
    for (int i = 0; i < k0 - 1; i++) ans[i] = 0.0;
    for (int q = 0; q < g->n_vvy; q++)
        ans[g->vvy[q].i - 1] += (beta * g->vvy[q].c) * (v[g->vvy[q].j - 1] * v[g->vvy[q].l - 1]);
}

/* udotgradu: the composed form of modes.M6.f90:14-20 (= modes.M4.f90:30-36,
 * which modes.M4.f90:14-28 expands by hand).  v, vx, vz, ans: KK entries
 * laid out u(0:K0-1) | v(1:K0-1) | w(0:K0-1). */
void orc_udotgradu(int k0, double beta, const double *v, const double *vx,
                   const double *vz, double *ans)
{
    const int K0 = k0, ov = k0, ow = 2 * k0 - 1;
    double a[KMAX], b[KMAX], c[KMAX];
    orc_nluw(K0, v, vx, a);
    orc_nluyv(K0, beta, v, v + ov, b);
    orc_nluw(K0, v + ow, vz, c);
    for (int i = 0; i < K0; i++) ans[i] = a[i] + b[i] + c[i];
    orc_nluv(K0, v, vx + ov, a);
    orc_nlvvy(K0, beta, v + ov, b);
    orc_nluv(K0, v + ow, vz + ov, c);
    for (int i = 0; i < K0 - 1; i++) ans[ov + i] = a[i] + b[i] + c[i];
    orc_nluw(K0, v, vx + ow, a);
    orc_nluyv(K0, beta, v + ow, v + ov, b);
    orc_nluw(K0, v + ow, vz + ow, c);
    for (int i = 0; i < K0; i++) ans[ow + i] = a[i] + b[i] + c[i];
}

/* epos: program/velocity.f90:115-126 */
double orc_epos(int k0, const double *u)
{
    double t[KMAX];
    double e;
    orc_nluw(k0, u, u, t);
    e = t[0];
    orc_nlvv(k0, u + k0, t);
    e = e + t[0];
    orc_nluw(k0, u + 2 * k0 - 1, u + 2 * k0 - 1, t);
    e = e + t[0];
    return e;
}

/* ------------------------------------------------------------------------ */
/* Model state                                                              */
/* ------------------------------------------------------------------------ */

struct orc {
    orc_params p;
    int K0, K, KK, MM, M, NN, P, Q;
    double beta;
    double *am1, *am2, *an1, *an2, *ak1, *ak2;
    double *lhs, *rhs;
    double *c_re, *c_im, *nl_re, *nl_im, *onl_re, *onl_im, *c2_re, *c2_im;
    int nst;
    fftplan plP, plQ;
};

static int nthreads_req = 0;
void orc_set_threads(int n) { nthreads_req = n; }

#define MIDX(o, k, m, n) ((size_t)(k) + (size_t)(o)->K * ((size_t)(m) + (size_t)(o)->M * (size_t)(n)))
#define SIDX(o, c, m, n) ((size_t)(c) + (size_t)(o)->KK * ((size_t)(m) + (size_t)(o)->M * (size_t)(n)))

/* var_precompute derivative tables: program/variables.f90:516-531 */
static void tables(orc *o)
{
    const double a = o->p.alpha, g = o->p.gamma, b = o->beta;
    for (int m = 0; m <= o->MM - 1; m++) {
        if (m > 0) {
            o->am1[o->M - m] = -a * m;
            o->am2[o->M - m] = -a * m * a * m;
        }
        o->am1[m] = a * m;
        o->am2[m] = -a * m * a * m;
    }
    for (int n = 0; n <= o->NN - 1; n++) {
        o->an2[n] = -g * n * g * n;
        o->an1[n] = g * n;
    }
    for (int k = 0; k <= o->K0 - 1; k++) {
        o->ak1[k] = ((k % 2) ? -1.0 : 1.0) * k * b;
        o->ak2[k] = -k * k * b * b;
    }
}

/* var_LHSRHS: program/variables.f90:565-597 (loop macro parallel.h:35-40) */
static void lhsrhs(orc *o, double *l, double *r)
{
    const double dt = o->p.dt, Re = o->p.re, hypo = o->p.hypo, drag = o->p.drag,
                 vdrag = o->p.vdrag;
    for (int n = 0; n < o->NN; n++)
        for (int m = 0; m < o->M; m++)
            for (int k = 0; k <= o->K0 - 1; k++) {
                int k1 = k + 1, k2 = k + o->K0; /* 1-based slots */
                double lap = (o->am2[m] + o->ak2[k]) + o->an2[n];
                double hlap = o->am2[m] + o->an2[n];
                double ilap16 = hypo / ((o->am2[m] + o->ak2[k]) + o->an2[n]);
                double tvl;
                if (o->p.dragall || k1 == 1)
                    tvl = 1.0 - dt * ((lap / Re + ilap16) - drag);
                else
                    tvl = 1.0 - dt * (lap / Re + ilap16);
                if (m == 0 && n == 0) {
                    l[MIDX(o, k1 - 1, m, n)] = tvl;
                    l[MIDX(o, k2 - 1, m, n)] = tvl;
                    r[MIDX(o, k1 - 1, m, n)] = 1.0;
                    r[MIDX(o, k2 - 1, m, n)] = 1.0;
                    l[MIDX(o, 0, m, n)] = 1.0;
                    r[MIDX(o, 0, m, n)] = 0.0;
                } else {
                    l[MIDX(o, k1 - 1, m, n)] = tvl * hlap;
                    if (o->p.dragall)
                        l[MIDX(o, k2 - 1, m, n)] = tvl * hlap * lap + dt * (vdrag - drag) * hlap * hlap;
                    else
                        l[MIDX(o, k2 - 1, m, n)] = tvl * hlap * lap;
                    r[MIDX(o, k1 - 1, m, n)] = hlap;
                    r[MIDX(o, k2 - 1, m, n)] = hlap * lap;
                }
            }
}

orc *orc_create(const orc_params *p)
{
    if (p->k0 < 2 || p->k0 > KMAX || p->mm < 2 || p->nn < 2) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->p = *p;
    o->K0 = p->k0;
    o->K = 2 * p->k0 - 1;
    o->KK = 3 * p->k0 - 1;
    o->MM = p->mm;
    o->M = 2 * (p->mm - 1) + 1;
    o->NN = p->nn;
    o->P = 3 * p->mm;
    o->Q = 3 * p->nn;
    o->beta = 3.1415926535897931 / 2.0; /* d_beta = d_PI/2d0, parameters.f90:60 */
    o->am1 = calloc((size_t)o->M, sizeof(double));
    o->am2 = calloc((size_t)o->M, sizeof(double));
    o->an1 = calloc((size_t)o->NN, sizeof(double));
    o->an2 = calloc((size_t)o->NN, sizeof(double));
    o->ak1 = calloc((size_t)o->K0, sizeof(double));
    o->ak2 = calloc((size_t)o->K0, sizeof(double));
    size_t nm = (size_t)o->K * o->M * o->NN;
    o->lhs = calloc(nm, sizeof(double));
    o->rhs = calloc(nm, sizeof(double));
    o->c_re = calloc(nm, sizeof(double));
    o->c_im = calloc(nm, sizeof(double));
    o->nl_re = calloc(nm, sizeof(double));
    o->nl_im = calloc(nm, sizeof(double));
    o->onl_re = calloc(nm, sizeof(double));
    o->onl_im = calloc(nm, sizeof(double));
    o->c2_re = calloc(nm, sizeof(double));
    o->c2_im = calloc(nm, sizeof(double));
    tables(o);
    lhsrhs(o, o->lhs, o->rhs);
    o->nst = 1;
    plan_init(&o->plP, o->P);
    plan_init(&o->plQ, o->Q);
    gt(o->K0);
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->am1); free(o->am2); free(o->an1); free(o->an2); free(o->ak1); free(o->ak2);
    free(o->lhs); free(o->rhs); free(o->c_re); free(o->c_im); free(o->nl_re); free(o->nl_im);
    free(o->onl_re); free(o->onl_im); free(o->c2_re); free(o->c2_im);
    plan_free(&o->plP);
    plan_free(&o->plQ);
    free(o);
}

void orc_set_state(orc *o, const double *re, const double *im)
{
    size_t nm = (size_t)o->K * o->M * o->NN;
    memcpy(o->c_re, re, nm * sizeof(double));
    memcpy(o->c_im, im, nm * sizeof(double));
    o->nst = 1; /* vel_precompute: program/velocity.f90:48 */
}

/* vel_imposesym (program/velocity.f90:68-74) changes vel_c between steps and
 * leaves vel_onl and nst as they are */
void orc_put_state(orc *o, const double *re, const double *im)
{
    size_t nm = (size_t)o->K * o->M * o->NN;
    memcpy(o->c_re, re, nm * sizeof(double));
    memcpy(o->c_im, im, nm * sizeof(double));
}

void orc_get_state(const orc *o, double *re, double *im)
{
    size_t nm = (size_t)o->K * o->M * o->NN;
    memcpy(re, o->c_re, nm * sizeof(double));
    memcpy(im, o->c_im, nm * sizeof(double));
}

void orc_lhsrhs(const orc *o, double *l, double *r)
{
    size_t nm = (size_t)o->K * o->M * o->NN;
    memcpy(l, o->lhs, nm * sizeof(double));
    memcpy(r, o->rhs, nm * sizeof(double));
}

/* var_mpt2spec: program/variables.f90:390-434 */
void orc_mpt2spec(const orc *o, const double *mre, const double *mim, double *sre, double *sim)
{
    const int K0 = o->K0;
    for (int n = 0; n < o->NN; n++)
        for (int m = 0; m < o->M; m++)
            for (int k = 0; k <= K0 - 1; k++) {
                int k1 = k, k2 = k + K0 - 1, k3 = k + 2 * K0 - 1; /* 0-based */
                const double an1 = o->an1[n], am1 = o->am1[m], ak1 = o->ak1[k];
                sre[SIDX(o, k1, m, n)] = -(-(mim[MIDX(o, k1, m, n)] * an1) + mim[MIDX(o, k2, m, n)] * ak1 * am1);
                sim[SIDX(o, k1, m, n)] = (-(mre[MIDX(o, k1, m, n)] * an1) + mre[MIDX(o, k2, m, n)] * ak1 * am1);
                sre[SIDX(o, k2, m, n)] = -(mre[MIDX(o, k2, m, n)] * (o->am2[m] + o->an2[n]));
                sim[SIDX(o, k2, m, n)] = -(mim[MIDX(o, k2, m, n)] * (o->am2[m] + o->an2[n]));
                sre[SIDX(o, k3, m, n)] = -(mim[MIDX(o, k1, m, n)] * am1 + mim[MIDX(o, k2, m, n)] * ak1 * an1);
                sim[SIDX(o, k3, m, n)] = (mre[MIDX(o, k1, m, n)] * am1 + mre[MIDX(o, k2, m, n)] * ak1 * an1);
            }
    /* no flux, mean modes */
    sre[SIDX(o, 0, 0, 0)] = 0.0;
    sim[SIDX(o, 0, 0, 0)] = 0.0;
    sre[SIDX(o, 2 * K0 - 1, 0, 0)] = 0.0;
    sim[SIDX(o, 2 * K0 - 1, 0, 0)] = 0.0;
    for (int k = 1; k <= K0 - 1; k++) {
        int k1 = k, k2 = k + K0 - 1, k3 = k + 2 * K0 - 1;
        sre[SIDX(o, k1, 0, 0)] = mre[MIDX(o, k1, 0, 0)];
        sim[SIDX(o, k1, 0, 0)] = mim[MIDX(o, k1, 0, 0)];
        sre[SIDX(o, k3, 0, 0)] = mre[MIDX(o, k2, 0, 0)];
        sim[SIDX(o, k3, 0, 0)] = mim[MIDX(o, k2, 0, 0)];
    }
}

/* vel_addlam: program/velocity.f90:76-82 */
void orc_addlam(const orc *o, double *sre)
{
    sre[SIDX(o, 1, 0, 0)] = sre[SIDX(o, 1, 0, 0)] + cos(o->p.theta);
    sre[SIDX(o, 2 * o->K0, 0, 0)] = sre[SIDX(o, 2 * o->K0, 0, 0)] + sin(o->p.theta);
}

/* var_spec_grad: program/variables.f90:600-612 */
void orc_spec_grad(const orc *o, const double *pre, const double *pim,
                   double *uxre, double *uxim, double *uzre, double *uzim)
{
    for (int n = 0; n < o->NN; n++)
        for (int m = 0; m < o->M; m++)
            for (int c = 0; c < o->KK; c++) {
                size_t i = SIDX(o, c, m, n);
                uxre[i] = -(pim[i] * o->am1[m]);
                uxim[i] = pre[i] * o->am1[m];
                uzre[i] = -(pim[i] * o->an1[n]);
                uzim[i] = pre[i] * o->an1[n];
            }
}

/* var_spec2mpt: program/variables.f90:440-480 */
void orc_spec2mpt(const orc *o, const double *sre, const double *sim, double *mre, double *mim)
{
    const int K0 = o->K0;
    for (int n = 0; n < o->NN; n++)
        for (int m = 0; m < o->M; m++)
            for (int k = 0; k <= K0 - 1; k++) {
                int k1 = k, k2 = k + K0 - 1, k3 = k + 2 * K0 - 1;
                const double an1 = o->an1[n], am1 = o->am1[m], ak1 = o->ak1[k];
                const double h = o->am2[m] + o->an2[n];
                mre[MIDX(o, k2, m, n)] = -(sim[SIDX(o, k1, m, n)] * am1 * ak1 + sim[SIDX(o, k3, m, n)] * an1 * ak1)
                                         + sre[SIDX(o, k2, m, n)] * h;
                mim[MIDX(o, k2, m, n)] = (sre[SIDX(o, k1, m, n)] * am1 * ak1 + sre[SIDX(o, k3, m, n)] * an1 * ak1)
                                         + sim[SIDX(o, k2, m, n)] * h;
                mre[MIDX(o, k1, m, n)] = (sim[SIDX(o, k3, m, n)] * am1 - sim[SIDX(o, k1, m, n)] * an1);
                mim[MIDX(o, k1, m, n)] = -(sre[SIDX(o, k3, m, n)] * am1 - sre[SIDX(o, k1, m, n)] * an1);
            }
    mre[MIDX(o, 0, 0, 0)] = 0.0;
    mim[MIDX(o, 0, 0, 0)] = 0.0;
    for (int k = 1; k <= K0 - 1; k++) {
        int k1 = k, k2 = k + K0 - 1, k3 = k + 2 * K0 - 1;
        mre[MIDX(o, k1, 0, 0)] = -sre[SIDX(o, k1, 0, 0)];
        mim[MIDX(o, k1, 0, 0)] = -sim[SIDX(o, k1, 0, 0)];
        mre[MIDX(o, k2, 0, 0)] = -sre[SIDX(o, k3, 0, 0)];
        mim[MIDX(o, k2, 0, 0)] = -sim[SIDX(o, k3, 0, 0)];
    }
}

/* ------------------------------------------------------------------------ */
/* Dealiased transforms: program/transform.fftw5.f90                        */
/* Intermediate T(c, n, x) = reference Ts(KK, 0:NN-1, 0:Mp-1) after T2Ts.   */
/* ------------------------------------------------------------------------ */

static int nthr(void)
{
#ifdef _OPENMP
    return nthreads_req > 0 ? nthreads_req : omp_get_max_threads();
#else
    return 1;
#endif
}

/* x part of tra_spec2phys (transform.fftw5.f90:68-78, plan_inp2mid sign +1,
 * :41-43) followed by tra_T2Ts (:211-218). */
static void x_inverse(orc *o, const double *sre, const double *sim, cplx *T)
{
    const int P = o->P, M = o->M, MM = o->MM, KK = o->KK, NN = o->NN;
#pragma omp parallel num_threads(nthr())
    {
        cplx *in = malloc(sizeof(cplx) * (size_t)P), *out = malloc(sizeof(cplx) * (size_t)P),
             *sc = malloc(sizeof(cplx) * (size_t)P);
#pragma omp for schedule(static)
        for (int n = 0; n < NN; n++)
            for (int c = 0; c < KK; c++) {
                for (int x = 0; x < P; x++) in[x] = 0;
                for (int m = -(M - MM); m <= MM - 1; m++) {
                    int mm = ((m % P) + P) % P, ms = ((m % M) + M) % M;
                    in[mm] = sre[SIDX(o, c, ms, n)] + I * sim[SIDX(o, c, ms, n)];
                }
                fft_exec(&o->plP, +1, in, out, sc);
                for (int x = 0; x < P; x++) T[(size_t)c + KK * ((size_t)n + (size_t)NN * x)] = out[x];
            }
        free(in); free(out); free(sc);
    }
}

/* z c2r of one x-row of T for all KK components (transform.fftw5.f90:194-199,
 * plan_mid2phy :54-55): phys row p[c + KK*z]. */
static void z_c2r_row(orc *o, const cplx *Trow, double *prow, cplx *in, cplx *out, cplx *sc)
{
    const int Q = o->Q, NN = o->NN, KK = o->KK, Na = Q / 2;
    for (int c = 0; c < KK; c++) {
        /* state_mid(:, NN:) = 0, state_mid(:, 0:NN-1) = Ts */
        for (int z = 0; z < Q; z++) in[z] = 0;
        for (int n = 0; n <= Na; n++) {
            cplx h = n < NN ? Trow[c + (size_t)KK * n] : 0;
            if (n == 0 || (2 * n == Q)) h = creal(h);
            in[n] = h;
            if (n > 0 && 2 * n != Q) in[Q - n] = conj(h);
        }
        fft_exec(&o->plQ, +1, in, out, sc);
        for (int z = 0; z < Q; z++) prow[c + (size_t)KK * z] = creal(out[z]);
    }
}

/* z r2c of one phys row after scaling (transform.fftw5.f90:243-247,
 * plan_phy2mid :56-57) into a T row (n < NN kept). */
static void z_r2c_row(orc *o, const double *prow, cplx *Trow, cplx *in, cplx *out, cplx *sc)
{
    const int Q = o->Q, NN = o->NN, KK = o->KK;
    const double scale_ = 1.0 / (double)(o->P * o->Q);
    for (int c = 0; c < KK; c++) {
        for (int z = 0; z < Q; z++) in[z] = scale_ * prow[c + (size_t)KK * z];
        fft_exec(&o->plQ, -1, in, out, sc);
        for (int n = 0; n < NN; n++) Trow[c + (size_t)KK * n] = out[n];
    }
}

/* x part of tra_phys2spec after tra_Ts2T (transform.fftw5.f90:249-258,
 * plan_mid2inp sign -1 :45-47), truncating to |m| <= MM-1. */
static void x_forward(orc *o, const cplx *T, double *sre, double *sim)
{
    const int P = o->P, M = o->M, MM = o->MM, KK = o->KK, NN = o->NN;
#pragma omp parallel num_threads(nthr())
    {
        cplx *in = malloc(sizeof(cplx) * (size_t)P), *out = malloc(sizeof(cplx) * (size_t)P),
             *sc = malloc(sizeof(cplx) * (size_t)P);
#pragma omp for schedule(static)
        for (int n = 0; n < NN; n++)
            for (int c = 0; c < KK; c++) {
                for (int x = 0; x < P; x++) in[x] = T[(size_t)c + KK * ((size_t)n + (size_t)NN * x)];
                fft_exec(&o->plP, -1, in, out, sc);
                for (int m = -(M - MM); m <= MM - 1; m++) {
                    int mm = ((m % P) + P) % P, ms = ((m % M) + M) % M;
                    sre[SIDX(o, c, ms, n)] = creal(out[mm]);
                    sim[SIDX(o, c, ms, n)] = cimag(out[mm]);
                }
            }
        free(in); free(out); free(sc);
    }
}

/* tra_spec2phys: transform.fftw5.f90:64-87 */
void orc_spec2phys(orc *o, const double *sre, const double *sim, double *phys)
{
    const int P = o->P, Q = o->Q, KK = o->KK, NN = o->NN;
    cplx *T = malloc(sizeof(cplx) * (size_t)KK * NN * P);
    x_inverse(o, sre, sim, T);
#pragma omp parallel num_threads(nthr())
    {
        cplx *in = malloc(sizeof(cplx) * (size_t)Q), *out = malloc(sizeof(cplx) * (size_t)Q),
             *sc = malloc(sizeof(cplx) * (size_t)Q);
#pragma omp for schedule(static)
        for (int x = 0; x < P; x++)
            z_c2r_row(o, T + (size_t)KK * NN * x, phys + (size_t)KK * Q * x, in, out, sc);
        free(in); free(out); free(sc);
    }
    free(T);
}

/* tra_phys2spec: transform.fftw5.f90:122-145 */
void orc_phys2spec(orc *o, const double *phys, double *sre, double *sim)
{
    const int P = o->P, Q = o->Q, KK = o->KK, NN = o->NN;
    cplx *T = malloc(sizeof(cplx) * (size_t)KK * NN * P);
#pragma omp parallel num_threads(nthr())
    {
        cplx *in = malloc(sizeof(cplx) * (size_t)Q), *out = malloc(sizeof(cplx) * (size_t)Q),
             *sc = malloc(sizeof(cplx) * (size_t)Q);
#pragma omp for schedule(static)
        for (int x = 0; x < P; x++)
            z_r2c_row(o, phys + (size_t)KK * Q * x, T + (size_t)KK * NN * x, in, out, sc);
        free(in); free(out); free(sc);
    }
    x_forward(o, T, sre, sim);
    free(T);
}

/* vel_nonlinear: program/velocity.f90:92-113.  The z transforms, the
 * pointwise udotgradu (:106-108) and the forward z transform are done one
 * x-row at a time; every row is independent, so this is the same arithmetic
 * as the reference's whole-array sequence. */
static void nonlinear(orc *o, double *nlre, double *nlim)
{
    const int P = o->P, Q = o->Q, KK = o->KK, NN = o->NN, M = o->M;
    size_t ns = (size_t)KK * M * NN;
    double *ure = malloc(ns * sizeof(double)), *uim = malloc(ns * sizeof(double));
    double *xre = malloc(ns * sizeof(double)), *xim = malloc(ns * sizeof(double));
    double *zre = malloc(ns * sizeof(double)), *zim = malloc(ns * sizeof(double));
    orc_mpt2spec(o, o->c_re, o->c_im, ure, uim);
    orc_addlam(o, ure);
    orc_spec_grad(o, ure, uim, xre, xim, zre, zim);
    size_t nt = (size_t)KK * NN * P;
    cplx *Tu = malloc(sizeof(cplx) * nt), *Tx = malloc(sizeof(cplx) * nt),
         *Tz = malloc(sizeof(cplx) * nt);
    x_inverse(o, ure, uim, Tu);
    x_inverse(o, xre, xim, Tx);
    x_inverse(o, zre, zim, Tz);
#pragma omp parallel num_threads(nthr())
    {
        cplx *in = malloc(sizeof(cplx) * (size_t)Q), *out = malloc(sizeof(cplx) * (size_t)Q),
             *sc = malloc(sizeof(cplx) * (size_t)Q);
        double *p = malloc(sizeof(double) * (size_t)KK * Q), *px = malloc(sizeof(double) * (size_t)KK * Q),
               *pz = malloc(sizeof(double) * (size_t)KK * Q), *ans = malloc(sizeof(double) * (size_t)KK * Q);
#pragma omp for schedule(static)
        for (int x = 0; x < P; x++) {
            z_c2r_row(o, Tu + (size_t)KK * NN * x, p, in, out, sc);
            z_c2r_row(o, Tx + (size_t)KK * NN * x, px, in, out, sc);
            z_c2r_row(o, Tz + (size_t)KK * NN * x, pz, in, out, sc);
            for (int z = 0; z < Q; z++)
                orc_udotgradu(o->K0, o->beta, p + (size_t)KK * z, px + (size_t)KK * z,
                              pz + (size_t)KK * z, ans + (size_t)KK * z);
            /* reuse Tu's row for the forward transform result */
            z_r2c_row(o, ans, Tu + (size_t)KK * NN * x, in, out, sc);
        }
        free(in); free(out); free(sc); free(p); free(px); free(pz); free(ans);
    }
    x_forward(o, Tu, ure, uim);
    orc_spec2mpt(o, ure, uim, nlre, nlim);
    free(Tu); free(Tx); free(Tz);
    free(ure); free(uim); free(xre); free(xim); free(zre); free(zim);
}

void orc_nonlinear(orc *o, double *nlre, double *nlim) { nonlinear(o, nlre, nlim); }

/* vel_TS: program/velocity.f90:21-41 (loop order parallel.h:35-40) */
static void vel_ts(orc *o)
{
    size_t nm = (size_t)o->K * o->M * o->NN;
    nonlinear(o, o->nl_re, o->nl_im);
    if (o->nst) {
        memcpy(o->onl_re, o->nl_re, nm * sizeof(double));
        memcpy(o->onl_im, o->nl_im, nm * sizeof(double));
        o->nst = 0;
    }
    const double dt = o->p.dt;
    for (size_t i = 0; i < nm; i++) {
        double r = o->rhs[i], l = o->lhs[i];
        double cr = r * o->c_re[i];
        double ci = r * o->c_im[i];
        cr = cr + dt / 2.0 * (3.0 * o->nl_re[i] - o->onl_re[i]);
        ci = ci + dt / 2.0 * (3.0 * o->nl_im[i] - o->onl_im[i]);
        o->c2_re[i] = cr / l;
        o->c2_im[i] = ci / l;
    }
    o->c2_re[0] = 0.0;
    memcpy(o->onl_re, o->nl_re, nm * sizeof(double));
    memcpy(o->onl_im, o->nl_im, nm * sizeof(double));
    memcpy(o->c_re, o->c2_re, nm * sizeof(double));
    memcpy(o->c_im, o->c2_im, nm * sizeof(double));
}

void orc_step(orc *o, int nsteps)
{
    for (int s = 0; s < nsteps; s++) vel_ts(o);
}

/* vel_mpt2phys: program/velocity.f90:84-90 */
void orc_mpt2phys(orc *o, const double *re, const double *im, double *phys)
{
    size_t ns = (size_t)o->KK * o->M * o->NN;
    double *sre = malloc(ns * sizeof(double)), *sim = malloc(ns * sizeof(double));
    orc_mpt2spec(o, re, im, sre, sim);
    orc_spec2phys(o, sre, sim, phys);
    free(sre); free(sim);
}

/* vel_mpt2phys(vel_c) + vel_energy: program/io.f90:98-99,
 * program/velocity.f90:144-162.  The sum runs x outer, z inner, sequentially,
 * as _loop_phy_begin does (parallel.h:25-28). */
double orc_energy(orc *o)
{
    const int P = o->P, Q = o->Q, KK = o->KK;
    double *phys = malloc(sizeof(double) * (size_t)KK * Q * P);
    double *ep = malloc(sizeof(double) * (size_t)Q * P);
    orc_mpt2phys(o, o->c_re, o->c_im, phys);
#pragma omp parallel for num_threads(nthr()) schedule(static)
    for (long i = 0; i < (long)Q * P; i++) ep[i] = orc_epos(o->K0, phys + (size_t)KK * i);
    double e = 0;
    for (long i = 0; i < (long)Q * P; i++) e = e + ep[i];
    e = e / (double)(P * Q);
    free(phys);
    free(ep);
    return e;
}
