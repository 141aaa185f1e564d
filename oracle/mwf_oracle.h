/*
 * mwf_oracle.h -- CPU restatement of the Model Waleffe Flow per-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library in
 * mwf_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it, and only to check or to time the CPU path; the product
 * library never links or calls it.
 *
 * Array layouts are those of the reference (Fortran column-major, first index
 * fastest), so a state can be handed to the oracle and to the HIP library
 * unchanged:
 *   mpt   re/im[k + K*(m + M*n)]        k < K = 2*K0-1, m < M = 2*MM-1, n < NN
 *   spec  re/im[c + KK*(m + M*n)]       c < KK = 3*K0-1
 *   phys  p[c + KK*(z + Q*x)]           z < Q = 3*NN, x < P = 3*MM
 * (program/variables.f90:8-20).
 */
#ifndef MWF_ORACLE_H
#define MWF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int k0;        /* wall-normal modes per velocity component (4 = M4, 6 = M6) */
    int mm, nn;    /* i_MM, i_NN */
    int dragall;   /* s_dragall */
    double alpha, gamma, re, dt, drag, vdrag, hypo, theta; /* theta in radians */
} orc_params;

typedef struct orc orc;

orc *orc_create(const orc_params *p);
void orc_destroy(orc *o);
void orc_set_threads(int n);

void orc_set_state(orc *o, const double *re, const double *im);
void orc_put_state(orc *o, const double *re, const double *im);
void orc_get_state(const orc *o, double *re, double *im);
void orc_step(orc *o, int nsteps);
double orc_energy(orc *o);
void orc_nonlinear(orc *o, double *nlre, double *nlim);
void orc_mpt2phys(orc *o, const double *re, const double *im, double *phys);

void orc_lhsrhs(const orc *o, double *lhs, double *rhs);
void orc_mpt2spec(const orc *o, const double *mre, const double *mim,
                  double *sre, double *sim);
void orc_addlam(const orc *o, double *sre);
void orc_spec2mpt(const orc *o, const double *sre, const double *sim,
                  double *mre, double *mim);
void orc_spec_grad(const orc *o, const double *pre, const double *pim,
                   double *uxre, double *uxim, double *uzre, double *uzim);
void orc_spec2phys(orc *o, const double *sre, const double *sim, double *phys);
void orc_phys2spec(orc *o, const double *phys, double *sre, double *sim);

/* Galerkin closures (program/modes.M4.f90, program/modes.M6.f90), any K0. */
void orc_udotgradu(int k0, double beta, const double *v, const double *vx,
                   const double *vz, double *ans);
void orc_nluw(int k0, const double *u, const double *w, double *ans);
void orc_nlvv(int k0, const double *v, double *ans);
void orc_nluv(int k0, const double *u, const double *v, double *ans);
void orc_nluyv(int k0, double beta, const double *u, const double *v, double *ans);
void orc_nlvvy(int k0, double beta, const double *v, double *ans);
double orc_epos(int k0, const double *u);

/* complex DFT of any length, unnormalised, sign +1 or -1 (FFTW convention) */
void orc_dft(int n, int sign, const double *in_reim, double *out_reim);

#ifdef __cplusplus
}
#endif
#endif
