/*
 * mwf.h -- C-ABI of libmwf.so, the MI355X timestepper for Model Waleffe Flow.
 *
 * The reference exposes its per-step work as Fortran module procedures acting
 * on module-global state (module velocity, program/velocity.f90:15-17,21-162,
 * and module transform, program/transform.fftw5.f90:33-145), selected at build
 * time by the Makefile's TRANSFORM/MODES variables (Makefile:6-7,68-80).  This
 * library replaces that work on one GPU.  The Fortran velocity variant in
 * mwf_amd/fortran/ binds these entry points with ISO_C_BINDING.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative MWF_E* code;
 *    mwf_last_error() returns a message for the calling thread's last error.
 *  - State arrays use the reference layout of type(mpt) (program/variables.f90
 *    :12-15): separate Re and Im planes, column-major (K, 0:M-1, 0:NN-1),
 *    element (k, m, n) at k + K*(m + M*n), K = 2*K0-1, M = 2*MM-1.
 *  - phys arrays use type(phys) (:8-10): (KK, 0:Q-1, 0:P-1), element (c, z, x)
 *    at c + KK*(z + Q*x), KK = 3*K0-1, P = 3*MM, Q = 3*NN.
 *  - Calls on one handle come from one host thread.  GPU work is queued on the
 *    handle's own HIP stream; entry points that return host data synchronise it.
 */
#ifndef MWF_H
#define MWF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MWF_OK 0
#define MWF_E_ARG -1      /* invalid argument or unsupported size */
#define MWF_E_HIP -2      /* HIP runtime error (message has the HIP error) */
#define MWF_E_NOMEM -3    /* device or host allocation failed */
#define MWF_E_STATE -4    /* call out of order (e.g. step before set_state) */

#define MWF_FP64 0        /* every kernel in FP64 */
#define MWF_MIXED 1       /* FFTs and products FP32; state, AB2, update, energy FP64 */

/* Model parameters: the compile-time constants of program/parameters.f90
 * (:7-46) as run-time values.  theta is in radians (d_theta, :61). */
typedef struct mwf_params {
    int32_t k0;         /* i_K0: 4 (modes.M4.f90) or 6 (modes.M6.f90) */
    int32_t mm;         /* i_MM, power of two, 16..1024 */
    int32_t nn;         /* i_NN, power of two, 16..1024 */
    int32_t dragall;    /* s_dragall */
    int32_t precision;  /* MWF_FP64 or MWF_MIXED */
    int32_t reserved;   /* must be 0 */
    double alpha;       /* d_alpha = 2 pi / Lx */
    double gamma;       /* d_gamma = 2 pi / Lz */
    double re;          /* d_Re */
    double dt;          /* d_dt */
    double drag;        /* d_drag */
    double vdrag;       /* d_vdrag */
    double hypo;        /* d_HYPO */
    double theta;       /* d_theta */
} mwf_params;

typedef struct mwf_handle mwf_handle;

/* vel_precompute + var_precompute + tra_precompute (program/velocity.f90:43-50,
 * program/variables.f90:488-533, program/transform.fftw5.f90:33-59): device
 * buffers, twiddle tables and launch plans.  The state starts at zero. */
int mwf_create(const mwf_params *p, mwf_handle **out);
int mwf_destroy(mwf_handle *h);

/* Upload vel_c (what io_load_mpt fills, program/io.f90:153,166-204); marks the
 * next step as the first (nst = .true., program/velocity.f90:48). */
int mwf_set_state(mwf_handle *h, const double *re, const double *im);
/* Download vel_c (what io_save_state reads, program/io.f90:257-290). */
int mwf_get_state(mwf_handle *h, double *re, double *im);
/* Replace vel_c in place, keeping the AB2 history vel_onl and nst: what
 * vel_imposesym's symmetry projections do to the state between steps
 * (program/velocity.f90:68-74, called each step from program/main.f90:22). */
int mwf_put_state(mwf_handle *h, const double *re, const double *im);

/* nsteps x vel_TS (program/velocity.f90:21-41): nonlinear term, AB2 (forward
 * Euler on the first step) and the diagonal implicit update. */
int mwf_step(mwf_handle *h, int32_t nsteps);

/* vel_mpt2phys(vel_c) + vel_energy (program/io.f90:98-99,
 * program/velocity.f90:84-90,144-162): perturbation energy of the current
 * state, reduced on the device (spectral Parseval sum, compensated). */
int mwf_energy(mwf_handle *h, double *e);

/* No counterpart in the reference: mwf_energy for callers that cannot
 * allocate or synchronise (code under HIP graph capture, tight serving
 * loops).  The perturbation energy of the current state is written to the
 * device double e_dev, asynchronously on the handle's stream, with the
 * per-block partials in a fixed 64 KB device array of the library (one per
 * GPU, held until the process ends) instead of handle memory: no allocation
 * and no synchronisation.  The array holds 8192 partials (any single-model
 * handle fits: at most 3072); larger needs fail with MWF_E_ARG.  Calls that
 * overlap on one GPU (handles on different streams) race on the array. */
int mwf_energy_noalloc(mwf_handle *h, double *e_dev);

/* vel_nonlinear (program/velocity.f90:92-113) of the current state, without
 * stepping: the mpt-form nonlinear term vel_nl. */
int mwf_nonlinear(mwf_handle *h, double *nlre, double *nlim);

/* vel_mpt2phys (program/velocity.f90:84-90): mpt -> phys for any host mpt. */
int mwf_mpt2phys(mwf_handle *h, const double *re, const double *im, double *phys);

/* tra_spec2phys (program/transform.fftw5.f90:64-87): the 3/2-dealiased
 * inverse transform of a type(spec) (re, im planes, (KK, M, NN) column-major,
 * program/variables.f90:17-20) to a type(phys) ((KK, Q, P), :8-10): x c2c of
 * length P with sign +1, then z c2r of length Q, both unnormalised.  Runs in
 * the handle's precision. */
int mwf_spec2phys(mwf_handle *h, const double *re, const double *im, double *phys);
/* tra_phys2spec (program/transform.fftw5.f90:122-145): scale by 1/(P Q),
 * z r2c, x c2c with sign -1, truncated to |m| < MM, n < NN. */
int mwf_phys2spec(mwf_handle *h, const double *phys, double *re, double *im);

/* AB2 history vel_onl (program/velocity.f90:16,24-27,39), for a warm restart
 * that continues AB2 instead of restarting with forward Euler: the state
 * files of program/io.f90:257-313 hold vel_c only.  get returns the history
 * in the mpt layout and first_step = 1 while the next step is a forward-Euler
 * step (nst); set (after mwf_set_state) installs a saved history and clears
 * nst, so that set_state + set_onl resumes bit-identically. */
int mwf_get_onl(mwf_handle *h, double *re, double *im, int32_t *first_step);
int mwf_set_onl(mwf_handle *h, const double *re, const double *im);

/* io_save_Uspec (program/io.f90:345-395) on the device: with d = |spec| of
 * var_mpt2spec(vel_c) over all wall-normal slots, m_max[|m|] (MM entries) and
 * n_max[n] (NN entries) are the maxima of d over the other two indices. */
int mwf_spectrum_max(mwf_handle *h, double *m_max, double *n_max);

/* vel_history(vel_mpt2phys(vel_c), y, ans) (program/velocity.f90:164-180,
 * program/io.f90:472-483) on the device: ans is (4, nh) column-major, per
 * sample j the z position ji*Lz/Q (ji = (j-1)*Q/nh + 1) and velU, velV, velW
 * (modes.M4.f90:83-108, modes.M6.f90:81-111) at x index 0 and height y.
 * 1 <= nh <= Q/2.  Only the 4*nh results cross PCIe. */
int mwf_vel_history(mwf_handle *h, int32_t nh, double y, double *ans);

/* turb_sample(vel_mpt2phys(vel_c)) (program/turb.f90:37-47) on the device:
 * out is (3, SN, SM) column-major, SM = P/dm, SN = Q/dn (parameters.f90:60-61),
 * holding eposC (program/velocity.f90:128-142) at x = m*dm, z = n*dn.  The
 * caller keeps the i_MT time history (turb.f90:43, the shift of tdata). */
int mwf_turb_sample(mwf_handle *h, int32_t dm, int32_t dn, double *out);

/* Turbulent fraction of io_writeVTK_txz (program/io.f90:550-617), as
 * utils/turbfrac.f90:14-38 drives it, on the device: of the physical field
 * vel_mpt2phys(vel_c), the number of points ix = 0, xs, ... < P,
 * iz = 0, zs, ... < Q with epos > ct (program/velocity.f90:115-126), divided
 * by (N_+1)(M_+1), M_ = P/xs, N_ = Q/zs (the reference's denominator, which
 * includes the periodic copies its VTK file repeats). */
int mwf_turb_fraction(mwf_handle *h, int32_t xs, int32_t zs, double ct, double *frac);

/* var_doubleX / var_quadX (program/variables.f90:147-185, utils/doubleX.f90,
 * utils/quadX.f90) on the device: dst (same k0, mm, nn; alpha = src alpha /
 * factor, a box factor times longer in x) receives src's state with mode m
 * moved to factor*m; modes beyond MM-1 are dropped.  factor is a power of
 * two in [2, mm].  dst's next step is a first (forward-Euler) step. */
int mwf_extend_x(mwf_handle *src, int32_t factor, mwf_handle *dst);

/* utils/interpstate.f90:30-31 on the device: dst state = (1 - d) a + d b,
 * for handles of equal (k0, mm, nn) (dst may be a or b). */
int mwf_blend(mwf_handle *a, mwf_handle *b, double d, mwf_handle *dst);

/* Autotuning, in place of tra_precompute's FFTW_PATIENT planning
 * (program/transform.fftw5.f90:33-59).  Process-wide mode for handles and
 * ensembles created afterwards: 0 (default) = rule-based plans; 1 = the first
 * handle or ensemble class of each (precision, k0, mm, nn) times its
 * candidate kernel plans at creation (a few steps each, on its zero state)
 * and keeps the fastest, and later ones of that size reuse the choice, so
 * results stay reproducible within a process; 2..65 = force plan choice
 * mode - 2 where that alternative exists (for tests and plan comparisons). */
int mwf_autotune(int32_t mode);
/* The plan a handle runs: bit 0 = phase-sequential z kernel instead of the
 * three-phase one, bit 1 = separate x-inverse and x-forward kernels instead
 * of the fused x kernel, bit 2 = three-phase z kernel with two points per
 * thread (768-thread blocks) instead of three (512), bit 3 = the flux
 * (conservative) form of vel_nonlinear's products: the same projected
 * (u.grad)u, with d/dx, d/dz taken in spectral space and d/dy moved onto the
 * wall-normal bases, equal to the advective form up to rounding; bit 4 (with
 * bit 3) = its z kernel compiled for two waves per SIMD instead of four;
 * bit 5 (advective form, NN = 1024; implies bit 1 in the mixed form) = the
 * phase-sequential z kernel with the three fold phases of an x-row in three
 * blocks, each writing its own NL partial, which the x kernel sums. */
int mwf_plan_info(mwf_handle *h, int32_t *choice);

/* Wait for all queued work of the handle. */
int mwf_sync(mwf_handle *h);

/* Message of the calling thread's last error ("" if none). */
const char *mwf_last_error(void);

/* Seeded Gaussian-spectrum initial condition, host only.  The reference ships
 * no IC generator (README.md:7-13 names a randIC utility absent from the
 * tree; program/variables.f90:233-286 draw unseeded uniforms); this one
 * defines the synthetic inputs:  SplitMix64(seed) + Box-Muller; per mode
 * (n outer, m, then spec slot) u, v, w coefficients
 *   scl * exp(-(kx^2 + kz^2)/k0^2) * (N(0,1) + i N(0,1))/sqrt(2),
 * kx = alpha*m (signed), kz = gamma*n; projected per wavenumber onto the
 * divergence-free toroidal/poloidal form; zero deviation at (0,0) (laminar
 * mean); n = 0 column made Hermitian; finally scaled so that the
 * perturbation energy (vel_energy, the mean square velocity) is scl^2.
 * Writes an mpt in the layout above. */
int mwf_ic_gaussian(const mwf_params *p, uint64_t seed, double scl, double k0,
                    double *re, double *im);

/* Per-kernel timing with HIP events on the handle's stream.  kernel ids:
 * 0 = x-inverse, 1 = z fused, 2 = x-forward + update, 3 = energy,
 * 4 = fused x-forward + update + next x-inverse.  total_ms is the
 * summed event time since profiling was enabled. */
int mwf_profile_enable(mwf_handle *h, int32_t on);
int mwf_profile_read(mwf_handle *h, int32_t kernel, double *total_ms, int64_t *count);

/* Algorithmic HBM bytes one launch of a kernel id moves (inputs read once,
 * outputs written once), for roofline accounting. */
int mwf_kernel_bytes(mwf_handle *h, int32_t kernel, double *bytes);

/* ------------------------------------------------------------------------
 * Ensembles: many independent models stepped together.  The reference runs
 * one model per executable (program/main.f90); an ensemble holds n members
 * with their own parameters (Re, box, dt, basis, size, precision), groups
 * them into size classes that step with one launch per kernel, and replays
 * the step sequence from a hipGraph.  Member states use the mpt layout of
 * that member's (k0, mm, nn).  Setting any member's state restarts AB2 with
 * forward Euler for all members (nst, program/velocity.f90:24-27).
 * ------------------------------------------------------------------------ */
typedef struct mwf_ens mwf_ens;

int mwf_ens_create(const mwf_params *members, int32_t n, mwf_ens **out);
int mwf_ens_destroy(mwf_ens *e);
int mwf_ens_set_state(mwf_ens *e, int32_t member, const double *re, const double *im);
int mwf_ens_get_state(mwf_ens *e, int32_t member, double *re, double *im);
/* nsteps x vel_TS for every member (hipGraph replay after the first step). */
int mwf_ens_step(mwf_ens *e, int32_t nsteps);
/* vel_energy of every member (n values). */
int mwf_ens_energy(mwf_ens *e, double *energies);
int mwf_ens_sync(mwf_ens *e);
/* Record every member's energy every `every` steps (0 = off) and the first
 * step at which it is <= threshold (relaminarisation; the reference's
 * io_write_energy kill, program/io.f90:462-465, uses an absolute 1e-10).
 * The samples are taken on the device inside the replayed step graphs; the
 * first one (step 0) is taken at this call. */
int mwf_ens_track(mwf_ens *e, int32_t every, double threshold);
/* Member's recorded energies (up to cap), their count, and the
 * relaminarisation step (-1 if none). */
int mwf_ens_trace(mwf_ens *e, int32_t member, double *energies, int32_t cap, int32_t *count,
                  int64_t *relam_step);
/* Add every member's turbulent fraction (as mwf_turb_fraction, with strides
 * xs, zs and threshold ct; utils/turbfrac.f90:14-38) to the tracking samples
 * of mwf_ens_track, which must be on.  Restarts the tracking. */
int mwf_ens_track_turb(mwf_ens *e, int32_t xs, int32_t zs, double ct);
/* Member's recorded turbulent fractions (up to cap) and their count. */
int mwf_ens_trace_turb(mwf_ens *e, int32_t member, double *fractions, int32_t cap, int32_t *count);
/* Number of size classes the members were grouped into. */
int mwf_ens_classes(mwf_ens *e, int32_t *nclasses);

#ifdef __cplusplus
}
#endif
#endif
