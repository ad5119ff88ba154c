/*
 * wslda_td.h -- C ABI of the MI355X td-W-SLDA step library (libwslda_td.so).
 *
 * Two layers, both plain C: raw device pointers, sizes, int return codes
 * (0 = success, otherwise a hipError_t value or one of the TDW_E* codes), no
 * C++ or torch types.
 *
 *  1. The generic layer, tdw_*: a context (lattice, functional, integration
 *     scheme, constants) created at run time, so one library serves 1D, 2D
 *     and 3D problems and any lattice size.
 *
 *  2. The drop-in layer: the extern "C" entry points of the reference's
 *     kernel interface with the same names, argument order and meaning
 *     (hpc-engine/pca_kernels.h:44-95, pca_derivative.h:12-22,
 *     pca_dens.h:5-19, wslda_cuda_utils.h:47-157).  They act on the context
 *     bound with tdw_bind(); a project build binds it from its predefines.h
 *     (python -m wslda_amd.build --project DIR).  The 1D and 2D codes of the
 *     reference declare
 *     apply_hamiltonian / calculate_densities with extra arguments under the
 *     same names (ccpca_kernels.h:74-77, cpca_kernels.h:74-77,
 *     ccpca_dens.h:5-11); those are exported here as *_1d / *_2d and mapped
 *     to the reference names when TDW_CODEDIM is 1 or 2.
 *
 * State layout (identical to the reference, pca.c:307): n quasiparticle
 * states are stored as [u_0 .. u_{n-1}, v_0 .. v_{n-1}], each a contiguous
 * array of N complex values (N = NX, NX*NY or NX*NY*NZ).
 * Density block: 12*N doubles [nu(2N) rho_a tau_a jax jay jaz rho_b tau_b
 * jbx jby jbz] (wslda_potdens.c:14-36).  Potential block: 12*N doubles
 * [V_a V_b delta(2N) alpha_a alpha_b A_a(3) A_b(3)] (wslda_potdens.c:41-70).
 */
#ifndef WSLDA_TD_H
#define WSLDA_TD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ABI of cufftDoubleComplex / C99 double complex (pca_kernels.h:6-8). */
typedef struct { double re, im; } tdw_complex;
typedef struct { float re, im; } tdw_complex32;

/* ids shared with the reference's predefines (pca_settings.h:63-67, 201-203, 223-224) */
#define TDW_SLDA 111
#define TDW_ASLDA 112
#define TDW_BDG 113
#define TDW_SLDAE 115
#define TDW_SPHERICAL_CUTOFF 88
#define TDW_CUBIC_CUTOFF 89
#define TDW_AB3AM4 34
#define TDW_AB4AM5 45
#define TDW_AB5AM5 55
/* precision of the operator application */
#define TDW_FP64 0
#define TDW_SPLIT 1   /* FP32 FFT and operator, FP64 state, history update and sums */

/* observable slots of compute_energy (pca_settings.h:146-161) */
#define TDW_EKIN 0
#define TDW_EPOT 1
#define TDW_EPAIR 2
#define TDW_ECURRENT 3
#define TDW_EPOTEXT 4
#define TDW_EPAIREXT 5
#define TDW_EVELEXT 6
#define TDW_NPARTA 7
#define TDW_NPARTB 8
#define TDW_LZA 9
#define TDW_LZB 10
#define TDW_NITEMS 11

/* MAX_USER_PARAMS of the reference's predefines.h: length of the params
 * array handed to the problem's callbacks (memcopy_const_params). */
#define TDW_MAX_USER_PARAMS 32
/* bytes per lattice point of the fftONE zone that create_cufftPlans adds
 * behind the kernels and derivatives zones of the caller's work area
 * (2 complex values per point) */
#define TDW_FFTONE_BYTES_PER_POINT 32

/* error codes (besides hipError_t values) */
#define TDW_EINVAL (-1001)  /* bad argument / unsupported combination   */
#define TDW_ENOCTX (-1002)  /* drop-in call before tdw_bind()           */
#define TDW_ESIZE (-1003)   /* lattice axis not a product of 2,3,4,8    */
#define TDW_ENOMEM (-1004)
#define TDW_EEXIST (-1005)  /* output file exists and overwrite is off   */
#define TDW_EIO (-1006)     /* file could not be opened / written       */
#define TDW_ESCRATCH (-1007) /* input larger than the no-allocation scratch */

/* External potential of the built-in problem: a sum of Gaussians
 * amp*s(t)*exp(wx*(x-cx)^2 + wy*(y-cy)^2 + wz*(z-cz)^2), a harmonic trap
 * 0.5*omega^2*r^2 and the smooth on/off switch s(t) of the testsuite's
 * problem-definition.h; x = DX*(ix-NX/2) etc.  A project build replaces it
 * with the user's v_ext() callback. */
#define TDW_MAX_GAUSS 4
typedef struct {
    int ngauss;
    double amp[TDW_MAX_GAUSS];
    double cx[TDW_MAX_GAUSS], cy[TDW_MAX_GAUSS], cz[TDW_MAX_GAUSS];
    double wx[TDW_MAX_GAUSS], wy[TDW_MAX_GAUSS], wz[TDW_MAX_GAUSS];
    double omega;
    int switch_on;
    double t_start, t_stop, t_switch;
    int flip_b;
} tdw_vext;

typedef struct {
    int dim;             /* 1, 2, 3 (CODEDIM)                                 */
    int nx, ny, nz;      /* NX, NY, NZ; for dim<3 the extra axes are transverse */
    double dx, dy, dz;   /* DX, DY, DZ                                        */
    int functional;      /* TDW_SLDA, TDW_ASLDA, TDW_BDG or TDW_SLDAE         */
    int regularization;  /* TDW_SPHERICAL_CUTOFF or TDW_CUBIC_CUTOFF          */
    int mf_correction;   /* INCLUDE_MF_CORRECTION_FROM_REG_SCHEME             */
    int scheme;          /* TDW_AB3AM4 / TDW_AB4AM5 / TDW_AB5AM5              */
    double mu_a, mu_b, ec, dt, t0; /* memcopy_const()                          */
    double sclgth;       /* memcopy_const_sclgth(), BdG scattering length      */
    double p_nmin, p_nmax, denseps; /* stabilisation densities, DENSEPSILON   */
    tdw_vext vext;
    int spinsymmetry;    /* 1: spin-symmetric mode (SPINSYMMETRY_MODE): only
                            E > 0 states are kept; densities sum both
                            components into spin b and copy it to a       */
    double kF;           /* reference kF (memcopy_const); n_F = kF^3/(3 pi^2) */
    double qfalpha, qfbeta, qfgamma; /* quantum-friction strengths (md.qf*) */
    int force_a1;        /* SLDAE built with SLDA_FORCE_A1: A = 1, the A1
                            parameter set, constant-mass (FAST) step       */
    int precision;       /* TDW_FP64 or TDW_SPLIT                              */
    /* Fermi-Dirac high-frequency filter of the mean fields (md.hkf_mode,
     * md.hkf_mu, md.hkf_T; pca_utils.c:107-109).  hkf_mode >= 1: V_a, V_b
     * (and, with a varying effective mass, alpha_a, alpha_b, A_a, A_b) are
     * filtered before every H application; >= 2: Delta too.  The drop-in
     * apply_hamiltonian filters V, Delta and A in place after the friction
     * terms (pca_kernels.cu:733-764, 786); the whole-step entry points also
     * filter alpha, as the driver does before alpha psi (pca.c:1448-1452). */
    int hkf_mode;
    double hkf_mu, hkf_T;
} tdw_params;

typedef struct tdw_ctx tdw_ctx;

/* ---------------------------------------------------------------------------
 * Generic layer.  All pointers are device pointers unless named h_*.
 * Work is queued on the context's stream; results are ordered with every
 * later call on the same context (as the reference's default stream is).
 * ------------------------------------------------------------------------- */
int tdw_create(const tdw_params *p, int nwf, int device, tdw_ctx **out);
int tdw_destroy(tdw_ctx *c);
int tdw_sync(tdw_ctx *c);
size_t tdw_npts(const tdw_ctx *c);
int tdw_set_params(tdw_ctx *c, const tdw_params *p);   /* constants, V_ext; lattice fixed */
/* The problem's user parameters (TDW_MAX_USER_PARAMS doubles, host pointer;
 * copied).  The built-in problem ignores them; a project build
 * (problem-definition.h compiled in) hands them to its callbacks as
 * params / dc_params.  memcopy_const_params is this call on the bound
 * context.  tdw_get_user_params copies them back. */
int tdw_set_user_params(tdw_ctx *c, const double *h_params);
int tdw_get_user_params(const tdw_ctx *c, double *h_params);
/* per-state data: occupation weights f(E_n) (n doubles), transverse momenta
 * ([ky|kz], 2n doubles for dim 1; kz, n doubles for dim 2; NULL for dim 3) and
 * degeneracies (n ints, dim 1).  Host pointers; copied. */
int tdw_set_states(tdw_ctx *c, const double *h_fbeta, const double *h_kt, const int *h_cnt);

int tdw_abm_step1(tdw_ctx *c, int n, tdw_complex *y, tdw_complex *f1, tdw_complex *f2,
                  tdw_complex *f3, tdw_complex *f4, tdw_complex *f5);
int tdw_abm_step4(tdw_ctx *c, int n, const tdw_complex *hy, tdw_complex *y, const tdw_complex *flast);
int tdw_normalize(tdw_ctx *c, int n, tdw_complex *wf, double *norms /* n, may be NULL */);
/* tdw_normalize that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The partial sums (and
 * the norms when norms is NULL) go to a fixed per-device array of
 * TDW_NOALLOC_SCRATCH_DOUBLES doubles set up by tdw_create; inputs needing
 * more (n * ceil(N / 4096) doubles, + n without norms) return TDW_ESCRATCH.
 * All calls on one device share that array: calls that overlap on a device
 * race.  Same kernels, same results as tdw_normalize. */
#define TDW_NOALLOC_SCRATCH_DOUBLES (1 << 20)
int tdw_normalize_noalloc(tdw_ctx *c, int n, tdw_complex *wf, double *norms /* n, may be NULL */);
int tdw_laplace(tdw_ctx *c, int m, const tdw_complex *in, tdw_complex *out);
int tdw_derivatives(tdw_ctx *c, int m, const tdw_complex *in, tdw_complex *ddx,
                    tdw_complex *ddy, tdw_complex *ddz, tdw_complex *lap);
int tdw_laplace_real(tdw_ctx *c, const double *f, double *lap);
/* calculate_densities (pca_dens.cu:220).  With gradients = 1 and all four
 * derivative arrays NULL, the gradients are formed inside the density passes
 * and not stored (3D lattices with the two-pass kernels; TDW_EINVAL
 * elsewhere). */
int tdw_densities(tdw_ctx *c, int n, const tdw_complex *wf, const tdw_complex *ddx,
                  const tdw_complex *ddy, const tdw_complex *ddz, const tdw_complex *lap,
                  const double *weights /* n or NULL */, double *dens, int gradients);
int tdw_density_tau(tdw_ctx *c, double *dens);
int tdw_potentials(tdw_ctx *c, int it, const double *dens, double *pots);
/* SLDA / BdG: out = H in - <H> in from in and its Laplacian (TDW_EINVAL for
 * ASLDA, whose H needs tdw_apply_h_full). */
int tdw_apply_h(tdw_ctx *c, int it, int n, const tdw_complex *in, tdw_complex *out,
                const tdw_complex *lap, const double *pots, const double *useqpe, double *qpe);
/* Any functional: the H application of apply_hamiltonian (pca_kernels.cu:674)
 * with the gradients ddx..ddz of in along the evolved axes and alap =
 * lap(alpha psi) (tdw_alpha_wf, then tdw_laplace; NULL: formed in a context
 * array).  ASLDA adds the effective-mass and current terms
 * (pca_kernels.cu:233-391 without FAST_CONST_EFFECTIVE_MASS_MODE); the
 * others ignore ddx..ddz and alap.  out may equal lap or alap. */
int tdw_apply_h_full(tdw_ctx *c, int it, int n, const tdw_complex *in, tdw_complex *out,
                     const tdw_complex *ddx, const tdw_complex *ddy, const tdw_complex *ddz,
                     const tdw_complex *lap, const tdw_complex *alap, const double *pots,
                     const double *useqpe, double *qpe);
/* out = alpha_a u, alpha_b v (multiply_wf_by_alpha, pca_kernels.cu:1016);
 * in place allowed. */
int tdw_alpha_wf(tdw_ctx *c, int n, const tdw_complex *in, const double *pots, tdw_complex *out);
/* <wf1_a|wf2_a> = dV sum (u1* u2 + v1* v2), real parts into re[a] and
 * imaginary parts into im[a] (either may be NULL; device arrays). */
int tdw_overlap(tdw_ctx *c, int n, const tdw_complex *wf1, const tdw_complex *wf2, double *re, double *im);
int tdw_energy(tdw_ctx *c, int it, const double *dens, const double *pots, double *h_e /* 11 */);
int tdw_vext_eval(tdw_ctx *c, int it, double *va, double *vb);

/* One PECE step of pca.c:1401-1570 on the arrays of a run.  f[0..4] are
 * fkm1..fkm5 and are rotated in place (pointer swaps) as pca.c:1543-1562
 * does.  grad_last: full derivatives (tau, j) on this step's corrector.
 * ASLDA computes the derivatives on both half-steps (the reference's path
 * without FAST_CONST_EFFECTIVE_MASS_MODE), so ddx..ddz of the evolved axes
 * must be set; grad_last changes nothing there.  On 3D lattices with the
 * two-pass kernels the ASLDA step forms the gradients inside its density
 * and H passes and leaves ddx..ddz untouched. */
typedef struct {
    int n;
    tdw_complex *wf, *lap, *ddx, *ddy, *ddz;
    tdw_complex *f[5];
    double *dens, *pots;
    double *qpe;          /* n: <H> of the last H application */
} tdw_state;
int tdw_step(tdw_ctx *c, tdw_state *s, int it, int grad_last);
/* Same step with the fused kernels (six passes over the state arrays instead
 * of the reference's ~57 touches); same results to rounding.  Between fused
 * steps the stored wf is unnormalised and f[0] holds H psi without the x
 * part of its kinetic term and before the <H> subtraction: the next fused
 * step, or tdw_finalize(), which every other entry point calls first,
 * completes both.  Densities, potentials and qpe are exact after every step.
 * ASLDA and SLDAE (varying effective mass) on 3D lattices of the two-pass
 * kernels run the folded step: the modular step's density and H passes with
 * the normalisations and the <H> subtraction carried as per-state scalars,
 * so between steps wf is unnormalised and f[0] lacks only the subtraction
 * (same completion rules); it has a split form too (below; its
 * measurement steps also use ddy and ddz as FP64 scratch).  Other lattices
 * run the modular step (tdw_fused_supported() says whether the
 * constant-mass fused passes run).
 *
 * Split precision (params.precision == TDW_SPLIT): the x / yz FFTs and the
 * kinetic term run in FP32 complex, and f[0..4] hold FP32 complex values
 * (8 bytes each; a buffer of 2*n*N*8 bytes is enough).  The state wf, every
 * ABM update of it (the AM partial corrector c is kept in FP64 in wf between
 * the predictor and the corrector, the prediction as an FP32 copy), the
 * densities, norms and <H> stay FP64; lap is used as FP32 scratch within a
 * step.  Gradient steps (grad_last) compute derivatives, tau, j and H psi in
 * FP64.  tdw_step and lattices without fused kernels return TDW_EINVAL in
 * this mode. */
int tdw_step_fused(tdw_ctx *c, tdw_state *s, int it, int grad_last);
int tdw_finalize(tdw_ctx *c, tdw_state *s);
int tdw_fused_supported(tdw_ctx *c);
/* pca.c:939-975: normalise, derivatives, densities, tau (potentials kept). */
int tdw_initial_measure(tdw_ctx *c, tdw_state *s);
/* Self-start of the ABM history (selfstart 1, pca.c:1086-1392): K = 3, 4, 5
 * steps (AB3AM4, AB4AM5, AB5AM5) of the midpoint rule with exp(-i (H -
 * <H>) dt) as its Taylor polynomial of order 4, 5, 5 (taylor_expansion_
 * contribution), the potentials of iteration 0 throughout, no quantum
 * friction.  Leaves wf advanced by K steps (the caller's iteration counter
 * is not advanced, as in the reference), f[0] = fkm1 = (H - <H>) psi of the
 * last self-start step and f[1..K-1] = that of the step before (the
 * reference's host shift copies slot j to j + 1 for increasing j), then the
 * derivatives, densities with tau and j, and the potentials of the new
 * state.  f[0..2] serve as scratch; the two kept entries live in f[3], f[4]
 * for AB5AM5, in f[3] and scratch0 for AB4AM5, in scratch0 and scratch1
 * for AB3AM4 (state-sized, may be NULL where not needed).  FP64 only. */
int tdw_selfstart(tdw_ctx *c, tdw_state *s, tdw_complex *scratch0, tdw_complex *scratch1);

/* td checkpoint files (hpc-engine/tdwslda_checkpoint.c:118-364), byte layout
 * of a single-rank run: header (t0, mu[2], ec, kF, nwf, eF, Effg, beta), then
 * fbeta, kz (dim <= 2) and ky (dim 1) sections, wf, fkm1..fkm3 [fkm4]
 * [fkm5], and the 12*N potentials.  Per-state sections are sized for nwf + 1
 * states, as tdwslda_checkpoint.c sizes them for one rank.  f[0..4] of the
 * state are written as fkm1..fkm5; save applies a pending fused-step fix-up
 * first.  FP64 runs only.
 *   tdw_checkpoint_nwf   -- load_nwf (tdwslda_checkpoint.c:22-59)
 *   tdw_checkpoint_load  -- load_checkpoint (:61-220), into s (s->n == nwf)
 *   tdw_checkpoint_save  -- save_checkpoint (:223-364)
 * h_kt: the state's transverse momenta, as passed to tdw_set_states. */
typedef struct {
    double mu_a, mu_b, ec, kF, eF, Effg, beta;
} tdw_checkpoint_scalars;
int tdw_checkpoint_nwf(const char *path, int *nwf);
int tdw_checkpoint_save(tdw_ctx *c, const char *path, tdw_state *s, double t0, const double *h_fbeta,
                        const double *h_kt, const tdw_checkpoint_scalars *sc);
int tdw_checkpoint_load(tdw_ctx *c, const char *path, tdw_state *s, double *t0, double *h_fbeta, double *h_kt,
                        tdw_checkpoint_scalars *sc);

/* Static-solver initial states (inittype 3, pca.c:628-725): the st-wslda-3d
 * output in <prefix>/ -- wf.info (or s3dpca.info), wf.%04d.info per file,
 * wf.%04d_{un,vn}.wdat and wf.%04d_en.dat (or s3dpca.%04d.{wfu,wfv,en}),
 * wf.pud (or s3dpca.pud), raw binary as the static solver writes them.
 *   tdw_st3d_info -- check_comptibility + read_checkpoint_info_pca +
 *                    scan_s3dpca_info_files: the run constants and the
 *                    number of states in each of the iogroups files
 *   tdw_st3d_load -- read_s3dpca_wf of states [lidx, uidx) into h_wf (host,
 *                    [u_lidx..u_uidx-1, v_lidx..v_uidx-1]), their
 *                    occupations f(E) into h_fbeta, the 12*N potentials into
 *                    h_pots (may be NULL) and mu, ec, kF, eF, beta and Effg
 *                    (0.6 N eF, N = N_a + N_b with spinsymmetry == 1, else
 *                    N_a, as pca.c:716-718) into sc; TDW_EINVAL when the
 *                    files' lattice is not p's.
 * Return codes: the reference readers' (negative from the wave-function
 * reader, 10002 / 10006 / 10007 / 10030-10033 from the info files). */
typedef struct {
    int nwf, nx, ny, nz;
    double dx, dy, dz, kF, mu_a, mu_b, ec, beta;
    int new_format;          /* 1: wf.* names, 0: s3dpca.* names */
} tdw_st_info;
int tdw_st3d_info(const char *prefix, int iogroups, tdw_st_info *info, int *nwf_per_file /* iogroups */);
int tdw_st3d_load(const tdw_params *p, const char *prefix, int iogroups, const int *nwf_per_file,
                  const tdw_st_info *info, int lidx, int uidx, int spinsymmetry, tdw_complex *h_wf,
                  double *h_fbeta, double *h_pots, tdw_checkpoint_scalars *sc);

/* W-data measurement output (lib/wdata/doc/README.md): <outprefix>.wtxt and
 * <outprefix>_<var>.wdat, one datablock appended per measurement.
 *   tdw_wdata_create   -- create_wdata_metadata (wslda_writevars.c:109-253):
 *                         variables from the writevar keys (density|rho,
 *                         current|j, tau, V, V_ext, velocity_ext, delta,
 *                         delta_ext, nu, alpha, A, "all", "default"; none ->
 *                         rho delta j, pca_utils.c:495-566), _b variables as
 *                         links when spinsymmetry; datadim = the code's dim
 *   tdw_wdata_setconst -- wdata_setconst (lib/wdata/c/wdata.c:1444-1462)
 *   tdw_wdata_begin    -- clear_files + write_wdata_metadata_file
 *                         (wslda_writevars.c:258-288, pca.c:1033-1038);
 *                         TDW_EEXIST if overwrite is 0 and a file exists
 *   tdw_wdata_write    -- write_measurments (wslda_writevars.c:298-452) of
 *                         the device densities / potentials at step it, then
 *                         the cycle count and the .wtxt rewrite
 *                         (pca.c:1623-1630).  V is written as V - V_ext(it)
 *                         (except at it == 0 for inittype 1..3), delta_ext
 *                         and velocity_ext from the problem's callbacks (zero
 *                         for the built-in problem, which has neither);
 *                         writeprec 'f' stores FP32 casts. */
typedef struct {
    const char *outprefix;   /* md.outprefix                                    */
    const char *writevar;    /* writevar keys, separated by spaces or commas   */
    char writeprec;          /* 'd' or 'f'                                      */
    int spinsymmetry;
    int overwrite;
    int inittype;
    double t0;               /* time of the first cycle                         */
    double dt;               /* time between cycles (timesteps * dt)            */
    const char *codename;    /* NULL: td-wslda-{1,2,3}d                         */
} tdw_wdata_params;
typedef struct tdw_wdata tdw_wdata;
int tdw_wdata_create(tdw_ctx *c, const tdw_wdata_params *p, tdw_wdata **out);
int tdw_wdata_setconst(tdw_wdata *w, const char *name, double value, const char *unit /* NULL: none */);
int tdw_wdata_begin(tdw_wdata *w);
int tdw_wdata_write(tdw_wdata *w, int it, const double *dens, const double *pots);
int tdw_wdata_cycles(const tdw_wdata *w);
int tdw_wdata_destroy(tdw_wdata *w);

/* ---------------------------------------------------------------------------
 * Adjoint (reverse-mode) time stepping.  No counterpart in the reference:
 * gradients of a final-time observable with respect to the external-potential
 * parameters, for SLDA in FP64 through the modular step's non-gradient form.
 *
 * tdw_adj_forward_step: tdw_step (grad_last = 0) that also records what the
 *   backward step needs: a copy of y at the step's start (y_copy), a copy of
 *   the new history entry (fnew_copy) and the tape fields below.
 * tdw_adj_backward_step: applies the transpose of that step's Jacobian.
 *   y0, f[0..h-1]: y and the history at the step's start; y1, fnew: y and
 *   the new history entry after it; pots_prev: the 12*N potentials the step
 *   started from.  a->ly holds dL/dy after the step on entry and dL/dy
 *   before it on return; a->lf[0..h-1] likewise for the history (rotated in
 *   place by pointer swaps); a->lpots (4*N: dL/d(Va, Vb, Re D, Im D)) holds
 *   the gradient of the potentials the step left on entry and of pots_prev
 *   on return.  dens_grad (4*N, or NULL): dL/d(rho_a, rho_b, Re nu, Im nu)
 *   of the densities the step leaves.  theta_grad accumulates
 *   dL/d(amp, cx, cy, cz, wx, wy, wz) of each Gaussian of params.vext.
 * ------------------------------------------------------------------------- */
typedef struct {
    int n;
    double *scal;     /* 4*n: norm^2 of y_p, <H> of y_p, norm^2 of y_c, <H> of y' */
    double *dens_p;   /* 12*N densities of the normalised prediction            */
    double *pots_p;   /* 12*N potentials after the predictor's update           */
    double *dens_c;   /* 12*N densities of the corrected state                  */
    double *pots_c;   /* 12*N potentials after the corrector's update           */
} tdw_tape_step;
typedef struct {
    int n;
    tdw_complex *ly;
    tdw_complex *lf[5];
    double *lpots;        /* 4*N */
    tdw_complex *w[5];    /* scratch, state-sized */
    double *lgrid;        /* scratch, 12*N */
    double *theta_grad;   /* 7*TDW_MAX_GAUSS, accumulated */
} tdw_adj_state;
int tdw_adj_forward_step(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_copy, tdw_complex *fnew_copy,
                         tdw_tape_step *t);
/* The taped step without copies, for a caller whose tape holds y and the
 * history entries themselves: y at the step's start is read from s->wf and
 * kept; y after the step is written to y_next and s->wf set to it; the new
 * history entry is written to f_next and becomes s->f[0] (the others shift
 * by pointer; no history entry is written); s->ddx is used as scratch for
 * the partial corrector.  y_next, f_next, s->ddx and s->lap must differ from
 * s->wf, s->f[0..h-1] and from each other. */
int tdw_adj_forward_step_to(tdw_ctx *c, tdw_state *s, int it, tdw_complex *y_next, tdw_complex *f_next,
                            tdw_tape_step *t);
/* The history entry taped step `it` wrote, rebuilt from y after that step
 * (y1) and the step's tape t: H y1 - <H> y1 with t->pots_c, the H
 * application the step used and the <H> it recorded (t->scal[3n..4n)), so a
 * tape may keep y alone and rebuild the entries during the backward sweep.
 * scratch: a state-sized array (the Laplacian or D_x of y1).  fnew and
 * scratch must differ from y1 and from each other. */
int tdw_adj_history_entry(tdw_ctx *c, int it, const tdw_tape_step *t, const tdw_complex *y1, tdw_complex *fnew,
                          tdw_complex *scratch);
int tdw_adj_backward_step(tdw_ctx *c, tdw_adj_state *a, int it, const tdw_complex *y0,
                          const tdw_complex *const *f, const tdw_complex *y1, const tdw_complex *fnew,
                          const tdw_tape_step *t, const double *pots_prev, const double *dens_grad);

/* Uniform SLDA solution and plane-wave BdG states (host arrays).
 * tdw_fbeta: fbeta (pca_uniform.h:70-78).
 * tdw_uniform_solve: solve_uniform_problem (pca_uniform.h:86-502); mu_rule 0
 *   is the chemical-potential update of pca_uniform.h:427-428, mu_rule 1 the
 *   one kept commented at :425-426 (the rule of the code that produced the
 *   testsuite's td-gauss goldens).  With p->functional == TDW_BDG it is
 *   solve_uniform_problem_bdg (pca_uniform.h:1105-1470; scattering length
 *   p->sclgth, mu_rule ignored).  Returns 1 if maxiter was reached.
 * tdw_uniform_states: create_uniform_wf (pca_uniform.h:636-850) and its 2D /
 *   1D twins (:2044-2287, :2321-2560): states [idxfrom, idxto) in the
 *   reference's order, energies E_n, transverse momenta and degeneracies. */
typedef struct {
    double n0_a, n0_b, mu_a, mu_b, V_a, V_b, alph_a, alph_b, kc, ec;
    double nu, delta, tau_a, tau_b, beta, S, ekin, epot, epair;
    int nwf, iters;
} tdw_uniform;
double tdw_fbeta(double E, double beta);
int tdw_uniform_solve(const tdw_params *p, double n0_a, double n0_b, double kc, double Tstart,
                      double Tstop, double dT, double eps, double scmix, double muchange,
                      int maxiter, int mu_rule, tdw_uniform *u);
int tdw_uniform_count(const tdw_params *p, const tdw_uniform *u);
int tdw_uniform_states(const tdw_params *p, const tdw_uniform *u, int idxfrom, int idxto,
                       tdw_complex *h_wf, double *h_En, double *h_kt, int *h_cnt);

/* The context's HIP stream (hipStream_t), for callers that order their own
 * work or record events around the library's. */
void *tdw_stream(tdw_ctx *c);
/* Per-kernel timing: while enabled, every kernel launch is bracketed by HIP
 * events on the context stream.  tdw_timing_report() synchronises, writes at
 * most max entries (name: 32 chars each, total ms, launch count) and resets
 * the counters; returns the number of entries. */
int tdw_timing(tdw_ctx *c, int enable);
int tdw_timing_report(tdw_ctx *c, int max, char *names, double *total_ms, int *count);

/* ---------------------------------------------------------------------------
 * Project builds.  `python -m wslda_amd.build --project DIR --codedim D`
 * compiles a problem's predefines.h and problem-definition.h (DIR) into the
 * library, as the reference compiles them into its kernel unit
 * (pca_kernels.cu:28-42): its v_ext, delta_ext, velocity_ext and
 * scattering_length callbacks, and with ENABLE_MODIFY_* its modify_potentials,
 * modify_energies and modify_densities, then drive the engine in place of
 * the built-in V_ext (params.vext is ignored).  The lattice and functional
 * are compiled in: tdw_create refuses others with TDW_EINVAL, and an
 * external velocity needs a varying-mass functional (ASLDA, SLDAE), whose
 * step carries the current terms it enters (TDW_EINVAL otherwise).  The
 * problem's user parameters and extra data reach its callbacks as params /
 * dc_params and extra_data (tdw_set_user_params / memcopy_const_params,
 * memcopy_extra_data).  The project library also exports the problem's own
 * host functions (include/wslda_td_project.h).
 * ------------------------------------------------------------------------- */
#define TDW_PROJ_V_EXT 0x01u             /* ENABLE_V_EXT               */
#define TDW_PROJ_DELTA_EXT 0x02u         /* ENABLE_DELTA_EXT           */
#define TDW_PROJ_VELOCITY_EXT 0x04u      /* ENABLE_VELOCITY_EXT        */
#define TDW_PROJ_MODIFY_POTENTIALS 0x08u /* ENABLE_MODIFY_POTENTIALS   */
#define TDW_PROJ_MODIFY_ENERGIES 0x10u   /* ENABLE_MODIFY_ENERGIES     */
#define TDW_PROJ_MODIFY_DENSITIES 0x20u  /* ENABLE_MODIFY_DENSITIES    */
/* 1 and the compiled settings (lattice, functional, integration and
 * regularisation schemes, stabilisation densities, DENSEPSILON, spin
 * symmetry, SLDA_FORCE_A1) written into *p, the TDW_PROJ_* switches into
 * *flags (either may be NULL) in a project library; 0 in the plain one. */
int tdw_project_info(tdw_params *p, unsigned *flags);

/* ---------------------------------------------------------------------------
 * Drop-in layer: the reference's entry points, acting on the bound context.
 * ------------------------------------------------------------------------- */
int tdw_bind(tdw_ctx *c);
tdw_ctx *tdw_bound(void);

int set_gpu(int deviceId);                                         /* wslda_cuda_utils.h:109-113 */
int gpu_malloc(size_t size, void **ptr);                           /* wslda_cuda_utils.h:47-57   */
int gpu_free(void *ptr);
int host_malloc_pl(size_t size, void **ptr);                       /* pinned host memory         */
int host_free_pl(void *ptr);
int memcopy_host2gpu(void *h_src, void *d_dst, size_t size);       /* wslda_cuda_utils.h:63-76   */
int memcopy_gpu2host(void *d_src, void *h_dst, size_t size);
int memcopy_gpu2gpu(void *d_src, void *d_dst, size_t size);
int memcopy_const(double mu_a, double mu_b, double ec, double t0, double dt, double kF); /* :21-36 */
int memcopy_const_params(double *params);                          /* :80-90: TDW_MAX_USER_PARAMS doubles onto the
                                                                      bound context (tdw_set_user_params) */
int memcopy_const_sclgth(double sclgth);                           /* :92-97 */

int create_cufftPlans(int batch_size, int nwfip, size_t *workSize); /* pca_derivative.h:12 */
int set_workspace_for_cufftPlan(void *workArea);                    /* pca_derivative.h:13 */
int compute_derivatives(int n, tdw_complex *wf, tdw_complex *wf_d_dx, tdw_complex *wf_d_dy,
                        tdw_complex *wf_d_dz, tdw_complex *wf_laplace, int nthreads); /* pca_derivative.cu:217 */
int compute_laplace(int n, tdw_complex *wf, tdw_complex *wf_laplace, int nthreads); /* pca_derivative.cu:332 */
int compute_laplace_real_f(double *f, double *laplace_f, int nthreads);            /* pca_derivative.cu:640 */
/* Real-field derivatives along the evolved axes (1D: x only, 2D: x, y; the
 * other arguments are ignored, as in ccpca_ / cpca_derivative.cu); the
 * Nyquist derivative is zero; outputs may alias inputs. */
int compute_gradient_real_f(double *f, double *df_dx, double *df_dy, double *df_dz,
                            int nthreads);                           /* pca_derivative.cu:451 */
int compute_derivative_real_vector_f(double *fx, double *fy, double *fz, double *dfx_dx,
                                     double *dfy_dy, double *dfz_dz,
                                     int nthreads);                  /* pca_derivative.cu:549 */
int compute_divergence_real_vector_f(double *fx, double *fy, double *fz, double *divf,
                                     int nthreads);                  /* pca_derivative.cu:735 */
/* out = F^-1[FD(k^2/2) F[in]], FD(e) = 1/(exp((e - fd_mu)/fd_T) + 1); out may be in */
int high_frequency_filter_d(double *in, double *out, double fd_mu, double fd_T,
                            int nthreads);                           /* pca_derivative.cu:829 */
int high_frequency_filter_c(tdw_complex *in, tdw_complex *out, double fd_mu, double fd_T,
                            int nthreads);                           /* pca_derivative.cu:906 */
int high_frequency_filter_massive_d(int n, double *in, double *out, double fd_mu, double fd_T,
                                    int nthreads);                   /* pca_derivative.cu:950 */

int calculate_densities(int n, tdw_complex *wf, tdw_complex *wf_d_dx, tdw_complex *wf_d_dy,
                        tdw_complex *wf_d_dz, tdw_complex *d_wf_laplace, double *d_fbetaEn,
                        double *weights, double *d_densities, int gradients_computed,
                        int nthreads);                              /* pca_dens.cu:220 */
int calculate_densities_1d(int n, tdw_complex *wf, tdw_complex *wf_d_dx, tdw_complex *d_wf_laplace,
                           double *d_kkyz, double *d_fbetaEn, double *weights, int *d_cnt,
                           double *d_densities, int gradients_computed, int nthreads); /* ccpca_dens.h:5-11 */
int calculate_densities_2d(int n, tdw_complex *wf, tdw_complex *wf_d_dx, tdw_complex *wf_d_dy,
                           tdw_complex *d_wf_laplace, double *d_kkz, double *d_fbetaEn,
                           double *weights, double *d_densities, int gradients_computed,
                           int nthreads);                            /* cpca_dens.h */
int density_caculate_tau(double *d_densities, int nthreads);         /* pca_dens.cu:300 */
/* Quantum-friction densities [U_a, U_b, D] (4 N doubles; the driver passes
 * d_densities + 12 N).  apply_hamiltonian with qfswitch > 0 reads them there,
 * adds the friction terms to V_a, V_b, Delta (strengths p.qfalpha,
 * p.qfbeta, p.qfgamma times qfswitch, qfgamma also times pccoeff, n_F from
 * p.kF / memcopy_const), applies H, restores the potentials and leaves the
 * unmodified V_a, V_b, Delta in those 4 N doubles, as the reference does. */
int calculate_quantum_friction_densities(int n, tdw_complex *wf, tdw_complex *d_wf_laplace,
                                         double *d_fbetaEn, double *weights, double *d_qf_densities,
                                         int nthreads);              /* pca_dens.cu:459 */
int calculate_quantum_friction_densities_1d(int n, tdw_complex *wf, tdw_complex *d_wf_laplace,
                                            double *kkyz, double *d_fbetaEn, double *weights,
                                            int *cnt, double *d_qf_densities,
                                            int nthreads);           /* ccpca_dens.cu:470 */
int calculate_quantum_friction_densities_2d(int n, tdw_complex *wf, tdw_complex *d_wf_laplace,
                                            double *kkz, double *d_fbetaEn, double *weights,
                                            double *d_qf_densities, int nthreads); /* cpca_dens.cu:471 */

int compute_potentials(int it, double *d_densities, double *d_potentials, double cccoeff,
                       int nthreads);                                /* pca_kernels.cu:81 */
int compute_energy(int it, double *d_densities, double *d_potentials, double *d_workarea,
                   int nthreads);                                    /* pca_kernels.cu:210 */
int apply_hamiltonian(int it, int n, tdw_complex *wf_in, tdw_complex *wf_out,
                      tdw_complex *wf_d_dx, tdw_complex *wf_d_dy, tdw_complex *wf_d_dz,
                      tdw_complex *wf_laplace, tdw_complex *alphawf_laplace,
                      double *d_densities, double *d_potentials, double qfswitch,
                      double *useqpe, double pccoeff, int nthreads); /* pca_kernels.cu:674 */
int apply_hamiltonian_1d(int it, int n, tdw_complex *wf_in, tdw_complex *wf_out,
                         tdw_complex *wf_d_dx, double *d_kkyz, tdw_complex *wf_laplace,
                         tdw_complex *alphawf_laplace, double *d_densities, double *d_potentials,
                         double qfswitch, double *useqpe, double pccoeff, int nthreads);
int apply_hamiltonian_2d(int it, int n, tdw_complex *wf_in, tdw_complex *wf_out,
                         tdw_complex *wf_d_dx, tdw_complex *wf_d_dy, double *d_kkz,
                         tdw_complex *wf_laplace, tdw_complex *alphawf_laplace,
                         double *d_densities, double *d_potentials, double qfswitch,
                         double *useqpe, double pccoeff, int nthreads);
int multiply_wf_by_alpha(int n, tdw_complex *wf_in, tdw_complex *wf_out, double *d_potentials,
                         int nthreads);                              /* pca_kernels.cu:1016 */
int normalize_wf(int n, tdw_complex *wf, int nthreads);              /* pca_kernels.cu:955 */
int abm_step1(int n, tdw_complex *ykm1, tdw_complex *fkm1, tdw_complex *fkm2, tdw_complex *fkm3,
              tdw_complex *fkm4, tdw_complex *fkm5, int abm_scheme, int nthreads); /* tdwslda_ode_integrator.h:106 */
int abm_step4(int n, tdw_complex *ykm1_in, tdw_complex *ykm1_out, tdw_complex *fkm_last,
              int abm_scheme, int nthreads);                         /* tdwslda_ode_integrator.h:195 */
int taylor_expansion_contribution(int it, double dt, int n, tdw_complex *wf_hpsi,
                                  tdw_complex *wf_update, tdw_complex *wf_contr,
                                  int nthreads);                     /* pca_kernels.cu:1076 */
int symmetrize_densities_device(double *d_densities);                /* pca_dens.cu:343 */
/* Re <wf1_a|wf2_a> = Re dV sum (u1* u2 + v1* v2) for a < n into
 * overlap_re[a] and, as the reference does (it copies the first reduced sum
 * into both outputs, pca_kernels.cu:899-907), into overlap_im[a] too (either
 * may be NULL; device arrays); workarea is unused.  tdw_overlap gives the
 * imaginary part. */
int compute_ovelap(int n, tdw_complex *wf1, tdw_complex *wf2, double *overlap_re,
                   double *overlap_im, double *workarea, int nthreads); /* pca_kernels.cu:877 */
/* sum of size device doubles into partial_sums[0] (mode 0 only) */
int local_reductionR(double *array, int size, double *partial_sums, int threads,
                     int mode);                                      /* wslda_cuda_utils.h:250 */
int memcopy_extra_data(size_t extra_data_size, void *extra_data);   /* wslda_cuda_utils.h:47 */
/* Views of the work area given to set_workspace_for_cufftPlan:
 * [kernels 24 n doubles][derivatives 8 n doubles][fft ONE]
 * (tdwslda_memory_management.cu:31-94) */
size_t mm_get_size_of_kernels_workspace(size_t n);
size_t mm_get_size_of_derivatives_workspace(size_t n);
size_t mm_get_size_of_total_workspace(size_t n);
void *mm_get_pointer_to_kernels_workspace(size_t n);
void *mm_get_pointer_to_derivatives_workspace(size_t n);
void *mm_get_pointer_to_total_workspace(size_t n);
void *mm_get_pointer_to_fftONE_workspace(size_t n);
int get_v_ext(int datadim, int spin, int it, double *data);          /* pca_kernels.cu:1106 (host data) */
/* st-wslda-1d starts (inittype 1; host).  tdw_kmode has the layout of the
 * driver's wslda_kmode (wslda_wavevectors.h:104-109): a transverse mode
 * (ky, kz) and its degeneracy.  tdw_st1d_kmodes lists the modes of p's
 * (ny, nz) grid as fill_wslda_kmodes_1d does (wslda_wavevectors.h:25-100;
 * kvecs may be NULL to count).  The readers take codedim 1 (each state
 * evolved once, h_kkyz = (ky..., kz...), h_cnt = degeneracy), 2 (each state
 * expanded over its mode's ky with kz >= 0, h_kkyz = kz; the scan needs a
 * bound context for the mode table) or 3 (each state expanded over its
 * mode's (ky, kz)).  read_stwslda1d_wf takes the lattice from the bound
 * context, whose dim must equal codedim. */
typedef struct {
    double ky, kz;
    int weight;
} tdw_kmode;
int tdw_st1d_kmodes(const tdw_params *p, tdw_kmode *kvecs, int max, int *cnt);
int scan_stwslda1d_info_files(const char *prefix, int codedim, int kvecs_to_consder, tdw_kmode *kvecs,
                              int *nwf, int *nwf_per_kyz);                    /* pca_io.h:845 */
int read_stwslda1d_wf(const char *prefix, int codedim, int kvecs_to_consder, tdw_kmode *kvecs,
                      int *nwf_per_kyz, int mylidx, int myuidx, tdw_complex *h_wavefun,
                      double *h_fbetaEn, double *h_kkyz, int *h_cnt);        /* pca_io.h:1177 */
/* st-wslda-3d readers of the driver (host; pca_io.h) */
int check_comptibility(char *prefix, int requested_dim, char *suffix);          /* pca_io.h:382 */
int read_checkpoint_info_pca(const char *file_name, int *nwf, int *nx, int *ny, int *nz, double *dx,
                             double *dy, double *dz, double *kF, double *mu, double *ec,
                             double *beta);                                   /* pca_io.h:340 */
int scan_s3dpca_info_files(const char *prefix, int number_of_files, int *nwf,
                           int *nwf_per_file);                                /* pca_io.h:1519 */
int read_s3dpca_wf(const char *prefix, int number_of_files, int *nwf_per_file, int mylidx, int myuidx,
                   tdw_complex *h_wavefun, double *h_fbetaEn);                /* pca_io.h:1562, N of the bound context */
int get_delta_ext(int datadim, int it, void *deltain, void *data);  /* pca_kernels.cu:1135 (device in, host out) */
int get_velocity_ext(int datadim, int spin, int it, double *data); /* pca_kernels.cu:1166 (host data, 3 blocks) */

#if defined(TDW_CODEDIM) && TDW_CODEDIM == 1
#define apply_hamiltonian apply_hamiltonian_1d
#define calculate_densities calculate_densities_1d
#elif defined(TDW_CODEDIM) && TDW_CODEDIM == 2
#define apply_hamiltonian apply_hamiltonian_2d
#define calculate_densities calculate_densities_2d
#endif

#ifdef __cplusplus
}
#endif
#endif
