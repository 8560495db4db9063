/*
 * pnfam_gpu -- C ABI of the MI355X (gfx950) pnFAM iteration engine.
 *
 * One context holds the omega-independent inputs of a pnFAM solve resident in
 * HBM (HFB basis tables, Bogoliubov matrices, couplings on the quadrature
 * grid, the operator in the qp basis) and solves batches of complex
 * frequencies.  All arrays cross the boundary as host pointers in the
 * reference's storage conventions:
 *   - state-major tables (nghl x dqp), column-major, as hfbtho_solution.f90:88-98;
 *   - block matrices as pnfam_type_blockmatrix.f90:19-26: one non-zero block per
 *     block row, blocks stored in increasing block-row order, column-major inside;
 *     maps are 0-based, -1 meaning "no block" (the Fortran side adds/subtracts 1).
 * Every entry point returns 0 on success and a negative PNFAM_GPU_E* code on
 * failure; pnfam_gpu_last_error() gives the message.  Entry points are not
 * re-entrant on one context; distinct contexts may be used from distinct
 * threads.
 */
#ifndef PNFAM_GPU_H
#define PNFAM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNFAM_GPU_OK 0
#define PNFAM_GPU_EINVAL -1   /* bad argument or inconsistent block structure */
#define PNFAM_GPU_EDEVICE -2  /* HIP runtime / device failure */
#define PNFAM_GPU_ESTATE -3   /* call out of order (e.g. solve before set_operator) */
#define PNFAM_GPU_ENOMEM -4   /* device allocation failed */

typedef struct pnfam_gpu_ctx pnfam_gpu_ctx;

/* omega-independent HFB / functional inputs.  Replaces the module state the
 * reference's calc_hamiltonian reads implicitly (hfb_solution: wf tables,
 * U/V/E, num_spin_up; pnfam_interaction: couplings), pnfam_setup.f90:276-321,
 * hfbtho_solution.f90:137-395, pnfam_interaction.f90:250-257. */
typedef struct {
  int32_t nb;                 /* blocks of the doubled basis (Omega>0 then TR copies) */
  int32_t dqp;                /* sum(db) */
  int32_t nghl;               /* quadrature points */
  const int32_t *db;          /* [nb] block sizes */
  const int32_t *num_spin_up; /* [nb] spin-up states lead each block */
  const int32_t *ns;          /* [dqp] 2*spin projection (+1/-1) */
  const double *wf, *wfdr, *wfdz, *wfdp, *wfd2_all; /* [nghl x dqp] column-major */
  const double *wdcori;       /* [nghl] inverse quadrature weight x Jacobian */
  const double *Ep, *En;      /* [dqp] qp energies (0 above the pairing window) */
  const double *Up, *Vp, *Un, *Vn; /* [sum db^2]; V maps block i <-> i +- nb/2 */
  const double *crho, *cs, *cpair, *cspair; /* [nghl] density-dependent couplings */
  const double *cscal;        /* [12] cdrho ctau ctj0 ctj1 ctj2 crdj cds ct cj cgs cf csdj */
} pnfam_gpu_basis;

typedef struct {
  int32_t max_iter;           /* /solver/ max_iter (pnfam_setup.f90:98-108) */
  double convergence_epsilon; /* max-abs residual threshold (1e-7) */
  int32_t broyden_history;    /* M; <0 disables mixing */
  double quench;              /* quench_residual_int; 0 skips the residual interaction */
  double alpha_mix;           /* linear mixing factor of the Broyden step */
  int32_t precision;          /* PNFAM_GPU_FP64, PNFAM_GPU_MIXED or PNFAM_GPU_MIXED_NOREFINE */
} pnfam_gpu_opts;

/* Precision of the residual interaction (the reference computes everything in
 * FP64; pnfam_solver.f90:93-221).
 *   FP64: every step in FP64.
 *   MIXED: the hamiltonian is linear in the amplitudes, so H(v_k) is carried as
 *     an FP64 sum of H applied to the increments v_k - v_{k-1}; each increment goes
 *     through the block transforms and the grid stage on split-BF16 MFMA with
 *     FP32 accumulation.  The residual, its max-norm, Broyden, G(w) and S stay
 *     FP64.  When a point first
 *     meets the tolerance, one full FP64 evaluation replaces the accumulated H
 *     and the point iterates on until it meets the tolerance again; the Broyden
 *     pair formed across that evaluation is zeroed (environment
 *     PNFAM_GPU_MIX_DROP=0 keeps it).
 *   MIXED_NOREFINE: as MIXED without the final FP64 evaluation; S then carries
 *     the low-precision error of the first increments (about 1e-6 relative).
 *   Both mixed modes keep the Broyden history in FP32 (environment
 *   PNFAM_GPU_MIX_HIST=fp64 keeps it in FP64); the residual stays FP64.
 *   In both mixed modes a point whose increment, judged by its last residual,
 *   is below 3e-3 of its largest amplitude (environment PNFAM_GPU_MIX_TAU)
 *   goes through single BF16 products instead of the three split products:
 *   its rounding error is scaled down by the increment.  The grid stage runs as
 *   GEMMs against resident pair tables or as factored kernels whose cost is per
 *   point; by default each pass takes the factored kernels for groups of at
 *   most PNFAM_GPU_MIX_AUTO_N (6) points (environment PNFAM_GPU_MIX_GRID =
 *   tables | factored | auto; PNFAM_GPU_MIX_TRIPROD = bf16 | fp64 for the
 *   block transforms). */
#define PNFAM_GPU_FP64 0
#define PNFAM_GPU_MIXED 1
#define PNFAM_GPU_MIXED_NOREFINE 2

/* Copies the basis to the device and builds the resident tables.
 * Replaces setup_pnfam's HFB + interaction set-up (pnfam_setup.f90:134-174). */
int pnfam_gpu_create(const pnfam_gpu_basis *b, pnfam_gpu_ctx **out);

/* Sets the one-body operator f (single-particle basis, block map f_ir2c) and
 * transforms it to the qp basis, Fqp = W_p^T F W_n (beta-) or W_n^T F W_p (beta+).
 * Replaces setup_extfield + the Fqp part of init_pnfam_solver
 * (pnfam_solver.f90:556-654, 364-372).  Returns nxy through *nxy_out. */
int pnfam_gpu_set_operator(pnfam_gpu_ctx *h, int32_t beta_minus, const int32_t *f_ir2c, const double *f_elem,
                           int64_t *nxy_out);

/* Solves npts frequencies together.  Per point: strength (re_s, im_s), the
 * iteration count in the reference's convention (>=0 converged after that many
 * steps, -1 interrupted at max_iter) and the last max-abs residual.  x_out, if
 * non-null, receives the final mixed amplitudes [npts][ReX, ReY, ImX, ImY].
 * Semantics of pnfam_solver.f90:93-221 (ifam) applied to each point. */
int pnfam_gpu_solve_batch(pnfam_gpu_ctx *h, int32_t npts, const double *re_w, const double *im_w,
                          const pnfam_gpu_opts *o, double *re_s, double *im_s, int32_t *iters, double *si_final,
                          double *x_out);

/* ---- per-unit entry points (one frequency's arrays, host in/out) -------- */

/* Block structure of the four quadrants of drho_sp and of X/Y (0-based maps):
 * maps[0..3][nb] = m11 (rho_pn), m12 (kappa+), m21 (kappa-), m22 (rho_np);
 * xy_maps[0..1][nb] = X (F20 layout), Y (F02 layout). */
int pnfam_gpu_structures(pnfam_gpu_ctx *h, int32_t *maps, int32_t *xy_maps);

/* Fqp = W^T f W for the operator set by pnfam_gpu_set_operator. */
int pnfam_gpu_get_fqp(pnfam_gpu_ctx *h, double *f20, double *f02);

/* qp -> sp transform of the amplitudes: drho_sp = W_p dR_qp W_n^T
 * (pnfam_solver.f90:144-150).  rsp: [8][nxy] re m11, im m11, re m12, ... im m22. */
int pnfam_gpu_qp2sp(pnfam_gpu_ctx *h, const double *x_re, const double *x_im, const double *y_re,
                    const double *y_im, double *rsp);

/* The 16-array calc_hamiltonian with the semantics of calc_dHsp_interface
 * (pnfam_setup.f90:56-71, pnfam_hamiltonian_blas.f90:51-74): in [8][nxy] =
 * re/im rho_pn, kappa+, rho_np, kappa- in the maps of pnfam_gpu_structures;
 * out [8][nxy] = re/im h_pn, Delta+, h_np, Delta- in the maps f, X, f, Y. */
int pnfam_gpu_calc_hamiltonian(pnfam_gpu_ctx *h, const double *in, double *out);

/* sp -> qp transform: dH_qp = W_p^T dH_sp W_n (pnfam_solver.f90:160-166);
 * hsp as the out array of pnfam_gpu_calc_hamiltonian. */
int pnfam_gpu_sp2qp(pnfam_gpu_ctx *h, const double *hsp, double *h20_re, double *h20_im, double *h02_re,
                    double *h02_im);

/* Selects the arithmetic of the per-unit entry points: PNFAM_GPU_FP64
 * (default) or PNFAM_GPU_MIXED, the split-BF16 kernels that the mixed solve
 * applies to its increments (grid stage of pnfam_gpu_calc_hamiltonian and
 * pnfam_gpu_densities in the form PNFAM_GPU_MIX_GRID selects, block transforms
 * of pnfam_gpu_qp2sp and pnfam_gpu_sp2qp).  No counterpart in the reference: a
 * parity hook for the mixed path. */
int pnfam_gpu_set_unit_precision(pnfam_gpu_ctx *h, int32_t precision);

/* Coordinate-space densities of one sector (diagnostic): dens [56][nghl],
 * 28 complex local densities (re, im interleaved by density) after the
 * quadrature weights are stripped (pnfam_hamiltonian_blas.f90:679-709). */
int pnfam_gpu_densities(pnfam_gpu_ctx *h, const double *in, int32_t sector, double *dens);

/* ---- measurement (no counterpart in the reference) ---------------------- */

/* Resets the per-kernel totals and turns HIP-event timing of every launch on
 * the engine stream on (enable != 0) or off. */
int pnfam_gpu_profile(pnfam_gpu_ctx *h, int32_t enable);

/* Per kernel: name (32 bytes each), total milliseconds, launches, processed
 * frequency-units (sum of active frequencies over launches) and the
 * algorithmic FLOP per frequency-unit of that kernel (0 for data movement). */
int pnfam_gpu_profile_read(pnfam_gpu_ctx *h, int32_t max_kernels, char *names, double *ms, int64_t *launches,
                           double *units, double *flops_per_unit, int32_t *n_out);

void pnfam_gpu_destroy(pnfam_gpu_ctx *h);
const char *pnfam_gpu_last_error(void);

/* Build description (arch, precision paths) for logs. */
const char *pnfam_gpu_version(void);

#ifdef __cplusplus
}
#endif
#endif
