/*
 * bdg_amd.h -- extensions of the BdG API that have no counterpart in the
 * reference header (include/bdg/bdg.h).  A program that only uses bdg.h
 * never needs this file; it gets the plain FP64 device solver.
 */
#ifndef BDG_AMD_H
#define BDG_AMD_H

#include "bdg/bdg.h"
#include "bdgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Solver precision policy. */
typedef enum {
  BDG_PREC_FP64 = 0,  /* every operator apply and block update in FP64 (default)   */
  BDG_PREC_MIXED = 1  /* FP64 basis, Grams, residuals, images and Rayleigh-Ritz;
                         the preconditioner in FP32 (optionally also the images
                         of new directions while the residual is large, see
                         bdg_set_mixed_params) */
} bdg_precision_t;

void bdg_set_precision(bdg_t *bdg, bdg_precision_t prec);

/* Form of the eigen-iteration.
 *   BDG_FORM_FIRST_ORDER (default): LOBPCG on the reference's first-order
 *     pencil K u = w v, M v = w u (reference src/solver.c:155-164: A = diag(K,
 *     M), B = Swap, T = [T_K; T_M]) with separate, orthonormal search spaces
 *     for u and v.  With BDG_PREC_MIXED: an FP32 phase (FP32 basis, images,
 *     operator applies; FP64 Grams and Rayleigh-Ritz) until the FP32 residual
 *     target or the FP32 rounding floor, then an FP64 phase from its vectors
 *     (FP64 sweep, FP64 iterations to tol if the sweep's residuals need them).
 *   BDG_FORM_PRODUCT: LOBPCG on (M K M) v = w^2 M v for the v half (the
 *     form used when K == M, and by the learned preconditioner).
 * Both give the same eigenpairs. */
typedef enum { BDG_FORM_FIRST_ORDER = 0, BDG_FORM_PRODUCT = 1 } bdg_formulation_t;
void bdg_set_formulation(bdg_t *bdg, bdg_formulation_t form);

/* Mixed precision, first-order form: the FP32 phase stops a column at
 * residual max(tol, fp32_tol) (default 1e-5), or when its residual has not
 * halved in stall_window iterations (default 40; 0: never) once below 1e-4. */
void bdg_set_fp32_phase(bdg_t *bdg, f64 fp32_tol, uint64_t stall_window);

/* First-order form: once at most max_active columns are still active, the
 * converged pairs are locked and the rest continue in a small block deflated
 * against them (default -1: max(2, sizeSub / 4); 0: never lock). */
void bdg_set_locking(bdg_t *bdg, int max_active);
/* First-order formulation: the small block that continues after locking also
 * carries this many guard columns, started from the Ritz pairs just beyond
 * the block (default 4).  They are iterated but need not converge and are
 * not returned; the last wanted pairs then converge at the rate set by the
 * gap to the first pair beyond the guards instead of the first beyond the
 * block. */
void bdg_set_lock_guards(bdg_t *bdg, int guards);

/* Device-resident parameter sweeps (default on): the modes of a solve stay
 * in device memory beside the host results (2 x S x nev values), so that
 * bdg_reuse_modes builds the reuse start block on the device and the next
 * bdg_solve starts from it without a host round trip (the sweep of reference
 * src/bdg.c:55-207).  Off: the modes live on the host only and the reuse
 * block is built there, as in the reference. */
void bdg_set_device_sweeps(bdg_t *bdg, int on);

/* Mixed precision, first-order form: the FP64 phase is the refinement sweep
 * alone (Rayleigh-Ritz on the FP64 images of the FP32 vectors and the FP64
 * residuals, no iterations).  Where the block's FP64 basis does not fit in
 * device memory (256^3 at block 128) the FP64 phase iterates the modes in
 * groups instead (bdg_solve_stats_t.fp64_group_size), and takes the sweep
 * alone only where not even a group fits; on = 1 forces the sweep. */
void bdg_set_fp64_sweep_only(bdg_t *bdg, int on);
/* Mixed precision, first-order form: run the FP64 phase in groups of
 * `group` modes (each deflated against the groups before it) even where the
 * whole block fits; 0 (default): only where it does not fit. */
void bdg_set_fp64_groups(bdg_t *bdg, int group);

/* Binary fast path beside the reference's text formats (reference
 * src/io.c:13-165 writes ~30 bytes of text per value): a 64-byte header
 * (magic "BDGBIN1", kind 1 wavefunction / 2 modes, complex flag, dim, N[3],
 * count) and raw little-endian FP64 values, bit-exact.  Modes files hold the
 * nev eigenvalues, then modes_u and modes_v (S x nev each, column-major).
 * Codes as the text functions: 0, -1 (file), -2 (format, or header not
 * matching this grid / variant), -3 (nothing to write). */
int bdg_save_wavefunction_bin(const bdg_t *bdg, const char *filename);
int bdg_load_wavefunction_bin(bdg_t *bdg, const char *filename);
int bdg_write_modes_bin(const bdg_t *bdg, const char *filename);

typedef struct {
  uint64_t iterations;          /* LOBPCG iterations                           */
  uint64_t column_applies;      /* (M then K) column applies, incl. start      */
  uint64_t column_applies_fp32; /* of which in FP32 (mixed policy)             */
  uint64_t image_refreshes;     /* FP64 re-applications of M, K to X (mixed)   */
  uint64_t restarts;            /* iterations that left the P block out        */
  uint64_t learned_iterations;  /* iterations whose W came from the network    */
  int learned_fallback_at;      /* iteration of the switch back to T_M T_K, or 0 */
  uint64_t graph_launches;      /* iterations whose device chain was a graph replay */
  uint64_t graph_captures;      /* graphs captured during the solve            */
  double seconds_total;         /* bdg_solve wall clock                        */
  double seconds_rr;            /* host Rayleigh-Ritz time                     */
  int status;                   /* 0 ok, else a BDGX_E* code or -100 (RR)     */
  uint64_t frozen_skips;        /* block columns the update left unwritten     */
  uint64_t iterations_fp32;     /* of the iterations, those of the FP32 phase  */
  double fp32_worst_residual;   /* worst residual when the FP32 phase ended    */
  int fp64_sweep_only;          /* mixed, first order: the FP64 phase was one  */
                                /* refinement sweep (Rayleigh-Ritz and         */
                                /* residuals on the FP32 vectors) because its  */
                                /* full basis did not fit in device memory     */
  int fp64_group_size;          /* mixed, first order: the FP64 phase refined  */
                                /* the modes in groups of this many (each     */
                                /* deflated against the ones before) because   */
                                /* the block's FP64 basis did not fit; 0: no  */
} bdg_solve_stats_t;

void bdg_get_solve_stats(const bdg_t *bdg, bdg_solve_stats_t *out);

/* Final relative residuals ||A x - w B x|| / (||x|| max(1,|w|)), nev values. */
const f64 *bdg_residuals(const bdg_t *bdg);

/* 1 for modes identified as zero modes (Goldstone), nev values. */
const int *bdg_zero_modes(const bdg_t *bdg);

/* Device context of the handle (valid after bdg_set_system). */
bdgx_ctx *bdg_device_context(bdg_t *bdg);

/* Host copies of the setup arrays (size entries; NULL before they exist). */
const f64 *bdg_local_term_K(const bdg_t *bdg);
const f64 *bdg_local_term_M(const bdg_t *bdg);
const f64 *bdg_precond_weight_K(const bdg_t *bdg);
const f64 *bdg_precond_weight_M(const bdg_t *bdg);
f64 bdg_mu(const bdg_t *bdg);

/* Mixed policy: new search directions get FP32 images while the worst
 * relative residual is above img_tol (default 0: never); meanwhile the images
 * of X are re-applied in FP64 every refresh_every iterations (default 20, 0:
 * only when the residual reaches img_tol). */
void bdg_set_mixed_params(bdg_t *bdg, f64 img_tol, uint64_t refresh_every);

/* Learned preconditioner for the search directions (real psi0): theta holds
 * BDGX_LP_NPARAM floats (layout in bdgx.h), copied; NULL switches back to the
 * analytic T_M T_K.  The solver falls back to T_M T_K for the rest of a solve
 * when the worst residual has not dropped below factor x its value `window`
 * iterations earlier (bdg_set_learned_fallback; default 30 and 0.5, window 0:
 * never). */
void bdg_set_learned_preconditioner(bdg_t *bdg, const float *theta);
void bdg_set_learned_fallback(bdg_t *bdg, int window, f64 factor);

/* 0 silent (default), 1 per-phase summary, 2 per-iteration trace (stderr). */
/* Replay the per-iteration device chain (preconditioner, images, Gram
 * columns) as captured HIP graphs (default on; 0 runs every launch eagerly).
 * Results are the same either way. */
void bdg_set_graph_capture(bdg_t *bdg, int on);

/* Converged columns whose Rayleigh-Ritz update moves them by less than tol
 * (relative, Euclidean, up to a phase) keep their vectors, and the block
 * update writes only the columns that change (default 1e-6; 0 updates every
 * column every iteration, as plain soft locking). */
void bdg_set_freeze(bdg_t *bdg, f64 tol);

/* Solve n independent, fully set-up problems concurrently: worker threads
 * (nthreads, 0 = one per problem) take problems in order and call bdg_solve
 * on each; every handle has its own HIP stream, so the device overlaps the
 * problems' kernels and the hosts' Rayleigh-Ritz steps run in parallel.
 * Per-problem return codes go to rc[n] (may be NULL); returns 0 when all
 * solves returned 0, else the first nonzero code.  Handles must be
 * distinct; results are read per handle as after bdg_solve. */
int bdg_solve_batch(bdg_t **probs, int n, int nthreads, int *rc);

void bdg_set_verbosity(bdg_t *bdg, int level);

/* Built-in host callbacks usable with bdg_set_trap / bdg_set_local_interactions
 * (for callers that cannot pass native function pointers cheaply):
 *   bdg_cb_harmonic_trap: param = const f64 omega[dim], V = 0.5 sum (w_d r_d)^2
 *   bdg_cb_contact:       param = const f64 *g,         U = g * density      */
f64 bdg_cb_harmonic_trap(uint64_t dim, const f64 *r, const void *param);
f64 bdg_cb_contact(const void *param, f64 density);

/* Fill X (n x sizeSub column-major, f64 or c64) with the start block the next
 * bdg_solve would use for the current init mode (does not consume a REUSE
 * buffer).  Needs bdg_set_system. */
void bdg_start_block(bdg_t *bdg, void *X);

/* Built-in bdg_init_fn for BDG_INIT_CUSTOM: u = v = xrand - 0.5 from
 * xorshift32 (reference src/bdg_internal.h:53-65), column by column, with
 * param = const uint32_t *seed.  When a problem's start block is this
 * callback, bdg_solve generates its v half directly on the device
 * (bdgx_fill_uniform_xorshift: the same values) instead of calling it. */
void bdg_init_uniform_xorshift(const bdg_t *bdg, void *X, uint64_t n, uint64_t sizeSub,
                               void *param);

#ifdef __cplusplus
}
#endif
#endif
