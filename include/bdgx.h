/*
 * bdgx.h -- thin C ABI over the HIP/gfx950 kernels of the BdG hot path.
 *
 * The reference applies each operator to one host vector at a time through
 * `void op(matmul_ctx_t *ctx, const T *x, T *y)` (reference
 * src/bdg_internal.h:159-215, adapters src/solver.c:26-62).  This layer
 * replaces that with block-wise device applies: one call transforms `ncols`
 * column vectors resident in HBM (column-major, leading dimension `ld`
 * elements) on the context's HIP stream.  Everything is plain C: opaque
 * handles, raw pointers and sizes.
 *
 * Element types follow the context's variant: real psi0 -> real vectors
 * (f64 for BDGX_FP64, f32 for BDGX_FP32); complex psi0 -> interleaved complex
 * (c64 / 2 x f32).  All calls return 0 on success and a negative BDGX_E*
 * code on failure (never exit()).
 */
#ifndef BDGX_H
#define BDGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdgx_ctx bdgx_ctx;

/* precisions */
enum { BDGX_FP64 = 0, BDGX_FP32 = 1 };

/* error codes */
enum {
  BDGX_OK = 0,
  BDGX_EINVAL = -1,   /* bad argument / unsupported shape */
  BDGX_EHIP = -2,     /* HIP runtime error */
  BDGX_ESTATE = -3,   /* required field not uploaded yet */
  BDGX_ENOMEM = -4,
  BDGX_ETOOBIG = -5   /* beyond the fixed scratch of an allocation-free entry */
};

/* operator codes; each replaces the reference function named beside it */
enum {
  BDGX_OP_KINETIC = 0,      /* kinetic_d/_z      src/kinetic_impl.inc:17-30      */
  BDGX_OP_K = 1,            /* matmulK_d/_z      src/matmulK_impl.inc:11-28      */
  BDGX_OP_M = 2,            /* matmulM_d/_z      src/matmulM_impl.inc:14-44      */
  BDGX_OP_LREP = 3,         /* matmulLrep_d/_z   src/matmulLrep_impl.inc:13-18   */
  BDGX_OP_SWAP = 4,         /* matmulSwap_d/_z   src/matmulSwap_impl.inc:14-26   */
  BDGX_OP_PRECOND_K = 5,    /* precondK_d/_z     src/precond_impl.inc:21-65      */
  BDGX_OP_PRECOND_M = 6,    /* precondM_d/_z     src/precond_impl.inc:67-111     */
  BDGX_OP_PRECOND_LREP = 7, /* precondLrep_d/_z  src/precond_impl.inc:113-118    */
  BDGX_OP_DIPOLAR_CONV = 8  /* dipolar_conv_d/_z src/dipolar_conv_impl.inc:22-52 */
};

/* ---- context: replaces matmul_ctx_alloc + matmul_ctx_set_system
 *      (reference src/matmul_ctx.c:21-42, 88-172) ------------------------ */
int bdgx_ctx_create(bdgx_ctx **out, uint64_t dim, const uint64_t *N,
                    const double *L, int complex_psi0);
void bdgx_ctx_destroy(bdgx_ctx *ctx);
uint64_t bdgx_ctx_size(const bdgx_ctx *ctx);   /* prod N            */
uint64_t bdgx_ctx_ksize(const bdgx_ctx *ctx);  /* spectrum length   */
const char *bdgx_error_string(int code);

/* ---- physics fields (host FP64 arrays; NULL leaves a field untouched).
 *      Replaces the ctx fields written by src/setup.c:17-177. ---------- */
int bdgx_set_fields(bdgx_ctx *ctx, const double *ltK, const double *ltM,
                    const void *wf, const double *pK, const double *pM,
                    double mu);

/* The preconditioner weights from the uploaded local terms, on the device:
 * pK = 1 / sqrt(max(floor, ltK + mu)), pM likewise from ltM (FP64, and the
 * FP32 copies converted from them); mu as given to bdgx_set_fields.  The
 * same values as host arrays passed as pK, pM (IEEE square root and
 * division), without their computation and upload on the host. */
int bdgx_derive_precond(bdgx_ctx *ctx, double floor);

/* Mark every physics field absent (allocations are kept); the next uploads
 * re-define the problem.  Device side of bdg_reset (reference src/bdg.c:55-84). */
int bdgx_reset_fields(bdgx_ctx *ctx);

/* Dipolar kernel on the device (reference src/dipolar.c:21-72). */
int bdgx_dipolar_set_kernel(bdgx_ctx *ctx, double g_ddi, const double *dir,
                            double cutoff_radius);
int bdgx_dipolar_get_kernel(bdgx_ctx *ctx, double *h_out /* ksize */);
/* g_ddi * Phi_dd into h_out[size] (reference src/dipolar_meanfield_impl.inc:20-55);
 * requires the wavefunction and the kernel. */
int bdgx_dipolar_meanfield(bdgx_ctx *ctx, double *h_out);

/* ---- operator applies ------------------------------------------------- */
/* Device block apply: dY[:, j] = op(dX[:, j]), j < ncols, on the ctx stream.
 * Column length is 2*size for LREP/SWAP/PRECOND_LREP, size otherwise. */
int bdgx_apply(bdgx_ctx *ctx, int op, int prec, const void *dX, void *dY,
               uint64_t ncols, uint64_t ld);
/* Host convenience (uploads, applies in FP64, downloads, synchronises). */
int bdgx_apply_host(bdgx_ctx *ctx, int op, const void *hX, void *hY,
                    uint64_t ncols, uint64_t ld);

/* ---- dense block kernels (tall-skinny, rows = nrows) ------------------- */
/* G (p x q, column-major FP64 or complex FP64, ldg = p) = A^H B, computed
 * with FP64 accumulation.  swap_b != 0 reads B with its two halves of
 * nrows/2 rows exchanged (gives A^H (Swap B) without materialising it). */
int bdgx_gram(bdgx_ctx *ctx, int prec, const void *dA, uint64_t lda, uint64_t p,
              const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows,
              int swap_b, double *dG);
/* bdgx_gram without allocating (no counterpart in the reference: an entry
 * point for callers that cannot allocate device memory -- code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and result as bdgx_gram, bit for bit; the per-chunk
 * partial sums live in one fixed device array of
 * BDGX_GRAM_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends)
 * instead of the context's scratch.  The partials of one call need
 * 8 * (chunks + ceil(chunks / 64)) * p * q * (2 if complex) bytes, chunks =
 * twice the number of compute units for tall real blocks with p <= 96 and
 * 32 < q <= 96 (the single-pass kernel), else the number of compute units
 * (256 on MI355X: p * q up to 8065 real with p <= 96 and 32 < q <= 96,
 * 16131 other real, 8065 complex); larger calls return BDGX_ETOOBIG.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (other
 * contexts or streams) race; calls in one stream order are safe.  Needs a
 * context created on that device (bdgx_ctx_create looks the array up). */
#define BDGX_GRAM_NOALLOC_SCRATCH_BYTES (32u << 20)
int bdgx_gram_noalloc(bdgx_ctx *ctx, int prec, const void *dA, uint64_t lda, uint64_t p,
                      const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b,
                      double *dG);
/* Two Grams with one left block read once (no reference counterpart; the
 * solver's V^H V_W and V^H M V_W, src/solver.c:262 through the external
 * iLOBPCG's Gram step): dG (p x 2q, column-major) = dA^H [dB1 | dB2], dB1
 * and dB2 q columns each.  Real contexts, 16-byte aligned blocks with leading
 * dimensions a multiple of 4 (else BDGX_EINVAL).  Each entry equals the
 * corresponding entry of the separate bdgx_gram calls bit for bit. */
int bdgx_gram_pair(bdgx_ctx *ctx, int prec, const void *dA, uint64_t lda, uint64_t p,
                   const void *dB1, uint64_t ldb1, const void *dB2, uint64_t ldb2, uint64_t q,
                   uint64_t nrows, double *dG);
/* dY[:, 0:m] = dA[:, 0:k] * C  (C is k x m column-major FP64 / complex FP64
 * in device memory, ldc = k); dY must not alias dA. */
int bdgx_combine(bdgx_ctx *ctx, int prec, const void *dA, uint64_t lda,
                 uint64_t k, const double *dC, void *dY, uint64_t ldy,
                 uint64_t m, uint64_t nrows);
/* dY[:, 0:m] = [dA1 | dA2] * C: k1 columns of dA1 (ld lda1) followed by k2
 * columns of dA2 (ld lda2); C is (k1 + k2) x m column-major in device memory
 * (no reference counterpart: the first-order solver's residuals
 * [K X_U | X_V] [R_A; -R_B D] and its orthonormalisation of new directions
 * against [X | P], which live in different blocks).  dY must alias neither
 * source.  k2 = 0 is bdgx_combine. */
int bdgx_combine2(bdgx_ctx *ctx, int prec, const void *dA1, uint64_t lda1, uint64_t k1,
                  const void *dA2, uint64_t lda2, uint64_t k2, const double *dC, void *dY,
                  uint64_t ldy, uint64_t m, uint64_t nrows);
#define BDGX_COMBINE_COLS_MAX 8
/* dY[:, cols[j]] = dA[:, 0:k] * C[:, j] for j < ncols <= BDGX_COMBINE_COLS_MAX
 * (C is k x ncols
 * in device memory, ldc = k; cols a host array of distinct column indices):
 * the update of a block in which most columns keep their vectors, written
 * only where it changes.  FP64 only; the sum over the k source columns runs
 * in a fixed order and a unit coefficient column copies its source column
 * exactly.  dY must not alias dA. */
int bdgx_combine_cols(bdgx_ctx *ctx, int prec, const void *dA, uint64_t lda,
                      uint64_t k, const double *dC, void *dY, uint64_t ldy,
                      const int *cols, uint64_t ncols, uint64_t nrows);
/* R[:, j] = AX[:, j] - lambda_j * X'[:, j] with X' = Swap(X) when swap_x,
 * else X.  dnorms2 (3 * ncols FP64, device) receives ||R_j||^2, ||X_j||^2 and
 * ||MX_j||^2 (the last only when dMX != NULL, else 0).  dR may be NULL (norms
 * only).  dAX, dX and dMX share the leading dimension ld. */
int bdgx_residual(bdgx_ctx *ctx, int prec, const void *dAX, const void *dX,
                  const void *dMX, uint64_t ld, uint64_t ncols, uint64_t nrows,
                  const double *dlambda, int swap_x, void *dR, uint64_t ldr,
                  double *dnorms2);
/* bdgx_residual on the columns cols[0..ncols) (a host array) of dAX, dX, dMX
 * only: output column j of dR and norms j of dnorms2 belong to source column
 * cols[j], with eigenvalue dlambda[j]. */
int bdgx_residual_cols(bdgx_ctx *ctx, int prec, const void *dAX, const void *dX,
                       const void *dMX, uint64_t ld, const int *cols, uint64_t ncols,
                       uint64_t nrows, const double *dlambda, int swap_x, void *dR,
                       uint64_t ldr, double *dnorms2);
/* dDst[:, j] = dSrc[:, idx[j]] for j < ncols (same precision, one launch);
 * idx is a host array.  dDst must not overlap the gathered source columns. */
int bdgx_gather_cols(bdgx_ctx *ctx, int prec, const void *dSrc, uint64_t lds,
                     const int *idx, uint64_t ncols, void *dDst, uint64_t ldd,
                     uint64_t nrows);

/* Uniform start block on the device (no reference counterpart: the values
 * of bdg_init_uniform_xorshift, include/bdg/bdg_amd.h, i.e. the reference's
 * xorshift32 + xrand, src/bdg_internal.h:53-65): element (i, j), i < nrows,
 * of the FP64 block dX (ld elements; complex ctx: interleaved, imaginary 0)
 * is xrand value number j * steps_per_col + i + 1 of the stream seeded with
 * `seed`, minus 0.5.  Columns start at jumped-ahead states (xorshift32 is
 * linear over GF(2)), so the block is generated in parallel yet equals the
 * sequential host fill bit for bit. */
int bdgx_fill_uniform_xorshift(bdgx_ctx *ctx, void *dX, uint64_t ld, uint64_t nrows,
                               uint64_t ncols, uint64_t steps_per_col, uint32_t seed);

/* Reuse start block of a parameter sweep, built on the device from the
 * device-resident modes of the last solve (replaces the host loops of
 * reference src/bdg.c:112-207, bdg_reuse_modes).  FP64 blocks, S = nrows
 * rows per half, complex ctx: interleaved.  Column j < nev:
 *   OU_j = U_j + sigma_j z0,  OV_j = V_j + sigma_j z1,
 *   sigma_j = noise_frac * ||[U_j; V_j]||,
 * (z0, z1) the Box-Muller pair sqrt(-2 ln max(u1, 1e-10)) (cos, sin)(2 pi u2)
 * of the xrand values u1, u2 number 2 (j S + i) + 1, + 2 of the xorshift32
 * stream seeded with `seed` (the reference uses 12345); complex: the noise
 * goes to the real parts.  Columns nev <= j < ncols: both halves
 * |z0| * |psi0_i| (|psi0| of the context's wavefunction, 1 where none is
 * set).  Values equal the sequential host recipe up to the rounding of the
 * device log / cos / sin and of the norm's summation order. */
int bdgx_reuse_block(bdgx_ctx *ctx, const void *dU, const void *dV, uint64_t ldm, uint64_t nev,
                     void *dOU, void *dOV, uint64_t ldo, uint64_t ncols, uint64_t nrows,
                     double noise_frac, uint32_t seed);

/* Element-type conversion FP64 <-> FP32 of an nrows x ncols block. */
int bdgx_convert(bdgx_ctx *ctx, int prec_src, const void *dsrc, uint64_t lds,
                 int prec_dst, void *ddst, uint64_t ldd, uint64_t nrows,
                 uint64_t ncols);

/* ---- psi0 by imaginary-time split-step (no reference counterpart: the
 *      reference only loads psi0, src/io.c:13-50).  Same step definition as
 *      the oracle's orc_itp_groundstate. ------------------------------- */
int bdgx_itp_groundstate(bdgx_ctx *ctx, const double *omega, double g,
                         double g_dd, const double *dir, double cutoff,
                         double dt, uint64_t nsteps, void *h_psi /* in/out */,
                         double *mu_out);

/* ---- memory, streams, timing ------------------------------------------ */
/* Device memory.  bdgx_free waits for the device (as hipFree does) and keeps
 * the block for the next bdgx_malloc of the same size on the same device
 * (up to 64 GiB kept per device; released when hipMalloc runs short). */
int bdgx_malloc(void **dptr, size_t bytes);
int bdgx_free(void *dptr);
/* Free and total device memory of the current device (hipMemGetInfo), the
 * blocks kept by bdgx_free counted as free. */
int bdgx_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Page-locked host memory (hipHostMalloc / hipHostFree): copies between it
 * and the device run at the full link rate. */
int bdgx_host_alloc(void **hptr, size_t bytes);
int bdgx_host_free(void *hptr);
/* Copies on the ctx stream; both return after the copy has completed. */
int bdgx_memcpy_h2d(bdgx_ctx *ctx, void *dst, const void *src, size_t bytes);
int bdgx_memcpy_d2h(bdgx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Two device-to-host copies of `bytes` each, after the work queued on the
 * ctx stream, on two streams at once (two copy engines); returns after both
 * have completed. */
int bdgx_memcpy_d2h_pair(bdgx_ctx *ctx, void *dst1, const void *src1, void *dst2,
                         const void *src2, size_t bytes);
/* Host-to-device copy queued on the ctx stream without waiting: src must be
 * page-locked (bdgx_host_alloc) and must not change until the stream has
 * passed the copy (a later bdgx_sync or synchronous copy on this ctx). */
int bdgx_memcpy_h2d_async(bdgx_ctx *ctx, void *dst, const void *src, size_t bytes);
int bdgx_memset(bdgx_ctx *ctx, void *dst, int value, size_t bytes);
int bdgx_sync(bdgx_ctx *ctx);
void *bdgx_stream(bdgx_ctx *ctx);          /* the ctx's hipStream_t */
/* HIP-event timer on the ctx stream: begin records, end records + waits and
 * returns the elapsed milliseconds. */
int bdgx_timer_begin(bdgx_ctx *ctx);
int bdgx_timer_end(bdgx_ctx *ctx, float *ms);
/* Bytes of device workspace the context currently holds. */
uint64_t bdgx_workspace_bytes(const bdgx_ctx *ctx);

/* Measurement hook (no reference counterpart): launches one FFT pass kind
 * `iters` times over `nbatch` spectrum batches (one kinetic channel per
 * batch) between HIP events on the ctx stream and reports the mean launch
 * time and the pass's algorithmic HBM bytes per launch.
 *   pass: 0 x-forward (r2c/c2c), 1 y-forward, 2 z fused (fwd, multiply, inv),
 *         3 y-inverse, 4 x-inverse + combine, 5 whole-plane x+y forward,
 *         6 whole-plane y+x inverse + combine (5 and 6 only for the grids
 *         that use them: real psi0, Nx = Ny in {32, 64, 128}; else EINVAL). */
enum { BDGX_PASS_XFWD = 0, BDGX_PASS_YFWD = 1, BDGX_PASS_Z = 2, BDGX_PASS_YINV = 3,
       BDGX_PASS_XINV = 4, BDGX_PASS_PLANE_FWD = 5, BDGX_PASS_PLANE_INV = 6 };
int bdgx_bench_pass(bdgx_ctx *ctx, int pass, int prec, uint64_t nbatch, int iters,
                    float *ms_per_launch, double *bytes_per_launch);

/* Launch-shape autotuner (no reference counterpart; it takes the place of
 * the reference's FFTW_PATIENT planning, src/matmul_ctx.c:134-171).  The
 * fixed-size FFT kernels (whole-plane forward / inverse, fixed z pass) and
 * the block combine have launch variants: plane 0 / 1 = one workgroup per
 * plane with 4 Nx / 8 Nx threads, 2 / 3 = as many workgroups as fit on the
 * device, each walking several planes, with 4 Nx / 8 Nx threads; z 0 = 256
 * (Nz = 128) or 128 threads, 1 = half of that; combine 0 = 4 rows per lane,
 * 1 = 2 rows per lane, 2 / 3 = persistent workgroups that stage their tile
 * of C in LDS once (2 / 4 rows per lane; where it fits in 64 KB and the
 * blocks are vector-aligned, else as 1 / 0); real Gram 0 / 1 / 2 / 3 =
 * tiles of up to 96 / 48 / 32 / 64 rows of p.  bdgx_autotune times every variant of
 * every kind for `prec` on `nbatch` columns (kinetic batches for the FFT
 * kinds, a 3 nbatch -> 2 nbatch combine and a 3 nbatch x nbatch Gram, the
 * solver's main shapes; 1 warm-up + 3 timed
 * launches each, HIP events on the ctx stream) and keeps the fastest; it
 * returns at once when `prec` is already tuned and `force` is 0.  Choices
 * are recorded process-wide per device, grid, psi0 kind, precision and
 * nbatch: with `force` 0 a new context of a recorded shape takes them without
 * timing.  The FFT kinds are skipped on grids that use the pencil passes.
 * The variants compute identical results (same operations in the same
 * order). */
enum { BDGX_TUNE_PLANE_FWD = 0, BDGX_TUNE_PLANE_INV = 1, BDGX_TUNE_Z = 2,
       BDGX_TUNE_COMBINE = 3, BDGX_TUNE_GRAM = 4 };
int bdgx_autotune(bdgx_ctx *ctx, int prec, uint64_t nbatch, int force);
int bdgx_get_launch_variant(const bdgx_ctx *ctx, int kind, int prec, int *variant);
int bdgx_set_launch_variant(bdgx_ctx *ctx, int kind, int prec, int variant);

/* Graph capture of device work on the ctx stream (no reference counterpart;
 * used by the solver to replay its per-iteration operator / Gram chain).
 * Between bdgx_graph_begin and bdgx_graph_end the bdgx calls on this ctx are
 * recorded, not run; calls that allocate, free, synchronise or copy through
 * host memory are not allowed in between (they fail, and the capture must
 * then be ended with out == NULL, which discards it).  The workspace must
 * not need to grow during capture: reserve it first (bdgx_reserve).  A graph
 * replays exactly the recorded kernels on the recorded pointers; launching
 * it after the ctx workspace has moved returns BDGX_ESTATE. */
typedef struct bdgx_graph bdgx_graph;
int bdgx_graph_begin(bdgx_ctx *ctx);
int bdgx_graph_end(bdgx_ctx *ctx, bdgx_graph **out);
int bdgx_graph_launch(bdgx_ctx *ctx, const bdgx_graph *g);
void bdgx_graph_destroy(bdgx_graph *g);
/* Grow the spectrum workspace and the small scratch to at least the given
 * sizes now (so that later calls of up to that size do not reallocate). */
int bdgx_reserve(bdgx_ctx *ctx, uint64_t spec_bytes, uint64_t scratch_bytes);
/* Spectrum bytes per column chunk of an operator apply (0 = 6 GiB, the
 * default).  A chunk's three passes run back to back, so a chunk small
 * enough for the 256 MB Infinity Cache keeps its spectrum there between the
 * forward, z and inverse passes. */
int bdgx_set_chunk_bytes(bdgx_ctx *ctx, uint64_t bytes);

/* ---- learned preconditioner (no reference counterpart; its analytic base
 *      is precondK/precondM, reference src/precond_impl.inc:21-111) --------
 *   N(r) = T'_M(V),  V = U + s0 sig sum_h v_h tanh(sum_nb w_{h,nb} U(x+nb)/(s0 sig) + d_h),
 *   U = T'_K(r),  sig = rms(r) per column,  s0 = mean(p_K^2) / safe(mu),
 *   T'_X v = a_X F^-1[g_X F(b_X v)],  a_X = p_X exp(alpha_X.phi),
 *   b_X = p_X exp(beta_X.phi),  phi = (rho, rho^2),  rho = psi0^2 / max psi0^2,
 *   g_X(k) = exp(interp(kappa_X, |k|)) / (S safe(mu + k^2/2))
 * theta (float, BDGX_LP_NPARAM): alpha_K[2] beta_K[2] alpha_M[2] beta_M[2]
 * kappa_K[NKNOT] kappa_M[NKNOT] w[HIDDEN][27] d[HIDDEN] v[HIDDEN]; the 27
 * stencil offsets in the order (dz+1)*9 + (dy+1)*3 + (dx+1), periodic.
 * Real psi0 only; FP32; columns with leading dimension size. */
enum { BDGX_LP_NKNOT = 16, BDGX_LP_HIDDEN = 16,
       BDGX_LP_NPARAM = 8 + 2 * BDGX_LP_NKNOT + 29 * BDGX_LP_HIDDEN };
/* Set the parameters (needs psi0 and the preconditioner weights uploaded). */
int bdgx_lp_set_params(bdgx_ctx *ctx, const float *theta);
/* dW = N(dR) for ncols FP32 columns (prec must be BDGX_FP32, dW != dR). */
int bdgx_lp_apply(bdgx_ctx *ctx, int prec, const void *dR, void *dW, uint64_t ncols,
                  uint64_t ld);
/* h_grad = d/dtheta sum(dGW * N(dR)) (device forward + adjoint, FP64 sums). */
int bdgx_lp_backward(bdgx_ctx *ctx, const float *dR, const float *dGW, uint64_t ncols,
                     uint64_t ld, float *h_grad);
/* One spectral layer alone, the first-order solver's per-half preconditioner
 * (the learned counterpart of precondK / precondM applied to R_U / R_V):
 * dW = T'_X dR, layer 0 = K, 1 = M (FP32, dW != dR). */
int bdgx_lp_apply_layer(bdgx_ctx *ctx, int layer, const void *dR, void *dW, uint64_t ncols,
                        uint64_t ld);
/* h_grad = d/dtheta sum(dGW * T'_X(dR)): nonzero only in alpha_X, beta_X,
 * kappa_X. */
int bdgx_lp_backward_layer(bdgx_ctx *ctx, int layer, const float *dR, const float *dGW,
                           uint64_t ncols, uint64_t ld, float *h_grad);

#ifdef __cplusplus
}
#endif
#endif /* BDGX_H */
