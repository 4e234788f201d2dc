/* Host-side definitions of the BdG handle (not installed). */
#ifndef BDG_AMD_HOST_H
#define BDG_AMD_HOST_H

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bdg/bdg.h"
#include "bdg/bdg_amd.h"
#include "bdgx.h"

/* setup-state flags (same order semantics as reference src/bdg_internal.h:25-44) */
enum {
  BDG_HAS_SYSTEM = 1u << 0,
  BDG_HAS_TRAP = 1u << 1,
  BDG_HAS_WF = 1u << 2,
  BDG_HAS_INTERACTIONS = 1u << 3,
  BDG_HAS_DIPOLAR = 1u << 4,
  BDG_HAS_MU = 1u << 5
};

/* Misuse exits with EXIT_FAILURE after a message, as the reference does
 * (reference src/bdg_log.h:29-37). */
#define BDG_FAIL(...)                                   \
  do {                                                  \
    fprintf(stderr, "ERROR [%s:%d] ", __FILE__, __LINE__); \
    fprintf(stderr, __VA_ARGS__);                       \
    fprintf(stderr, "\n");                              \
    exit(EXIT_FAILURE);                                 \
  } while (0)

#define BDG_REQUIRE(b, flag, who)                                          \
  do {                                                                     \
    if (0 == ((b)->state & (flag)))                                        \
      BDG_FAIL("%s: prerequisite %s not satisfied", who, #flag);          \
  } while (0)

#define BDG_FORBID(b, flag, who)                                           \
  do {                                                                     \
    if (0 != ((b)->state & (flag)))                                        \
      BDG_FAIL("%s: %s already set (double-call forbidden)", who, #flag); \
  } while (0)

struct bdg_t {
  /* grid */
  uint64_t dim, N[3];
  double L[3];
  uint64_t size, k_size;
  int complex_psi0;
  uint32_t state;
  double *kx2[3]; /* per-axis k_d^2 (reference src/matmul_ctx.c:102-113) */
  /* physics (host, FP64) */
  double *ltK, *ltM, *pK, *pM;
  void *wf;
  double mu, g_ddi;
  int dipolar;
  /* device */
  bdgx_ctx *dev;
  /* solver parameters (defaults reference src/bdg.c:23-27) */
  uint64_t nev, sizeSub, maxIter;
  double tol;
  int precision;
  double img_tol;
  uint64_t refresh_every;
  float *lp_theta; /* learned preconditioner parameters (NULL: analytic) */
  int lp_window;
  double lp_factor;
  int graphs; /* replay the per-iteration device chain as HIP graphs (default 1) */
  int formulation;    /* bdg_formulation_t                                       */
  double fp32_tol;    /* mixed, first order: residual at which the FP32 phase    */
  int fp32_stall;     /*   stops; columns also stop after this many iterations  */
                      /*   without halving their residual (0: never)            */
  int lock_guards;    /* first order: guard columns after locking              */
  int lock_max;       /* first order: lock converged pairs once at most this   */
                      /*   many columns are active (0: never, < 0: automatic)   */
  double freeze_tol; /* converged columns moved less than this keep their vectors */
  int verbosity;
  /* results */
  double *eigvals;
  void *modes_u, *modes_v;
  int modes_pinned; /* modes_u / modes_v page-locked (from the pinned pool) */
  size_t modes_bytes; /* bytes of each of modes_u / modes_v */
  uint64_t converged;
  double *residuals;
  int *zero_modes;
  bdg_solve_stats_t stats;
  /* start-block strategy (survives reset) */
  bdg_init_mode_t init_mode;
  bdg_init_fn custom_init_fn;
  void *custom_init_param;
  void *reuse_buf;
  uint64_t reuse_n, reuse_cols;
  /* device-resident sweeps: the modes of the last solve kept on the device
   * (S x nev per half, FP64; freed with the results), and the reuse start
   * block built there from them (S x reuse_cols per half; survives reset,
   * consumed by the next solve) */
  void *d_modes_u, *d_modes_v;
  void *d_reuse_u, *d_reuse_v;
  int device_sweeps; /* keep d_modes after a solve (default 1) */
  int sweep_only;    /* mixed, first order: FP64 phase = refinement sweep only */
  int fp64_groups;   /* mixed, first order: FP64 phase in groups of this size  */
};

/* xorshift32 / uniform (0, 1] (reference src/bdg_internal.h:53-65) */
static inline uint32_t bdg_xorshift32(uint32_t *s) {
  uint32_t x = *s;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  *s = x;
  return x;
}
static inline double bdg_xrand(uint32_t *s) {
  return (double)(bdg_xorshift32(s) & 0x7FFFFFFFu) / (double)0x7FFFFFFF;
}

/* start block: fills X (n x sizeSub column-major, f64 or interleaved c64) */
void bdg_fill_start_block(bdg_t *b, void *X);

/* device LOBPCG on (MKM, M); see ilobpcg.c */
typedef struct {
  uint64_t S;
  int m, nev, cplx;
  int maxiter;
  double tol;
  double zero_tol;
  int verbosity;
  int gram_refresh;      /* full Gram recompute every n iterations (0: always)     */
  int w_prec;            /* FP32: mixed policy (T in FP32, images while res high)   */
  double img_tol;        /* mixed: FP32 images of W while the worst residual > this */
  int refresh_every;     /* mixed: FP64 images of X every n iterations meanwhile    */
  int mixed_min_cols;    /* FP32 directions only when at least this many active   */
  int k_equals_m;        /* K == M: standard form K^2 v = w^2 v (B = I)            */
  int learned;           /* W from the learned network (bdgx_lp_apply, FP32)      */
  int lp_window;         /* learned: back to T_M T_K when the worst residual has  */
  double lp_factor;      /*   not dropped below lp_factor x its value lp_window   */
                         /*   iterations earlier                                   */
  int graphs;            /* replay the per-iteration device chain as HIP graphs   */
  double freeze_tol;     /* converged columns whose update is below this keep     */
                         /*   their vectors (0: every column updated)             */
} lobpcg_cfg;

typedef struct {
  double *w2;      /* m */
  int *zflag;      /* m */
  double *res;     /* m */
  int iters, restarts, precond_fallbacks;
  uint64_t applies, applies_fp32, refreshes;
  int learned_its, learned_fallback_at;
  uint64_t graph_launches, graph_captures;
  double capture_seconds; /* host time spent capturing and instantiating graphs */
  uint64_t frozen_skips;  /* block columns the update left unwritten */
  double rr_seconds;
  int status;
} lobpcg_result;

/* Runs from the FP64 device start block dX0 (S x m, ld S).  On success *dX
 * (device, FP64, S x m) holds the final block; the caller frees it with
 * bdgx_free. */
int bdg_lobpcg_device(bdgx_ctx *ctx, const lobpcg_cfg *cfg, const void *dX0, lobpcg_result *res,
                      void **dX);

/* device LOBPCG on the first-order pencil (K u = w v, M v = w u); see lrep.c */
typedef struct {
  uint64_t S;
  int m, nev, cplx;
  int maxiter;
  double tol;          /* a column is done when its residual is below this       */
  int stall_window;    /* ... or (0: never) when its residual has not halved in  */
  double stall_below;  /*   stall_window iterations and is below stall_below     */
  int verbosity;
  int lock_max;        /* lock the converged pairs once at most this many       */
                       /*   columns are active (0: never), see lrep.c           */
  int lock_guards;     /* extra Ritz pairs beyond the block that the small      */
                       /*   block after locking carries as guard columns        */
  const void *dUL, *dVL; /* locked pairs (S x nl, phase precision): new         */
  int nl;                /*   directions are deflated against them              */
  int learned;           /* FP32 phase: W_U = T'_K R_U, W_V = T'_M R_V from the   */
                         /*   learned spectral layers (bdgx_lp_apply_layer)     */
  int lp_window;         /* ... handed back to precondK / precondM for the rest  */
  double lp_factor;      /*   of the solve once the worst residual is above      */
                         /*   lp_factor x its value lp_window iterations earlier */
  void (*on_iterating)(void *); /* called once (may be NULL) when the block's     */
  void *on_iterating_arg;       /*   allocations are done and the iterations start */
} lrep_cfg;

typedef struct {
  double *w2;  /* m (negative: a pair with imaginary w) */
  double *res; /* m */
  int iters;
  uint64_t applies; /* operator applies (K or M) on single columns */
  double rr_seconds;
  double worst;
  int status;
  int learned_its;         /* iterations whose W came from the learned layers */
  int learned_fallback_at; /* iteration of the hand-back, or 0                */
} lrep_result;

/* One phase at storage precision prec, from the blocks dXU0, dXV0 (S x m,
 * ld S, precision prec0); on success the Ritz vectors u, v (u negated for
 * w^2 < 0, so that M v = w u with w = -|w|) are written to dXU, dXV
 * (precision prec0; they may be the input blocks). */
int bdg_lrep_device(bdgx_ctx *ctx, const lrep_cfg *cfg, int prec, const void *dXU0,
                    const void *dXV0, int prec0, lrep_result *res, void *dXU, void *dXV);

double bdg_wtime(void);

#endif
