/*
 * Device LOBPCG on the first-order BdG pencil (host driver in C).
 *
 * The reference hands A = diag(K, M), B = Swap and T = [T_K; T_M] to an
 * external iLOBPCG (reference src/solver.c:147-164, 262).  This solver works
 * on that first-order form, K u = w v, M v = w u, with separate search
 * spaces for the two halves (locally optimal block CG of Bai and Li for the
 * linear-response eigenproblem):
 *
 *   U half: Z_U = [X_U | P_U | W_U] (S x k_U) and its image K Z_U;
 *   V half: Z_V = [X_V | P_V | W_V] (S x k_V) and its image M Z_V;
 *   residuals R_U = K u - s|w| v, R_V = M v - |w| u of the Ritz pairs
 *   u = X_U R_A, v = X_V R_B (s = sign w^2); W_U = T_K R_U, W_V = T_M R_V;
 *   Rayleigh-Ritz from G_K = Z_U^H K Z_U, G_M = Z_V^H M Z_V, G_W = Z_U^H Z_V
 *   (dn_lrep_rr), FP64 on the host.
 *
 * Each half's basis is kept (nearly) orthonormal: new directions W are
 * orthonormalised against [X | P] on the device before their images are
 * formed (two passes, dn_orth_coef), and X, P are updated with orthonormal
 * coefficients [Q1 | Q2] (dn_lrep_update: A = Q1 R_A, Q2 the new P
 * directions).  No update then multiplies a stored vector or image by a
 * large coefficient, so the rounding of FP32 storage and FP32 images is
 * never amplified: the same iteration runs with FP32 basis, images and
 * operator applies (Grams accumulate in FP64).
 *
 * Locking.  Once at most cfg->lock_max columns are still active, the
 * converged Ritz pairs (u_c, v_c) are locked and the active ones continue in
 * a small block (a nested call with those pairs as its deflation set).
 * Eigenvectors of the pencil with different w^2 are biorthogonal
 * (u_i^H v_j = 0: the u are left and the v right eigenvectors of K M), so
 * the remaining pairs lie in {u : V_L^H u = 0} x {v : U_L^H v = 0}; the new
 * directions are projected there obliquely,
 *   W_U <- W_U - U_L (V_L^H U_L)^-1 V_L^H W_U,
 *   W_V <- W_V - V_L (U_L^H V_L)^-1 U_L^H W_V,
 * and the tail iterations touch only the active columns and one pass over
 * the locked ones, instead of the whole block.
 *
 * One call is one phase at one storage precision, started from given X
 * blocks; the mixed policy runs an FP32 phase to the residual FP32 can
 * reach, then an FP64 phase started from its X blocks, whose first
 * Rayleigh-Ritz (FP64 images of the FP32 vectors) is the FP64 refinement
 * sweep and whose iterations, if any, take the residual to the tolerance.
 */
#include <complex.h>

#include "bdg_host.h"
#include "dense.h"

#define TRY(x)         \
  do {                 \
    rc = (x);          \
    if (rc) goto done; \
  } while (0)

static size_t lr_es(int prec, int cplx) {
  return (size_t)(prec == BDGX_FP64 ? 8 : 4) * (cplx ? 2 : 1);
}

static char *lr_col(void *base, uint64_t col, uint64_t S, size_t es) {
  return (char *)base + (size_t)col * S * es;
}

typedef struct {
  void *Z[2], *Y[2]; /* basis and its image, two generations */
  int np;            /* P columns of the current generation */
} lr_half;

/* host copy of a device Gram */
static int lr_gram(bdgx_ctx *ctx, int prec, const void *A, uint64_t p, const void *Bm, uint64_t q,
                   uint64_t S, double *dG, double *hG, int w) {
  int rc = bdgx_gram(ctx, prec, A, S, p, Bm, S, q, S, 0, dG);
  if (!rc) rc = bdgx_memcpy_d2h(ctx, hG, dG, sizeof(double) * p * q * w);
  return rc;
}

/* Queued uploads of small host matrices (coefficients) through page-locked
 * staging: each piece is copied into the next free part of the staging
 * region and its copy queued on the stream, so the host never waits for an
 * upload.  The region is reused once the stream has drained: every download
 * (lr_get) and every bdgx_sync drains it. */
typedef struct {
  double *h;  /* page-locked */
  size_t cap; /* doubles */
  size_t off;
} lr_stage;

static int lr_put(bdgx_ctx *ctx, lr_stage *st, double *dDst, const double *src, size_t n) {
  if (!st->h || n > st->cap) return bdgx_memcpy_h2d(ctx, dDst, src, sizeof(double) * n);
  if (st->off + n > st->cap) {
    const int rc = bdgx_sync(ctx);
    if (rc) return rc;
    st->off = 0;
  }
  double *p = st->h + st->off;
  memcpy(p, src, sizeof(double) * n);
  st->off += (n + 31) & ~(size_t)31;
  return bdgx_memcpy_h2d_async(ctx, dDst, p, sizeof(double) * n);
}

static int lr_get(bdgx_ctx *ctx, lr_stage *st, double *hDst, const double *dSrc, size_t n) {
  const int rc = bdgx_memcpy_d2h(ctx, hDst, dSrc, sizeof(double) * n);
  st->off = 0;
  return rc;
}

/* C = [-Binv G; I] ((nl + nw) x nw) from G = Lp^H W (nl x nw): the
 * coefficients of the oblique deflation W - Lo Binv (Lp^H W) over [Lo | W] */
static void lr_deflate_coef(int nl, int nw, int cplx, const double *Binv, const double *G,
                            double *C) {
  const int w = cplx ? 2 : 1;
  memset(C, 0, sizeof(double) * (size_t)(nl + nw) * nw * w);
  for (int j = 0; j < nw; ++j) {
    for (int i = 0; i < nl; ++i) {
      double complex acc = 0.0;
      for (int t = 0; t < nl; ++t) {
        const size_t eb = (size_t)i + (size_t)t * nl, eg = (size_t)t + (size_t)j * nl;
        const double complex bv = cplx ? Binv[2 * eb] + I * Binv[2 * eb + 1] : Binv[eb];
        const double complex gv = cplx ? G[2 * eg] + I * G[2 * eg + 1] : G[eg];
        acc += bv * gv;
      }
      const size_t e = (size_t)i + (size_t)j * (nl + nw);
      C[e * w] = -creal(acc);
      if (cplx) C[e * w + 1] = -cimag(acc);
    }
    C[((size_t)(nl + j) + (size_t)j * (nl + nw)) * w] = 1.0;
  }
}

/* Smallest fraction of a column's norm that survives the projection out of
 * span(Q), Q orthonormal, from GQW = Q^H W (nq x nw) and GWW = W^H W */
static double lr_kept_min(int nq, int nw, int w, const double *hQW, const double *hWW) {
  double kept_min = 1.0;
  for (int j = 0; j < nw; ++j) {
    const double n2 = hWW[((size_t)j + (size_t)j * nw) * w];
    double q2 = 0.0;
    for (int i = 0; i < nq * w; ++i) q2 += hQW[(size_t)j * nq * w + i] * hQW[(size_t)j * nq * w + i];
    if (n2 > 0) kept_min = fmin(kept_min, sqrt(fmax(0.0, 1.0 - q2 / n2)));
  }
  return kept_min;
}

/* Second orthonormalisation pass over the r columns W' = Z[:, nq:nq+r] (the
 * first pass cancelled most of a column, which costs orthogonality in
 * proportion): [Q | W']^H W' in one Gram, W' <- [Q | W'] C through the temp
 * Wt.  *ncols = the columns kept. */
static int lr_orth2(bdgx_ctx *ctx, int prec, int cplx, void *Z, int nq, int r, void *Wt,
                    uint64_t S, double *dG, double *dC, double *hG, double *hC, int *ncols) {
  const int w = cplx ? 2 : 1;
  const size_t es = lr_es(prec, cplx);
  int rc = lr_gram(ctx, prec, Z, (uint64_t)(nq + r), lr_col(Z, (uint64_t)nq, S, es),
                   (uint64_t)r, S, dG, hG, w);
  if (rc) return rc;
  double *tmp = hG + (size_t)(nq + r) * r * w;
  double *gq = tmp, *gw = tmp + (size_t)nq * r * w;
  for (int j = 0; j < r; ++j) {
    memcpy(gq + (size_t)j * nq * w, hG + (size_t)j * (nq + r) * w, sizeof(double) * nq * w);
    memcpy(gw + (size_t)j * r * w, hG + ((size_t)j * (nq + r) + nq) * w, sizeof(double) * r * w);
  }
  const int r2 = dn_orth_coef(nq, r, cplx, gq, gw, 1e-2, hC);
  *ncols = r2 > 0 ? r2 : 0;
  if (r2 <= 0) return 0;
  rc = bdgx_memcpy_h2d(ctx, dC, hC, sizeof(double) * (size_t)(nq + r) * r2 * w);
  if (!rc) rc = bdgx_combine(ctx, prec, Z, S, (uint64_t)(nq + r), dC, Wt, S, (uint64_t)r2, S);
  if (!rc)
    rc = bdgx_convert(ctx, prec, Wt, S, prec, lr_col(Z, (uint64_t)nq, S, es), S, S,
                      (uint64_t)r2);
  return rc;
}

/* Orthonormalise the nw columns Wt (temp) against Q = Z[:, 0:nq] and write
 * the result to Z[:, nq:nq+r].  A second pass (against the first's output,
 * which then sits in Z) runs when the first cancelled most of a column (kept
 * less than 1/8 of its norm): the projection then loses orthogonality in
 * proportion.  Returns 0 or an error; *ncols = r. */
static int lr_orth(bdgx_ctx *ctx, int prec, int cplx, void *Z, int nq, void *Wt, int nw,
                   uint64_t S, double drop, double *dG, double *dC, double *hG, double *hC,
                   int *ncols) {
  const int w = cplx ? 2 : 1;
  const size_t es = lr_es(prec, cplx);
  int rc = 0;
  *ncols = 0;
  if (nw == 0) return 0;
  double *hQW = hG, *hWW = hG + (size_t)nq * nw * w;
  if (nq) TRY(lr_gram(ctx, prec, Z, (uint64_t)nq, Wt, (uint64_t)nw, S, dG, hQW, w));
  TRY(lr_gram(ctx, prec, Wt, (uint64_t)nw, Wt, (uint64_t)nw, S, dG, hWW, w));
  const double kept_min = lr_kept_min(nq, nw, w, hQW, hWW);
  const int r = dn_orth_coef(nq, nw, cplx, hQW, hWW, drop, hC);
  if (r <= 0) return 0;
  TRY(bdgx_memcpy_h2d(ctx, dC, hC, sizeof(double) * (size_t)(nq + nw) * r * w));
  TRY(bdgx_combine2(ctx, prec, Z, S, (uint64_t)nq, Wt, S, (uint64_t)nw, dC,
                    lr_col(Z, (uint64_t)nq, S, es), S, (uint64_t)r, S));
  *ncols = r;
  if (kept_min > 0.125) return 0;
  return lr_orth2(ctx, prec, cplx, Z, nq, r, Wt, S, dG, dC, hG, hC, ncols);
done:
  return rc;
}


/* Assemble a ku x kv Gram (ld ku) from its projected top-left block P
 * (nu x nv, ld nu), the new columns Cn (ku x (kv - nv), ld ku: Z_a^H B_new)
 * and the new rows given conjugate-transposed, Rn (kv x (ku - nu), ld kv:
 * Z_b^H A_new); for a Hermitian Gram (herm) Rn is Cn itself. */
static void lr_assemble(int ku, int kv, int nu, int nv, int w, const double *P, const double *Cn,
                        const double *Rn, double *out) {
  for (int j = 0; j < kv; ++j)
    for (int i = 0; i < ku; ++i) {
      double *d = out + ((size_t)i + (size_t)j * ku) * w;
      const double *s;
      int cj = 0;
      if (j >= nv) {
        s = Cn + ((size_t)i + (size_t)(j - nv) * ku) * w;
      } else if (i >= nu) {
        s = Rn + ((size_t)j + (size_t)(i - nu) * kv) * w;
        cj = 1;
      } else {
        s = P + ((size_t)i + (size_t)j * nu) * w;
      }
      d[0] = s[0];
      if (w == 2) d[1] = cj ? -s[1] : s[1];
    }
}

/* Ainv = A^-1 (n x n, complex interleaved when cplx) by Gauss-Jordan with
 * partial pivoting; returns 0, or -1 when A is singular to working precision */
static int lr_invert(int n, int cplx, const double *A, double *Ainv) {
  double complex *a = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  double complex *b = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  int bad = 0;
  double amax = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t e = (size_t)i + (size_t)j * n;
      a[e] = cplx ? A[2 * e] + I * A[2 * e + 1] : A[e];
      b[e] = i == j ? 1.0 : 0.0;
      amax = fmax(amax, cabs(a[e]));
    }
  for (int c = 0; c < n && !bad; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (cabs(a[r + (size_t)c * n]) > cabs(a[p + (size_t)c * n])) p = r;
    if (!(cabs(a[p + (size_t)c * n]) > 1e-13 * amax)) {
      bad = 1;
      break;
    }
    if (p != c)
      for (int j = 0; j < n; ++j) {
        double complex t = a[c + (size_t)j * n];
        a[c + (size_t)j * n] = a[p + (size_t)j * n];
        a[p + (size_t)j * n] = t;
        t = b[c + (size_t)j * n];
        b[c + (size_t)j * n] = b[p + (size_t)j * n];
        b[p + (size_t)j * n] = t;
      }
    const double complex d = 1.0 / a[c + (size_t)c * n];
    for (int j = 0; j < n; ++j) {
      a[c + (size_t)j * n] *= d;
      b[c + (size_t)j * n] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double complex f = a[r + (size_t)c * n];
      if (f == 0) continue;
      for (int j = 0; j < n; ++j) {
        a[r + (size_t)j * n] -= f * a[c + (size_t)j * n];
        b[r + (size_t)j * n] -= f * b[c + (size_t)j * n];
      }
    }
  }
  for (size_t e = 0; e < (size_t)n * n && !bad; ++e) {
    if (cplx) {
      Ainv[2 * e] = creal(b[e]);
      Ainv[2 * e + 1] = cimag(b[e]);
    } else {
      Ainv[e] = creal(b[e]);
    }
  }
  free(a);
  free(b);
  return bad ? -1 : 0;
}

/* Oblique deflation of the nw columns W against the locked pairs:
 * Wout = W - Lo Binv (Lp^H W), Lp (S x nl) the partner vectors, Lo the own
 * half's locked vectors, Binv = (Lp^H Lo)^-1 (host, nl x nl). */
static int lr_deflate(bdgx_ctx *ctx, int prec, int cplx, const void *Lo, const void *Lp, int nl,
                      const double *Binv, const void *W, int nw, void *Wout, uint64_t S,
                      double *dG, double *dC, double *hG, double *hC) {
  const int w = cplx ? 2 : 1;
  int rc = lr_gram(ctx, prec, Lp, (uint64_t)nl, W, (uint64_t)nw, S, dG, hG, w);
  if (rc) return rc;
  lr_deflate_coef(nl, nw, cplx, Binv, hG, hC);
  rc = bdgx_memcpy_h2d(ctx, dC, hC, sizeof(double) * (size_t)(nl + nw) * nw * w);
  if (!rc)
    rc = bdgx_combine2(ctx, prec, Lo, S, (uint64_t)nl, W, S, (uint64_t)nw, dC, Wout, S,
                       (uint64_t)nw, S);
  return rc;
}

int bdg_lrep_device(bdgx_ctx *ctx, const lrep_cfg *cfg, int prec, const void *dXU0,
                    const void *dXV0, int prec0, lrep_result *res, void *dXU, void *dXV) {
  const uint64_t S = cfg->S;
  const int m = cfg->m, nev = cfg->nev, cplx = cfg->cplx;
  const int w = cplx ? 2 : 1;
  const size_t es = lr_es(prec, cplx);
  /* basis [X | P | W] of at most kmax columns: the host Grams and the
   * coefficient scratch (dG, dC, hG, hC: the update's, the residuals' and
   * the deflation's matrices at fixed offsets) are sized from it.  The
   * device basis of a phase of no iterations (the refinement sweep:
   * Rayleigh-Ritz of the start block and its residuals) holds X only:
   * kcols columns. */
  const int kmax = 3 * m;
  const int kcols = cfg->maxiter > 0 ? kmax : m;
  const int opT[2] = {BDGX_OP_PRECOND_K, BDGX_OP_PRECOND_M};
  const int opA[2] = {BDGX_OP_K, BDGX_OP_M};
  /* full device Gram recompute period (FP32: the projected Grams of rounded
   * vectors drift from their stored values faster) */
  const int refresh = prec == BDGX_FP64 ? 25 : 8;
  int rc = 0, it = 0, g = 0;
  lr_half H[2];
  memset(H, 0, sizeof(H));
  void *dR[2] = {NULL, NULL}, *dWt[2] = {NULL, NULL};
  double *dG = NULL, *dC = NULL, *dZero = NULL, *dN2 = NULL;
  const size_t gsz = (size_t)kmax * kmax * w;
  /* scratch Grams and coefficients also hold the locked pairs' (nl x nl)
   * and deflation ((nl + m) x m) matrices */
  const int kbuf = kmax > cfg->nl + m ? kmax : cfg->nl + m;
  const size_t bsz = (size_t)kbuf * kbuf * w;
  double *hGK = (double *)calloc(gsz, sizeof(double));
  double *hGM = (double *)calloc(gsz, sizeof(double));
  double *hGW = (double *)calloc(gsz, sizeof(double));
  double *hP[3] = {(double *)calloc(gsz, sizeof(double)), (double *)calloc(gsz, sizeof(double)),
                   (double *)calloc(gsz, sizeof(double))};
  double *hG = (double *)calloc(4 * bsz, sizeof(double));
  double *hC = (double *)calloc(4 * bsz, sizeof(double));
  double *hA = (double *)calloc((size_t)kmax * m * w, sizeof(double));
  double *hB = (double *)calloc((size_t)kmax * m * w, sizeof(double));
  double *hCQ[2] = {(double *)calloc(gsz, sizeof(double)), (double *)calloc(gsz, sizeof(double))};
  double *hR[2] = {(double *)calloc((size_t)m * m * w, sizeof(double)),
                   (double *)calloc((size_t)m * m * w, sizeof(double))};
  double *w2 = (double *)calloc((size_t)m, sizeof(double));
  double *hn2 = (double *)calloc(3 * (size_t)m, sizeof(double));
  double *resv = (double *)calloc((size_t)m, sizeof(double));
  double *best = (double *)calloc((size_t)m, sizeof(double));
  int *best_it = (int *)calloc((size_t)m, sizeof(int));
  int *conv = (int *)calloc((size_t)m, sizeof(int));
  int *act = (int *)calloc((size_t)m, sizeof(int));
  int *wact = (int *)calloc((size_t)m, sizeof(int)); /* columns whose W is in the basis */
  int *sel = (int *)calloc((size_t)m, sizeof(int));
  int k[2] = {0, 0}, nw[2] = {0, 0}, nwact = 0, nact = 0, last_full = -1000, since_full = 0;
  int kprev[2] = {0, 0}; /* basis size of the last Rayleigh-Ritz */
  int gvalid = 0; /* hGK/hGM/hGW hold the Grams of the current basis */
  double *hBinv[2] = {NULL, NULL};
  double *gA = NULL, *gB = NULL, *gw2 = NULL; /* guard Ritz pairs at locking */
  lr_stage st = {NULL, 0, 0};
  void *dUL = NULL, *dVL = NULL, *dU0 = NULL, *dV0 = NULL, *dOut = NULL;
  double trr = 0.0;
  const double t0 = bdg_wtime();
  res->iters = 0;
  res->status = 0;
  for (int j = 0; j < m; ++j) best[j] = HUGE_VAL;
  /* learned per-half preconditioner in the FP32 phase, with its stall rule
   * over the worst residual of the wanted pairs */
  int lp_on = cfg->learned && prec == BDGX_FP32;
  double *whist = lp_on ? (double *)calloc((size_t)cfg->maxiter + 2, sizeof(double)) : NULL;

  TRY(bdgx_autotune(ctx, prec, (uint64_t)(m < 32 ? m : 32), 0));
  for (int h = 0; h < 2; ++h) {
    for (int q = 0; q < 2; ++q) {
      TRY(bdgx_malloc(&H[h].Z[q], (size_t)S * kcols * es));
      TRY(bdgx_malloc(&H[h].Y[q], (size_t)S * kcols * es));
    }
    TRY(bdgx_malloc(&dR[h], (size_t)S * m * es));
    TRY(bdgx_malloc(&dWt[h], (size_t)S * m * es));
  }
  TRY(bdgx_malloc((void **)&dG, sizeof(double) * 4 * bsz));
  TRY(bdgx_malloc((void **)&dC, sizeof(double) * 4 * bsz));
  st.cap = 8 * bsz;
  if (bdgx_host_alloc((void **)&st.h, sizeof(double) * st.cap)) st.h = NULL; /* then synchronous */
  TRY(bdgx_malloc((void **)&dZero, sizeof(double) * (size_t)m));
  TRY(bdgx_malloc((void **)&dN2, sizeof(double) * 3 * (size_t)m));
  TRY(bdgx_memset(ctx, dZero, 0, sizeof(double) * (size_t)m));
  /* locked pairs: (V_L^H U_L)^-1 for the U half, its adjoint for the V half */
  const void *Lown[2] = {cfg->dUL, cfg->dVL}, *Lpar[2] = {cfg->dVL, cfg->dUL};
  if (cfg->nl > 0) {
    const int nl = cfg->nl;
    hBinv[0] = (double *)calloc((size_t)nl * nl * w, sizeof(double));
    hBinv[1] = (double *)calloc((size_t)nl * nl * w, sizeof(double));
    TRY(lr_gram(ctx, prec, cfg->dVL, (uint64_t)nl, cfg->dUL, (uint64_t)nl, S, dG, hG, w));
    if (lr_invert(nl, cplx, hG, hBinv[0])) {
      rc = -102; /* locked pairs not biorthogonally independent */
      goto done;
    }
    for (int j = 0; j < nl; ++j)
      for (int i = 0; i < nl; ++i) {
        const size_t a = (size_t)i + (size_t)j * nl, b = (size_t)j + (size_t)i * nl;
        hBinv[1][a * w] = hBinv[0][b * w];
        if (cplx) hBinv[1][a * w + 1] = -hBinv[0][b * w + 1];
      }
  }

  /* start: X_U, X_V from the given blocks (converted, deflated against the
   * locked pairs), orthonormalised */
  {
    const void *src[2] = {dXU0, dXV0};
    for (int h = 0; h < 2; ++h) {
      void *x0 = dWt[h];
      TRY(bdgx_convert(ctx, prec0, src[h], S, prec, dWt[h], S, S, (uint64_t)m));
      if (cfg->nl > 0) {
        TRY(lr_deflate(ctx, prec, cplx, Lown[h], Lpar[h], cfg->nl, hBinv[h], dWt[h], m, dR[h], S,
                       dG, dC, hG, hC));
        x0 = dR[h];
      }
      int r = 0;
      TRY(lr_orth(ctx, prec, cplx, H[h].Z[0], 0, x0, m, S, 1e-8, dG, dC, hG, hC, &r));
      if (r < m) {
        rc = -101; /* start block numerically dependent */
        goto done;
      }
      TRY(bdgx_apply(ctx, opA[h], prec, H[h].Z[0], H[h].Y[0], (uint64_t)m, S));
      k[h] = m;
    }
    res->applies += 2 * (uint64_t)m;
  }

  if (cfg->on_iterating) cfg->on_iterating(cfg->on_iterating_arg);
  for (it = 0; it <= cfg->maxiter; ++it) {
    /* Grams of the current basis: all on the device at the start and every
     * `refresh` iterations, else the new W columns only */
    if (!gvalid || ++since_full >= refresh) {
      /* three Grams, one download */
      const size_t nK = (size_t)k[0] * k[0] * w, nM = (size_t)k[1] * k[1] * w;
      const size_t nW = (size_t)k[0] * k[1] * w;
      TRY(bdgx_gram(ctx, prec, H[0].Z[g], S, (uint64_t)k[0], H[0].Y[g], S, (uint64_t)k[0], S, 0,
                    dG));
      TRY(bdgx_gram(ctx, prec, H[1].Z[g], S, (uint64_t)k[1], H[1].Y[g], S, (uint64_t)k[1], S, 0,
                    dG + nK));
      TRY(bdgx_gram(ctx, prec, H[0].Z[g], S, (uint64_t)k[0], H[1].Z[g], S, (uint64_t)k[1], S, 0,
                    dG + nK + nM));
      TRY(lr_get(ctx, &st, hG, dG, nK + nM + nW));
      memcpy(hGK, hG, sizeof(double) * nK);
      memcpy(hGM, hG + nK, sizeof(double) * nM);
      memcpy(hGW, hG + nK + nM, sizeof(double) * nW);
      since_full = 0;
    } else {
      /* hP[0..2]: projected [X | P] blocks (n0 x n0, n1 x n1, n0 x n1) */
      const int n0 = k[0] - nw[0], n1 = k[1] - nw[1];
      const size_t c0 = (size_t)k[0] * nw[0] * w, c1 = (size_t)k[1] * nw[1] * w;
      double *gKc = hG, *gMc = gKc + c0, *gWc = gMc + c1, *gWr = gWc + (size_t)k[0] * nw[1] * w;
      const void *W0 = lr_col(H[0].Z[g], (uint64_t)n0, S, es);
      const void *W1 = lr_col(H[1].Z[g], (uint64_t)n1, S, es);
      const void *KW0 = lr_col(H[0].Y[g], (uint64_t)n0, S, es);
      const void *MW1 = lr_col(H[1].Y[g], (uint64_t)n1, S, es);
      int paired = 0;
      if (!cplx && nw[0] == nw[1] && nw[0] > 0) {
        /* Z_U^H [K W_U | W_V] and Z_V^H [M W_V | W_U]: each basis read once;
         * both into dG, one download */
        const int q = nw[0];
        const size_t n0p = (size_t)k[0] * 2 * q, n1p = (size_t)k[1] * 2 * q;
        double *tmp = hC;
        if (bdgx_gram_pair(ctx, prec, H[0].Z[g], S, (uint64_t)k[0], KW0, S, W1, S, (uint64_t)q,
                           S, dG) == 0 &&
            bdgx_gram_pair(ctx, prec, H[1].Z[g], S, (uint64_t)k[1], MW1, S, W0, S, (uint64_t)q,
                           S, dG + n0p) == 0) {
          TRY(lr_get(ctx, &st, tmp, dG, n0p + n1p));
          memcpy(gKc, tmp, sizeof(double) * (size_t)k[0] * q);
          memcpy(gWc, tmp + (size_t)k[0] * q, sizeof(double) * (size_t)k[0] * q);
          memcpy(gMc, tmp + n0p, sizeof(double) * (size_t)k[1] * q);
          memcpy(gWr, tmp + n0p + (size_t)k[1] * q, sizeof(double) * (size_t)k[1] * q);
          paired = 1;
        }
      }
      if (!paired) {
        /* the four W-column Grams into dG side by side, one download */
        const size_t o1 = c0, o2 = o1 + (size_t)k[1] * nw[0] * w, o3 = o2 + c1;
        const size_t ntot = o3 + (size_t)k[0] * nw[1] * w;
        if (nw[0]) {
          TRY(bdgx_gram(ctx, prec, H[0].Z[g], S, (uint64_t)k[0], KW0, S, (uint64_t)nw[0], S, 0,
                        dG));
          TRY(bdgx_gram(ctx, prec, H[1].Z[g], S, (uint64_t)k[1], W0, S, (uint64_t)nw[0], S, 0,
                        dG + o1));
        }
        if (nw[1]) {
          TRY(bdgx_gram(ctx, prec, H[1].Z[g], S, (uint64_t)k[1], MW1, S, (uint64_t)nw[1], S, 0,
                        dG + o2));
          TRY(bdgx_gram(ctx, prec, H[0].Z[g], S, (uint64_t)k[0], W1, S, (uint64_t)nw[1], S, 0,
                        dG + o3));
        }
        double *tmp = hC;
        TRY(lr_get(ctx, &st, tmp, dG, ntot));
        memcpy(gKc, tmp, sizeof(double) * c0);
        memcpy(gWr, tmp + o1, sizeof(double) * (size_t)k[1] * nw[0] * w);
        memcpy(gMc, tmp + o2, sizeof(double) * c1);
        memcpy(gWc, tmp + o3, sizeof(double) * (size_t)k[0] * nw[1] * w);
      }
      lr_assemble(k[0], k[0], n0, n0, w, hP[0], gKc, gKc, hGK);
      lr_assemble(k[1], k[1], n1, n1, w, hP[1], gMc, gMc, hGM);
      lr_assemble(k[0], k[1], n0, n1, w, hP[2], gWc, gWr, hGW);
    }
    gvalid = 1;
    double tr = bdg_wtime();
    const int got = dn_lrep_rr(k[0], k[1], m, cplx, hGK, hGM, hGW, w2, hA, hB);
    if (got < m) {
      rc = -100;
      goto done;
    }
    /* orthonormal update coefficients; the new P from the columns whose W
     * is in this basis */
    int nc[2];
    nc[0] = dn_lrep_update(k[0], m, cplx, hA, wact, nwact, 1e-3, hCQ[0], hR[0]);
    nc[1] = dn_lrep_update(k[1], m, cplx, hB, wact, nwact, 1e-3, hCQ[1], hR[1]);
    if (nc[0] < m || nc[1] < m) {
      rc = -100;
      goto done;
    }
    trr += bdg_wtime() - tr;
    /* the update: both halves' coefficients in one upload (each half's
     * block at its own offset of dC), then the four combines, queued */
    {
      const size_t off1 = (size_t)k[0] * nc[0] * w;
      TRY(lr_put(ctx, &st, dC, hCQ[0], off1));
      TRY(lr_put(ctx, &st, dC + off1, hCQ[1], (size_t)k[1] * nc[1] * w));
      for (int h = 0; h < 2; ++h) {
        const double *dCh = dC + (h ? off1 : 0);
        TRY(bdgx_combine(ctx, prec, H[h].Z[g], S, (uint64_t)k[h], dCh, H[h].Z[g ^ 1], S,
                         (uint64_t)nc[h], S));
        TRY(bdgx_combine(ctx, prec, H[h].Y[g], S, (uint64_t)k[h], dCh, H[h].Y[g ^ 1], S,
                         (uint64_t)nc[h], S));
      }
    }
    /* the Grams of the next [X | P], projected on the host while the device
     * runs the update */
    tr = bdg_wtime();
    dn_project(k[0], k[0], nc[0], nc[0], cplx, hGK, hCQ[0], hCQ[0], hP[0]);
    dn_project(k[1], k[1], nc[1], nc[1], cplx, hGM, hCQ[1], hCQ[1], hP[1]);
    dn_project(k[0], k[1], nc[0], nc[1], cplx, hGW, hCQ[0], hCQ[1], hP[2]);
    trr += bdg_wtime() - tr;
    for (int h = 0; h < 2; ++h) {
      H[h].np = nc[h] - m;
      kprev[h] = k[h];
      k[h] = nc[h];
      nw[h] = 0;
    }
    g ^= 1;
    /* residuals of the Ritz pairs u = X_U R_A, v = X_V R_B:
     *   R_U = [K X_U | X_V] [R_A; -R_B s|w|],  R_V = [M X_V | X_U] [R_B; -R_A |w|]
     * for the selected columns sel[0..nsel), packed in that order: all
     * columns (the active ones first) every 10 iterations, else the active
     * ones (the others keep the residual of the last full evaluation) */
    int full = it - last_full >= 10 || nact == 0;
    int nsel = 0;
    for (int pass = 0; pass < 2; ++pass) {
      nsel = 0;
      if (full) {
        for (int j = 0; j < m; ++j)
          if (!conv[j]) sel[nsel++] = j;
        for (int j = 0; j < m; ++j)
          if (conv[j]) sel[nsel++] = j;
      } else {
        for (int i = 0; i < nact; ++i) sel[nsel++] = act[i];
      }
      for (int h = 0; h < 2; ++h) {
        const double *Rs = hR[h], *Ro = hR[h ^ 1];
        double *Cc = hC;
        memset(Cc, 0, sizeof(double) * 2 * (size_t)m * nsel * w);
        for (int c = 0; c < nsel; ++c) {
          const int j = sel[c];
          const double wt = sqrt(fabs(w2[j]));
          const double f = (h == 0 && w2[j] < 0) ? -wt : wt;
          for (int i = 0; i < m; ++i)
            for (int q = 0; q < w; ++q) {
              Cc[((size_t)i + (size_t)c * 2 * m) * w + q] = Rs[((size_t)i + (size_t)j * m) * w + q];
              Cc[((size_t)(m + i) + (size_t)c * 2 * m) * w + q] =
                  -f * Ro[((size_t)i + (size_t)j * m) * w + q];
            }
        }
        /* each half's coefficients in a region of dC of its own, after the
         * update's (at most 2 x 3m x 2m) */
        double *dCh = dC + (size_t)(6 + h) * 2 * m * m * w;
        TRY(lr_put(ctx, &st, dCh, Cc, 2 * (size_t)m * nsel * w));
        TRY(bdgx_combine2(ctx, prec, H[h].Y[g], S, (uint64_t)m, H[h ^ 1].Z[g], S, (uint64_t)m,
                          dCh, dR[h], S, (uint64_t)nsel, S));
      }
      TRY(bdgx_residual(ctx, prec, dR[0], dR[1], NULL, S, (uint64_t)nsel, S, dZero, 0, NULL, S,
                        dN2));
      TRY(lr_get(ctx, &st, hn2, dN2, 3 * (size_t)nsel));
      /* ||A x - w B x|| / (||x|| max(1, |w|)), ||x||^2 from the coefficients
       * (orthonormal bases) */
      for (int c = 0; c < nsel; ++c) {
        const int j = sel[c];
        double xn2 = 0.0;
        for (int h = 0; h < 2; ++h)
          for (int i = 0; i < m * w; ++i)
            xn2 += hR[h][(size_t)j * m * w + i] * hR[h][(size_t)j * m * w + i];
        const double wt = sqrt(fabs(w2[j]));
        resv[j] = sqrt(hn2[3 * c] + hn2[3 * c + 1]) / (sqrt(xn2) * fmax(1.0, wt));
        if (resv[j] < 0.5 * best[j]) {
          best[j] = resv[j];
          best_it[j] = it;
        }
        /* FP32 phase: a column also leaves when its residual stopped falling
         * (the rounding floor of FP32 images) */
        conv[j] = resv[j] < cfg->tol ||
                  (cfg->stall_window > 0 && resv[j] < cfg->stall_below &&
                   it - best_it[j] >= cfg->stall_window);
      }
      if (full) last_full = it;
      nact = 0;
      for (int j = 0; j < m; ++j)
        if (!conv[j]) act[nact++] = j;
      int allc = 1;
      for (int j = 0; j < nev; ++j)
        if (!conv[j]) allc = 0;
      /* a partial evaluation that finds everything converged, or after which
       * the converged pairs would be locked, is confirmed by a full one: the
       * other columns' residuals can be up to 10 iterations old, and a pair
       * that has drifted since must neither end the solve nor be locked
       * (locked pairs are final) */
      const int will_lock = !allc && cfg->lock_max > 0 && cfg->nl == 0 && nact <= cfg->lock_max &&
                            nact < m && it < cfg->maxiter;
      if ((allc || will_lock) && !full) {
        full = 1;
        continue;
      }
      break;
    }
    double worst = 0.0;
    int allc = 1;
    for (int j = 0; j < nev; ++j) {
      worst = fmax(worst, resv[j]);
      if (!conv[j]) allc = 0;
    }
    if (cfg->verbosity > 1 || (cfg->verbosity == 1 && it % 10 == 0))
      fprintf(stderr, "[bdg] %s it %d t %.3f w0 %.10g wlast %.10g worst res %.3e active %d\n",
              prec == BDGX_FP64 ? "fp64" : "fp32", it, bdg_wtime() - t0,
              w2[0] < 0 ? -sqrt(-w2[0]) : sqrt(w2[0]),
              w2[nev - 1] < 0 ? -sqrt(-w2[nev - 1]) : sqrt(w2[nev - 1]), worst, nact);
    res->worst = worst;
    if (lp_on) {
      whist[it] = worst;
      if (cfg->lp_window > 0 && it >= cfg->lp_window &&
          worst > cfg->lp_factor * whist[it - cfg->lp_window]) {
        lp_on = 0;
        res->learned_fallback_at = it;
        if (cfg->verbosity > 0)
          fprintf(stderr, "[bdg] fp32 it %d: learned preconditioner stalled, precondK/M from here\n",
                  it);
      }
    }
    if (!allc && cfg->lock_max > 0 && cfg->nl == 0 && nact <= cfg->lock_max && nact < m &&
        it < cfg->maxiter) {
      /* lock the converged pairs; the active ones continue in a nested
       * small-block solve deflated against them */
      const int nc = m - nact;
      int *cidx = (int *)malloc(sizeof(int) * (size_t)nc);
      int ncv = 0, nev_in = 0;
      for (int j = 0; j < m; ++j)
        if (conv[j]) cidx[ncv++] = j;
      for (int i = 0; i < nact; ++i)
        if (act[i] < nev) ++nev_in;
      TRY(bdgx_malloc(&dUL, (size_t)S * nc * es));
      TRY(bdgx_malloc(&dVL, (size_t)S * nc * es));
      /* guard columns for the small block: Ritz pairs m .. m + ng - 1 of
       * this iteration's Rayleigh-Ritz, whose basis (now the other buffer)
       * and Grams are still intact */
      int ng = cfg->lock_guards;
      if (ng > kprev[0] - m) ng = kprev[0] - m;
      if (ng > kprev[1] - m) ng = kprev[1] - m;
      if (ng < 0) ng = 0;
      if (ng > 0) {
        gA = (double *)calloc((size_t)kprev[0] * (m + ng) * w, sizeof(double));
        gB = (double *)calloc((size_t)kprev[1] * (m + ng) * w, sizeof(double));
        gw2 = (double *)calloc((size_t)(m + ng), sizeof(double));
        const int got2 = dn_lrep_rr(kprev[0], kprev[1], m + ng, cplx, hGK, hGM, hGW, gw2, gA, gB);
        ng = got2 > m ? got2 - m : 0;
      }
      const int mi = nact + ng;
      TRY(bdgx_malloc(&dU0, (size_t)S * mi * es));
      TRY(bdgx_malloc(&dV0, (size_t)S * mi * es));
      void *dstL[2] = {dUL, dVL}, *dstA[2] = {dU0, dV0};
      for (int h = 0; h < 2; ++h) {
        /* Ritz vectors X R[:, C] and X R[:, J] */
        for (int pass = 0; pass < 2; ++pass) {
          const int *cols = pass == 0 ? cidx : act;
          const int nn = pass == 0 ? nc : nact;
          for (int c = 0; c < nn; ++c)
            memcpy(hC + (size_t)c * m * w, hR[h] + (size_t)cols[c] * m * w,
                   sizeof(double) * (size_t)m * w);
          TRY(bdgx_memcpy_h2d(ctx, dC, hC, sizeof(double) * (size_t)m * nn * w));
          TRY(bdgx_combine(ctx, prec, H[h].Z[g], S, (uint64_t)m, dC, pass == 0 ? dstL[h] : dstA[h],
                           S, (uint64_t)nn, S));
        }
        if (ng > 0) {
          /* guards after the active columns: old basis x coefficients m.. */
          const double *gc = (h == 0 ? gA : gB) + (size_t)kprev[h] * m * w;
          TRY(bdgx_memcpy_h2d(ctx, dC, gc, sizeof(double) * (size_t)kprev[h] * ng * w));
          TRY(bdgx_combine(ctx, prec, H[h].Z[g ^ 1], S, (uint64_t)kprev[h], dC,
                           lr_col(dstA[h], (uint64_t)nact, S, es), S, (uint64_t)ng, S));
        }
      }
      TRY(bdgx_sync(ctx));
      for (int h = 0; h < 2; ++h)
        for (int q = 0; q < 2; ++q) {
          bdgx_free(H[h].Z[q]);
          bdgx_free(H[h].Y[q]);
          H[h].Z[q] = H[h].Y[q] = NULL;
        }
      lrep_cfg icfg = *cfg;
      icfg.nev = nev_in > 0 ? nev_in : 1;
      icfg.maxiter = cfg->maxiter - it - 1;
      icfg.m = mi;
      icfg.lock_max = 0;
      icfg.lock_guards = 0;
      icfg.dUL = dUL;
      icfg.dVL = dVL;
      icfg.nl = nc;
      icfg.learned = lp_on;
      icfg.on_iterating = NULL;
      lrep_result ires;
      memset(&ires, 0, sizeof(ires));
      double *iw2 = (double *)calloc((size_t)mi, sizeof(double));
      double *ires_v = (double *)calloc((size_t)mi, sizeof(double));
      ires.w2 = iw2;
      ires.res = ires_v;
      if (cfg->verbosity > 0)
        fprintf(stderr, "[bdg] %s it %d: %d pairs locked, %d continue (+%d guards)\n",
                prec == BDGX_FP64 ? "fp64" : "fp32", it, nc, nact, ng);
      rc = bdg_lrep_device(ctx, &icfg, prec, dU0, dV0, prec, &ires, dU0, dV0);
      res->applies += ires.applies;
      res->learned_its += ires.learned_its;
      if (ires.learned_fallback_at && !res->learned_fallback_at)
        res->learned_fallback_at = it + 1 + ires.learned_fallback_at;
      trr += ires.rr_seconds;
      res->iters = it + 1 + ires.iters;
      if (!rc) {
        /* merged pairs, sorted by w^2: locked ones (u negated for w^2 < 0)
         * and the nested solve's (already in output form) */
        int *src = (int *)malloc(sizeof(int) * (size_t)m);   /* < nc: locked, else nested */
        double *fw2 = (double *)malloc(sizeof(double) * (size_t)m);
        double *fres = (double *)malloc(sizeof(double) * (size_t)m);
        for (int c = 0; c < nc; ++c) {
          src[c] = c;
          fw2[c] = w2[cidx[c]];
          fres[c] = resv[cidx[c]];
        }
        for (int i = 0; i < nact; ++i) {
          src[nc + i] = nc + i;
          fw2[nc + i] = iw2[i];
          fres[nc + i] = ires_v[i];
        }
        for (int a = 1; a < m; ++a) /* insertion sort by w^2 (stable) */
          for (int b = a; b > 0 && fw2[b] < fw2[b - 1]; --b) {
            double t = fw2[b];
            fw2[b] = fw2[b - 1];
            fw2[b - 1] = t;
            t = fres[b];
            fres[b] = fres[b - 1];
            fres[b - 1] = t;
            const int ti = src[b];
            src[b] = src[b - 1];
            src[b - 1] = ti;
          }
        void *dst[2] = {dXU, dXV};
        if (prec0 != prec) TRY(bdgx_malloc(&dOut, (size_t)S * m * es));
        for (int h = 0; h < 2; ++h) {
          memset(hC, 0, sizeof(double) * (size_t)m * m * w);
          for (int p = 0; p < m; ++p) {
            const int sc = src[p];
            const double sg = (h == 0 && sc < nc && fw2[p] < 0) ? -1.0 : 1.0;
            hC[((size_t)sc + (size_t)p * m) * w] = sg;
          }
          TRY(bdgx_memcpy_h2d(ctx, dC, hC, sizeof(double) * (size_t)m * m * w));
          TRY(bdgx_combine2(ctx, prec, dstL[h], S, (uint64_t)nc, dstA[h], S, (uint64_t)nact, dC,
                            prec0 == prec ? dst[h] : dOut, S, (uint64_t)m, S));
          if (prec0 != prec) TRY(bdgx_convert(ctx, prec, dOut, S, prec0, dst[h], S, S, (uint64_t)m));
        }
        worst = 0.0;
        for (int j = 0; j < m; ++j) {
          res->w2[j] = fw2[j];
          res->res[j] = fres[j];
          if (j < nev) worst = fmax(worst, fres[j]);
        }
        res->worst = worst;
        TRY(bdgx_sync(ctx));
        free(src);
        free(fw2);
        free(fres);
      }
      free(iw2);
      free(ires_v);
      free(cidx);
      goto done;
    }
    if (allc || it == cfg->maxiter) break;
    /* new directions of the active columns, orthonormalised against
     * [X | P], and their images.  The residual of act[i] sits in column
     * pos[i] of dR (the evaluation order); when that is not i the columns
     * are gathered first. */
    int packed = 1;
    for (int i = 0; i < nact; ++i) {
      int p = 0;
      while (p < nsel && sel[p] != act[i]) ++p;
      wact[i] = p; /* position, reused below as the column list */
      if (p != i) packed = 0;
    }
    /* both halves go through each phase together, so each phase waits for
     * the device once: preconditioner applies, deflation against the
     * locked pairs (one download), orthonormalisation against [X | P] (one
     * download), images */
    void *wsrc[2], *wtmp[2];
    for (int h = 0; h < 2; ++h) {
      wsrc[h] = dR[h];
      wtmp[h] = dWt[h];
      if (!packed) {
        TRY(bdgx_gather_cols(ctx, prec, dR[h], S, wact, (uint64_t)nact, dWt[h], S, S));
        wsrc[h] = dWt[h];
        wtmp[h] = dR[h];
      }
      TRY(lp_on ? bdgx_lp_apply_layer(ctx, h, wsrc[h], wtmp[h], (uint64_t)nact, S)
                : bdgx_apply(ctx, opT[h], prec, wsrc[h], wtmp[h], (uint64_t)nact, S));
    }
    if (lp_on) ++res->learned_its;
    const int nl = cfg->nl;
    if (nl > 0) {
      /* W into the deflated subspace (the residual in wsrc is consumed) */
      const size_t ngl = (size_t)nl * nact * w, ncf = (size_t)(nl + nact) * nact * w;
      for (int h = 0; h < 2; ++h)
        TRY(bdgx_gram(ctx, prec, Lpar[h], S, (uint64_t)nl, wtmp[h], S, (uint64_t)nact, S, 0,
                      dG + h * ngl));
      TRY(lr_get(ctx, &st, hG, dG, 2 * ngl));
      for (int h = 0; h < 2; ++h) {
        lr_deflate_coef(nl, nact, cplx, hBinv[h], hG + h * ngl, hC);
        TRY(lr_put(ctx, &st, dC + h * ncf, hC, ncf));
        TRY(bdgx_combine2(ctx, prec, Lown[h], S, (uint64_t)nl, wtmp[h], S, (uint64_t)nact,
                          dC + h * ncf, wsrc[h], S, (uint64_t)nact, S));
        void *t = wtmp[h];
        wtmp[h] = wsrc[h];
        wsrc[h] = t;
      }
    }
    {
      /* [X | P]^H W and W^H W of both halves, one download */
      int nq[2], r[2];
      double kept[2];
      size_t oQ[2], oW[2], tot = 0;
      for (int h = 0; h < 2; ++h) {
        nq[h] = k[h];
        oQ[h] = tot;
        tot += (size_t)nq[h] * nact * w;
        oW[h] = tot;
        tot += (size_t)nact * nact * w;
        TRY(bdgx_gram(ctx, prec, H[h].Z[g], S, (uint64_t)nq[h], wtmp[h], S, (uint64_t)nact, S, 0,
                      dG + oQ[h]));
        TRY(bdgx_gram(ctx, prec, wtmp[h], S, (uint64_t)nact, wtmp[h], S, (uint64_t)nact, S, 0,
                      dG + oW[h]));
      }
      TRY(lr_get(ctx, &st, hG, dG, tot));
      /* a region of dC after the deflation's */
      const size_t odef = nl > 0 ? 2 * (size_t)(nl + nact) * nact * w : 0;
      const size_t o8 = (size_t)8 * m * m * w;
      double *dCo = dC + (odef > o8 ? odef : o8);
      for (int h = 0; h < 2; ++h) {
        kept[h] = lr_kept_min(nq[h], nact, w, hG + oQ[h], hG + oW[h]);
        r[h] = dn_orth_coef(nq[h], nact, cplx, hG + oQ[h], hG + oW[h], 1e-3, hC);
        if (r[h] <= 0) {
          r[h] = 0;
          continue;
        }
        const size_t nco = (size_t)(nq[h] + nact) * r[h] * w;
        TRY(lr_put(ctx, &st, dCo, hC, nco));
        TRY(bdgx_combine2(ctx, prec, H[h].Z[g], S, (uint64_t)nq[h], wtmp[h], S, (uint64_t)nact,
                          dCo, lr_col(H[h].Z[g], (uint64_t)nq[h], S, es), S, (uint64_t)r[h],
                          S));
        dCo += nco;
      }
      for (int h = 0; h < 2; ++h) {
        if (r[h] > 0 && kept[h] <= 0.125)
          TRY(lr_orth2(ctx, prec, cplx, H[h].Z[g], nq[h], r[h], wtmp[h], S, dG, dC, hG, hC,
                       &r[h]));
        if (r[h] > 0)
          TRY(bdgx_apply(ctx, opA[h], prec, lr_col(H[h].Z[g], (uint64_t)nq[h], S, es),
                         lr_col(H[h].Y[g], (uint64_t)nq[h], S, es), (uint64_t)r[h], S));
        nw[h] = r[h];
        k[h] = nq[h] + r[h];
        res->applies += (uint64_t)r[h];
      }
    }
    nwact = nact;
    memcpy(wact, act, sizeof(int) * (size_t)nact);
    res->iters = it + 1;
  }
  /* output: u = X_U R_A, v = X_V R_B (u negated for w^2 < 0, so that
   * M v = w u with the reported w = -|w|) into dXU, dXV (prec0) */
  {
    for (int j = 0; j < m; ++j)
      if (w2[j] < 0)
        for (int i = 0; i < m * w; ++i) hR[0][(size_t)j * m * w + i] = -hR[0][(size_t)j * m * w + i];
    void *dst[2] = {dXU, dXV};
    for (int h = 0; h < 2; ++h) {
      TRY(bdgx_memcpy_h2d(ctx, dC, hR[h], sizeof(double) * (size_t)m * m * w));
      if (prec0 == prec) {
        TRY(bdgx_combine(ctx, prec, H[h].Z[g], S, (uint64_t)m, dC, dst[h], S, (uint64_t)m, S));
      } else {
        TRY(bdgx_combine(ctx, prec, H[h].Z[g], S, (uint64_t)m, dC, dWt[h], S, (uint64_t)m, S));
        TRY(bdgx_convert(ctx, prec, dWt[h], S, prec0, dst[h], S, S, (uint64_t)m));
      }
    }
    for (int j = 0; j < m; ++j) {
      res->w2[j] = w2[j];
      res->res[j] = resv[j];
    }
    TRY(bdgx_sync(ctx));
  }
done:
  res->rr_seconds += trr;
  res->status = rc;
  for (int h = 0; h < 2; ++h) {
    for (int q = 0; q < 2; ++q) {
      bdgx_free(H[h].Z[q]);
      bdgx_free(H[h].Y[q]);
    }
    bdgx_free(dR[h]);
    bdgx_free(dWt[h]);
    free(hCQ[h]);
    free(hR[h]);
  }
  for (int q = 0; q < 3; ++q) free(hP[q]);
  free(hBinv[0]);
  free(hBinv[1]);
  free(gA);
  free(gB);
  free(gw2);
  if (st.h) {
    (void)bdgx_sync(ctx); /* queued uploads read the staging region */
    bdgx_host_free(st.h);
  }
  bdgx_free(dUL);
  bdgx_free(dVL);
  bdgx_free(dU0);
  bdgx_free(dV0);
  bdgx_free(dOut);
  bdgx_free(dG);
  bdgx_free(dC);
  bdgx_free(dZero);
  bdgx_free(dN2);
  free(whist);
  free(hGK);
  free(hGM);
  free(hGW);
  free(hG);
  free(hC);
  free(hA);
  free(hB);
  free(w2);
  free(hn2);
  free(resv);
  free(best);
  free(best_it);
  free(conv);
  free(act);
  free(wact);
  free(sel);
  return rc;
}
