/*
 * Device block LOBPCG for the BdG eigenproblem (host driver in C).
 *
 * The reference hands A = diag(K, M), B = Swap and T = [T_K; T_M] to an
 * external iLOBPCG (reference src/solver.c:147-164, 262).  Here the same
 * eigenproblem is solved in its Hermitian-definite form
 *     (M K M) v = w^2 M v,   u = M v / w
 * (K u = w v, M v = w u), on the v half only:
 *
 *   basis V = [X | W | P] (S x k, ld S) in HBM with images MV, KMV;
 *   W = T_M T_K R (the reference's precondK / precondM), R = KMX - w^2 X;
 *   Gram matrices V^H V, V^H MV, MV^H KMV on the device (FP64 accumulation),
 *   Rayleigh-Ritz on the host in FP64 (dense.c), block updates on the device.
 *
 * Soft locking: converged columns get no W / P.  When [X, W, P] is
 * numerically dependent the P block is dropped for that iteration.
 */
#include <complex.h>
#include <time.h>

#include "bdg_host.h"
#include "dense.h"

double bdg_wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
  void *V, *MV, *KMV;
} blk3;

static size_t elem_bytes(int prec, int cplx) {
  return (size_t)(prec == BDGX_FP64 ? 8 : 4) * (cplx ? 2 : 1);
}

#define TRY(x)               \
  do {                       \
    rc = (x);                \
    if (rc) goto done;       \
  } while (0)

static int alloc3(blk3 *b, size_t bytes) {
  memset(b, 0, sizeof(*b));
  int rc = bdgx_malloc(&b->V, bytes);
  if (!rc) rc = bdgx_malloc(&b->MV, bytes);
  if (!rc) rc = bdgx_malloc(&b->KMV, bytes);
  return rc;
}

static void free3(blk3 *b) {
  bdgx_free(b->V);
  bdgx_free(b->MV);
  bdgx_free(b->KMV);
  memset(b, 0, sizeof(*b));
}

static char *colp(void *base, uint64_t col, uint64_t S, size_t es) {
  return (char *)base + (size_t)col * S * es;
}

/* y <- A[:, 0:k] * C for the three blocks */
static int combine3(bdgx_ctx *ctx, int prec, const blk3 *src, uint64_t col0, uint64_t k,
                    const double *dC, blk3 *dst, uint64_t dcol0, uint64_t m, uint64_t S,
                    size_t es) {
  int rc = bdgx_combine(ctx, prec, colp(src->V, col0, S, es), S, k, dC,
                        colp(dst->V, dcol0, S, es), S, m, S);
  if (!rc)
    rc = bdgx_combine(ctx, prec, colp(src->MV, col0, S, es), S, k, dC,
                      colp(dst->MV, dcol0, S, es), S, m, S);
  if (!rc)
    rc = bdgx_combine(ctx, prec, colp(src->KMV, col0, S, es), S, k, dC,
                      colp(dst->KMV, dcol0, S, es), S, m, S);
  return rc;
}

/* column copy (same precision) */
static int copy_col(bdgx_ctx *ctx, int prec, const void *src, void *dst, uint64_t S) {
  return bdgx_convert(ctx, prec, src, S, prec, dst, S, S, 1);
}

/* Gram set on V[:, 0:k]: GV = V^H V, GM = V^H MV, GA = MV^H KMV (host out). */
static int grams(bdgx_ctx *ctx, int prec, const blk3 *b, uint64_t k, uint64_t S, int cplx,
                 double *dG, double *hGV, double *hGM, double *hGA) {
  const size_t gsz = (size_t)k * k * (cplx ? 2 : 1);
  int rc = bdgx_gram(ctx, prec, b->V, S, k, b->V, S, k, S, 0, dG);
  if (!rc) rc = bdgx_gram(ctx, prec, b->V, S, k, b->MV, S, k, S, 0, dG + gsz);
  if (!rc) rc = bdgx_gram(ctx, prec, b->MV, S, k, b->KMV, S, k, S, 0, dG + 2 * gsz);
  if (rc) return rc;
  double *tmp = (double *)malloc(sizeof(double) * 3 * gsz);
  rc = bdgx_memcpy_d2h(ctx, tmp, dG, sizeof(double) * 3 * gsz);
  if (!rc) {
    memcpy(hGV, tmp, sizeof(double) * gsz);
    memcpy(hGM, tmp + gsz, sizeof(double) * gsz);
    memcpy(hGA, tmp + 2 * gsz, sizeof(double) * gsz);
  }
  free(tmp);
  return rc;
}

/* extract the leading kk x kk block of a k x k (complex-aware) matrix */
static void lead_block(const double *G, int k, int kk, int cplx, double *out) {
  const int w = cplx ? 2 : 1;
  for (int j = 0; j < kk; ++j)
    for (int i = 0; i < kk; ++i)
      for (int q = 0; q < w; ++q) out[(i + (size_t)j * kk) * w + q] = G[(i + (size_t)j * k) * w + q];
}

/* Cholesky pivot test of the diagonally scaled Gram (real or complex) */
static int gram_ok(const double *G, int k, int cplx) {
  const int w = cplx ? 2 : 1;
  double complex *A = (double complex *)malloc(sizeof(double complex) * (size_t)k * k);
  double *d = (double *)malloc(sizeof(double) * (size_t)k);
  int ok = 1;
  for (int i = 0; i < k; ++i) {
    const double g = G[(i + (size_t)i * k) * w];
    if (!(g > 0)) ok = 0;
    d[i] = g > 0 ? 1.0 / sqrt(g) : 0.0;
  }
  for (int j = 0; j < k && ok; ++j)
    for (int i = 0; i < k; ++i) {
      const double re = G[(i + (size_t)j * k) * w];
      const double im = cplx ? G[(i + (size_t)j * k) * w + 1] : 0.0;
      A[i + (size_t)j * k] = (re + I * im) * d[i] * d[j];
    }
  for (int j = 0; j < k && ok; ++j) {
    double djj = creal(A[j + (size_t)j * k]);
    for (int p = 0; p < j; ++p) djj -= creal(A[j + (size_t)p * k] * conj(A[j + (size_t)p * k]));
    if (!(djj >= 1e-10)) {
      ok = 0;
      break;
    }
    const double l = sqrt(djj);
    A[j + (size_t)j * k] = l;
    for (int i = j + 1; i < k; ++i) {
      double complex s = A[i + (size_t)j * k];
      for (int p = 0; p < j; ++p) s -= A[i + (size_t)p * k] * conj(A[j + (size_t)p * k]);
      A[i + (size_t)j * k] = s / l;
    }
  }
  free(A);
  free(d);
  return ok;
}

static double signed_sqrt(double x) { return x >= 0 ? sqrt(x) : -sqrt(-x); }

int bdg_lobpcg_device(bdgx_ctx *ctx, int prec, const lobpcg_cfg *cfg, const void *dX0,
                      lobpcg_result *res, void **dXout, void **dMXout) {
  const uint64_t S = cfg->S;
  const int m = cfg->m, nev = cfg->nev, cplx = cfg->cplx;
  const int w = cplx ? 2 : 1;
  const size_t es = elem_bytes(prec, cplx);
  const int kmax = 3 * m;
  int rc = 0;
  blk3 A, B, P;
  memset(&A, 0, sizeof(A));
  memset(&B, 0, sizeof(B));
  memset(&P, 0, sizeof(P));
  void *dR = NULL, *dTmp = NULL;
  double *dG = NULL, *dC = NULL, *dlam = NULL, *dn2 = NULL;
  const size_t gsz = (size_t)kmax * kmax * w;
  double *hGV = (double *)malloc(sizeof(double) * gsz);
  double *hGM = (double *)malloc(sizeof(double) * gsz);
  double *hGA = (double *)malloc(sizeof(double) * gsz);
  double *hT1 = (double *)malloc(sizeof(double) * gsz);
  double *hT2 = (double *)malloc(sizeof(double) * gsz);
  double *hT3 = (double *)malloc(sizeof(double) * gsz);
  double *hY = (double *)malloc(sizeof(double) * (size_t)kmax * m * w);
  double *hYp = (double *)malloc(sizeof(double) * (size_t)kmax * m * w);
  double *w2 = (double *)calloc((size_t)m, sizeof(double));
  int *zf = (int *)calloc((size_t)m, sizeof(int));
  double *hn2 = (double *)malloc(sizeof(double) * 3 * (size_t)m);
  double *resv = (double *)malloc(sizeof(double) * (size_t)m);
  int *conv = (int *)calloc((size_t)m, sizeof(int));
  int *act = (int *)malloc(sizeof(int) * (size_t)m);
  int haveP = 0, it = 0;
  double trr = 0.0;
  res->status = 0;
  res->iters = 0;
  res->restarts = 0;
  res->precond_fallbacks = 0;
  res->applies = 0;

  TRY(alloc3(&A, (size_t)S * kmax * es));
  TRY(alloc3(&B, (size_t)S * kmax * es));
  TRY(alloc3(&P, (size_t)S * m * es));
  TRY(bdgx_malloc(&dR, (size_t)S * m * es));
  TRY(bdgx_malloc(&dTmp, (size_t)S * m * es));
  TRY(bdgx_malloc((void **)&dG, sizeof(double) * 3 * gsz));
  TRY(bdgx_malloc((void **)&dC, sizeof(double) * (size_t)kmax * m * w));
  TRY(bdgx_malloc((void **)&dlam, sizeof(double) * (size_t)m));
  TRY(bdgx_malloc((void **)&dn2, sizeof(double) * 3 * (size_t)m));

  /* start: X = X0, MX = M X, KMX = K MX, Rayleigh-Ritz on X alone */
  TRY(bdgx_convert(ctx, prec, dX0, S, prec, A.V, S, S, (uint64_t)m));
  TRY(bdgx_apply(ctx, BDGX_OP_M, prec, A.V, A.MV, (uint64_t)m, S));
  TRY(bdgx_apply(ctx, BDGX_OP_K, prec, A.MV, A.KMV, (uint64_t)m, S));
  res->applies += (uint64_t)m;
  TRY(grams(ctx, prec, &A, (uint64_t)m, S, cplx, dG, hGV, hGM, hGA));
  {
    const double t0 = bdg_wtime();
    const int got = dn_rayleigh_ritz(m, m, cplx, hGV, hGM, hGA, w2, hY, zf);
    trr += bdg_wtime() - t0;
    if (got < m) {
      rc = -100;
      goto done;
    }
  }
  TRY(bdgx_memcpy_h2d(ctx, dC, hY, sizeof(double) * (size_t)m * m * w));
  TRY(combine3(ctx, prec, &A, 0, (uint64_t)m, dC, &B, 0, (uint64_t)m, S, es));
  {
    blk3 t = A;
    A = B;
    B = t;
  }

  for (it = 1; it <= cfg->maxiter; ++it) {
    /* residuals R = KMX - w2 X and norms */
    TRY(bdgx_memcpy_h2d(ctx, dlam, w2, sizeof(double) * (size_t)m));
    TRY(bdgx_residual(ctx, prec, A.KMV, A.V, A.MV, S, (uint64_t)m, S, dlam, 0, dR, S, dn2));
    TRY(bdgx_memcpy_d2h(ctx, hn2, dn2, sizeof(double) * 3 * (size_t)m));
    double scale = 1.0;
    for (int j = 0; j < nev && j < m; ++j) scale = fmax(scale, fabs(signed_sqrt(w2[j])));
    int nact = 0, allc = 1;
    double worst = 0.0;
    for (int j = 0; j < m; ++j) {
      const double wj = signed_sqrt(w2[j]);
      const double aw = fmax(fabs(wj), 1e-300);
      const double rn = sqrt(hn2[3 * j]), xn = sqrt(hn2[3 * j + 1]), mn = sqrt(hn2[3 * j + 2]);
      const double xnorm = sqrt(xn * xn + (mn / aw) * (mn / aw));
      resv[j] = (rn / aw) / (xnorm * fmax(1.0, fabs(wj)));
      const int zero = zf[j] || fabs(wj) < cfg->zero_tol * scale;
      conv[j] = zero || resv[j] < cfg->tol;
      if (j < nev) {
        if (!conv[j]) allc = 0;
        if (!zero) worst = fmax(worst, resv[j]);
      }
      if (!conv[j]) act[nact++] = j;
    }
    if (cfg->verbosity > 1 || (cfg->verbosity == 1 && it % 10 == 0))
      fprintf(stderr, "[bdg] it %d prec %s w0 %.10g wlast %.10g worst res %.3e active %d\n", it,
              prec == BDGX_FP64 ? "fp64" : "fp32", signed_sqrt(w2[0]),
              signed_sqrt(w2[nev - 1]), worst, nact);
    if (allc) break;
    if (it == cfg->maxiter) break;
    /* W = T_M T_K R[:, act] into A.V[:, m : m+nact] */
    for (int a = 0; a < nact; ++a)
      TRY(copy_col(ctx, prec, colp(dR, (uint64_t)act[a], S, es), colp(dTmp, (uint64_t)a, S, es), S));
    TRY(bdgx_apply(ctx, BDGX_OP_PRECOND_K, prec, dTmp, dR, (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_PRECOND_M, prec, dR, colp(A.V, (uint64_t)m, S, es),
                   (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_M, prec, colp(A.V, (uint64_t)m, S, es),
                   colp(A.MV, (uint64_t)m, S, es), (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_K, prec, colp(A.MV, (uint64_t)m, S, es),
                   colp(A.KMV, (uint64_t)m, S, es), (uint64_t)nact, S));
    res->applies += (uint64_t)nact;
    int k = m + nact;
    if (haveP) {
      for (int a = 0; a < nact; ++a) {
        const uint64_t src = (uint64_t)act[a], dst = (uint64_t)(m + nact + a);
        TRY(copy_col(ctx, prec, colp(P.V, src, S, es), colp(A.V, dst, S, es), S));
        TRY(copy_col(ctx, prec, colp(P.MV, src, S, es), colp(A.MV, dst, S, es), S));
        TRY(copy_col(ctx, prec, colp(P.KMV, src, S, es), colp(A.KMV, dst, S, es), S));
      }
      k += nact;
    }
    TRY(grams(ctx, prec, &A, (uint64_t)k, S, cplx, dG, hGV, hGM, hGA));
    double t0 = bdg_wtime();
    if (haveP && !gram_ok(hGV, k, cplx)) {
      /* [X, W, P] numerically dependent: drop P for this iteration */
      const int kk = m + nact;
      lead_block(hGV, k, kk, cplx, hT1);
      lead_block(hGM, k, kk, cplx, hT2);
      lead_block(hGA, k, kk, cplx, hT3);
      memcpy(hGV, hT1, sizeof(double) * (size_t)kk * kk * w);
      memcpy(hGM, hT2, sizeof(double) * (size_t)kk * kk * w);
      memcpy(hGA, hT3, sizeof(double) * (size_t)kk * kk * w);
      k = kk;
      res->restarts++;
    }
    if (!gram_ok(hGV, k, cplx)) {
      /* [X, W] still dependent: the preconditioner collapsed the residuals
       * (e.g. a 1e8 gain on one mode); use W = R for this iteration */
      trr += bdg_wtime() - t0;
      k = m + nact;
      for (int a = 0; a < nact; ++a)
        TRY(copy_col(ctx, prec, colp(dTmp, (uint64_t)a, S, es),
                     colp(A.V, (uint64_t)(m + a), S, es), S));
      TRY(bdgx_apply(ctx, BDGX_OP_M, prec, colp(A.V, (uint64_t)m, S, es),
                     colp(A.MV, (uint64_t)m, S, es), (uint64_t)nact, S));
      TRY(bdgx_apply(ctx, BDGX_OP_K, prec, colp(A.MV, (uint64_t)m, S, es),
                     colp(A.KMV, (uint64_t)m, S, es), (uint64_t)nact, S));
      res->applies += (uint64_t)nact;
      res->precond_fallbacks++;
      TRY(grams(ctx, prec, &A, (uint64_t)k, S, cplx, dG, hGV, hGM, hGA));
      t0 = bdg_wtime();
    }
    const int got = dn_rayleigh_ritz(k, m, cplx, hGV, hGM, hGA, w2, hY, zf);
    trr += bdg_wtime() - t0;
    if (got < m) {
      rc = -100;
      goto done;
    }
    /* X' = V C, P' = V[:, m:k] C[m:k, :] */
    TRY(bdgx_memcpy_h2d(ctx, dC, hY, sizeof(double) * (size_t)k * m * w));
    TRY(combine3(ctx, prec, &A, 0, (uint64_t)k, dC, &B, 0, (uint64_t)m, S, es));
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < k - m; ++i)
        for (int q = 0; q < w; ++q)
          hYp[(i + (size_t)j * (k - m)) * w + q] = hY[((m + i) + (size_t)j * k) * w + q];
    TRY(bdgx_memcpy_h2d(ctx, dC, hYp, sizeof(double) * (size_t)(k - m) * m * w));
    TRY(combine3(ctx, prec, &A, (uint64_t)m, (uint64_t)(k - m), dC, &P, 0, (uint64_t)m, S, es));
    haveP = 1;
    {
      blk3 t = A;
      A = B;
      B = t;
    }
  }
  if (it > cfg->maxiter) it = cfg->maxiter;
  TRY(bdgx_sync(ctx));
  /* hand back X and MX (copy into fresh buffers) */
  {
    void *x = NULL, *mx = NULL;
    TRY(bdgx_malloc(&x, (size_t)S * m * es));
    rc = bdgx_malloc(&mx, (size_t)S * m * es);
    if (!rc) rc = bdgx_convert(ctx, prec, A.V, S, prec, x, S, S, (uint64_t)m);
    if (!rc) rc = bdgx_convert(ctx, prec, A.MV, S, prec, mx, S, S, (uint64_t)m);
    if (!rc) rc = bdgx_sync(ctx);
    if (rc) {
      bdgx_free(x);
      bdgx_free(mx);
      goto done;
    }
    *dXout = x;
    *dMXout = mx;
  }
  for (int j = 0; j < m; ++j) {
    res->w2[j] = w2[j];
    res->zflag[j] = zf[j];
    res->res[j] = resv[j];
  }
done:
  res->iters = it;
  res->rr_seconds = trr;
  res->status = rc;
  free3(&A);
  free3(&B);
  free3(&P);
  bdgx_free(dR);
  bdgx_free(dTmp);
  bdgx_free(dG);
  bdgx_free(dC);
  bdgx_free(dlam);
  bdgx_free(dn2);
  free(hGV);
  free(hGM);
  free(hGA);
  free(hT1);
  free(hT2);
  free(hT3);
  free(hY);
  free(hYp);
  free(w2);
  free(zf);
  free(hn2);
  free(resv);
  free(conv);
  free(act);
  return rc;
}
