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
 *   Gram matrices V^H V, V^H MV, MV^H KMV with FP64 accumulation;
 *   Rayleigh-Ritz on the host in FP64 (dense.c), block updates on the device.
 *
 * Gram reuse: X and P are linear combinations of the previous basis, so their
 * Gram blocks follow from the previous Gram matrices and the Rayleigh-Ritz
 * coefficients (Q^H G Q); only the blocks that involve the new W are formed
 * on the device.  A full device recompute runs every gram_refresh iterations.
 *
 * Soft locking: converged columns get no W / P.  When [X, W, P] is
 * numerically dependent the P block is dropped for that iteration; when
 * [X, W] still is, W falls back to the plain residual.
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

#define TRY(x)         \
  do {                 \
    rc = (x);          \
    if (rc) goto done; \
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

/* dst[:, dcol0 : dcol0+m] = src[:, col0 : col0+k] * C, for the three blocks */
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

/* ---- small host matrix helpers (complex-aware, column-major) ------------ */

typedef struct {
  int w; /* 1 real, 2 complex */
} mctx;

static inline double complex getz(const double *A, int ld, int i, int j, int w) {
  const double *p = A + ((size_t)i + (size_t)j * ld) * w;
  return w == 2 ? p[0] + I * p[1] : p[0];
}

static inline void setz(double *A, int ld, int i, int j, int w, double complex v) {
  double *p = A + ((size_t)i + (size_t)j * ld) * w;
  p[0] = creal(v);
  if (w == 2) p[1] = cimag(v);
}

/* out (q x q) = Q^H G Q, G k x k, Q k x q */
static void qhgq(const double *G, int k, const double *Q, int q, int w, double *out) {
  double complex *T = (double complex *)malloc(sizeof(double complex) * (size_t)k * q);
  for (int j = 0; j < q; ++j)
    for (int i = 0; i < k; ++i) {
      double complex s = 0;
      for (int p = 0; p < k; ++p) s += getz(G, k, i, p, w) * getz(Q, k, p, j, w);
      T[i + (size_t)j * k] = s;
    }
  for (int j = 0; j < q; ++j)
    for (int i = 0; i < q; ++i) {
      double complex s = 0;
      for (int p = 0; p < k; ++p) s += conj(getz(Q, k, p, i, w)) * T[p + (size_t)j * k];
      setz(out, q, i, j, w, s);
    }
  free(T);
}

/* Cholesky pivot test of the diagonally scaled Gram (real or complex) */
static int gram_ok(const double *G, int k, int w) {
  double complex *A = (double complex *)malloc(sizeof(double complex) * (size_t)k * k);
  double *d = (double *)malloc(sizeof(double) * (size_t)k);
  int ok = 1;
  for (int i = 0; i < k; ++i) {
    const double g = creal(getz(G, k, i, i, w));
    if (!(g > 0)) ok = 0;
    d[i] = g > 0 ? 1.0 / sqrt(g) : 0.0;
  }
  for (int j = 0; j < k && ok; ++j)
    for (int i = 0; i < k; ++i) A[i + (size_t)j * k] = getz(G, k, i, j, w) * d[i] * d[j];
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

static void lead_block(const double *G, int k, int kk, int w, double *out) {
  for (int j = 0; j < kk; ++j)
    for (int i = 0; i < kk; ++i) setz(out, kk, i, j, w, getz(G, k, i, j, w));
}

static double signed_sqrt(double x) { return x >= 0 ? sqrt(x) : -sqrt(-x); }

/* ---- Gram sets --------------------------------------------------------- */

typedef struct {
  double *V, *M, *A; /* k x k each (complex interleaved when w = 2) */
} gram3;

/* full device Grams on V[:, 0:k] */
static int grams_full(bdgx_ctx *ctx, int prec, const blk3 *b, int k, uint64_t S, int w,
                      double *dG, gram3 *h) {
  const size_t gsz = (size_t)k * k * w;
  int rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, b->V, S, (uint64_t)k, S, 0, dG);
  if (!rc) rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, b->MV, S, (uint64_t)k, S, 0, dG + gsz);
  if (!rc)
    rc = bdgx_gram(ctx, prec, b->MV, S, (uint64_t)k, b->KMV, S, (uint64_t)k, S, 0, dG + 2 * gsz);
  if (rc) return rc;
  double *tmp = (double *)malloc(sizeof(double) * 3 * gsz);
  rc = bdgx_memcpy_d2h(ctx, tmp, dG, sizeof(double) * 3 * gsz);
  if (!rc) {
    memcpy(h->V, tmp, sizeof(double) * gsz);
    memcpy(h->M, tmp + gsz, sizeof(double) * gsz);
    memcpy(h->A, tmp + 2 * gsz, sizeof(double) * gsz);
  }
  free(tmp);
  return rc;
}

/* device Grams of V[:, 0:k] against the columns [c0, c0+nc) of V; written
 * into columns c0.. of the full k x k sets, and mirrored into the rows. */
static int grams_cols(bdgx_ctx *ctx, int prec, const blk3 *b, int k, int c0, int nc, uint64_t S,
                      int w, size_t es, double *dG, gram3 *h) {
  const size_t csz = (size_t)k * nc * w;
  int rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, colp(b->V, (uint64_t)c0, S, es), S,
                     (uint64_t)nc, S, 0, dG);
  if (!rc)
    rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, colp(b->MV, (uint64_t)c0, S, es), S,
                   (uint64_t)nc, S, 0, dG + csz);
  if (!rc)
    rc = bdgx_gram(ctx, prec, b->MV, S, (uint64_t)k, colp(b->KMV, (uint64_t)c0, S, es), S,
                   (uint64_t)nc, S, 0, dG + 2 * csz);
  if (rc) return rc;
  double *tmp = (double *)malloc(sizeof(double) * 3 * csz);
  rc = bdgx_memcpy_d2h(ctx, tmp, dG, sizeof(double) * 3 * csz);
  if (!rc) {
    double *dst[3] = {h->V, h->M, h->A};
    for (int g = 0; g < 3; ++g)
      for (int j = 0; j < nc; ++j)
        for (int i = 0; i < k; ++i) {
          const double complex v = getz(tmp + g * csz, k, i, j, w);
          setz(dst[g], k, i, c0 + j, w, v);
          setz(dst[g], k, c0 + j, i, w, conj(v));
        }
  }
  free(tmp);
  return rc;
}

int bdg_lobpcg_device(bdgx_ctx *ctx, int prec, const lobpcg_cfg *cfg, const void *dX0,
                      const void *dP0, lobpcg_result *res, void **dXout, void **dPout) {
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
  gram3 G, Gp;
  G.V = (double *)calloc(gsz, sizeof(double));
  G.M = (double *)calloc(gsz, sizeof(double));
  G.A = (double *)calloc(gsz, sizeof(double));
  Gp.V = (double *)calloc(gsz, sizeof(double));
  Gp.M = (double *)calloc(gsz, sizeof(double));
  Gp.A = (double *)calloc(gsz, sizeof(double));
  double *hT = (double *)malloc(sizeof(double) * gsz);
  double *hQ = (double *)malloc(sizeof(double) * (size_t)kmax * kmax * w);
  double *hY = (double *)malloc(sizeof(double) * (size_t)kmax * m * w);
  double *hYp = (double *)malloc(sizeof(double) * (size_t)kmax * m * w);
  double *w2 = (double *)calloc((size_t)m, sizeof(double));
  int *zf = (int *)calloc((size_t)m, sizeof(int));
  double *hn2 = (double *)malloc(sizeof(double) * 3 * (size_t)m);
  double *resv = (double *)malloc(sizeof(double) * (size_t)m);
  int *conv = (int *)calloc((size_t)m, sizeof(int));
  int *act = (int *)malloc(sizeof(int) * (size_t)m);
  int haveP = 0, it = 0, kprev = 0, gvalid = 0, since_full = 0, since_img = 0;
  double trr = 0.0, best = 1e300;
  int best_it = 0;
  res->status = 0;
  res->iters = 0;
  res->restarts = 0;
  res->precond_fallbacks = 0;
  res->stalled = 0;
  res->applies = 0;

  TRY(alloc3(&A, (size_t)S * kmax * es));
  TRY(alloc3(&B, (size_t)S * kmax * es));
  TRY(alloc3(&P, (size_t)S * m * es));
  TRY(bdgx_malloc(&dR, (size_t)S * m * es));
  TRY(bdgx_malloc(&dTmp, (size_t)S * m * es));
  TRY(bdgx_malloc((void **)&dG, sizeof(double) * 3 * gsz));
  TRY(bdgx_malloc((void **)&dC, sizeof(double) * (size_t)kmax * kmax * w));
  TRY(bdgx_malloc((void **)&dlam, sizeof(double) * (size_t)m));
  TRY(bdgx_malloc((void **)&dn2, sizeof(double) * 3 * (size_t)m));

  /* start: basis [X0] or [X0 | P0], images by M then K, Rayleigh-Ritz */
  {
    const int k0 = dP0 ? 2 * m : m;
    TRY(bdgx_convert(ctx, prec, dX0, S, prec, A.V, S, S, (uint64_t)m));
    if (dP0) TRY(bdgx_convert(ctx, prec, dP0, S, prec, colp(A.V, (uint64_t)m, S, es), S, S, (uint64_t)m));
    TRY(bdgx_apply(ctx, BDGX_OP_M, prec, A.V, A.MV, (uint64_t)k0, S));
    TRY(bdgx_apply(ctx, BDGX_OP_K, prec, A.MV, A.KMV, (uint64_t)k0, S));
    res->applies += (uint64_t)k0;
    TRY(grams_full(ctx, prec, &A, k0, S, w, dG, &G));
    int k = k0;
    if (dP0 && !gram_ok(G.V, k0, w)) {
      lead_block(G.V, k0, m, w, hT);
      memcpy(G.V, hT, sizeof(double) * (size_t)m * m * w);
      lead_block(G.M, k0, m, w, hT);
      memcpy(G.M, hT, sizeof(double) * (size_t)m * m * w);
      lead_block(G.A, k0, m, w, hT);
      memcpy(G.A, hT, sizeof(double) * (size_t)m * m * w);
      k = m;
    }
    const double t0 = bdg_wtime();
    const int got = dn_rayleigh_ritz(k, m, cplx, G.V, G.M, G.A, w2, hY, zf);
    trr += bdg_wtime() - t0;
    if (got < m) {
      rc = -100;
      goto done;
    }
    TRY(bdgx_memcpy_h2d(ctx, dC, hY, sizeof(double) * (size_t)k * m * w));
    TRY(combine3(ctx, prec, &A, 0, (uint64_t)k, dC, &B, 0, (uint64_t)m, S, es));
    if (k > m) {
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < k - m; ++i)
          for (int q = 0; q < w; ++q)
            hYp[(i + (size_t)j * (k - m)) * w + q] = hY[((m + i) + (size_t)j * k) * w + q];
      TRY(bdgx_memcpy_h2d(ctx, dC, hYp, sizeof(double) * (size_t)(k - m) * m * w));
      TRY(combine3(ctx, prec, &A, (uint64_t)m, (uint64_t)(k - m), dC, &P, 0, (uint64_t)m, S, es));
      haveP = 1;
    }
    memcpy(Gp.V, G.V, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.M, G.M, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.A, G.A, sizeof(double) * (size_t)k * k * w);
    kprev = k;
    gvalid = 1;
    {
      blk3 t = A;
      A = B;
      B = t;
    }
  }

  for (it = 1; it <= cfg->maxiter; ++it) {
    /* optional periodic re-application of M, K to X (removes the drift of
     * images kept by linear combination; P is dropped at that point) */
    if (cfg->image_refresh > 0 && ++since_img >= cfg->image_refresh) {
      since_img = 0;
      TRY(bdgx_apply(ctx, BDGX_OP_M, prec, A.V, A.MV, (uint64_t)m, S));
      TRY(bdgx_apply(ctx, BDGX_OP_K, prec, A.MV, A.KMV, (uint64_t)m, S));
      res->applies += (uint64_t)m;
      gvalid = 0;
    }
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
    if (cfg->stall_window > 0) {
      if (worst < cfg->stall_factor * best) {
        best = worst;
        best_it = it;
      } else if (it - best_it >= cfg->stall_window || worst > 10.0 * best) {
        res->stalled = 1;
        break;
      }
    }
    /* W = T_M T_K R[:, act] into A.V[:, m : m+nact] */
    TRY(bdgx_gather_cols(ctx, prec, dR, S, act, (uint64_t)nact, dTmp, S, S));
    TRY(bdgx_apply(ctx, BDGX_OP_PRECOND_K, prec, dTmp, dR, (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_PRECOND_M, prec, dR, colp(A.V, (uint64_t)m, S, es),
                   (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_M, prec, colp(A.V, (uint64_t)m, S, es),
                   colp(A.MV, (uint64_t)m, S, es), (uint64_t)nact, S));
    TRY(bdgx_apply(ctx, BDGX_OP_K, prec, colp(A.MV, (uint64_t)m, S, es),
                   colp(A.KMV, (uint64_t)m, S, es), (uint64_t)nact, S));
    res->applies += (uint64_t)nact;
    const int np = haveP ? nact : 0;
    int k = m + nact + np;
    if (np) {
      TRY(bdgx_gather_cols(ctx, prec, P.V, S, act, (uint64_t)nact,
                           colp(A.V, (uint64_t)(m + nact), S, es), S, S));
      TRY(bdgx_gather_cols(ctx, prec, P.MV, S, act, (uint64_t)nact,
                           colp(A.MV, (uint64_t)(m + nact), S, es), S, S));
      TRY(bdgx_gather_cols(ctx, prec, P.KMV, S, act, (uint64_t)nact,
                           colp(A.KMV, (uint64_t)(m + nact), S, es), S, S));
    }
    /* Gram set of [X | W | P_act] */
    const int full = !gvalid || cfg->gram_refresh <= 0 || ++since_full >= cfg->gram_refresh;
    if (full) {
      since_full = 0;
      TRY(grams_full(ctx, prec, &A, k, S, w, dG, &G));
    } else {
      /* X = V_prev C, P_act = V_prev[:, m:] C[m:, act]: blocks from Q^H G_prev Q */
      const int q = m + np;
      for (int j = 0; j < q; ++j)
        for (int i = 0; i < kprev; ++i) {
          double complex v = 0;
          if (j < m)
            v = getz(hY, kprev, i, j, w);
          else if (i >= m)
            v = getz(hY, kprev, i, act[j - m], w);
          setz(hQ, kprev, i, j, w, v);
        }
      double *gp[3] = {Gp.V, Gp.M, Gp.A}, *gn[3] = {G.V, G.M, G.A};
      for (int g = 0; g < 3; ++g) {
        qhgq(gp[g], kprev, hQ, q, w, hT);
        for (int j = 0; j < q; ++j)
          for (int i = 0; i < q; ++i) {
            const int ii = i < m ? i : i + nact, jj = j < m ? j : j + nact;
            setz(gn[g], k, ii, jj, w, getz(hT, q, i, j, w));
          }
      }
      TRY(grams_cols(ctx, prec, &A, k, m, nact, S, w, es, dG, &G));
    }
    double t0 = bdg_wtime();
    if (np && !gram_ok(G.V, k, w)) {
      /* [X, W, P] numerically dependent: drop P for this iteration */
      const int kk = m + nact;
      double *gn[3] = {G.V, G.M, G.A};
      for (int g = 0; g < 3; ++g) {
        lead_block(gn[g], k, kk, w, hT);
        memcpy(gn[g], hT, sizeof(double) * (size_t)kk * kk * w);
      }
      k = kk;
      res->restarts++;
    }
    if (!gram_ok(G.V, k, w)) {
      /* [X, W] still dependent: the preconditioner collapsed the residuals
       * (e.g. a 1e8 gain on one mode); use W = R for this iteration */
      trr += bdg_wtime() - t0;
      k = m + nact;
      TRY(bdgx_convert(ctx, prec, dTmp, S, prec, colp(A.V, (uint64_t)m, S, es), S, S,
                       (uint64_t)nact));
      TRY(bdgx_apply(ctx, BDGX_OP_M, prec, colp(A.V, (uint64_t)m, S, es),
                     colp(A.MV, (uint64_t)m, S, es), (uint64_t)nact, S));
      TRY(bdgx_apply(ctx, BDGX_OP_K, prec, colp(A.MV, (uint64_t)m, S, es),
                     colp(A.KMV, (uint64_t)m, S, es), (uint64_t)nact, S));
      res->applies += (uint64_t)nact;
      res->precond_fallbacks++;
      TRY(grams_full(ctx, prec, &A, k, S, w, dG, &G));
      since_full = 0;
      t0 = bdg_wtime();
    }
    const int got = dn_rayleigh_ritz(k, m, cplx, G.V, G.M, G.A, w2, hY, zf);
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
    memcpy(Gp.V, G.V, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.M, G.M, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.A, G.A, sizeof(double) * (size_t)k * k * w);
    kprev = k;
    gvalid = 1;
    {
      blk3 t = A;
      A = B;
      B = t;
    }
  }
  if (it > cfg->maxiter) it = cfg->maxiter;
  TRY(bdgx_sync(ctx));
  /* hand back X and P (copies into fresh buffers) */
  {
    void *x = NULL, *p = NULL;
    TRY(bdgx_malloc(&x, (size_t)S * m * es));
    rc = bdgx_malloc(&p, (size_t)S * m * es);
    if (!rc) rc = bdgx_convert(ctx, prec, A.V, S, prec, x, S, S, (uint64_t)m);
    if (!rc && haveP) rc = bdgx_convert(ctx, prec, P.V, S, prec, p, S, S, (uint64_t)m);
    if (!rc && !haveP) {
      bdgx_free(p);
      p = NULL;
    }
    if (!rc) rc = bdgx_sync(ctx);
    if (rc) {
      bdgx_free(x);
      bdgx_free(p);
      goto done;
    }
    *dXout = x;
    *dPout = p;
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
  free(G.V);
  free(G.M);
  free(G.A);
  free(Gp.V);
  free(Gp.M);
  free(Gp.A);
  free(hT);
  free(hQ);
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
