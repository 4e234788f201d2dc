/*
 * Device block LOBPCG for the BdG eigenproblem (host driver in C).
 *
 * The reference hands A = diag(K, M), B = Swap and T = [T_K; T_M] to an
 * external iLOBPCG (reference src/solver.c:147-164, 262).  Here the same
 * eigenproblem is solved in its Hermitian-definite form
 *     (M K M) v = w^2 M v,   u = M v / w
 * (K u = w v, M v = w u), on the v half only:
 *
 *   basis V = [X | P | W] (S x k, ld S, FP64) in HBM with images MV, KMV;
 *   W = T_M T_K R (the reference's precondK / precondM), R = KMX - w^2 X;
 *   Gram matrices V^H V, V^H MV, MV^H KMV with FP64 accumulation;
 *   Rayleigh-Ritz on the host in FP64 (dense.c), block updates on the device.
 *
 * One combine per block produces the next X (all m columns) and the next P
 * (the active columns only) side by side, so they land at [X | P] of the next
 * basis with no copies.
 *
 * Gram reuse: X and P are linear combinations of the previous basis, so their
 * Gram blocks follow from the previous Gram matrices and the combine
 * coefficients (Q^H G Q); only the blocks that involve the new W are formed
 * on the device.  A full device recompute runs every gram_refresh iterations.
 *
 * Soft locking: converged columns get no W / P.  When [X, P, W] is
 * numerically dependent the P block is left out of that Rayleigh-Ritz; when
 * [X, W] still is, W falls back to the plain residual.
 *
 * When K == M (no interactions, no dipolar term) M is as singular as K on
 * the Goldstone direction and the product form divides rounding noise by a
 * vanishing M-norm; that case runs the same iteration on the standard form
 * K^2 v = w^2 v (B = I): images K V and K^2 V, Grams V^H V and V^H K^2 V.
 *
 * Mixed precision (w_prec = FP32): the basis, its Grams, the residuals and
 * the Rayleigh-Ritz stay FP64.  T_K and T_M always run in FP32 (W is only a
 * search direction, so its rounding changes which direction is searched, not
 * the accuracy of the basis).  Optionally (img_tol > 0), while the worst
 * residual is above img_tol, the images M W, K M W are FP32 too, of W
 * projected out of span(X); their error enters the carried images of X in
 * proportion to the W coefficients, and the images of X are re-applied in
 * FP64 every refresh_every iterations.  Once the worst residual reaches
 * img_tol the images are FP64 for the rest of the solve,
 * and convergence is only accepted on FP64 images.  FP32 images are off by
 * default: rounding noise in an FP32 image is amplified by the stiffness of
 * K M (about 1e-3 relative at 128^3), which stops interior modes from
 * converging while they are on, so they cost more FP64 iterations later
 * than they save.
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

/* Solve G Y = Z for Hermitian positive definite G (n x n) and n x r right-hand
 * sides, in place in Z, by Cholesky.  Returns 0, or -1 when a pivot is not
 * positive. */
static int chol_solve(const double *G, int n, int w, double *Z, int r) {
  double complex *L = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  int ok = 0;
  for (int j = 0; j < n && !ok; ++j) {
    double complex d = getz(G, n, j, j, w);
    for (int p = 0; p < j; ++p) d -= L[j + (size_t)p * n] * conj(L[j + (size_t)p * n]);
    if (!(creal(d) > 0)) {
      ok = -1;
      break;
    }
    const double l = sqrt(creal(d));
    L[j + (size_t)j * n] = l;
    for (int i = j + 1; i < n; ++i) {
      double complex s = getz(G, n, i, j, w);
      for (int p = 0; p < j; ++p) s -= L[i + (size_t)p * n] * conj(L[j + (size_t)p * n]);
      L[i + (size_t)j * n] = s / l;
    }
  }
  if (!ok)
    for (int c = 0; c < r; ++c) {
      /* forward L y = z, backward L^H x = y */
      for (int i = 0; i < n; ++i) {
        double complex s = getz(Z, n, i, c, w);
        for (int p = 0; p < i; ++p) s -= L[i + (size_t)p * n] * getz(Z, n, p, c, w);
        setz(Z, n, i, c, w, s / L[i + (size_t)i * n]);
      }
      for (int i = n - 1; i >= 0; --i) {
        double complex s = getz(Z, n, i, c, w);
        for (int p = i + 1; p < n; ++p) s -= conj(L[p + (size_t)i * n]) * getz(Z, n, p, c, w);
        setz(Z, n, i, c, w, s / L[i + (size_t)i * n]);
      }
    }
  free(L);
  return ok;
}

static double signed_sqrt(double x) { return x >= 0 ? sqrt(x) : -sqrt(-x); }

/* ---- Gram sets --------------------------------------------------------- */

typedef struct {
  double *V, *M, *A; /* k x k each (complex interleaved when w = 2) */
} gram3;

/* Gram blocks of [X | P] = V_prev [C_X | C_P] (the first q columns of the
 * new basis) from the previous Grams: Q^H G_prev Q, written into the
 * top-left q x q blocks of the new k x k sets */
static void project_grams(const gram3 *gp, int kprev, const double *hCp, int q, int k, int w,
                          double *hT, gram3 *gn) {
  const double *src[3] = {gp->V, gp->M, gp->A};
  double *dst[3] = {gn->V, gn->M, gn->A};
  for (int g = 0; g < 3; ++g) {
    dn_qhgq(kprev, q, w == 2, src[g], hCp, hT);
    for (int j = 0; j < q; ++j)
      memcpy(dst[g] + (size_t)j * k * w, hT + (size_t)j * q * w, sizeof(double) * (size_t)q * w);
  }
}

/* Frozen columns.  A cluster J of converged columns with equal Ritz values
 * (relative difference <= 1e-11 in w^2: a degenerate multiplet, which the
 * Rayleigh-Ritz may rotate freely, or a single column) whose new Ritz
 * vectors V c_j all lie within tol of span(X_J) (relative Euclidean norm of
 * the part of V c_j outside the rows J, from the Euclidean Gram) keeps its
 * vectors: c_j := e_j, and w^2_j becomes the Rayleigh quotient of X_j; the
 * residuals are then the ones last measured.  fz[j] = 1 marks the frozen
 * columns; returns their number. */
static double outside_part(const gram3 *G, int k, int w, const double *c, int j0, int j1,
                           double *d) {
  /* ||V d||^2 / ||V c||^2 with d = c outside the rows [j0, j1) */
  double qd = 0.0, qc = 0.0;
  if (w == 1) {
    for (int i = 0; i < k; ++i) d[i] = (i >= j0 && i < j1) ? 0.0 : c[i];
    for (int l = 0; l < k; ++l) {
      const double *g = G->V + (size_t)l * k;
      double td = 0.0, tc = 0.0;
      for (int i = 0; i < k; ++i) {
        td += g[i] * d[i];
        tc += g[i] * c[i];
      }
      qd += d[l] * td;
      qc += c[l] * tc;
    }
  } else {
    double complex *dz = (double complex *)d;
    const double complex *cz = (const double complex *)c;
    const double complex *gz = (const double complex *)G->V;
    for (int i = 0; i < k; ++i) dz[i] = (i >= j0 && i < j1) ? 0.0 : cz[i];
    double complex sd = 0.0, sc = 0.0;
    for (int l = 0; l < k; ++l) {
      const double complex *g = gz + (size_t)l * k;
      double complex td = 0.0, tc = 0.0;
      for (int i = 0; i < k; ++i) {
        td += conj(dz[i]) * g[i];
        tc += conj(cz[i]) * g[i];
      }
      sd += td * dz[l];
      sc += tc * cz[l];
    }
    qd = creal(sd);
    qc = creal(sc);
  }
  return qc > 0 ? qd / qc : 1.0;
}

static int freeze_cols(const gram3 *G, int k, int m, int w, const int *conv, double tol,
                       double *hC, double *w2, int *fz) {
  int n = 0;
  double *d = (double *)malloc(sizeof(double) * (size_t)k * w);
  for (int j = 0; j < m; ++j) fz[j] = 0;
  for (int j0 = 0; j0 < m;) {
    int j1 = j0 + 1;
    while (j1 < m && fabs(w2[j1] - w2[j0]) <= 1e-11 * fmax(1.0, fabs(w2[j0]))) ++j1;
    int ok = 1;
    for (int j = j0; j < j1 && ok; ++j)
      ok = conv[j] && outside_part(G, k, w, hC + (size_t)j * k * w, j0, j1, d) <= tol * tol;
    for (int j = j0; j < j1 && ok; ++j) {
      const double gm = creal(getz(G->M, k, j, j, w));
      if (!(gm != 0)) ok = 0;
    }
    if (ok) {
      for (int j = j0; j < j1; ++j) {
        double *c = hC + (size_t)j * k * w;
        memset(c, 0, sizeof(double) * (size_t)k * w);
        c[(size_t)j * w] = 1.0;
        w2[j] = creal(getz(G->A, k, j, j, w)) / creal(getz(G->M, k, j, j, w));
        fz[j] = 1;
        ++n;
      }
    }
    j0 = j1;
  }
  free(d);
  return n;
}

/* full device Grams on V[:, 0:k].  Product form: V^H V, V^H MV, MV^H KMV.
 * Standard form (std != 0, the "M" images hold K V and the "KM" images K^2 V):
 * V^H V, V^H V, V^H K^2 V. */
static int grams_full_dev(bdgx_ctx *ctx, int prec, int std, const blk3 *b, int k, uint64_t S,
                          int w, double *dG) {
  const size_t gsz = (size_t)k * k * w;
  int rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, b->V, S, (uint64_t)k, S, 0, dG);
  if (!rc && !std)
    rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, b->MV, S, (uint64_t)k, S, 0, dG + gsz);
  if (!rc)
    rc = bdgx_gram(ctx, prec, std ? b->V : b->MV, S, (uint64_t)k, b->KMV, S, (uint64_t)k, S, 0,
                   dG + 2 * gsz);
  return rc;
}

/* Host staging of the per-iteration small copies (eigenvalues and
 * coefficients up, norms and Grams down).  Page-locked when available: the
 * uploads are then queued without waiting for the device, so the next
 * kernels follow the running ones directly.  Each buffer is rewritten only
 * after a later synchronous download, by which time its upload has run. */
typedef struct {
  double *lam, *C, *n2, *G;
  void *block;
  int pinned;
} stage_t;

static int stage_init(stage_t *sg, int m, int kmax, int w) {
  const size_t nl = (size_t)m, nC = (size_t)kmax * 2 * m * w, nn = 3 * (size_t)m,
               nG = 3 * (size_t)kmax * kmax * w, tot = nl + nC + nn + nG;
  sg->pinned = bdgx_host_alloc(&sg->block, sizeof(double) * tot) == 0;
  if (!sg->pinned) sg->block = malloc(sizeof(double) * tot);
  if (!sg->block) return -1;
  sg->lam = (double *)sg->block;
  sg->C = sg->lam + nl;
  sg->n2 = sg->C + nC;
  sg->G = sg->n2 + nn;
  return 0;
}

static void stage_free(stage_t *sg) {
  if (!sg->block) return;
  if (sg->pinned)
    bdgx_host_free(sg->block);
  else
    free(sg->block);
  sg->block = NULL;
}

/* upload bytes of src through the staging buffer buf (src may be buf) */
static int stage_up(bdgx_ctx *ctx, const stage_t *sg, double *buf, void *dst, const double *src,
                    size_t bytes) {
  if (src != buf) memcpy(buf, src, bytes);
  return sg->pinned ? bdgx_memcpy_h2d_async(ctx, dst, buf, bytes)
                    : bdgx_memcpy_h2d(ctx, dst, buf, bytes);
}

static int grams_full_host(bdgx_ctx *ctx, const stage_t *sg, int std, int k, int w,
                           const double *dG, gram3 *h) {
  const size_t gsz = (size_t)k * k * w;
  double *tmp = sg->G;
  int rc = bdgx_memcpy_d2h(ctx, tmp, dG, sizeof(double) * 3 * gsz);
  if (!rc) {
    memcpy(h->V, tmp, sizeof(double) * gsz);
    memcpy(h->M, std ? tmp : tmp + gsz, sizeof(double) * gsz);
    memcpy(h->A, tmp + 2 * gsz, sizeof(double) * gsz);
  }
  return rc;
}

static int grams_full(bdgx_ctx *ctx, const stage_t *sg, int prec, int std, const blk3 *b, int k,
                      uint64_t S, int w, double *dG, gram3 *h) {
  int rc = grams_full_dev(ctx, prec, std, b, k, S, w, dG);
  return rc ? rc : grams_full_host(ctx, sg, std, k, w, dG, h);
}

/* device Grams of V[:, 0:k] against the columns [c0, c0+nc) of V; written
 * into columns c0.. of the full k x k sets, and mirrored into the rows. */
static int grams_cols_dev(bdgx_ctx *ctx, int prec, int std, const blk3 *b, int k, int c0, int nc,
                          uint64_t S, int w, size_t es, double *dG) {
  const size_t csz = (size_t)k * nc * w;
  int rc = BDGX_EINVAL;
  /* few new columns: V^H W and V^H M W in one pass over V */
  if (!std && w == 1 && 2 * nc <= 32)
    rc = bdgx_gram_pair(ctx, prec, b->V, S, (uint64_t)k, colp(b->V, (uint64_t)c0, S, es), S,
                        colp(b->MV, (uint64_t)c0, S, es), S, (uint64_t)nc, S, dG);
  if (rc) {
    rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, colp(b->V, (uint64_t)c0, S, es), S,
                   (uint64_t)nc, S, 0, dG);
    if (!rc && !std)
      rc = bdgx_gram(ctx, prec, b->V, S, (uint64_t)k, colp(b->MV, (uint64_t)c0, S, es), S,
                     (uint64_t)nc, S, 0, dG + csz);
  }
  if (!rc)
    rc = bdgx_gram(ctx, prec, std ? b->V : b->MV, S, (uint64_t)k,
                   colp(b->KMV, (uint64_t)c0, S, es), S, (uint64_t)nc, S, 0, dG + 2 * csz);
  return rc;
}

static int grams_cols_host(bdgx_ctx *ctx, const stage_t *sg, int std, int k, int c0, int nc,
                           int w, const double *dG, gram3 *h) {
  const size_t csz = (size_t)k * nc * w;
  double *tmp = sg->G;
  int rc = bdgx_memcpy_d2h(ctx, tmp, dG, sizeof(double) * 3 * csz);
  if (!rc) {
    if (std) memcpy(tmp + csz, tmp, sizeof(double) * csz);
    double *dst[3] = {h->V, h->M, h->A};
    for (int g = 0; g < 3; ++g)
      for (int j = 0; j < nc; ++j)
        for (int i = 0; i < k; ++i) {
          const double complex v = getz(tmp + g * csz, k, i, j, w);
          setz(dst[g], k, i, c0 + j, w, v);
          setz(dst[g], k, c0 + j, i, w, conj(v));
        }
  }
  return rc;
}

/* residual norms of X and the convergence decisions */
typedef struct {
  double *resv;
  int *conv, *act;
  int *rpos;      /* column of dR holding the residual of act[i] */
  int *sel;       /* columns evaluated by a partial evaluation */
  double *lsel;   /* their w^2 */
  int nact, allc;
  double worst;
} res_state;

/* Residuals R = KMX - w^2 X and the convergence decisions.  Full (nsel = 0):
 * every column, dR column j = column j.  Partial (nsel > 0): only the
 * columns sel[0..nsel), packed into dR columns 0..nsel-1; the other columns
 * keep the state of the last full evaluation (they had converged). */
static int eval_residuals(bdgx_ctx *ctx, const lobpcg_cfg *cfg, const stage_t *sg,
                          const blk3 *A, const double *w2, const int *zf, double *dlam,
                          double *dn2, void *dR, res_state *st, int nsel) {
  const uint64_t S = cfg->S;
  const int m = cfg->m, nev = cfg->nev;
  double *hn2 = sg->n2;
  int rc;
  if (nsel == 0) {
    rc = stage_up(ctx, sg, sg->lam, dlam, w2, sizeof(double) * (size_t)m);
    if (!rc)
      rc = bdgx_residual(ctx, BDGX_FP64, A->KMV, A->V, A->MV, S, (uint64_t)m, S, dlam, 0, dR, S,
                         dn2);
    if (!rc) rc = bdgx_memcpy_d2h(ctx, hn2, dn2, sizeof(double) * 3 * (size_t)m);
  } else {
    for (int i = 0; i < nsel; ++i) sg->lam[i] = w2[st->sel[i]];
    rc = stage_up(ctx, sg, sg->lam, dlam, sg->lam, sizeof(double) * (size_t)nsel);
    if (!rc)
      rc = bdgx_residual_cols(ctx, BDGX_FP64, A->KMV, A->V, A->MV, S, st->sel, (uint64_t)nsel, S,
                              dlam, 0, dR, S, dn2);
    if (!rc) rc = bdgx_memcpy_d2h(ctx, hn2, dn2, sizeof(double) * 3 * (size_t)nsel);
  }
  if (rc) return rc;
  double scale = 1.0;
  for (int j = 0; j < nev && j < m; ++j) scale = fmax(scale, fabs(signed_sqrt(w2[j])));
  const int ncols = nsel ? nsel : m;
  for (int i = 0; i < ncols; ++i) {
    const int j = nsel ? st->sel[i] : i;
    const double wj = signed_sqrt(w2[j]);
    const double aw = fmax(fabs(wj), 1e-300);
    const double rn = sqrt(hn2[3 * i]), xn = sqrt(hn2[3 * i + 1]), mn = sqrt(hn2[3 * i + 2]);
    if (fabs(wj) < cfg->zero_tol * scale) {
      /* |w| too small for the pencil residual's 1/|w| scaling: absolute
       * residual of the product form, ||K M x - w^2 x|| / (||x|| scale^2) */
      st->resv[j] = rn / (fmax(xn, 1e-300) * scale * scale);
    } else {
      const double xnorm = sqrt(xn * xn + (mn / aw) * (mn / aw));
      st->resv[j] = (rn / aw) / (xnorm * fmax(1.0, fabs(wj)));
    }
  }
  st->nact = 0;
  st->allc = 1;
  st->worst = 0.0;
  int si = 0;
  for (int j = 0; j < m; ++j) {
    /* only the numerical null space of the M-Gram (zf) bypasses the test;
     * a soft mode (|w| below zero_tol * scale) must meet it in absolute form */
    const int evaluated = !nsel || (si < nsel && st->sel[si] == j);
    if (evaluated) st->conv[j] = zf[j] || st->resv[j] < cfg->tol;
    if (j < nev) {
      if (!st->conv[j]) st->allc = 0;
      if (!zf[j]) st->worst = fmax(st->worst, st->resv[j]);
    }
    if (!st->conv[j]) {
      st->rpos[st->nact] = nsel ? si : j;
      st->act[st->nact++] = j;
    }
    if (nsel && evaluated) ++si;
  }
  return 0;
}

/* re-apply M then K (FP64) to the X block of A */
static int refresh_images(bdgx_ctx *ctx, const lobpcg_cfg *cfg, blk3 *A, lobpcg_result *res) {
  const int opM = cfg->k_equals_m ? BDGX_OP_K : BDGX_OP_M;
  int rc = bdgx_apply(ctx, opM, BDGX_FP64, A->V, A->MV, (uint64_t)cfg->m, cfg->S);
  if (!rc) rc = bdgx_apply(ctx, BDGX_OP_K, BDGX_FP64, A->MV, A->KMV, (uint64_t)cfg->m, cfg->S);
  res->applies += (uint64_t)cfg->m;
  res->refreshes++;
  return rc;
}

/* symmetric sub-block G[sel, sel] of a k x k set */
static void sub_block(const double *G, int k, const int *sel, int ks, int w, double *out) {
  for (int j = 0; j < ks; ++j)
    for (int i = 0; i < ks; ++i) setz(out, ks, i, j, w, getz(G, k, sel[i], sel[j], w));
}

/* ---- the per-iteration device chain ------------------------------------ */

/* W into column wc.. of A.V from the residual columns Rsrc: the learned
 * network (FP32), T_M T_K in FP64, or T_M T_K in FP32 (mixed) */
typedef struct {
  const void *rsrc;
  void *other, *f1, *f2;
  int wc, nact, k, full, lp, mixed, std, opM, w;
  uint64_t S;
  size_t es;
  double *dG;
} chain_args;

static int chain_w(bdgx_ctx *ctx, const blk3 *A, const chain_args *c) {
  const uint64_t S = c->S, n = (uint64_t)c->nact;
  void *wdst = colp(A->V, (uint64_t)c->wc, S, c->es);
  int rc;
  if (c->lp) {
    rc = bdgx_convert(ctx, BDGX_FP64, c->rsrc, S, BDGX_FP32, c->f1, S, S, n);
    if (!rc) rc = bdgx_lp_apply(ctx, BDGX_FP32, c->f1, c->f2, n, S);
    if (!rc) rc = bdgx_convert(ctx, BDGX_FP32, c->f2, S, BDGX_FP64, wdst, S, S, n);
  } else if (!c->mixed) {
    rc = bdgx_apply(ctx, BDGX_OP_PRECOND_K, BDGX_FP64, c->rsrc, c->other, n, S);
    if (!rc) rc = bdgx_apply(ctx, BDGX_OP_PRECOND_M, BDGX_FP64, c->other, wdst, n, S);
  } else {
    /* T_K, T_M in FP32: W is a search direction, its rounding only changes
     * which direction is searched, not the accuracy of the basis */
    rc = bdgx_convert(ctx, BDGX_FP64, c->rsrc, S, BDGX_FP32, c->f1, S, S, n);
    if (!rc) rc = bdgx_apply(ctx, BDGX_OP_PRECOND_K, BDGX_FP32, c->f1, c->f2, n, S);
    if (!rc) rc = bdgx_apply(ctx, BDGX_OP_PRECOND_M, BDGX_FP32, c->f2, c->f1, n, S);
    if (!rc) rc = bdgx_convert(ctx, BDGX_FP32, c->f1, S, BDGX_FP64, wdst, S, S, n);
  }
  return rc;
}

/* FP64 images M W, K M W of the W columns */
static int chain_images(bdgx_ctx *ctx, const blk3 *A, const chain_args *c) {
  const uint64_t S = c->S, n = (uint64_t)c->nact, wc = (uint64_t)c->wc;
  int rc = bdgx_apply(ctx, c->opM, BDGX_FP64, colp(A->V, wc, S, c->es), colp(A->MV, wc, S, c->es),
                      n, S);
  if (!rc)
    rc = bdgx_apply(ctx, BDGX_OP_K, BDGX_FP64, colp(A->MV, wc, S, c->es),
                    colp(A->KMV, wc, S, c->es), n, S);
  return rc;
}

static int chain_grams(bdgx_ctx *ctx, const blk3 *A, const chain_args *c) {
  if (c->full) return grams_full_dev(ctx, BDGX_FP64, c->std, A, c->k, c->S, c->w, c->dG);
  return grams_cols_dev(ctx, BDGX_FP64, c->std, A, c->k, c->wc, c->nact, c->S, c->w, c->es,
                        c->dG);
}

static int chain_all(bdgx_ctx *ctx, const blk3 *A, const chain_args *c) {
  int rc = chain_w(ctx, A, c);
  if (!rc) rc = chain_images(ctx, A, c);
  if (!rc) rc = chain_grams(ctx, A, c);
  return rc;
}

/* Graph cache: the chain W -> images -> Grams is captured once per distinct
 * argument set (basis buffer, residual source, column counts, Gram kind,
 * preconditioner) after one eager run of that set, and replayed from then
 * on.  Launch is refused when the ctx workspace moved after the capture; the
 * slot is then captured again. */
#define NGSLOT 8
typedef struct {
  const void *v, *r;
  int wc, nact, k, full, lp;
} gkey;
typedef struct {
  gkey key;
  bdgx_graph *g;
  int used, seen;
  unsigned long stamp;
} gslot;

static int gkey_eq(const gkey *a, const gkey *b) {
  return a->v == b->v && a->r == b->r && a->wc == b->wc && a->nact == b->nact &&
         a->k == b->k && a->full == b->full && a->lp == b->lp;
}

static int chain_run(bdgx_ctx *ctx, const blk3 *A, const chain_args *c, int use_graph,
                     gslot *slots, unsigned long *clock, lobpcg_result *res) {
  if (!use_graph) return chain_all(ctx, A, c);
  gkey key;
  memset(&key, 0, sizeof(key));
  key.v = A->V;
  key.r = c->rsrc;
  key.wc = c->wc;
  key.nact = c->nact;
  key.k = c->k;
  key.full = c->full;
  key.lp = c->lp;
  gslot *sl = NULL;
  for (int i = 0; i < NGSLOT; ++i)
    if (slots[i].used && gkey_eq(&slots[i].key, &key)) sl = &slots[i];
  if (!sl) { /* least recently used slot */
    sl = &slots[0];
    for (int i = 1; i < NGSLOT; ++i)
      if (!slots[i].used || (sl->used && slots[i].stamp < sl->stamp)) sl = &slots[i];
    bdgx_graph_destroy(sl->g);
    memset(sl, 0, sizeof(*sl));
    sl->key = key;
    sl->used = 1;
  }
  sl->stamp = ++*clock;
  if (sl->g) {
    const int rc = bdgx_graph_launch(ctx, sl->g);
    if (rc != BDGX_ESTATE) {
      if (!rc) res->graph_launches++;
      return rc;
    }
    bdgx_graph_destroy(sl->g);
    sl->g = NULL;
    sl->seen = 0;
  }
  if (sl->seen++ == 0) return chain_all(ctx, A, c); /* first use: eager (sizes workspaces) */
  const double tc0 = bdg_wtime();
  if (bdgx_graph_begin(ctx)) return chain_all(ctx, A, c);
  int rc = chain_all(ctx, A, c);
  if (rc) {
    (void)bdgx_graph_end(ctx, NULL);
    return chain_all(ctx, A, c);
  }
  if (bdgx_graph_end(ctx, &sl->g) || !sl->g) {
    sl->g = NULL;
    return chain_all(ctx, A, c);
  }
  res->graph_captures++;
  res->capture_seconds += bdg_wtime() - tc0;
  rc = bdgx_graph_launch(ctx, sl->g);
  if (!rc) res->graph_launches++;
  return rc;
}

int bdg_lobpcg_device(bdgx_ctx *ctx, const lobpcg_cfg *cfg, const void *dX0, lobpcg_result *res,
                      void **dXout) {
  const uint64_t S = cfg->S;
  const int m = cfg->m, nev = cfg->nev, cplx = cfg->cplx;
  const int w = cplx ? 2 : 1;
  const int mixed = cfg->w_prec == BDGX_FP32;
  /* standard form when K == M: (K^2) v = w^2 v with B = I; the "M" images
   * then hold K V */
  const int std = cfg->k_equals_m;
  const int opM = std ? BDGX_OP_K : BDGX_OP_M;
  const size_t es = elem_bytes(BDGX_FP64, cplx), es32 = elem_bytes(BDGX_FP32, cplx);
  const int kmax = 3 * m;
  int rc = 0;
  blk3 A, B;
  memset(&A, 0, sizeof(A));
  memset(&B, 0, sizeof(B));
  void *dR = NULL, *dTmp = NULL, *f1 = NULL, *f2 = NULL, *f3 = NULL;
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
  double *hS[3];
  for (int g = 0; g < 3; ++g) hS[g] = (double *)malloc(sizeof(double) * gsz);
  double *hY = (double *)calloc((size_t)kmax * m * w, sizeof(double));
  double *hYs = (double *)malloc(sizeof(double) * (size_t)kmax * m * w);
  double *hCn = (double *)calloc((size_t)kmax * 2 * m * w, sizeof(double)); /* [X' | P'] */
  double *hCp = (double *)calloc((size_t)kmax * 2 * m * w, sizeof(double)); /* previous */
  double *w2 = (double *)calloc((size_t)m, sizeof(double));
  int *zf = (int *)calloc((size_t)m, sizeof(int));
  int *sel = (int *)malloc(sizeof(int) * (size_t)kmax);
  res_state st;
  st.resv = (double *)malloc(sizeof(double) * (size_t)m);
  st.conv = (int *)calloc((size_t)m, sizeof(int));
  st.act = (int *)malloc(sizeof(int) * (size_t)m);
  st.rpos = (int *)malloc(sizeof(int) * (size_t)m);
  st.sel = (int *)malloc(sizeof(int) * (size_t)m);
  st.lsel = (double *)malloc(sizeof(double) * (size_t)m);
  int last_full = 0; /* iteration of the last full residual evaluation */
  int it = 0, np = 0, kprev = 0, ncp = 0, gvalid = 0, since_full = 0, fresh = 1;
  double trr = 0.0;
  int refresh_it = 0;
  int img_done = 0; /* the FP32-image phase ends for good at img_tol */
  const double t_start = bdg_wtime();
  res->status = 0;
  res->iters = 0;
  res->restarts = 0;
  res->precond_fallbacks = 0;
  res->applies = 0;
  res->applies_fp32 = 0;
  res->refreshes = 0;
  res->learned_its = 0;
  res->learned_fallback_at = 0;
  int lp_on = cfg->learned;
  int nfrozen = 0; /* columns frozen by the last update */
  gslot gslots[NGSLOT];
  unsigned long gclock = 0;
  memset(gslots, 0, sizeof(gslots));
  res->graph_launches = 0;
  res->graph_captures = 0;
  res->capture_seconds = 0.0;
  double *worst_hist = (double *)calloc((size_t)cfg->maxiter + 2, sizeof(double));
  /* frozen columns: fz marks this iteration's; same[j] says that both block
   * buffers hold the same column j (then its update need not be written) */
  int *fz = (int *)calloc((size_t)m, sizeof(int));
  unsigned char *same = (unsigned char *)calloc((size_t)m, 1);
  int *wcols = (int *)malloc(sizeof(int) * (size_t)(2 * m));
  double *hCw = (double *)malloc(sizeof(double) * (size_t)kmax * BDGX_COMBINE_COLS_MAX * w);
  res->frozen_skips = 0;
  stage_t sg;
  memset(&sg, 0, sizeof(sg));
  if (stage_init(&sg, m, kmax, w)) {
    rc = BDGX_ENOMEM;
    goto done;
  }

  /* launch shapes of the FFT kernels for this grid and block width (once per
   * context and precision) */
  TRY(bdgx_autotune(ctx, BDGX_FP64, (uint64_t)(m < 32 ? m : 32), 0));
  if (mixed || cfg->learned) TRY(bdgx_autotune(ctx, BDGX_FP32, (uint64_t)(m < 32 ? m : 32), 0));
  TRY(alloc3(&A, (size_t)S * kmax * es));
  TRY(alloc3(&B, (size_t)S * kmax * es));
  TRY(bdgx_malloc(&dR, (size_t)S * m * es));
  TRY(bdgx_malloc(&dTmp, (size_t)S * m * es));
  if (mixed || cfg->learned) {
    TRY(bdgx_malloc(&f1, (size_t)S * m * es32));
    TRY(bdgx_malloc(&f2, (size_t)S * m * es32));
    TRY(bdgx_malloc(&f3, (size_t)S * m * es32));
  }
  TRY(bdgx_malloc((void **)&dG, sizeof(double) * 3 * gsz));
  TRY(bdgx_malloc((void **)&dC, sizeof(double) * (size_t)kmax * 2 * m * w));
  TRY(bdgx_malloc((void **)&dlam, sizeof(double) * (size_t)m));
  TRY(bdgx_malloc((void **)&dn2, sizeof(double) * 3 * (size_t)m));

  /* start: basis [X0], images by M then K (FP64), Rayleigh-Ritz, X = X0 Y */
  {
    TRY(bdgx_convert(ctx, BDGX_FP64, dX0, S, BDGX_FP64, A.V, S, S, (uint64_t)m));
    TRY(bdgx_apply(ctx, opM, BDGX_FP64, A.V, A.MV, (uint64_t)m, S));
    TRY(bdgx_apply(ctx, BDGX_OP_K, BDGX_FP64, A.MV, A.KMV, (uint64_t)m, S));
    res->applies += (uint64_t)m;
    TRY(grams_full(ctx, &sg, BDGX_FP64, std, &A, m, S, w, dG, &G));
    const double t0 = bdg_wtime();
    const int got = dn_rayleigh_ritz(m, m, cplx, G.V, G.M, G.A, w2, hY, zf);
    trr += bdg_wtime() - t0;
    if (got < m) {
      rc = -100;
      goto done;
    }
    TRY(bdgx_memcpy_h2d(ctx, dC, hY, sizeof(double) * (size_t)m * m * w));
    TRY(combine3(ctx, BDGX_FP64, &A, 0, (uint64_t)m, dC, &B, 0, (uint64_t)m, S, es));
    memcpy(Gp.V, G.V, sizeof(double) * (size_t)m * m * w);
    memcpy(Gp.M, G.M, sizeof(double) * (size_t)m * m * w);
    memcpy(Gp.A, G.A, sizeof(double) * (size_t)m * m * w);
    memcpy(hCp, hY, sizeof(double) * (size_t)m * m * w);
    kprev = m;
    ncp = m;
    np = 0;
    gvalid = 1;
    blk3 t = A;
    A = B;
    B = t;
  }

  for (it = 1; it <= cfg->maxiter; ++it) {
    /* few active columns: evaluate only those, with a full evaluation every
     * 10 iterations and before convergence is accepted (converged columns
     * still move a little under the Rayleigh-Ritz) */
    int nsel = 0;
    if (it > 1 && st.nact > 0 && st.nact <= m / 4 && it - last_full < 10) {
      nsel = st.nact;
      for (int i = 0; i < nsel; ++i) st.sel[i] = st.act[i];
    }
    TRY(eval_residuals(ctx, cfg, &sg, &A, w2, zf, dlam, dn2, dR, &st, nsel));
    if (nsel == 0) {
      last_full = it;
    } else if (st.allc) {
      TRY(eval_residuals(ctx, cfg, &sg, &A, w2, zf, dlam, dn2, dR, &st, 0));
      last_full = it;
    }
    /* mixed: FP64 images of X every refresh_every iterations while new
     * directions get FP32 images, when the residual reaches img_tol (FP64
     * images from then on), and before convergence is accepted */
    if (mixed && !fresh &&
        (st.allc || st.worst <= cfg->img_tol ||
         (cfg->refresh_every > 0 && it - refresh_it >= cfg->refresh_every))) {
      if (st.allc || st.worst <= cfg->img_tol) img_done = 1;
      TRY(refresh_images(ctx, cfg, &A, res));
      memset(same, 0, (size_t)m);
      fresh = 1;
      gvalid = 0;
      refresh_it = it;
      TRY(eval_residuals(ctx, cfg, &sg, &A, w2, zf, dlam, dn2, dR, &st, 0));
      last_full = it;
    }
    const int nact = st.nact;
    const int *act = st.act;
    if (cfg->verbosity > 1 || (cfg->verbosity == 1 && it % 10 == 0)) {
      fprintf(stderr, "[bdg] it %d t %.3f w0 %.10g wlast %.10g worst res %.3e active %d%s", it,
              bdg_wtime() - t_start, signed_sqrt(w2[0]), signed_sqrt(w2[nev - 1]), st.worst,
              nact, fresh ? "" : " (fp32 images)");
      if (nfrozen) fprintf(stderr, " frozen %d", nfrozen);
      if (nact <= 8) {
        fprintf(stderr, " [");
        for (int j = 0; j < nact; ++j)
          fprintf(stderr, "%s%d:%.4g/%.2e", j ? " " : "", act[j], signed_sqrt(w2[act[j]]),
                  st.resv[act[j]]);
        fprintf(stderr, "]");
      }
      fprintf(stderr, "\n");
    }
    if (st.allc) break;
    if (it == cfg->maxiter) break;
    worst_hist[it] = st.worst;
    if (lp_on && cfg->lp_window > 0 && it > cfg->lp_window &&
        !(st.worst < cfg->lp_factor * worst_hist[it - cfg->lp_window])) {
      lp_on = 0;  /* the network stopped helping: analytic T_M T_K from here on */
      res->learned_fallback_at = it;
      np = 0;     /* and restart the search-direction history */
      if (cfg->verbosity > 0)
        fprintf(stderr, "[bdg] it %d: learned preconditioner stalled, back to T_M T_K\n", it);
    }

    /* W = T_M T_K R[:, act] at columns [m + np, m + np + nact) */
    const int wc = m + np;
    const int k = m + np + nact;
    /* the active residual columns, packed: dR itself when they already are */
    const void *Rsrc = dR;
    int packed = 1;
    for (int i = 0; i < nact; ++i)
      if (st.rpos[i] != i) packed = 0;
    if (!packed) {
      TRY(bdgx_gather_cols(ctx, BDGX_FP64, dR, S, st.rpos, (uint64_t)nact, dTmp, S, S));
      Rsrc = dTmp;
    }
    void *other = (Rsrc == dR) ? dTmp : dR;
    const int full = !gvalid || cfg->gram_refresh <= 0 || ++since_full >= cfg->gram_refresh;
    if (mixed && st.worst <= cfg->img_tol) img_done = 1;
    const int img32 = mixed && !img_done && cfg->img_tol > 0 && nact >= cfg->mixed_min_cols &&
                      st.worst > cfg->img_tol;
    void *wdst = colp(A.V, (uint64_t)wc, S, es);
    chain_args ca;
    ca.rsrc = Rsrc;
    ca.other = other;
    ca.f1 = f1;
    ca.f2 = f2;
    ca.wc = wc;
    ca.nact = nact;
    ca.k = k;
    ca.full = full;
    ca.lp = lp_on;
    ca.mixed = mixed;
    ca.std = std;
    ca.opM = opM;
    ca.w = w;
    ca.S = S;
    ca.es = es;
    ca.dG = dG;
    if (lp_on) res->learned_its++;
    if (!img32) {
      /* W, FP64 images, Gram columns: one device chain (graph-replayed);
       * the host projects the [X | P] Gram blocks while it runs */
      TRY(chain_run(ctx, &A, &ca, cfg->graphs, gslots, &gclock, res));
      if (!full) project_grams(&Gp, kprev, hCp, m + np, k, w, hT, &G);
    } else {
      if (!full) project_grams(&Gp, kprev, hCp, m + np, k, w, hT, &G);
      TRY(chain_w(ctx, &A, &ca));
      /* FP32 images of W, with W first projected out of span(X) (Euclidean,
       * FP64) so that the image error is relative to the new part of W */
      const size_t zsz = (size_t)m * nact * w;
      TRY(bdgx_gram(ctx, BDGX_FP64, A.V, S, (uint64_t)m, wdst, S, (uint64_t)nact, S, 0, dG));
      if (full)
        TRY(bdgx_gram(ctx, BDGX_FP64, A.V, S, (uint64_t)m, A.V, S, (uint64_t)m, S, 0, dG + zsz));
      TRY(bdgx_memcpy_d2h(ctx, hT, dG, sizeof(double) * (zsz + (full ? (size_t)m * m * w : 0))));
      if (full) {
        memcpy(hS[1], hT + zsz, sizeof(double) * (size_t)m * m * w);
      } else {
        for (int j = 0; j < m; ++j)
          for (int i = 0; i < m; ++i) setz(hS[1], m, i, j, w, getz(G.V, k, i, j, w));
      }
      if (chol_solve(hS[1], m, w, hT, nact) == 0) {
        /* W <- [X | P | W] [-Y; 0; I] into `other`, then back into the basis */
        memset(hCn, 0, sizeof(double) * (size_t)k * nact * w);
        for (int j = 0; j < nact; ++j) {
          for (int i = 0; i < m; ++i) setz(hCn, k, i, j, w, -getz(hT, m, i, j, w));
          setz(hCn, k, wc + j, j, w, 1.0);
        }
        TRY(bdgx_memcpy_h2d(ctx, dC, hCn, sizeof(double) * (size_t)k * nact * w));
        TRY(bdgx_combine(ctx, BDGX_FP64, A.V, S, (uint64_t)k, dC, other, S, (uint64_t)nact, S));
        TRY(bdgx_convert(ctx, BDGX_FP64, other, S, BDGX_FP64, wdst, S, S, (uint64_t)nact));
        TRY(bdgx_convert(ctx, BDGX_FP64, other, S, BDGX_FP32, f1, S, S, (uint64_t)nact));
      }
      TRY(bdgx_apply(ctx, opM, BDGX_FP32, f1, f2, (uint64_t)nact, S));
      TRY(bdgx_apply(ctx, BDGX_OP_K, BDGX_FP32, f2, f3, (uint64_t)nact, S));
      TRY(bdgx_convert(ctx, BDGX_FP32, f2, S, BDGX_FP64, colp(A.MV, (uint64_t)wc, S, es), S, S,
                       (uint64_t)nact));
      TRY(bdgx_convert(ctx, BDGX_FP32, f3, S, BDGX_FP64, colp(A.KMV, (uint64_t)wc, S, es), S, S,
                       (uint64_t)nact));
      res->applies_fp32 += (uint64_t)nact;
      fresh = 0;
      TRY(chain_grams(ctx, &A, &ca));
    }
    res->applies += (uint64_t)nact;

    /* Gram set of [X | P | W] on the host */
    if (full) {
      since_full = 0;
      TRY(grams_full_host(ctx, &sg, std, k, w, dG, &G));
    } else {
      TRY(grams_cols_host(ctx, &sg, std, k, wc, nact, w, dG, &G));
    }
    /* Rayleigh-Ritz on a column selection: all, or without P when [X P W]
     * is numerically dependent */
    double t0 = bdg_wtime();
    int ks = 0;
    for (int i = 0; i < k; ++i) sel[ks++] = i;
    sub_block(G.V, k, sel, ks, w, hS[0]);
    if (np && !dn_chol_ok(ks, cplx, hS[0], 1e-10)) {
      ks = 0;
      for (int i = 0; i < m; ++i) sel[ks++] = i;
      for (int i = wc; i < k; ++i) sel[ks++] = i;
      sub_block(G.V, k, sel, ks, w, hS[0]);
      res->restarts++;
    }
    if (!dn_chol_ok(ks, cplx, hS[0], 1e-10)) {
      /* [X, W] still dependent: the preconditioner collapsed the residuals
       * (e.g. a 1e8 gain on one mode); use W = R for this iteration */
      trr += bdg_wtime() - t0;
      TRY(bdgx_convert(ctx, BDGX_FP64, Rsrc, S, BDGX_FP64, colp(A.V, (uint64_t)wc, S, es), S, S,
                       (uint64_t)nact));
      TRY(bdgx_apply(ctx, opM, BDGX_FP64, colp(A.V, (uint64_t)wc, S, es),
                     colp(A.MV, (uint64_t)wc, S, es), (uint64_t)nact, S));
      TRY(bdgx_apply(ctx, BDGX_OP_K, BDGX_FP64, colp(A.MV, (uint64_t)wc, S, es),
                     colp(A.KMV, (uint64_t)wc, S, es), (uint64_t)nact, S));
      res->applies += (uint64_t)nact;
      res->precond_fallbacks++;
      TRY(grams_full(ctx, &sg, BDGX_FP64, std, &A, k, S, w, dG, &G));
      since_full = 0;
      ks = 0;
      for (int i = 0; i < m; ++i) sel[ks++] = i;
      for (int i = wc; i < k; ++i) sel[ks++] = i;
      sub_block(G.V, k, sel, ks, w, hS[0]);
      t0 = bdg_wtime();
    }
    sub_block(G.M, k, sel, ks, w, hS[1]);
    sub_block(G.A, k, sel, ks, w, hS[2]);
    const int got = dn_rayleigh_ritz(ks, m, cplx, hS[0], hS[1], hS[2], w2, hYs, zf);
    trr += bdg_wtime() - t0;
    if (got < m) {
      rc = -100;
      goto done;
    }
    /* coefficients on the full basis: [C_X | C_P], C_P = rows m.. of the
     * active columns of C_X (P' = [P W] part of the update) */
    const int nc = m + nact;
    memset(hCn, 0, sizeof(double) * (size_t)k * nc * w);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < ks; ++i) setz(hCn, k, sel[i], j, w, getz(hYs, ks, i, j, w));
    for (int j = 0; j < nact; ++j)
      for (int i = m; i < k; ++i) setz(hCn, k, i, m + j, w, getz(hCn, k, i, act[j], w));
    nfrozen = cfg->freeze_tol > 0 ? freeze_cols(&G, k, m, w, st.conv, cfg->freeze_tol, hCn,
                                                w2, fz)
                                   : 0;
    /* write the columns that change, and the frozen ones that the other
     * buffer does not hold yet (a unit coefficient column copies exactly) */
    int nw = 0;
    for (int j = 0; j < nc; ++j)
      if (j >= m || !fz[j] || !same[j]) wcols[nw++] = j;
    if (nw < nc && nw <= BDGX_COMBINE_COLS_MAX) {
      for (int o = 0; o < nw; ++o)
        memcpy(hCw + (size_t)o * k * w, hCn + (size_t)wcols[o] * k * w,
               sizeof(double) * (size_t)k * w);
      TRY(stage_up(ctx, &sg, sg.C, dC, hCw, sizeof(double) * (size_t)k * nw * w));
      TRY(bdgx_combine_cols(ctx, BDGX_FP64, A.V, S, (uint64_t)k, dC, B.V, S, wcols,
                            (uint64_t)nw, S));
      TRY(bdgx_combine_cols(ctx, BDGX_FP64, A.MV, S, (uint64_t)k, dC, B.MV, S, wcols,
                            (uint64_t)nw, S));
      TRY(bdgx_combine_cols(ctx, BDGX_FP64, A.KMV, S, (uint64_t)k, dC, B.KMV, S, wcols,
                            (uint64_t)nw, S));
      res->frozen_skips += (uint64_t)(nc - nw);
    } else {
      TRY(stage_up(ctx, &sg, sg.C, dC, hCn, sizeof(double) * (size_t)k * nc * w));
      TRY(combine3(ctx, BDGX_FP64, &A, 0, (uint64_t)k, dC, &B, 0, (uint64_t)nc, S, es));
    }
    for (int j = 0; j < m; ++j) same[j] = (unsigned char)fz[j];
    memcpy(Gp.V, G.V, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.M, G.M, sizeof(double) * (size_t)k * k * w);
    memcpy(Gp.A, G.A, sizeof(double) * (size_t)k * k * w);
    memcpy(hCp, hCn, sizeof(double) * (size_t)k * nc * w);
    kprev = k;
    ncp = nc;
    np = nact;
    gvalid = 1;
    blk3 t = A;
    A = B;
    B = t;
  }
  (void)ncp;
  if (it > cfg->maxiter) it = cfg->maxiter;
  TRY(bdgx_sync(ctx));
  {
    void *x = NULL;
    TRY(bdgx_malloc(&x, (size_t)S * m * es));
    rc = bdgx_convert(ctx, BDGX_FP64, A.V, S, BDGX_FP64, x, S, S, (uint64_t)m);
    if (!rc) rc = bdgx_sync(ctx);
    if (rc) {
      bdgx_free(x);
      goto done;
    }
    *dXout = x;
  }
  for (int j = 0; j < m; ++j) {
    res->w2[j] = w2[j];
    res->zflag[j] = zf[j];
    res->res[j] = st.resv[j];
  }
done:
  for (int i = 0; i < NGSLOT; ++i) bdgx_graph_destroy(gslots[i].g);
  res->iters = it;
  res->rr_seconds = trr;
  res->status = rc;
  free3(&A);
  free3(&B);
  bdgx_free(dR);
  bdgx_free(dTmp);
  bdgx_free(f1);
  bdgx_free(f2);
  bdgx_free(f3);
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
  for (int g = 0; g < 3; ++g) free(hS[g]);
  free(hY);
  free(hYs);
  free(hCn);
  free(hCp);
  free(w2);
  free(zf);
  stage_free(&sg);
  free(sel);
  free(st.resv);
  free(st.conv);
  free(st.act);
  free(st.rpos);
  free(st.sel);
  free(st.lsel);
  free(worst_hist);
  free(fz);
  free(same);
  free(wcols);
  free(hCw);
  return rc;
}
