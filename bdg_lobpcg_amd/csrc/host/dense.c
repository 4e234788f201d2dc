/*
 * Small dense FP64 linear algebra for the host side of the iLOBPCG step.
 *
 * dn_syev: Householder reduction of a symmetric matrix to tridiagonal form
 * followed by the implicit symmetric QR iteration with Wilkinson shifts and
 * Givens chasing (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.3.1
 * and 8.3.3), accumulating the orthogonal transforms.
 *
 * dn_heev: Hermitian H = A + iB through the real symmetric embedding
 * [[A, -B], [B, A]], whose spectrum is H's with every eigenvalue doubled;
 * one complex eigenvector per pair is recovered by complex Gram-Schmidt
 * inside each (doubled) eigenvalue cluster.
 *
 * The Rayleigh-Ritz routine itself lives in rr_impl.inc (compiled once for
 * real and once for complex scalars).
 */
#include "dense.h"

#include <complex.h>
#include <float.h>
#include <immintrin.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* x . y in four independent 4-wide FMA chains (a single running sum makes
 * the dot products of the Gram-Schmidt steps latency bound) */
static inline double vdot(int n, const double *x, const double *y) {
  __m256d s0 = _mm256_setzero_pd(), s1 = s0, s2 = s0, s3 = s0;
  int i = 0;
  for (; i + 16 <= n; i += 16) {
    s0 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i), _mm256_loadu_pd(y + i), s0);
    s1 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i + 4), _mm256_loadu_pd(y + i + 4), s1);
    s2 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i + 8), _mm256_loadu_pd(y + i + 8), s2);
    s3 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i + 12), _mm256_loadu_pd(y + i + 12), s3);
  }
  for (; i + 4 <= n; i += 4)
    s0 = _mm256_fmadd_pd(_mm256_loadu_pd(x + i), _mm256_loadu_pd(y + i), s0);
  const __m256d s = _mm256_add_pd(_mm256_add_pd(s0, s1), _mm256_add_pd(s2, s3));
  const __m128d h = _mm_add_pd(_mm256_castpd256_pd128(s), _mm256_extractf128_pd(s, 1));
  double r = _mm_cvtsd_f64(_mm_add_sd(h, _mm_unpackhi_pd(h, h)));
  for (; i < n; ++i) r += x[i] * y[i];
  return r;
}

/* ---------------------------------------------------------------------- */
static void tridiagonalize(int n, double *A, double *Q, double *d, double *e) {
  double *v = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n);
  double *qv = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n * n; ++i) Q[i] = 0.0;
  for (int i = 0; i < n; ++i) Q[i + i * n] = 1.0;
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - j - 1; /* length of x = A[j+1.., j] */
    double nx = 0.0;
    for (int i = 0; i < m; ++i) {
      v[i] = A[(j + 1 + i) + j * n];
      nx += v[i] * v[i];
    }
    nx = sqrt(nx);
    if (nx == 0.0) continue;
    const double alpha = v[0] > 0 ? -nx : nx;
    v[0] -= alpha;
    double nv = 0.0;
    for (int i = 0; i < m; ++i) nv += v[i] * v[i];
    nv = sqrt(nv);
    if (nv == 0.0) continue;
    for (int i = 0; i < m; ++i) v[i] /= nv;
    /* p = A22 v */
    for (int i = 0; i < m; ++i) p[i] = 0.0;
    for (int c = 0; c < m; ++c) {
      const double vc = v[c];
      const double *col = A + (j + 1) + (j + 1 + c) * n;
      for (int i = 0; i < m; ++i) p[i] += col[i] * vc;
    }
    double K = 0.0;
    for (int i = 0; i < m; ++i) K += v[i] * p[i];
    /* w = 2p - 2Kv ; A22 -= v w^T + w v^T */
    for (int i = 0; i < m; ++i) p[i] = 2.0 * p[i] - 2.0 * K * v[i];
    for (int c = 0; c < m; ++c) {
      double *col = A + (j + 1) + (j + 1 + c) * n;
      const double vc = v[c], wc = p[c];
      for (int i = 0; i < m; ++i) col[i] -= v[i] * wc + p[i] * vc;
    }
    A[(j + 1) + j * n] = alpha;
    A[j + (j + 1) * n] = alpha;
    for (int i = 1; i < m; ++i) {
      A[(j + 1 + i) + j * n] = 0.0;
      A[j + (j + 1 + i) * n] = 0.0;
    }
    /* Q[:, j+1..] -= 2 (Q[:, j+1..] v) v^T, column by column (contiguous) */
    for (int r = 0; r < n; ++r) qv[r] = 0.0;
    for (int i = 0; i < m; ++i) {
      const double vi = v[i];
      const double *qc = Q + (size_t)(j + 1 + i) * n;
      for (int r = 0; r < n; ++r) qv[r] += qc[r] * vi;
    }
    for (int i = 0; i < m; ++i) {
      const double vi = 2.0 * v[i];
      double *qc = Q + (size_t)(j + 1 + i) * n;
      for (int r = 0; r < n; ++r) qc[r] -= qv[r] * vi;
    }
  }
  for (int i = 0; i < n; ++i) d[i] = A[i + i * n];
  for (int i = 0; i + 1 < n; ++i) e[i] = A[(i + 1) + i * n];
  if (n > 0) e[n - 1] = 0.0;
  free(v);
  free(p);
  free(qv);
}

/* sqrt(a^2 + b^2): the plain formula where it cannot overflow or lose all
 * digits to underflow, hypot otherwise (hypot costs several times more) */
static inline double hyp(double a, double b) {
  const double m = fmax(fabs(a), fabs(b));
  if (m < 1e150 && m > 1e-150) return sqrt(a * a + b * b);
  return hypot(a, b);
}

/* implicit QR on the tridiagonal (d, e) with vectors accumulated in Z (or
 * eigenvalues only when Z is NULL) */
static int tridiag_qr(int n, double *d, double *e, double *Z) {
  int q = n - 1;
  int guard = 0;
  while (q > 0) {
    for (int i = 0; i < q; ++i)
      if (fabs(e[i]) <= DBL_EPSILON * (fabs(d[i]) + fabs(d[i + 1]))) e[i] = 0.0;
    while (q > 0 && e[q - 1] == 0.0) --q;
    if (q == 0) break;
    int p = q - 1;
    while (p > 0 && e[p - 1] != 0.0) --p;
    if (++guard > 60 * n) return -1;
    /* Wilkinson shift from the trailing 2x2 of the block [p, q] */
    const double dl = 0.5 * (d[q - 1] - d[q]);
    const double eq = e[q - 1];
    const double den = dl + (dl >= 0 ? 1.0 : -1.0) * hyp(dl, eq);
    const double mu = den != 0.0 ? d[q] - eq * eq / den : d[q];
    double x = d[p] - mu, z = e[p];
    double bulge = 0.0;
    for (int k = p; k < q; ++k) {
      const double r = hyp(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = z / r;
      }
      if (k > p) e[k - 1] = r;
      const double a = d[k], b = d[k + 1], o = e[k];
      d[k] = c * c * a + 2.0 * c * s * o + s * s * b;
      d[k + 1] = s * s * a - 2.0 * c * s * o + c * c * b;
      e[k] = c * s * (b - a) + (c * c - s * s) * o;
      if (k + 1 < q) {
        bulge = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      if (Z) {
        /* columns k, k+1 of Z rotated, four rows per FMA pair (this vector
         * work overlaps the latency-bound scalar chain of the next steps) */
        double *zk = Z + (size_t)k * n, *zk1 = Z + (size_t)(k + 1) * n;
        const __m256d vc = _mm256_set1_pd(c), vs = _mm256_set1_pd(s);
        int i = 0;
        for (; i + 4 <= n; i += 4) {
          const __m256d u = _mm256_loadu_pd(zk + i), w = _mm256_loadu_pd(zk1 + i);
          _mm256_storeu_pd(zk + i, _mm256_fmadd_pd(vc, u, _mm256_mul_pd(vs, w)));
          _mm256_storeu_pd(zk1 + i, _mm256_fnmadd_pd(vs, u, _mm256_mul_pd(vc, w)));
        }
        for (; i < n; ++i) {
          const double u = zk[i], w = zk1[i];
          zk[i] = c * u + s * w;
          zk1[i] = -s * u + c * w;
        }
      }
      x = e[k];
      z = bulge;
    }
  }
  return 0;
}

static void sort_eig(int n, double *w, double *V, int ldv, int vlen) {
  /* selection sort (n is small); V may be NULL */
  double *tmp = V ? (double *)malloc(sizeof(double) * (size_t)vlen) : NULL;
  for (int i = 0; i < n; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      const double t = w[i];
      w[i] = w[k];
      w[k] = t;
      if (V) {
        memcpy(tmp, V + (size_t)i * ldv, sizeof(double) * vlen);
        memcpy(V + (size_t)i * ldv, V + (size_t)k * ldv, sizeof(double) * vlen);
        memcpy(V + (size_t)k * ldv, tmp, sizeof(double) * vlen);
      }
    }
  }
  free(tmp);
}

/* LU factorisation with partial pivoting of the tridiagonal T - s I (the
 * layout of LAPACK's dgttrf: dl multipliers, du / du2 the upper diagonals
 * of U, sw[i] = 1 when rows i and i+1 were exchanged; d returns the
 * reciprocals of U's diagonal); pivots smaller than tiny are replaced by
 * +-tiny */
static void trid_lu(int n, const double *dd, const double *e, double s, double tiny, double *dl,
                    double *d, double *du, double *du2, int *sw) {
  for (int i = 0; i < n; ++i) d[i] = dd[i] - s;
  for (int i = 0; i + 1 < n; ++i) {
    dl[i] = e[i];
    du[i] = e[i];
    du2[i] = 0.0;
    sw[i] = 0;
  }
  for (int i = 0; i + 1 < n; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (fabs(d[i]) < tiny) d[i] = d[i] < 0 ? -tiny : tiny;
      const double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du2[i];
      }
      sw[i] = 1;
    }
  }
  if (fabs(d[n - 1]) < tiny) d[n - 1] = d[n - 1] < 0 ? -tiny : tiny;
  for (int i = 0; i < n; ++i) d[i] = 1.0 / d[i];
}

/* x <- (T - s I)^-1 x with the factors of trid_lu (the row exchanges as
 * selects, not branches: their pattern is data dependent) */
static void trid_solve(int n, const double *dl, const double *rd, const double *du,
                       const double *du2, const int *sw, double *x) {
  for (int i = 0; i + 1 < n; ++i) {
    const double a = x[i], b = x[i + 1];
    const double lo = sw[i] ? b : a, hi = sw[i] ? a : b;
    x[i] = lo;
    x[i + 1] = hi - dl[i] * lo;
  }
  x[n - 1] *= rd[n - 1];
  if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) * rd[n - 2];
  for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) * rd[i];
}

int dn_trid_low(int n, const double *d0, const double *e0, int m, double *w, double *Y) {
  if (n <= 0 || m <= 0) return 0;
  if (m > n) m = n;
  double *d = (double *)malloc(sizeof(double) * (size_t)n);
  double *e = (double *)malloc(sizeof(double) * (size_t)n);
  double *lam = (double *)malloc(sizeof(double) * (size_t)n);
  double *f[4];
  for (int q = 0; q < 4; ++q) f[q] = (double *)malloc(sizeof(double) * (size_t)n);
  int *sw = (int *)malloc(sizeof(int) * (size_t)n);
  memcpy(d, d0, sizeof(double) * (size_t)n);
  for (int i = 0; i + 1 < n; ++i) e[i] = e0[i];
  e[n - 1] = 0.0;
  double tn = 0.0; /* 1-norm of T */
  for (int i = 0; i < n; ++i)
    tn = fmax(tn, fabs(d0[i]) + (i > 0 ? fabs(e0[i - 1]) : 0.0) + (i + 1 < n ? fabs(e0[i]) : 0.0));
  if (!(tn > 0)) tn = 1.0;
  int rc = tridiag_qr(n, d, e, NULL);
  memcpy(lam, d, sizeof(double) * (size_t)n);
  sort_eig(n, lam, NULL, 0, 0);
  const double ortol = 1e-3 * tn, pertol = 10.0 * DBL_EPSILON * tn, tiny = DBL_EPSILON * tn;
  int c0 = 0;          /* first vector of the current cluster */
  double sprev = 0.0;  /* previous shift */
  uint32_t seed = 0x9e3779b9u;
  for (int j = 0; j < m; ++j) {
    double s = lam[j];
    if (j > 0 && lam[j] - lam[j - 1] > ortol) c0 = j;
    if (j > c0 && s - sprev < pertol) s = sprev + pertol; /* distinct shifts in a cluster */
    sprev = s;
    trid_lu(n, d0, e0, s, tiny, f[0], f[1], f[2], f[3], sw);
    double *y = Y + (size_t)j * n;
    for (int i = 0; i < n; ++i) { /* fixed pseudo-random start */
      seed ^= seed << 13;
      seed ^= seed >> 17;
      seed ^= seed << 5;
      y[i] = (double)(seed & 0xffffff) / 16777216.0 - 0.5;
    }
    for (int itr = 0; itr < 3; ++itr) {
      trid_solve(n, f[0], f[1], f[2], f[3], sw, y);
      for (int pass = 0; pass < 2; ++pass) /* against the cluster's earlier vectors */
        for (int c = c0; c < j; ++c) {
          const double *yc = Y + (size_t)c * n;
          const double dot = vdot(n, yc, y);
#pragma omp simd
          for (int i = 0; i < n; ++i) y[i] -= dot * yc[i];
        }
      const double nrm = sqrt(vdot(n, y, y));
      if (!(nrm > 0)) {
        rc = -1;
        break;
      }
      for (int i = 0; i < n; ++i) y[i] /= nrm;
    }
    w[j] = lam[j];
  }
  free(d);
  free(e);
  free(lam);
  for (int q = 0; q < 4; ++q) free(f[q]);
  free(sw);
  return rc;
}

int dn_syev(int n, double *A, double *w) {
  if (n <= 0) return 0;
  double *Q = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *e = (double *)malloc(sizeof(double) * (size_t)n);
  /* symmetrise from the lower triangle */
  for (int c = 0; c < n; ++c)
    for (int r = c + 1; r < n; ++r) {
      const double s = 0.5 * (A[r + c * n] + A[c + r * n]);
      A[r + c * n] = s;
      A[c + r * n] = s;
    }
  tridiagonalize(n, A, Q, w, e);
  const int rc = tridiag_qr(n, w, e, Q);
  sort_eig(n, w, Q, n, n);
  memcpy(A, Q, sizeof(double) * (size_t)n * n);
  free(Q);
  free(e);
  return rc;
}

int dn_heev(int n, double *A, double *w) {
  if (n <= 0) return 0;
  const int N2 = 2 * n;
  double *M = (double *)malloc(sizeof(double) * (size_t)N2 * N2);
  double *w2 = (double *)malloc(sizeof(double) * (size_t)N2);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) {
      /* Hermitian part: average with the conjugate transpose */
      const double re = 0.5 * (A[2 * (r + c * n)] + A[2 * (c + r * n)]);
      const double im = 0.5 * (A[2 * (r + c * n) + 1] - A[2 * (c + r * n) + 1]);
      M[r + c * N2] = re;
      M[(r + n) + (c + n) * N2] = re;
      M[(r + n) + c * N2] = im;
      M[r + (c + n) * N2] = -im;
    }
  int rc = dn_syev(N2, M, w2);
  double scale = 0.0;
  for (int i = 0; i < N2; ++i) scale = fmax(scale, fabs(w2[i]));
  const double tol = 1e-9 * (scale > 0 ? scale : 1.0);
  int nout = 0;
  double complex *acc = (double complex *)malloc(sizeof(double complex) * (size_t)n * n);
  double complex *z = (double complex *)malloc(sizeof(double complex) * (size_t)n);
  int i = 0;
  while (i < N2 && nout < n) {
    int j = i + 1;
    while (j < N2 && fabs(w2[j] - w2[i]) <= tol) ++j;
    const int cl_start = nout;
    const int want = (j - i + 1) / 2;
    for (int t = i; t < j && nout - cl_start < want && nout < n; ++t) {
      for (int r = 0; r < n; ++r) z[r] = M[r + (size_t)t * N2] + I * M[(r + n) + (size_t)t * N2];
      for (int pass = 0; pass < 2; ++pass)
        for (int a = cl_start; a < nout; ++a) {
          double complex dot = 0;
          for (int r = 0; r < n; ++r) dot += conj(acc[r + (size_t)a * n]) * z[r];
          for (int r = 0; r < n; ++r) z[r] -= dot * acc[r + (size_t)a * n];
        }
      double nz = 0.0;
      for (int r = 0; r < n; ++r) nz += creal(z[r] * conj(z[r]));
      nz = sqrt(nz);
      if (nz < 0.3) continue;
      for (int r = 0; r < n; ++r) acc[r + (size_t)nout * n] = z[r] / nz;
      w[nout] = w2[t];
      ++nout;
    }
    i = j;
  }
  if (nout < n) rc = -2;
  for (int c = 0; c < nout; ++c)
    for (int r = 0; r < n; ++r) {
      A[2 * (r + (size_t)c * n)] = creal(acc[r + (size_t)c * n]);
      A[2 * (r + (size_t)c * n) + 1] = cimag(acc[r + (size_t)c * n]);
    }
  free(M);
  free(w2);
  free(acc);
  free(z);
  return rc;
}

/* ---- Rayleigh-Ritz, real and complex instantiations ------------------ */
#define RR_REAL
#include "rr_impl.inc"
#undef RR_REAL
#define RR_CPLX
#include "rr_impl.inc"
#undef RR_CPLX

int dn_rayleigh_ritz(int k, int m, int cplx, const double *GV, const double *GM,
                     const double *GA, double *w2, double *Y, int *zflags) {
  if (cplx) return rr_z(k, m, GV, GM, GA, w2, Y, zflags);
  return rr_d(k, m, GV, GM, GA, w2, Y, zflags);
}

void dn_qhgq(int k, int q, int cplx, const double *G, const double *Q, double *out) {
  if (cplx)
    qhgq_z(k, q, (const double complex *)G, (const double complex *)Q, (double complex *)out);
  else
    qhgq_d(k, q, G, Q, out);
}

void dn_project(int ku, int kv, int nu, int nv, int cplx, const double *G, const double *Qu,
                const double *Qv, double *out) {
  if (cplx)
    project_z(ku, kv, nu, nv, (const double complex *)G, (const double complex *)Qu,
              (const double complex *)Qv, (double complex *)out);
  else
    project_d(ku, kv, nu, nv, G, Qu, Qv, out);
}

int dn_chol_ok(int k, int cplx, const double *G, double tau) {
  return cplx ? chol_ok_z(k, (const double complex *)G, tau) : chol_ok_d(k, G, tau);
}

int dn_lrep_rr(int ku, int kv, int m, int cplx, const double *GK, const double *GM,
               const double *GW, double *w2, double *A, double *B) {
  if (cplx)
    return lrep_rr_z(ku, kv, m, (const double complex *)GK, (const double complex *)GM,
                     (const double complex *)GW, w2, (double complex *)A, (double complex *)B);
  return lrep_rr_d(ku, kv, m, GK, GM, GW, w2, A, B);
}

int dn_eig_low(int n, int m, int cplx, const double *G, double *w, double *X) {
  if (cplx) return eig_low_z(n, m, (const double complex *)G, w, (double complex *)X);
  return eig_low_d(n, m, G, w, X);
}

int dn_lrep_update(int k, int m, int cplx, const double *A, const int *act, int nact,
                   double drop, double *CQ, double *R11) {
  if (cplx)
    return lrep_update_z(k, m, (const double complex *)A, act, nact, drop, (double complex *)CQ,
                         (double complex *)R11);
  return lrep_update_d(k, m, A, act, nact, drop, CQ, R11);
}

int dn_orth_coef(int nq, int nw, int cplx, const double *GQW, const double *GWW, double drop,
                 double *C) {
  if (cplx)
    return orth_coef_z(nq, nw, (const double complex *)GQW, (const double complex *)GWW, drop,
                       (double complex *)C);
  return orth_coef_d(nq, nw, GQW, GWW, drop, C);
}
