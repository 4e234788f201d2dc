/*
 * bdg_oracle.c -- FP64 CPU oracle (test infrastructure only; see header).
 *
 * Every operator below restates the reference algorithm for one vector and
 * cites the reference file:line it follows.  The FFT is this file's own
 * mixed-radix Stockham transform with FFTW's conventions (unnormalised,
 * forward sign -1, r2c keeps kx in [0, Nx/2], c2r ignores the imaginary parts
 * of the kx = 0 and kx = Nx/2 bins).
 */
#include "bdg_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;

/* =====================================================================
 * 1D mixed-radix Stockham FFT
 * ===================================================================== */

typedef struct {
  int64_t n;
  int nf;
  int fac[64];
  cplx *tw; /* tw[m] = exp(-2 pi i m / n) */
} plan1;

static void plan1_init(plan1 *p, int64_t n) {
  p->n = n;
  p->nf = 0;
  int64_t m = n;
  const int pref[] = {4, 2, 3, 5};
  for (int t = 0; t < 4; t++) {
    while (m % pref[t] == 0 && m > 1) {
      p->fac[p->nf++] = pref[t];
      m /= pref[t];
    }
  }
  for (int64_t f = 7; m > 1; f += 2) {
    while (m % f == 0) {
      p->fac[p->nf++] = (int)f;
      m /= f;
    }
  }
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  for (int64_t k = 0; k < n; k++) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    p->tw[k] = cos(a) + I * sin(a);
  }
}

static void plan1_free(plan1 *p) {
  free(p->tw);
  p->tw = NULL;
}

/* twiddle exp(sign * 2 pi i m / n) */
static inline cplx twv(const plan1 *p, int64_t m, int sign) {
  const cplx w = p->tw[m % p->n];
  return sign < 0 ? w : conj(w);
}

/* In-place 1D transform of a (length n, contiguous); work has n entries. */
static void fft1(const plan1 *p, cplx *a, cplx *work, int sign) {
  const int64_t n = p->n;
  if (n <= 1) return;
  cplx *src = a, *dst = work;
  int64_t Ns = 1;
  for (int s = 0; s < p->nf; s++) {
    const int R = p->fac[s];
    const int64_t nr = n / R;
    const int64_t step = n / (Ns * R); /* twiddle index unit */
    for (int64_t j = 0; j < nr; j++) {
      const int64_t k = j % Ns;
      cplx v[64];
      cplx *vv = v;
      cplx *heap = NULL;
      if (R > 64) {
        heap = (cplx *)malloc(sizeof(cplx) * (size_t)R);
        vv = heap;
      }
      for (int r = 0; r < R; r++) {
        vv[r] = src[j + r * nr];
        if (r && k) vv[r] *= twv(p, (int64_t)r * k * step, sign);
      }
      /* DFT of length R */
      if (R == 2) {
        const cplx t = vv[0];
        vv[0] = t + vv[1];
        vv[1] = t - vv[1];
      } else if (R == 4) {
        const cplx a0 = vv[0] + vv[2], a1 = vv[0] - vv[2];
        const cplx b0 = vv[1] + vv[3];
        cplx b1 = vv[1] - vv[3];
        b1 = (sign < 0) ? (-I * b1) : (I * b1);
        vv[0] = a0 + b0;
        vv[2] = a0 - b0;
        vv[1] = a1 + b1;
        vv[3] = a1 - b1;
      } else {
        cplx out[64];
        cplx *oo = out;
        cplx *heap2 = NULL;
        if (R > 64) {
          heap2 = (cplx *)malloc(sizeof(cplx) * (size_t)R);
          oo = heap2;
        }
        const int64_t unit = n / R; /* exp(sign 2 pi i q r / R) = tw[q r n/R] */
        for (int q = 0; q < R; q++) {
          cplx acc = 0;
          for (int r = 0; r < R; r++)
            acc += vv[r] * twv(p, ((int64_t)q * r % R) * unit, sign);
          oo[q] = acc;
        }
        for (int q = 0; q < R; q++) vv[q] = oo[q];
        free(heap2);
      }
      const int64_t base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[base + r * Ns] = vv[r];
      free(heap);
    }
    cplx *t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)n);
}

/* Transform along axis `ax` of a 3D complex array with row length nx0
 * (elements per x-row of the stored array), extents e[0..2]. */
static void fft_axis(const plan1 *p, cplx *data, int ax, const int64_t *e,
                     int sign) {
  const int64_t n = e[ax];
  if (n <= 1) return;
  int64_t stride = 1, count = 1;
  for (int d = 0; d < ax; d++) stride *= e[d];
  for (int d = 0; d < 3; d++)
    if (d != ax) count *= e[d];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *wrk = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < count; c++) {
      /* pencil c -> base offset */
      int64_t base;
      if (ax == 0) {
        base = c * n;
      } else if (ax == 1) {
        const int64_t i0 = c % e[0], i2 = c / e[0];
        base = i0 + i2 * e[0] * e[1];
      } else {
        base = c; /* i0 + e0*i1 */
      }
      for (int64_t i = 0; i < n; i++) buf[i] = data[base + i * stride];
      fft1(p, buf, wrk, sign);
      for (int64_t i = 0; i < n; i++) data[base + i * stride] = buf[i];
    }
    free(buf);
    free(wrk);
  }
}

static void dims3(int dim, const int64_t *N, int64_t *e) {
  e[0] = e[1] = e[2] = 1;
  for (int d = 0; d < dim && d < 3; d++) e[d] = N[d];
}

typedef struct {
  plan1 p[3];   /* per axis, full lengths */
} plans3;

static void plans3_init(plans3 *P, const int64_t *e) {
  for (int d = 0; d < 3; d++) plan1_init(&P->p[d], e[d]);
}
static void plans3_free(plans3 *P) {
  for (int d = 0; d < 3; d++) plan1_free(&P->p[d]);
}

static void c2c_exec(const plans3 *P, const int64_t *e, cplx *data, int sign) {
  for (int ax = 0; ax < 3; ax++) fft_axis(&P->p[ax], data, ax, e, sign);
}

/* r2c: each x-row transformed as a complex sequence, first Nx/2+1 kept,
 * then the remaining axes on the half spectrum. */
static void r2c_exec(const plans3 *P, const int64_t *e, const double *in,
                     cplx *out) {
  const int64_t nx = e[0], nk = nx / 2 + 1, rows = e[1] * e[2];
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
    cplx *wrk = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
      for (int64_t i = 0; i < nx; i++) buf[i] = in[r * nx + i];
      fft1(&P->p[0], buf, wrk, -1);
      for (int64_t k = 0; k < nk; k++) out[r * nk + k] = buf[k];
    }
    free(buf);
    free(wrk);
  }
  const int64_t eh[3] = {nk, e[1], e[2]};
  fft_axis(&P->p[1], out, 1, eh, -1);
  fft_axis(&P->p[2], out, 2, eh, -1);
}

/* c2r: inverse along z, y on the half spectrum (destroys `in`), then each
 * x-row rebuilt as a Hermitian sequence (imaginary parts of the DC and
 * Nyquist bins dropped) and inverse-transformed; real part kept. */
static void c2r_exec(const plans3 *P, const int64_t *e, cplx *in, double *out) {
  const int64_t nx = e[0], nk = nx / 2 + 1, rows = e[1] * e[2];
  const int64_t eh[3] = {nk, e[1], e[2]};
  fft_axis(&P->p[2], in, 2, eh, +1);
  fft_axis(&P->p[1], in, 1, eh, +1);
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
    cplx *wrk = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
    for (int64_t r = 0; r < rows; r++) {
      const cplx *h = in + r * nk;
      for (int64_t k = 0; k < nx; k++) {
        if (k < nk)
          buf[k] = h[k];
        else
          buf[k] = conj(h[nx - k]);
      }
      buf[0] = creal(buf[0]);
      if (nx % 2 == 0) buf[nx / 2] = creal(buf[nx / 2]);
      fft1(&P->p[0], buf, wrk, +1);
      for (int64_t i = 0; i < nx; i++) out[r * nx + i] = creal(buf[i]);
    }
    free(buf);
    free(wrk);
  }
}

int orc_fft_c2c(int dim, const int64_t *N, double *data, int sign) {
  int64_t e[3];
  dims3(dim, N, e);
  plans3 P;
  plans3_init(&P, e);
  c2c_exec(&P, e, (cplx *)data, sign < 0 ? -1 : 1);
  plans3_free(&P);
  return 0;
}

int orc_fft_r2c(int dim, const int64_t *N, const double *in, double *out) {
  int64_t e[3];
  dims3(dim, N, e);
  plans3 P;
  plans3_init(&P, e);
  r2c_exec(&P, e, in, (cplx *)out);
  plans3_free(&P);
  return 0;
}

int orc_fft_c2r(int dim, const int64_t *N, double *in, double *out) {
  int64_t e[3];
  dims3(dim, N, e);
  plans3 P;
  plans3_init(&P, e);
  c2r_exec(&P, e, (cplx *)in, out);
  plans3_free(&P);
  return 0;
}

/* =====================================================================
 * Context (reference src/matmul_ctx.c:21-172)
 * ===================================================================== */

struct orc_ctx {
  int dim, cplx;
  int64_t N[3], e[3];
  double L[3];
  int64_t size, k_size, n0k;
  double *kx2[3];
  double *k2;
  double *ltK, *ltM, *pK, *pM, *U;
  void *wf;
  double mu, g_ddi;
  int dipolar;
  plans3 P;
  cplx *f_wrk;  /* k_size */
  void *c_wrk1; /* max(size, k_size) complex */
  void *c_wrk2;
};

orc_ctx *orc_ctx_new(int dim, const int64_t *N, const double *L,
                     int complex_psi0) {
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->dim = dim;
  c->cplx = complex_psi0 ? 1 : 0;
  c->size = 1;
  for (int d = 0; d < 3; d++) {
    c->N[d] = d < dim ? N[d] : 1;
    c->L[d] = d < dim ? L[d] : 1.0;
    c->e[d] = c->N[d];
  }
  for (int d = 0; d < dim; d++) c->size *= N[d];
  /* reference src/matmul_ctx.c:94-100 */
  c->n0k = c->cplx ? c->N[0] : c->N[0] / 2 + 1;
  c->k_size = c->n0k * c->N[1] * c->N[2];
  /* reference src/matmul_ctx.c:102-113: k_d(j) = dk*j for j <= N/2, else dk*(j-N) */
  for (int d = 0; d < 3; d++) {
    c->kx2[d] = (double *)calloc((size_t)c->N[d], sizeof(double));
    if (d >= dim) continue;
    const double dk = 2.0 * M_PI / c->L[d];
    for (int64_t j = 0; j < c->N[d]; j++) {
      const double kj = (j <= c->N[d] / 2) ? dk * (double)j : dk * (double)(j - c->N[d]);
      c->kx2[d][j] = kj * kj;
    }
  }
  /* reference src/matmul_ctx.c:115-132 */
  c->k2 = (double *)calloc((size_t)c->k_size, sizeof(double));
  for (int64_t iz = 0; iz < c->N[2]; iz++)
    for (int64_t iy = 0; iy < c->N[1]; iy++)
      for (int64_t ix = 0; ix < c->n0k; ix++)
        c->k2[(iz * c->N[1] + iy) * c->n0k + ix] =
            c->kx2[0][ix] + c->kx2[1][iy] + c->kx2[2][iz];
  c->ltK = (double *)calloc((size_t)c->size, sizeof(double));
  c->ltM = (double *)calloc((size_t)c->size, sizeof(double));
  plans3_init(&c->P, c->e);
  const int64_t w = c->size > c->k_size ? c->size : c->k_size;
  c->f_wrk = (cplx *)calloc((size_t)c->k_size, sizeof(cplx));
  c->c_wrk1 = calloc((size_t)w, sizeof(cplx));
  c->c_wrk2 = calloc((size_t)w, sizeof(cplx));
  return c;
}

void orc_ctx_free(orc_ctx *c) {
  if (!c) return;
  for (int d = 0; d < 3; d++) free(c->kx2[d]);
  free(c->k2);
  free(c->ltK);
  free(c->ltM);
  free(c->pK);
  free(c->pM);
  free(c->U);
  free(c->wf);
  free(c->f_wrk);
  free(c->c_wrk1);
  free(c->c_wrk2);
  plans3_free(&c->P);
  free(c);
}

int64_t orc_ctx_size(const orc_ctx *c) { return c->size; }
int64_t orc_ctx_ksize(const orc_ctx *c) { return c->k_size; }
void orc_ctx_get_k2(const orc_ctx *c, double *out) {
  memcpy(out, c->k2, sizeof(double) * (size_t)c->k_size);
}
void orc_ctx_get_kx2(const orc_ctx *c, int d, double *out) {
  memcpy(out, c->kx2[d], sizeof(double) * (size_t)c->N[d]);
}

void orc_ctx_set_fields(orc_ctx *c, const double *ltK, const double *ltM,
                        const double *wf, const double *pK, const double *pM,
                        double mu) {
  const size_t s = (size_t)c->size;
  if (ltK) memcpy(c->ltK, ltK, s * sizeof(double));
  if (ltM) memcpy(c->ltM, ltM, s * sizeof(double));
  if (wf) {
    const size_t bytes = s * sizeof(double) * (c->cplx ? 2 : 1);
    free(c->wf);
    c->wf = malloc(bytes);
    memcpy(c->wf, wf, bytes);
  }
  if (pK) {
    if (!c->pK) c->pK = (double *)malloc(s * sizeof(double));
    memcpy(c->pK, pK, s * sizeof(double));
  }
  if (pM) {
    if (!c->pM) c->pM = (double *)malloc(s * sizeof(double));
    memcpy(c->pM, pM, s * sizeof(double));
  }
  c->mu = mu;
}

/* ---- FFT helpers (reference src/fft_helpers_impl.inc:24-82) ---------- */

static void fwd(orc_ctx *c, const void *x) { /* x -> f_wrk */
  if (c->cplx) {
    memcpy(c->f_wrk, x, sizeof(cplx) * (size_t)c->size);
    c2c_exec(&c->P, c->e, c->f_wrk, -1);
  } else {
    r2c_exec(&c->P, c->e, (const double *)x, c->f_wrk);
  }
}

static void bwd(orc_ctx *c, void *y) { /* f_wrk -> y (f_wrk destroyed) */
  if (c->cplx) {
    c2c_exec(&c->P, c->e, c->f_wrk, +1);
    memcpy(y, c->f_wrk, sizeof(cplx) * (size_t)c->size);
  } else {
    c2r_exec(&c->P, c->e, c->f_wrk, (double *)y);
  }
}

/* =====================================================================
 * Operators
 * ===================================================================== */

/* reference src/kinetic_impl.inc:17-30: y = IFFT(k^2 * 0.5/size * FFT x) */
void orc_kinetic(orc_ctx *c, const double *x, double *y) {
  const double norm = 0.5 / (double)c->size;
  fwd(c, x);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c->k_size; i++) c->f_wrk[i] *= c->k2[i] * norm;
  bwd(c, y);
}

/* reference src/dipolar_conv_impl.inc:22-52:
 * out = g * wf * IFFT(U/size * FFT(conj(wf) * v)) */
void orc_dipolar_conv(orc_ctx *c, const double *v, double *out) {
  const double norm = 1.0 / (double)c->size, g = c->g_ddi;
  const int64_t S = c->size;
  if (c->cplx) {
    const cplx *wf = (const cplx *)c->wf, *vv = (const cplx *)v;
    cplx *buf = (cplx *)c->c_wrk1;
    for (int64_t i = 0; i < S; i++) buf[i] = conj(wf[i]) * vv[i];
    fwd(c, buf);
  } else {
    const double *wf = (const double *)c->wf;
    double *buf = (double *)c->c_wrk1;
    for (int64_t i = 0; i < S; i++) buf[i] = wf[i] * v[i];
    fwd(c, buf);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c->k_size; i++) c->f_wrk[i] *= c->U[i] * norm;
  if (c->cplx) {
    cplx *buf = (cplx *)c->c_wrk1;
    bwd(c, buf);
    const cplx *wf = (const cplx *)c->wf;
    cplx *o = (cplx *)out;
    for (int64_t i = 0; i < S; i++) o[i] = g * wf[i] * buf[i];
  } else {
    double *buf = (double *)c->c_wrk1;
    bwd(c, buf);
    const double *wf = (const double *)c->wf;
    for (int64_t i = 0; i < S; i++) out[i] = g * wf[i] * buf[i];
  }
}

/* reference src/matmulK_impl.inc:11-28: y = kinetic(x) + ltK * x */
void orc_matmulK(orc_ctx *c, const double *x, double *y) {
  orc_kinetic(c, x, y);
  const int64_t S = c->size;
  if (c->cplx) {
    const cplx *xx = (const cplx *)x;
    cplx *yy = (cplx *)y;
    for (int64_t i = 0; i < S; i++) yy[i] += c->ltK[i] * xx[i];
  } else {
    for (int64_t i = 0; i < S; i++) y[i] += c->ltK[i] * x[i];
  }
}

/* reference src/matmulM_impl.inc:14-44:
 * y = kinetic(x) + ltM * x (+ 2 * dipolar_conv(x) if dipolar and dim == 3) */
void orc_matmulM(orc_ctx *c, const double *x, double *y) {
  orc_kinetic(c, x, y);
  const int64_t S = c->size;
  const int dip = c->dipolar && c->dim == 3;
  if (c->cplx) {
    const cplx *xx = (const cplx *)x;
    cplx *yy = (cplx *)y;
    cplx *dc = (cplx *)c->c_wrk2;
    if (dip) orc_dipolar_conv(c, x, (double *)dc);
    for (int64_t i = 0; i < S; i++)
      yy[i] += c->ltM[i] * xx[i] + (dip ? 2.0 * dc[i] : 0.0);
  } else {
    double *dc = (double *)c->c_wrk2;
    if (dip) orc_dipolar_conv(c, x, dc);
    for (int64_t i = 0; i < S; i++)
      y[i] += c->ltM[i] * x[i] + (dip ? 2.0 * dc[i] : 0.0);
  }
}

static int64_t cw(const orc_ctx *c) { return c->cplx ? 2 : 1; } /* doubles per element */

/* reference src/matmulLrep_impl.inc:13-18: [u; v] -> [K u; M v] */
void orc_matmulLrep(orc_ctx *c, const double *x, double *y) {
  const int64_t off = c->size * cw(c);
  orc_matmulK(c, x, y);
  orc_matmulM(c, x + off, y + off);
}

/* reference src/matmulSwap_impl.inc:14-26: [u; v] -> [v; u] */
void orc_matmulSwap(orc_ctx *c, const double *x, double *y) {
  const int64_t off = c->size * cw(c);
  memcpy(y, x + off, sizeof(double) * (size_t)off);
  memcpy(y + off, x, sizeof(double) * (size_t)off);
}

/* reference src/bdg_internal.h:47-50 */
static inline double safe_val(double x) { return fabs(x) < 1e-8 ? 1e-8 : x; }

/* reference src/precond_impl.inc:21-111:
 * y = p * IFFT( FFT(p * x) / (size * safe_val(mu + k^2/2)) ) */
static void precond_common(orc_ctx *c, const double *p, const double *x,
                           double *y) {
  const int64_t S = c->size;
  if (c->cplx) {
    cplx *buf = (cplx *)c->c_wrk1;
    const cplx *xx = (const cplx *)x;
    for (int64_t i = 0; i < S; i++) buf[i] = p[i] * xx[i];
  } else {
    double *buf = (double *)c->c_wrk1;
    for (int64_t i = 0; i < S; i++) buf[i] = p[i] * x[i];
  }
  fwd(c, c->c_wrk1);
  const double inv = 1.0 / (double)S;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c->k_size; i++)
    c->f_wrk[i] *= inv / safe_val(c->mu + 0.5 * c->k2[i]);
  bwd(c, y);
  if (c->cplx) {
    cplx *yy = (cplx *)y;
    for (int64_t i = 0; i < S; i++) yy[i] *= p[i];
  } else {
    for (int64_t i = 0; i < S; i++) y[i] *= p[i];
  }
}

void orc_precondK(orc_ctx *c, const double *x, double *y) {
  precond_common(c, c->pK, x, y);
}
void orc_precondM(orc_ctx *c, const double *x, double *y) {
  precond_common(c, c->pM, x, y);
}
/* reference src/precond_impl.inc:113-118 */
void orc_precondLrep(orc_ctx *c, const double *x, double *y) {
  const int64_t off = c->size * cw(c);
  orc_precondK(c, x, y);
  orc_precondM(c, x + off, y + off);
}

void orc_apply_block(orc_ctx *c, int op, const double *X, double *Y,
                     int64_t ncols, int64_t ld) {
  const int64_t w = cw(c);
  for (int64_t j = 0; j < ncols; j++) {
    const double *x = X + j * ld * w;
    double *y = Y + j * ld * w;
    switch (op) {
      case 0: orc_kinetic(c, x, y); break;
      case 1: orc_matmulK(c, x, y); break;
      case 2: orc_matmulM(c, x, y); break;
      case 3: orc_matmulLrep(c, x, y); break;
      case 4: orc_matmulSwap(c, x, y); break;
      case 5: orc_precondK(c, x, y); break;
      case 6: orc_precondM(c, x, y); break;
      case 7: orc_precondLrep(c, x, y); break;
      case 8: orc_dipolar_conv(c, x, y); break;
      default: break;
    }
  }
}

/* =====================================================================
 * Dipolar kernel and mean field
 * ===================================================================== */

/* reference src/dipolar.c:21-72:
 * U(k) = (3 (k.d)^2 / (|k|^2 |d|^2) - 1) * (1 + 3 cos(kR)/(kR)^2 - 3 sin(kR)/(kR)^3),
 * U = 0 where |k|^2 < 1e-50; the complex layout wraps ix > Nx/2. */
static void dipolar_kernel_fill(const orc_ctx *c, const double *dir, double R,
                                double *U) {
  const int64_t Nx = c->N[0], Ny = c->N[1], Nz = c->N[2], n0k = c->n0k;
  const double dkx = 2.0 * M_PI / c->L[0], dky = 2.0 * M_PI / c->L[1],
               dkz = 2.0 * M_PI / c->L[2];
  const double d2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
  for (int64_t iz = 0; iz < Nz; iz++) {
    const double kz = dkz * (double)(iz <= Nz / 2 ? iz : iz - Nz);
    for (int64_t iy = 0; iy < Ny; iy++) {
      const double ky = dky * (double)(iy <= Ny / 2 ? iy : iy - Ny);
      for (int64_t ix = 0; ix < n0k; ix++) {
        const double kx = dkx * (double)((c->cplx && ix > Nx / 2) ? ix - Nx : ix);
        const int64_t idx = (iz * Ny + iy) * n0k + ix;
        const double ks = kx * kx + ky * ky + kz * kz;
        if (ks < 1e-50) {
          U[idx] = 0.0;
          continue;
        }
        const double kd = kx * dir[0] + ky * dir[1] + kz * dir[2];
        const double c2 = kd * kd / (ks * d2);
        const double kR = sqrt(ks) * R, kR2 = kR * kR;
        const double f = 1.0 + 3.0 * cos(kR) / kR2 - 3.0 * sin(kR) / (kR2 * kR);
        U[idx] = (3.0 * c2 - 1.0) * f;
      }
    }
  }
}

void orc_dipolar_set_kernel(orc_ctx *c, double g_ddi, const double *dir,
                            double R) {
  c->g_ddi = g_ddi;
  if (!c->U) c->U = (double *)calloc((size_t)c->k_size, sizeof(double));
  dipolar_kernel_fill(c, dir, R, c->U);
  c->dipolar = 1;
}

void orc_dipolar_get_kernel(const orc_ctx *c, double *out) {
  if (c->U) memcpy(out, c->U, sizeof(double) * (size_t)c->k_size);
}

/* reference src/dipolar_meanfield_impl.inc:20-55:
 * g * Re IFFT(U/size * FFT(|wf|^2)) */
void orc_dipolar_meanfield(orc_ctx *c, double *gphi) {
  const int64_t S = c->size;
  const double norm = 1.0 / (double)S;
  if (c->cplx) {
    const cplx *wf = (const cplx *)c->wf;
    cplx *buf = (cplx *)c->c_wrk1;
    for (int64_t i = 0; i < S; i++) {
      const double a = cabs(wf[i]);
      buf[i] = a * a;
    }
    fwd(c, buf);
  } else {
    const double *wf = (const double *)c->wf;
    double *buf = (double *)c->c_wrk1;
    for (int64_t i = 0; i < S; i++) buf[i] = fabs(wf[i]) * fabs(wf[i]);
    fwd(c, buf);
  }
  for (int64_t i = 0; i < c->k_size; i++) c->f_wrk[i] *= c->U[i] * norm;
  if (c->cplx) {
    cplx *buf = (cplx *)c->c_wrk1;
    bwd(c, buf);
    for (int64_t i = 0; i < S; i++) gphi[i] = c->g_ddi * creal(buf[i]);
  } else {
    double *buf = (double *)c->c_wrk1;
    bwd(c, buf);
    for (int64_t i = 0; i < S; i++) gphi[i] = c->g_ddi * buf[i];
  }
}

/* =====================================================================
 * Setup arrays (reference src/setup.c)
 * ===================================================================== */

/* Grid coordinate r_d = (i - N/2) * L/N (reference src/setup.c:35,68-72). */
static inline double coord(const orc_ctx *c, int d, int64_t i) {
  return ((double)i - 0.5 * (double)c->N[d]) * c->L[d] / (double)c->N[d];
}

void orc_setup_trap_harmonic(const orc_ctx *c, const double *om, double *ltK,
                             double *ltM) {
  for (int64_t iz = 0; iz < c->N[2]; iz++)
    for (int64_t iy = 0; iy < c->N[1]; iy++)
      for (int64_t ix = 0; ix < c->N[0]; ix++) {
        const double r[3] = {coord(c, 0, ix), coord(c, 1, iy), coord(c, 2, iz)};
        double V = 0.0;
        for (int d = 0; d < c->dim; d++) V += 0.5 * om[d] * om[d] * r[d] * r[d];
        const int64_t idx = (iz * c->N[1] + iy) * c->N[0] + ix;
        ltK[idx] += V;
        ltM[idx] += V;
      }
}

/* reference src/setup.c:119-145 with U_intK = gK*n, U_intM = gM*n */
void orc_setup_contact(const orc_ctx *c, const double *wf, double gK, double gM,
                       double *ltK, double *ltM) {
  for (int64_t i = 0; i < c->size; i++) {
    const double n = c->cplx ? wf[2 * i] * wf[2 * i] + wf[2 * i + 1] * wf[2 * i + 1]
                             : wf[i] * wf[i];
    const double uK = gK * n, uM = gM * n;
    ltK[i] += uK;
    ltM[i] += uK + 2.0 * uM;
  }
}

/* reference src/setup.c:150-177 */
void orc_setup_mu(int64_t S, double mu, double *ltK, double *ltM, double *pK,
                  double *pM) {
  for (int64_t i = 0; i < S; i++) {
    pK[i] = 1.0 / sqrt(fmax(1e-8, ltK[i]));
    pM[i] = 1.0 / sqrt(fmax(1e-8, ltM[i]));
  }
  for (int64_t i = 0; i < S; i++) {
    ltK[i] -= mu;
    ltM[i] -= mu;
  }
}

/* =====================================================================
 * Imaginary-time split-step ground state
 *
 * One step (dt):
 *   n = |psi|^2, Phi = Re IFFT(U * FFT(n)) / S       (if g_dd != 0)
 *   psi *= exp(-dt/2 * (V + g n + g_dd Phi))
 *   psi  = IFFT(exp(-dt k^2/2) * FFT(psi)) / S       (c2c, all k)
 *   n, Phi recomputed from the new psi
 *   psi *= exp(-dt/2 * (V + g n + g_dd Phi))
 *   psi /= sqrt(sum |psi|^2 dV)
 * After the last step mu = sum conj(psi) (T + V + g n + g_dd Phi) psi dV
 * with T = IFFT(k^2/2 * FFT(psi)) / S.  V is the harmonic trap
 * 0.5 sum_d (omega_d r_d)^2 on the grid of reference src/setup.c.
 * The dipolar kernel uses the layout of the context (real: r2c half spectrum,
 * complex: full).  psi is real (size) or interleaved complex (2*size).
 * ===================================================================== */

static void potential_eff(const orc_ctx *c, const double *V, double g,
                          double gdd, const double *U, const cplx *psi,
                          double *veff, cplx *tmp, double *tmpr) {
  const int64_t S = c->size;
  for (int64_t i = 0; i < S; i++) {
    const double n = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
    veff[i] = V[i] + g * n;
  }
  if (gdd != 0.0 && U) {
    const double inv = 1.0 / (double)S;
    if (c->cplx) {
      for (int64_t i = 0; i < S; i++)
        tmp[i] = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
      c2c_exec(&c->P, c->e, tmp, -1);
      for (int64_t i = 0; i < S; i++) tmp[i] *= U[i] * inv;
      c2c_exec(&c->P, c->e, tmp, +1);
      for (int64_t i = 0; i < S; i++) veff[i] += gdd * creal(tmp[i]);
    } else {
      for (int64_t i = 0; i < S; i++)
        tmpr[i] = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
      r2c_exec(&c->P, c->e, tmpr, tmp);
      for (int64_t i = 0; i < c->k_size; i++) tmp[i] *= U[i] * inv;
      c2r_exec(&c->P, c->e, tmp, tmpr);
      for (int64_t i = 0; i < S; i++) veff[i] += gdd * tmpr[i];
    }
  }
}

int orc_itp_groundstate(const orc_ctx *c, const double *omega, double g,
                        double g_dd, const double *dir, double cutoff,
                        double dt, int64_t nsteps, double *psi_io,
                        double *mu_out) {
  const int64_t S = c->size;
  double dV = 1.0;
  for (int d = 0; d < c->dim; d++) dV *= c->L[d] / (double)c->N[d];
  double *V = (double *)calloc((size_t)S, sizeof(double));
  double *dummy = (double *)calloc((size_t)S, sizeof(double));
  orc_setup_trap_harmonic(c, omega, V, dummy);
  double *U = NULL;
  if (g_dd != 0.0) {
    U = (double *)calloc((size_t)c->k_size, sizeof(double));
    dipolar_kernel_fill(c, dir, cutoff, U);
  }
  /* full c2c kinetic factor over all k */
  double *ek = (double *)malloc(sizeof(double) * (size_t)S);
  double *k2full = (double *)malloc(sizeof(double) * (size_t)S);
  for (int64_t iz = 0; iz < c->N[2]; iz++)
    for (int64_t iy = 0; iy < c->N[1]; iy++)
      for (int64_t ix = 0; ix < c->N[0]; ix++) {
        const int64_t i = (iz * c->N[1] + iy) * c->N[0] + ix;
        k2full[i] = c->kx2[0][ix] + c->kx2[1][iy] + c->kx2[2][iz];
        ek[i] = exp(-dt * 0.5 * k2full[i]) / (double)S;
      }
  cplx *psi = (cplx *)malloc(sizeof(cplx) * (size_t)S);
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(S > c->k_size ? S : c->k_size));
  double *tmpr = (double *)malloc(sizeof(double) * (size_t)S);
  double *veff = (double *)malloc(sizeof(double) * (size_t)S);
  for (int64_t i = 0; i < S; i++)
    psi[i] = c->cplx ? psi_io[2 * i] + I * psi_io[2 * i + 1] : psi_io[i];
  /* initial normalisation */
  {
    double nrm = 0.0;
    for (int64_t i = 0; i < S; i++) nrm += creal(psi[i] * conj(psi[i]));
    const double s = 1.0 / sqrt(nrm * dV);
    for (int64_t i = 0; i < S; i++) psi[i] *= s;
  }
  for (int64_t step = 0; step < nsteps; step++) {
    potential_eff(c, V, g, g_dd, U, psi, veff, tmp, tmpr);
    for (int64_t i = 0; i < S; i++) psi[i] *= exp(-0.5 * dt * veff[i]);
    c2c_exec(&c->P, c->e, psi, -1);
    for (int64_t i = 0; i < S; i++) psi[i] *= ek[i];
    c2c_exec(&c->P, c->e, psi, +1);
    if (!c->cplx)
      for (int64_t i = 0; i < S; i++) psi[i] = creal(psi[i]);
    potential_eff(c, V, g, g_dd, U, psi, veff, tmp, tmpr);
    double nrm = 0.0;
    for (int64_t i = 0; i < S; i++) {
      psi[i] *= exp(-0.5 * dt * veff[i]);
      nrm += creal(psi[i] * conj(psi[i]));
    }
    const double s = 1.0 / sqrt(nrm * dV);
    for (int64_t i = 0; i < S; i++) psi[i] *= s;
  }
  /* chemical potential */
  potential_eff(c, V, g, g_dd, U, psi, veff, tmp, tmpr);
  for (int64_t i = 0; i < S; i++) tmp[i] = psi[i];
  c2c_exec(&c->P, c->e, tmp, -1);
  for (int64_t i = 0; i < S; i++) tmp[i] *= 0.5 * k2full[i] / (double)S;
  c2c_exec(&c->P, c->e, tmp, +1);
  double mu = 0.0;
  for (int64_t i = 0; i < S; i++)
    mu += creal(conj(psi[i]) * (tmp[i] + veff[i] * psi[i]));
  mu *= dV;
  if (mu_out) *mu_out = mu;
  for (int64_t i = 0; i < S; i++) {
    if (c->cplx) {
      psi_io[2 * i] = creal(psi[i]);
      psi_io[2 * i + 1] = cimag(psi[i]);
    } else {
      psi_io[i] = creal(psi[i]);
    }
  }
  free(V);
  free(dummy);
  free(U);
  free(ek);
  free(k2full);
  free(psi);
  free(tmp);
  free(tmpr);
  free(veff);
  return 0;
}

/* =====================================================================
 * Start blocks
 * ===================================================================== */

/* reference src/bdg_internal.h:53-65 */
uint32_t orc_xorshift32(uint32_t *s) {
  uint32_t x = *s;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  *s = x;
  return x;
}

double orc_xrand(uint32_t *s) {
  return (double)(orc_xorshift32(s) & 0x7FFFFFFFu) / (double)0x7FFFFFFF;
}

void orc_start_block_uniform(const orc_ctx *c, uint32_t seed, int64_t ncols,
                             double *X) {
  const int64_t S = c->size, n = 2 * S;
  uint32_t st = seed;
  for (int64_t j = 0; j < ncols; j++)
    for (int64_t i = 0; i < S; i++) {
      const double v = orc_xrand(&st) - 0.5;
      if (c->cplx) {
        X[2 * (i + j * n)] = v;
        X[2 * (i + j * n) + 1] = 0.0;
        X[2 * (i + S + j * n)] = v;
        X[2 * (i + S + j * n) + 1] = 0.0;
      } else {
        X[i + j * n] = v;
        X[i + S + j * n] = v;
      }
    }
}

/* Reuse block of a parameter sweep, reference src/bdg.c:112-207
 * (bdg_reuse_modes): S rows per half, modes U, V (S x nev, FP64 or
 * interleaved complex), wf (S values, or NULL), output X (2S x m, [u; v]
 * column-major).  One xorshift32 stream from `seed`, two xrand values per
 * row, columns in order: columns j < nev get the modes plus
 * sigma_j (z0, z1), sigma_j = noise ||[u_j; v_j]|| (complex: noise on the
 * real parts); columns nev <= j < m get |z0| |wf| in both halves. */
void orc_reuse_block(int64_t S, int64_t nev, int64_t m, int cplx, const double *U,
                     const double *V, const double *wf, double noise, uint32_t seed,
                     double *X) {
  const int64_t n = 2 * S, w = cplx ? 2 : 1;
  uint32_t st = seed;
  for (int64_t j = 0; j < m; j++) {
    double sigma = 0.0;
    if (j < nev) {
      double cn = 0.0;
      for (int64_t i = 0; i < S * w; i++)
        cn += U[j * S * w + i] * U[j * S * w + i] + V[j * S * w + i] * V[j * S * w + i];
      sigma = noise * sqrt(cn);
    }
    for (int64_t i = 0; i < S; i++) {
      double u1 = orc_xrand(&st);
      const double u2 = orc_xrand(&st);
      u1 = fmax(u1, 1e-10);
      const double r = sqrt(-2.0 * log(u1));
      const double z0 = r * cos(2.0 * M_PI * u2), z1 = r * sin(2.0 * M_PI * u2);
      double *xu = X + (j * n + i) * w, *xv = X + (j * n + S + i) * w;
      if (j < nev) {
        const double *u = U + (j * S + i) * w, *v = V + (j * S + i) * w;
        xu[0] = u[0] + sigma * z0;
        xv[0] = v[0] + sigma * z1;
        if (cplx) {
          xu[1] = u[1];
          xv[1] = v[1];
        }
      } else {
        double a = 1.0;
        if (wf) a = cplx ? hypot(wf[2 * i], wf[2 * i + 1]) : fabs(wf[i]);
        xu[0] = xv[0] = fabs(z0) * a;
        if (cplx) xu[1] = xv[1] = 0.0;
      }
    }
  }
}

int orc_num_threads(void) { return omp_get_max_threads(); }
