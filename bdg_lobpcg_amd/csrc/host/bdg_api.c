/*
 * Host implementation of the public BdG API (include/bdg/bdg.h).
 *
 * Lifecycle, setup state machine and result ownership follow the reference
 * (src/bdg.c, src/setup.c, src/solver.c); the heavy work runs on the GPU
 * through the bdgx C ABI: the dipolar kernel and mean field at setup, and
 * the whole eigen-iteration in bdg_solve.
 */
#define _DEFAULT_SOURCE /* madvise, MADV_HUGEPAGE */
#include <pthread.h>
#include <sys/mman.h>
#include <complex.h>
#include <inttypes.h>

#include "bdg_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) BDG_FAIL("out of host memory (%zu x %zu bytes)", n, s);
  return p;
}

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) BDG_FAIL("out of host memory (%zu bytes)", n);
  return p;
}

/* large host blocks (start block, result modes): 2 MiB aligned and marked for
 * transparent huge pages, so that first touch faults once per 2 MiB instead
 * of once per 4 KiB page (a 1 GB block otherwise costs ~0.25 s of faults);
 * freed with free() */
static void *xmalloc_big(size_t n) {
  const size_t align = (size_t)2 << 20;
  if (n < align) return xmalloc(n);
  void *p = NULL;
  if (posix_memalign(&p, align, n) != 0 || !p)
    BDG_FAIL("out of host memory (%zu bytes)", n);
  (void)madvise(p, n, MADV_HUGEPAGE);
  return p;
}

/* The two mode arrays of a solve, allocated on a helper thread while the
 * device iterates: page-locked (the final download then runs at the link
 * rate instead of through the runtime's staging of pageable memory), or, if
 * that fails, huge-page memory touched page by page. */
typedef struct {
  size_t bytes;
  void *u, *v;
  int pinned;
} result_prefault;

/* Page-locked host blocks of finished results, kept for later solves of
 * this process: page-locking takes the runtime's memory lock for as long as
 * the pages take to pin (tens of ms per GB, now and then seconds), which
 * stalls the solver's own HIP calls on the other thread; a block taken from
 * here needs no pinning.  At most PIN_POOL blocks of at most PIN_POOL_MAX
 * bytes each are kept; the rest are released. */
#define PIN_POOL 4
#define PIN_POOL_MAX ((size_t)4 << 30)
static struct {
  void *p;
  size_t bytes;
} pin_pool[PIN_POOL];
static pthread_mutex_t pin_mu = PTHREAD_MUTEX_INITIALIZER;

/* the smallest kept block of at least `bytes`, or NULL */
static void *pin_take(size_t bytes) {
  pthread_mutex_lock(&pin_mu);
  int best = -1;
  for (int i = 0; i < PIN_POOL; ++i)
    if (pin_pool[i].p && pin_pool[i].bytes >= bytes &&
        (best < 0 || pin_pool[i].bytes < pin_pool[best].bytes))
      best = i;
  void *p = NULL;
  if (best >= 0) {
    p = pin_pool[best].p;
    pin_pool[best].p = NULL;
    pin_pool[best].bytes = 0;
  }
  pthread_mutex_unlock(&pin_mu);
  return p;
}

/* keep a page-locked block (a full pool gives up its smallest block for a
 * larger one) or release it */
static void pin_give(void *p, size_t bytes) {
  if (!p) return;
  void *drop = p;
  if (bytes <= PIN_POOL_MAX) {
    pthread_mutex_lock(&pin_mu);
    int slot = -1;
    for (int i = 0; i < PIN_POOL && slot < 0; ++i)
      if (!pin_pool[i].p) slot = i;
    if (slot < 0) {
      int small = 0;
      for (int i = 1; i < PIN_POOL; ++i)
        if (pin_pool[i].bytes < pin_pool[small].bytes) small = i;
      if (pin_pool[small].bytes < bytes) {
        slot = small;
        drop = pin_pool[small].p;
      }
    } else {
      drop = NULL;
    }
    if (slot >= 0) {
      pin_pool[slot].p = p;
      pin_pool[slot].bytes = bytes;
    }
    pthread_mutex_unlock(&pin_mu);
  }
  if (drop) (void)bdgx_host_free(drop);
}

static void touch_pages(void *p, size_t bytes) {
  volatile char *c = (volatile char *)p;
  for (size_t o = 0; o < bytes; o += 4096) c[o] = 0;
}

static void *prefault_main(void *arg) {
  result_prefault *r = (result_prefault *)arg;
  if ((r->u || bdgx_host_alloc(&r->u, r->bytes) == BDGX_OK) &&
      (r->v || bdgx_host_alloc(&r->v, r->bytes) == BDGX_OK)) {
    r->pinned = 1;
    return NULL;
  }
  pin_give(r->u, r->bytes);
  pin_give(r->v, r->bytes);
  r->u = NULL;
  r->v = NULL;
  r->u = xmalloc_big(r->bytes);
  touch_pages(r->u, r->bytes);
  r->v = xmalloc_big(r->bytes);
  touch_pages(r->v, r->bytes);
  return NULL;
}

/* the helper thread, started once the solver's own device allocations are
 * done (page-locking the host arrays holds the runtime's memory lock:
 * started earlier, it stalled the solver's first allocations by ~28 ms) */
typedef struct {
  result_prefault pf;
  pthread_t th;
  int on; /* 0 not started, 1 running or done, -1 could not start */
} prefault_job;

static void prefault_start(void *arg) {
  prefault_job *j = (prefault_job *)arg;
  if (j->on) return;
  /* blocks kept from earlier solves first: no thread when both are there */
  j->pf.u = pin_take(j->pf.bytes);
  j->pf.v = pin_take(j->pf.bytes);
  if (j->pf.u && j->pf.v) {
    j->pf.pinned = 1;
    j->on = 2;
    return;
  }
  j->on = pthread_create(&j->th, NULL, prefault_main, &j->pf) == 0 ? 1 : -1;
  if (j->on < 0) {
    pin_give(j->pf.u, j->pf.bytes);
    pin_give(j->pf.v, j->pf.bytes);
    j->pf.u = j->pf.v = NULL;
  }
}

static void release_prefault(result_prefault *r) {
  if (r->pinned) {
    pin_give(r->u, r->bytes);
    pin_give(r->v, r->bytes);
  } else {
    free(r->u);
    free(r->v);
  }
  r->u = r->v = NULL;
}

static void freep(void **p) {
  free(*p);
  *p = NULL;
}

/* ---- lifecycle (reference src/bdg.c:17-84) ------------------------------ */

bdg_t *bdg_alloc(uint64_t dim, const uint64_t *N, const f64 *L, int complex_psi0) {
  if (dim < 1 || dim > 3 || !N || !L) BDG_FAIL("bdg_alloc: invalid grid (dim %" PRIu64 ")", dim);
  bdg_t *b = (bdg_t *)xcalloc(1, sizeof(bdg_t));
  b->dim = dim;
  b->complex_psi0 = complex_psi0 ? 1 : 0;
  b->size = 1;
  for (int d = 0; d < 3; ++d) {
    b->N[d] = d < (int)dim ? N[d] : 1;
    b->L[d] = d < (int)dim ? L[d] : 1.0;
  }
  for (uint64_t d = 0; d < dim; ++d) b->size *= N[d];
  b->ltK = (double *)xcalloc(b->size, sizeof(double));
  b->ltM = (double *)xcalloc(b->size, sizeof(double));
  b->nev = 3;
  b->sizeSub = 6;
  b->maxIter = 500;
  b->tol = 1e-6;
  b->precision = BDG_PREC_FP64;
  b->img_tol = 0.0;
  b->refresh_every = 20;
  b->lp_window = 30;
  b->lp_factor = 0.5;
  b->graphs = 1;
  b->freeze_tol = 1e-6;
  b->formulation = BDG_FORM_FIRST_ORDER;
  b->fp32_tol = 1e-5;
  b->fp32_stall = 40;
  b->lock_max = -1;
  b->lock_guards = 4;
  b->device_sweeps = 1;
  return b;
}

static void free_results(bdg_t *b) {
  freep((void **)&b->eigvals);
  if (b->modes_pinned) {
    pin_give(b->modes_u, b->modes_bytes);
    pin_give(b->modes_v, b->modes_bytes);
    b->modes_u = b->modes_v = NULL;
    b->modes_pinned = 0;
  }
  freep(&b->modes_u);
  freep(&b->modes_v);
  freep((void **)&b->residuals);
  freep((void **)&b->zero_modes);
  (void)bdgx_free(b->d_modes_u);
  (void)bdgx_free(b->d_modes_v);
  b->d_modes_u = b->d_modes_v = NULL;
  b->converged = 0;
}

static void free_device_reuse(bdg_t *b) {
  (void)bdgx_free(b->d_reuse_u);
  (void)bdgx_free(b->d_reuse_v);
  b->d_reuse_u = b->d_reuse_v = NULL;
}

void bdg_free(bdg_t **pb) {
  if (!pb || !*pb) return;
  bdg_t *b = *pb;
  free_results(b);
  free_device_reuse(b);
  freep(&b->reuse_buf);
  for (int d = 0; d < 3; ++d) freep((void **)&b->kx2[d]);
  freep((void **)&b->ltK);
  freep((void **)&b->ltM);
  freep((void **)&b->pK);
  freep((void **)&b->pM);
  freep(&b->wf);
  freep((void **)&b->lp_theta);
  if (b->dev) bdgx_ctx_destroy(b->dev);
  free(b);
  *pb = NULL;
}

void bdg_reset(bdg_t *b) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_reset");
  freep(&b->wf);
  freep((void **)&b->pK);
  freep((void **)&b->pM);
  memset(b->ltK, 0, sizeof(double) * b->size);
  memset(b->ltM, 0, sizeof(double) * b->size);
  b->mu = 0.0;
  b->g_ddi = 0.0;
  b->dipolar = 0;
  free_results(b);
  /* the device context keeps grid, plans and workspaces; its physics fields
   * are marked absent and re-uploaded by the next bdg_set_dipolar / bdg_solve */
  if (b->dev && bdgx_reset_fields(b->dev) != BDGX_OK) BDG_FAIL("bdg_reset: device reset failed");
  b->state = BDG_HAS_SYSTEM;
}

/* ---- setup ---------------------------------------------------------------- */

void bdg_set_system(bdg_t *b) {
  /* k-tables (reference src/matmul_ctx.c:102-113) */
  b->k_size = b->complex_psi0 ? b->size : (b->N[0] / 2 + 1) * b->N[1] * b->N[2];
  for (int d = 0; d < 3; ++d) {
    freep((void **)&b->kx2[d]);
    b->kx2[d] = (double *)xcalloc(b->N[d], sizeof(double));
    if (d >= (int)b->dim) continue;
    const double dk = 2.0 * M_PI / b->L[d];
    for (uint64_t j = 0; j < b->N[d]; ++j) {
      const double kj = (j <= b->N[d] / 2) ? dk * (double)j : dk * ((double)j - (double)b->N[d]);
      b->kx2[d][j] = kj * kj;
    }
  }
  /* the device context is created on first device use (ensure_device) */
  b->state |= BDG_HAS_SYSTEM;
}

/* Create the device context on first use: lets the host-side setup run (and
 * be tested) on machines without a GPU. */
static int ensure_device(bdg_t *b) {
  if (b->dev) return BDGX_OK;
  return bdgx_ctx_create(&b->dev, b->dim, b->N, b->L, b->complex_psi0);
}

/* reference src/setup.c:17-84: r_d = (i - N_d/2) L_d / N_d, additive */
void bdg_set_trap(bdg_t *b, f64 (*V)(uint64_t, const f64 *, const void *), const void *param) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_set_trap");
  if (!V) {
    fprintf(stderr, "WARN bdg_set_trap: V_trap is NULL, skipping\n");
    return;
  }
  for (uint64_t iz = 0; iz < b->N[2]; ++iz)
    for (uint64_t iy = 0; iy < b->N[1]; ++iy)
      for (uint64_t ix = 0; ix < b->N[0]; ++ix) {
        const uint64_t idx[3] = {ix, iy, iz};
        f64 r[3];
        for (uint64_t d = 0; d < b->dim; ++d)
          r[d] = ((double)idx[d] - 0.5 * (double)b->N[d]) * b->L[d] / (double)b->N[d];
        const uint64_t i = (iz * b->N[1] + iy) * b->N[0] + ix;
        const double v = V(b->dim, r, param);
        b->ltK[i] += v;
        b->ltM[i] += v;
      }
  b->state |= BDG_HAS_TRAP;
}

/* reference src/setup.c:89-114 (copies) */
void bdg_set_wavefunction(bdg_t *b, const void *wf, uint64_t wf_size) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_set_wavefunction");
  if (!wf) BDG_FAIL("bdg_set_wavefunction: wf is NULL");
  if (wf_size != b->size)
    BDG_FAIL("bdg_set_wavefunction: wf_size (%" PRIu64 ") != grid size (%" PRIu64 ")", wf_size,
             b->size);
  freep(&b->wf);
  const size_t bytes = b->size * sizeof(double) * (b->complex_psi0 ? 2 : 1);
  b->wf = xcalloc(1, bytes);
  memcpy(b->wf, wf, bytes);
  b->state |= BDG_HAS_WF;
}

static double density_at(const bdg_t *b, uint64_t i) {
  if (b->complex_psi0) {
    const double *w = (const double *)b->wf;
    return w[2 * i] * w[2 * i] + w[2 * i + 1] * w[2 * i + 1];
  }
  const double p = ((const double *)b->wf)[i];
  return p * p;
}

/* reference src/setup.c:119-145: K += U_K(n), M += U_K(n) + 2 U_M(n) */
void bdg_set_local_interactions(bdg_t *b, f64 (*UK)(const void *, f64),
                                f64 (*UM)(const void *, f64), const void *param) {
  BDG_REQUIRE(b, BDG_HAS_WF, "bdg_set_local_interactions");
  BDG_FORBID(b, BDG_HAS_INTERACTIONS, "bdg_set_local_interactions");
  for (uint64_t i = 0; i < b->size; ++i) {
    const double n = density_at(b, i);
    const double uK = UK(param, n), uM = UM(param, n);
    b->ltK[i] += uK;
    b->ltM[i] += uK + 2.0 * uM;
  }
  b->state |= BDG_HAS_INTERACTIONS;
}

/* reference src/setup.c:150-177: p = 1/sqrt(max(1e-8, lt)) before lt -= mu */
void bdg_set_mu(bdg_t *b, f64 mu) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_set_mu");
  if (!b->pK) b->pK = (double *)xcalloc(b->size, sizeof(double));
  if (!b->pM) b->pM = (double *)xcalloc(b->size, sizeof(double));
  for (uint64_t i = 0; i < b->size; ++i) {
    b->pK[i] = 1.0 / sqrt(fmax(1e-8, b->ltK[i]));
    b->pM[i] = 1.0 / sqrt(fmax(1e-8, b->ltM[i]));
  }
  for (uint64_t i = 0; i < b->size; ++i) {
    b->ltK[i] -= mu;
    b->ltM[i] -= mu;
  }
  b->mu = mu;
  b->state |= BDG_HAS_MU;
}

/* reference src/setup.c:182-197 + src/dipolar.c: kernel and mean field, here
 * computed on the device */
void bdg_set_dipolar(bdg_t *b, f64 g_ddi, const f64 *dir, f64 cutoff) {
  BDG_REQUIRE(b, BDG_HAS_WF, "bdg_set_dipolar");
  BDG_FORBID(b, BDG_HAS_MU, "bdg_set_dipolar");
  if (3 != b->dim)
    BDG_FAIL("bdg_set_dipolar: dipolar interactions require dim == 3 (got %" PRIu64 ")", b->dim);
  int rc = ensure_device(b);
  if (!rc) rc = bdgx_set_fields(b->dev, NULL, NULL, b->wf, NULL, NULL, 0.0);
  if (!rc) rc = bdgx_dipolar_set_kernel(b->dev, g_ddi, dir, cutoff);
  double *gphi = (double *)xcalloc(b->size, sizeof(double));
  if (!rc) rc = bdgx_dipolar_meanfield(b->dev, gphi);
  if (rc) BDG_FAIL("bdg_set_dipolar: device error: %s", bdgx_error_string(rc));
  for (uint64_t i = 0; i < b->size; ++i) {
    b->ltK[i] += gphi[i];
    b->ltM[i] += gphi[i];
  }
  free(gphi);
  b->g_ddi = g_ddi;
  b->dipolar = 1;
  b->state |= BDG_HAS_DIPOLAR;
}

void bdg_set_solver_params(bdg_t *b, uint64_t nev, uint64_t sizeSub, uint64_t maxIter, f64 tol) {
  b->nev = nev;
  b->sizeSub = sizeSub;
  b->maxIter = maxIter;
  b->tol = tol;
}

void bdg_set_init_mode(bdg_t *b, bdg_init_mode_t mode, bdg_init_fn fn, void *param) {
  b->init_mode = mode;
  b->custom_init_fn = fn;
  b->custom_init_param = param;
}

/* reference src/bdg.c:112-207: modes + gaussian noise (seed 12345, Box-Muller);
 * extra columns |gaussian| * |psi0| */
int bdg_reuse_modes(bdg_t *b, f64 noise_frac) {
  if (!b->modes_u || !b->modes_v) return -3;
  const uint64_t S = b->size, n = 2 * S, nev = b->nev, m = b->sizeSub;
  const int c = b->complex_psi0;
  freep(&b->reuse_buf);
  free_device_reuse(b);
  if (b->dev && b->d_modes_u && b->d_modes_v) {
    /* device-resident sweep: the block is built next to the modes and
     * stays on the device for the next solve */
    const size_t es = sizeof(double) * (c ? 2 : 1);
    int rc = bdgx_malloc(&b->d_reuse_u, S * m * es);
    if (!rc) rc = bdgx_malloc(&b->d_reuse_v, S * m * es);
    if (!rc)
      rc = bdgx_reuse_block(b->dev, b->d_modes_u, b->d_modes_v, S, nev < m ? nev : m,
                            b->d_reuse_u, b->d_reuse_v, S, m, S, noise_frac, 12345u);
    if (rc) BDG_FAIL("bdg_reuse_modes: device reuse block failed: %s", bdgx_error_string(rc));
    b->reuse_n = n;
    b->reuse_cols = m;
    b->init_mode = BDG_INIT_REUSE;
    return 0;
  }
  b->reuse_buf = xcalloc(n * m, sizeof(double) * (c ? 2 : 1));
  b->reuse_n = n;
  b->reuse_cols = m;
  uint32_t seed = 12345;
  double *buf = (double *)b->reuse_buf;
  const double *u = (const double *)b->modes_u, *v = (const double *)b->modes_v;
  for (uint64_t j = 0; j < nev && j < m; ++j) {
    double cn = 0.0;
    for (uint64_t i = 0; i < S; ++i) {
      if (c) {
        const double *uu = u + 2 * (j * S + i), *vv = v + 2 * (j * S + i);
        cn += uu[0] * uu[0] + uu[1] * uu[1] + vv[0] * vv[0] + vv[1] * vv[1];
      } else {
        cn += u[j * S + i] * u[j * S + i] + v[j * S + i] * v[j * S + i];
      }
    }
    const double sigma = noise_frac * sqrt(cn);
    for (uint64_t i = 0; i < S; ++i) {
      double u1 = bdg_xrand(&seed);
      const double u2 = bdg_xrand(&seed);
      u1 = fmax(u1, 1e-10);
      const double z0 = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      const double z1 = sqrt(-2.0 * log(u1)) * sin(2.0 * M_PI * u2);
      if (c) {
        buf[2 * (i + j * n)] = u[2 * (j * S + i)] + sigma * z0;
        buf[2 * (i + j * n) + 1] = u[2 * (j * S + i) + 1];
        buf[2 * (i + S + j * n)] = v[2 * (j * S + i)] + sigma * z1;
        buf[2 * (i + S + j * n) + 1] = v[2 * (j * S + i) + 1];
      } else {
        buf[i + j * n] = u[j * S + i] + sigma * z0;
        buf[i + S + j * n] = v[j * S + i] + sigma * z1;
      }
    }
  }
  for (uint64_t j = nev; j < m; ++j)
    for (uint64_t i = 0; i < S; ++i) {
      double u1 = bdg_xrand(&seed);
      const double u2 = bdg_xrand(&seed);
      u1 = fmax(u1, 1e-10);
      const double wf = b->wf ? sqrt(density_at(b, i)) : 1.0;
      const double val = fabs(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)) * wf;
      if (c) {
        buf[2 * (i + j * n)] = val;
        buf[2 * (i + j * n) + 1] = 0.0;
        buf[2 * (i + S + j * n)] = val;
        buf[2 * (i + S + j * n) + 1] = 0.0;
      } else {
        buf[i + j * n] = val;
        buf[i + S + j * n] = val;
      }
    }
  b->init_mode = BDG_INIT_REUSE;
  return 0;
}

/* ---- start blocks (reference src/solver.c:79-259, 322-414) -------------- */

typedef struct {
  double k2;
  uint64_t idx[3];
} kvec_t;

static int kvec_cmp(const void *a, const void *b) {
  const double x = ((const kvec_t *)a)->k2, y = ((const kvec_t *)b)->k2;
  return (x > y) - (x < y);
}

static kvec_t *lowest_kvecs(const bdg_t *b, bdg_geom_hint_t hint, uint64_t need, uint64_t *nout) {
  uint64_t range[3] = {1, 1, 1};
  if (hint == BDG_GEOM_ELONGATED) {
    uint64_t ld = 0;
    for (uint64_t d = 1; d < b->dim; ++d)
      if (b->L[d] > b->L[ld]) ld = d;
    range[ld] = b->N[ld] < 2 * need ? b->N[ld] : 2 * need;
  } else if (hint == BDG_GEOM_RING) {
    for (uint64_t d = 0; d < b->dim && d < 2; ++d) range[d] = b->N[d] < 2 * need ? b->N[d] : 2 * need;
  } else {
    for (uint64_t d = 0; d < b->dim; ++d) range[d] = b->N[d] < 2 * need ? b->N[d] : 2 * need;
  }
  const uint64_t total = range[0] * range[1] * range[2];
  kvec_t *e = (kvec_t *)xcalloc(total, sizeof(kvec_t));
  uint64_t cnt = 0;
  for (uint64_t iz = 0; iz < range[2]; ++iz)
    for (uint64_t iy = 0; iy < range[1]; ++iy)
      for (uint64_t ix = 0; ix < range[0]; ++ix) {
        e[cnt].k2 = b->kx2[0][ix] + (b->dim > 1 ? b->kx2[1][iy] : 0.0) +
                    (b->dim > 2 ? b->kx2[2][iz] : 0.0);
        e[cnt].idx[0] = ix;
        e[cnt].idx[1] = iy;
        e[cnt].idx[2] = iz;
        ++cnt;
      }
  qsort(e, cnt, sizeof(kvec_t), kvec_cmp);
  *nout = cnt < need ? cnt : need;
  return e;
}

static double wf_abs(const bdg_t *b, uint64_t i) { return b->wf ? sqrt(density_at(b, i)) : 1.0; }

static void put(const bdg_t *b, void *X, uint64_t row, uint64_t col, double val) {
  const uint64_t n = 2 * b->size;
  if (b->complex_psi0) {
    double *x = (double *)X;
    x[2 * (row + col * n)] = val;
    x[2 * (row + col * n) + 1] = 0.0;
  } else {
    ((double *)X)[row + col * n] = val;
  }
}

void bdg_fill_start_block(bdg_t *b, void *X) {
  const uint64_t S = b->size, n = 2 * S, m = b->sizeSub;
  const size_t es = sizeof(double) * (b->complex_psi0 ? 2 : 1);
  switch (b->init_mode) {
    case BDG_INIT_REUSE:
      if (b->reuse_buf && b->reuse_n == n && b->reuse_cols == m) {
        memcpy(X, b->reuse_buf, n * m * es);
        freep(&b->reuse_buf);
        b->init_mode = BDG_INIT_WF_WEIGHTED;
        return;
      }
      /* fall through */
    case BDG_INIT_WF_WEIGHTED: {
      /* |gaussian| * |psi0|, u = v.  Both variants use xorshift32 seed 42 here
       * (the reference's real variant seeds rand_r with time(NULL),
       * src/solver.c:178-182, which cannot be reproduced). */
      uint32_t seed = 42;
      for (uint64_t j = 0; j < m; ++j)
        for (uint64_t i = 0; i < S; ++i) {
          double u1 = bdg_xrand(&seed);
          const double u2 = bdg_xrand(&seed);
          u1 = fmax(u1, 1e-10);
          const double val = fabs(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)) * wf_abs(b, i);
          put(b, X, i, j, val);
          put(b, X, i + S, j, val);
        }
      return;
    }
    case BDG_INIT_CUSTOM:
      if (b->custom_init_fn) b->custom_init_fn(b, X, n, m, b->custom_init_param);
      return;
    case BDG_INIT_PLANEWAVE:
    default: {
      uint32_t seed = 42;
      const bdg_geom_hint_t hint =
          b->custom_init_param ? (bdg_geom_hint_t)(intptr_t)b->custom_init_param : BDG_GEOM_AUTO;
      const uint64_t need = (m + 1) / 2;
      uint64_t nk = 0;
      kvec_t *kv = lowest_kvecs(b, hint, need, &nk);
      for (uint64_t j = 0; j < m; ++j) {
        const uint64_t kidx = (j + 1) / 2;
        const int use_sin = (j % 2) == 1;
        if (kidx >= nk) {
          for (uint64_t i = 0; i < S; ++i) {
            const double val = (bdg_xrand(&seed) - 0.5) * wf_abs(b, i);
            put(b, X, i, j, val);
            put(b, X, i + S, j, val);
          }
          continue;
        }
        double freq[3] = {0, 0, 0};
        for (uint64_t d = 0; d < b->dim; ++d) {
          const int64_t f = kv[kidx].idx[d] <= b->N[d] / 2 ? (int64_t)kv[kidx].idx[d]
                                                           : (int64_t)kv[kidx].idx[d] - (int64_t)b->N[d];
          freq[d] = 2.0 * M_PI * (double)f / b->L[d];
        }
        for (uint64_t i = 0; i < S; ++i) {
          const uint64_t id[3] = {i % b->N[0], (i / b->N[0]) % b->N[1], i / (b->N[0] * b->N[1])};
          double kr = 0.0;
          for (uint64_t d = 0; d < b->dim; ++d) kr += freq[d] * (double)id[d] * b->L[d] / (double)b->N[d];
          const double pw = use_sin ? sin(kr) : cos(kr);
          const double pert = 1e-4 * (bdg_xrand(&seed) - 0.5);
          const double val = (pw + pert) * wf_abs(b, i);
          put(b, X, i, j, val);
          put(b, X, i + S, j, val);
        }
      }
      free(kv);
      return;
    }
  }
}

void bdg_start_block(bdg_t *b, void *X) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_start_block");
  const bdg_init_mode_t mode = b->init_mode;
  void *keep = b->reuse_buf;
  if (mode == BDG_INIT_REUSE && b->d_reuse_u && b->reuse_n == 2 * b->size &&
      b->reuse_cols == b->sizeSub) {
    /* the device-resident block, column by column into [u; v] */
    const size_t es = sizeof(double) * (b->complex_psi0 ? 2 : 1);
    const uint64_t S = b->size;
    int rc = 0;
    for (uint64_t j = 0; j < b->sizeSub && !rc; ++j) {
      rc = bdgx_memcpy_d2h(b->dev, (char *)X + 2 * j * S * es,
                           (const char *)b->d_reuse_u + j * S * es, S * es);
      if (!rc)
        rc = bdgx_memcpy_d2h(b->dev, (char *)X + (2 * j + 1) * S * es,
                             (const char *)b->d_reuse_v + j * S * es, S * es);
    }
    if (rc) BDG_FAIL("bdg_start_block: download failed: %s", bdgx_error_string(rc));
    return;
  }
  if (mode == BDG_INIT_REUSE && keep) {
    /* copy without consuming the buffer */
    const size_t es = sizeof(double) * (b->complex_psi0 ? 2 : 1);
    if (b->reuse_n == 2 * b->size && b->reuse_cols == b->sizeSub) {
      memcpy(X, keep, 2 * b->size * b->sizeSub * es);
      return;
    }
  }
  bdg_fill_start_block(b, X);
  b->init_mode = mode;
}

/* ---- solve ---------------------------------------------------------------- */

/* The solver's preconditioner weights: the reference's
 * p = 1/sqrt(max(1e-8, lt)) (src/setup.c:165-166) with the pre-mu local term
 * lt clamped below at mu instead of 1e-8.  Where an attractive dipolar mean
 * field drives lt to zero or below, the reference weights reach 1e4 and every
 * preconditioned residual collapses onto those grid points; inside the
 * condensate lt ~ mu, so the clamp only changes those points.  The operator
 * API (bdgx_apply PRECOND_*) keeps the reference weights. */
static int upload_fields(bdg_t *b, int learned) {
  const int drc = ensure_device(b);
  if (drc) return drc;
  /* the preconditioner weights 1 / sqrt(max(floor, lt + mu)) derived on
   * the device from the uploaded local terms */
  const double floor_v = b->mu > 1e-8 ? b->mu : 1e-8;
  int rc = bdgx_set_fields(b->dev, b->ltK, b->ltM, b->wf, NULL, NULL, b->mu);
  if (!rc) rc = bdgx_derive_precond(b->dev, floor_v);
  /* the network's tables only when the solve will use it (real psi0 and
   * K != M); otherwise the analytic T_M T_K runs and theta is ignored */
  if (!rc && learned) rc = bdgx_lp_set_params(b->dev, b->lp_theta);
  return rc;
}

/* K == M: no dipolar term and identical local terms (the solver then runs
 * the standard form) */
static int k_equals_m(const bdg_t *b) {
  return !b->dipolar && memcmp(b->ltK, b->ltM, sizeof(double) * b->size) == 0;
}

/* the learned network replaces T_M T_K only for real psi0 with K != M */
static int learned_applies(const bdg_t *b) {
  return b->lp_theta != NULL && !b->complex_psi0 && (b->state & BDG_HAS_WF) && !k_equals_m(b);
}

/* bdg_solve on the first-order pencil (lrep.c): one FP64 phase, or (mixed)
 * an FP32 phase followed by an FP64 phase from its vectors.  dV0 is the v
 * half of the start block (u = v for the built-in blocks; the reference's
 * start blocks set u = v too, src/solver.c:166-259). */
/* frees a start block pair (u and v may share one buffer) */
static void free_start(void **pU0, void **pV0) {
  if (*pU0 && *pU0 != *pV0) bdgx_free(*pU0);
  bdgx_free(*pV0);
  *pU0 = *pV0 = NULL;
}

/* Device bytes of one phase's working set: per half the basis [X | P | W]
 * and its images, double-buffered (4 S x 3m), and the residual and
 * new-direction blocks (2 S x m) -- lrep.c's allocations. */
static double phase_bytes(uint64_t S, uint64_t m, size_t es) {
  return 2.0 * (4.0 * 3.0 * (double)m + 2.0 * (double)m) * (double)S * (double)es;
}

/* The FP64 phase in groups, where the block's FP64 basis does not fit the
 * device: modes g0 .. g0+G-1 plus guard columns start from the FP32 phase's
 * vectors (columns g0 .. of dU / dV, refined in place) and are iterated in a
 * block deflated against the groups already refined (columns 0 .. g0-1,
 * locked), as the solver's own locking does with converged pairs. */
static int fp64_groups(bdg_t *b, const lrep_cfg *base, int G, void *dU, void *dV, double *w2,
                       double *resv, lrep_result *acc) {
  const uint64_t S = b->size;
  const int m = (int)b->sizeSub, nev = (int)b->nev;
  const size_t es = sizeof(double) * (b->complex_psi0 ? 2 : 1);
  int rc = 0;
  for (int g0 = 0; g0 < nev && !rc; g0 += G) {
    const int gn = nev - g0 < G ? nev - g0 : G;
    int ng = base->lock_guards > 0 ? base->lock_guards : 0;
    if (g0 + gn + ng > m) ng = m - g0 - gn;
    const int mb = gn + ng;
    lrep_cfg cfg = *base;
    cfg.m = mb;
    cfg.nev = gn;
    cfg.lock_max = 0;
    cfg.lock_guards = 0;
    cfg.nl = g0;
    cfg.dUL = g0 ? dU : NULL;
    cfg.dVL = g0 ? dV : NULL;
    cfg.maxiter = base->maxiter - acc->iters;
    if (cfg.maxiter < 1) cfg.maxiter = 1;
    lrep_result r;
    memset(&r, 0, sizeof(r));
    double *gw2 = (double *)xcalloc((size_t)mb, sizeof(double));
    double *gres = (double *)xcalloc((size_t)mb, sizeof(double));
    r.w2 = gw2;
    r.res = gres;
    void *u0 = (char *)dU + (size_t)g0 * S * es, *v0 = (char *)dV + (size_t)g0 * S * es;
    rc = bdg_lrep_device(b->dev, &cfg, BDGX_FP64, u0, v0, BDGX_FP64, &r, u0, v0);
    for (int j = 0; j < gn; ++j) {
      w2[g0 + j] = gw2[j];
      resv[g0 + j] = gres[j];
    }
    acc->iters += r.iters;
    acc->applies += r.applies;
    acc->rr_seconds += r.rr_seconds;
    acc->worst = fmax(acc->worst, r.worst);
    if (b->verbosity > 0)
      fprintf(stderr, "[bdg] fp64 group %d..%d (+%d guards, %d locked): %d iterations, worst "
              "residual %.3e, status %d\n", g0, g0 + gn - 1, ng, g0, r.iters, r.worst, rc);
    free(gw2);
    free(gres);
  }
  return rc;
}

static int solve_first_order(bdg_t *b, void **pU0, void **pV0, int learned, double t_start) {
  const uint64_t S = b->size, m = b->sizeSub, nev = b->nev;
  const int c = b->complex_psi0;
  const size_t es64 = sizeof(double) * (c ? 2 : 1), es32 = sizeof(float) * (c ? 2 : 1);
  const int mixed = b->precision == BDG_PREC_MIXED;
  void *dU = NULL, *dV = NULL;
  int rc = 0;
  lrep_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.S = S;
  cfg.m = (int)m;
  cfg.nev = (int)nev;
  cfg.cplx = c;
  cfg.verbosity = b->verbosity;
  /* automatic locking: once at most max(2, m / 4) columns are active (at
   * 128^3 with m = 32 the FP64 solve took 128 iterations, 0.96 s, against
   * 172, 1.08 s when locking at m / 8: the small deflated block converges
   * the last pairs in fewer, cheaper iterations) */
  cfg.lock_max = b->lock_max >= 0 ? b->lock_max : (int)(m / 4 > 2 ? m / 4 : 2);
  cfg.lock_guards = b->lock_guards;
  /* the learned spectral layers replace precondK / precondM in the FP32
   * phase (mixed precision only; the FP64 phase keeps the analytic ones) */
  cfg.learned = learned;
  cfg.lp_window = b->lp_window;
  cfg.lp_factor = b->lp_factor;
  lrep_result r32, r64;
  memset(&r32, 0, sizeof(r32));
  memset(&r64, 0, sizeof(r64));
  double *w2 = (double *)xcalloc(m, sizeof(double)), *resv = (double *)xcalloc(m, sizeof(double));
  r32.w2 = r64.w2 = w2;
  r32.res = r64.res = resv;
  int it32 = 0;
  /* the host arrays of the modes, page-locked on a helper thread while the
   * device iterates (their download then runs at the link rate: pageable
   * memory took ~0.15 s for the 1 GB of 32 modes at 128^3) */
  prefault_job pj;
  memset(&pj, 0, sizeof(pj));
  pj.pf.bytes = S * nev * es64;
  cfg.on_iterating = prefault_start;
  cfg.on_iterating_arg = &pj;
  if (mixed) {
    /* FP32 phase on FP32 copies of the start halves, which then receive its
     * vectors; the FP64 start block is released first, so that at the
     * largest grids the phase has the memory to itself */
    void *sU = NULL, *sV = NULL;
    rc = bdgx_malloc(&sU, S * m * es32);
    if (!rc) rc = bdgx_malloc(&sV, S * m * es32);
    if (!rc) rc = bdgx_convert(b->dev, BDGX_FP64, *pU0, S, BDGX_FP32, sU, S, S, m);
    if (!rc) rc = bdgx_convert(b->dev, BDGX_FP64, *pV0, S, BDGX_FP32, sV, S, S, m);
    if (!rc) rc = bdgx_sync(b->dev);
    free_start(pU0, pV0);
    cfg.maxiter = (int)b->maxIter;
    cfg.tol = fmax(b->tol, b->fp32_tol);
    cfg.stall_window = b->fp32_stall;
    cfg.stall_below = 1e-4;
    if (!rc) rc = bdg_lrep_device(b->dev, &cfg, BDGX_FP32, sU, sV, BDGX_FP32, &r32, sU, sV);
    it32 = r32.iters;
    b->stats.iterations_fp32 = (uint64_t)it32;
    b->stats.fp32_worst_residual = r32.worst;
    b->stats.column_applies_fp32 = r32.applies;
    b->stats.learned_iterations = (uint64_t)r32.learned_its;
    b->stats.learned_fallback_at = r32.learned_fallback_at;
    if (b->verbosity > 0)
      fprintf(stderr, "[bdg] fp32 phase: %d iterations, worst residual %.3e, status %d\n", it32,
              r32.worst, rc);
    /* the FP64 phase starts from the FP32 vectors */
    if (!rc) rc = bdgx_malloc(&dU, S * m * es64);
    if (!rc) rc = bdgx_malloc(&dV, S * m * es64);
    if (!rc) rc = bdgx_convert(b->dev, BDGX_FP32, sU, S, BDGX_FP64, dU, S, S, m);
    if (!rc) rc = bdgx_convert(b->dev, BDGX_FP32, sV, S, BDGX_FP64, dV, S, S, m);
    if (!rc) rc = bdgx_sync(b->dev);
    bdgx_free(sU);
    bdgx_free(sV);
  } else {
    rc = bdgx_malloc(&dU, S * m * es64);
    if (!rc) rc = bdgx_malloc(&dV, S * m * es64);
  }
  if (!rc) {
    cfg.maxiter = (int)b->maxIter - it32;
    if (cfg.maxiter < 0) cfg.maxiter = 0;
    cfg.tol = b->tol;
    cfg.stall_window = 0;
    /* where the FP64 basis does not fit the device (256^3 at block 128:
     * ~480 GB), the FP64 phase of the mixed policy iterates the modes in
     * groups (the largest of m/2, m/4, ... whose basis fits), each deflated
     * against the groups before it; where not even that fits, it is the
     * refinement sweep alone (Rayleigh-Ritz on the FP64 images of the FP32
     * vectors and the FP64 residuals, no iterations) */
    size_t fr = 0, tot = 0;
    int group = 0;
    if (mixed && b->sweep_only) {
      cfg.maxiter = 0;
      b->stats.fp64_sweep_only = 1;
    } else if (mixed && b->fp64_groups > 0 && (uint64_t)b->fp64_groups < nev) {
      group = b->fp64_groups;
      b->stats.fp64_group_size = group;
    } else if (cfg.maxiter > 0 && bdgx_mem_info(&fr, &tot) == 0 &&
               phase_bytes(S, m, es64) + 2e9 > (double)fr) {
      if (!mixed) {
        if (b->verbosity > 0)
          fprintf(stderr, "[bdg] fp64 solve needs %.1f GB, %.1f GB free\n",
                  phase_bytes(S, m, es64) / 1e9, (double)fr / 1e9);
        rc = BDGX_ENOMEM;
      } else {
        const int ng = b->lock_guards > 0 ? b->lock_guards : 0;
        int G = (int)m / 2;
        /* 10% and 8 GB of headroom for the FP64 spectrum workspace, the
         * Gram partials and the allocator's granularity */
        while (G >= 8 && 1.1 * phase_bytes(S, (uint64_t)(G + ng), es64) + 8e9 > (double)fr) G /= 2;
        if (G >= 8 && (uint64_t)G < nev) {
          group = G;
          b->stats.fp64_group_size = G;
        } else {
          cfg.maxiter = 0;
          b->stats.fp64_sweep_only = 1;
        }
        if (b->verbosity > 0)
          fprintf(stderr, "[bdg] fp64 basis needs %.1f GB, %.1f GB free: %s %d\n",
                  phase_bytes(S, m, es64) / 1e9, (double)fr / 1e9,
                  group ? "groups of" : "refinement sweep only", group);
      }
    }
    if (!rc && group) rc = fp64_groups(b, &cfg, group, dU, dV, w2, resv, &r64);
  }
  if (!rc && !b->stats.fp64_group_size)
    rc = mixed ? bdg_lrep_device(b->dev, &cfg, BDGX_FP64, dU, dV, BDGX_FP64, &r64, dU, dV)
               : bdg_lrep_device(b->dev, &cfg, BDGX_FP64, *pU0, *pV0, BDGX_FP64, &r64, dU, dV);
  free_start(pU0, pV0);
  b->stats.iterations = (uint64_t)(it32 + r64.iters);
  b->stats.column_applies = r32.applies + r64.applies;
  b->stats.seconds_rr = r32.rr_seconds + r64.rr_seconds;
  if (b->verbosity > 0)
    fprintf(stderr, "[bdg] %s first-order solve: %d + %d iterations, %.3f s, status %d\n",
            b->precision == BDG_PREC_MIXED ? "mixed" : "fp64", it32, r64.iters,
            bdg_wtime() - t_start, rc);
  /* results: w = signed sqrt(w2); [u; v] scaled to unit norm on the device */
  const double t_iter = bdg_wtime();
  double *dzero = NULL, *dn2 = NULL, *dcu = NULL;
  double *hn2 = (double *)xcalloc(3 * nev, sizeof(double));
  double *cz = (double *)xcalloc(nev * nev * (c ? 2 : 1), sizeof(double));
  void *dOutU = NULL, *dOutV = NULL;
  if (!rc) rc = bdgx_malloc((void **)&dzero, sizeof(double) * nev);
  if (!rc) rc = bdgx_malloc((void **)&dn2, sizeof(double) * 3 * nev);
  if (!rc) rc = bdgx_memset(b->dev, dzero, 0, sizeof(double) * nev);
  if (!rc) rc = bdgx_residual(b->dev, BDGX_FP64, dU, dV, NULL, S, nev, S, dzero, 0, NULL, S, dn2);
  if (!rc) rc = bdgx_memcpy_d2h(b->dev, hn2, dn2, sizeof(double) * 3 * nev);
  if (!rc) {
    b->eigvals = (double *)xcalloc(nev, sizeof(double));
    b->residuals = (double *)xcalloc(nev, sizeof(double));
    b->zero_modes = (int *)xcalloc(nev, sizeof(int));
    double scale = 1.0;
    for (uint64_t j = 0; j < nev; ++j) {
      b->eigvals[j] = w2[j] >= 0 ? sqrt(w2[j]) : -sqrt(-w2[j]);
      scale = fmax(scale, fabs(b->eigvals[j]));
    }
    uint64_t conv = 0;
    for (uint64_t j = 0; j < nev; ++j) {
      b->residuals[j] = resv[j];
      b->zero_modes[j] = fabs(b->eigvals[j]) < 1e-3 * scale;
      if (resv[j] < b->tol) ++conv;
      const double nrm2 = hn2[3 * j] + hn2[3 * j + 1];
      cz[(j + j * nev) * (c ? 2 : 1)] = nrm2 > 0 ? 1.0 / sqrt(nrm2) : 1.0;
    }
    b->converged = conv;
    rc = bdgx_malloc((void **)&dcu, sizeof(double) * nev * nev * (c ? 2 : 1));
    if (!rc) rc = bdgx_memcpy_h2d(b->dev, dcu, cz, sizeof(double) * nev * nev * (c ? 2 : 1));
  }
  if (!rc) rc = bdgx_malloc(&dOutU, S * nev * es64);
  if (!rc) rc = bdgx_malloc(&dOutV, S * nev * es64);
  if (!rc) rc = bdgx_combine(b->dev, BDGX_FP64, dU, S, nev, dcu, dOutU, S, nev, S);
  if (!rc) rc = bdgx_combine(b->dev, BDGX_FP64, dV, S, nev, dcu, dOutV, S, nev, S);
  if (pj.on == 1) pthread_join(pj.th, NULL);
  if (!rc) {
    if (pj.pf.u && pj.pf.v) {
      b->modes_u = pj.pf.u;
      b->modes_v = pj.pf.v;
      b->modes_pinned = pj.pf.pinned;
      b->modes_bytes = pj.pf.bytes;
      pj.pf.u = pj.pf.v = NULL;
    } else {
      b->modes_u = xmalloc_big(S * nev * es64);
      b->modes_v = xmalloc_big(S * nev * es64);
    }
    rc = bdgx_memcpy_d2h_pair(b->dev, b->modes_u, dOutU, b->modes_v, dOutV, S * nev * es64);
  }
  release_prefault(&pj.pf);
  bdgx_free(dzero);
  bdgx_free(dn2);
  bdgx_free(dcu);
  if (!rc && b->device_sweeps) {  /* the modes stay on the device for a sweep's reuse block */
    b->d_modes_u = dOutU;
    b->d_modes_v = dOutV;
  } else {
    bdgx_free(dOutU);
    bdgx_free(dOutV);
  }
  bdgx_free(dU);
  bdgx_free(dV);
  free(hn2);
  free(cz);
  free(w2);
  free(resv);
  b->stats.status = rc;
  b->stats.seconds_total = bdg_wtime() - t_start;
  if (b->verbosity > 0)
    fprintf(stderr, "[bdg] results: %.1f ms (host arrays %s)\n", 1e3 * (bdg_wtime() - t_iter),
            b->modes_pinned ? "page-locked" : "pageable");
  return rc;
}

int bdg_solve(bdg_t *b) {
  BDG_REQUIRE(b, BDG_HAS_SYSTEM, "bdg_solve");
  BDG_REQUIRE(b, BDG_HAS_MU, "bdg_solve");
  const double t_start = bdg_wtime();
  free_results(b);
  memset(&b->stats, 0, sizeof(b->stats));
  const uint64_t S = b->size, n = 2 * S, m = b->sizeSub;
  uint64_t nev = b->nev;
  if (m < nev || nev == 0) {
    b->stats.status = BDGX_EINVAL;
    return BDGX_EINVAL;
  }
  const int c = b->complex_psi0;
  const size_t es64 = sizeof(double) * (c ? 2 : 1);
  const int learned = learned_applies(b);
  int rc = upload_fields(b, learned);
  if (rc) {
    b->stats.status = rc;
    return rc;
  }
  const double ts_fields = bdg_wtime();
  void *dV0 = NULL, *dU0 = NULL; /* dU0: u half where it differs from v (reuse) */
  double ts_block = ts_fields;
  if (b->init_mode == BDG_INIT_REUSE && b->d_reuse_u && b->d_reuse_v && b->reuse_n == n &&
      b->reuse_cols == m) {
    /* device-resident sweep: the reuse block built on the device by
     * bdg_reuse_modes, consumed here (as the host block is) */
    dU0 = b->d_reuse_u;
    dV0 = b->d_reuse_v;
    b->d_reuse_u = b->d_reuse_v = NULL;
    b->init_mode = BDG_INIT_WF_WEIGHTED;
  } else {
    free_device_reuse(b);
    rc = bdgx_malloc(&dV0, S * m * es64);
  }
  if (dU0) {
    /* start block already on the device */
  } else if (!rc && b->init_mode == BDG_INIT_CUSTOM && b->custom_init_fn == bdg_init_uniform_xorshift) {
    /* the built-in uniform block (u = v): its v half generated on the
     * device, the same values as the host fill */
    const uint32_t seed = b->custom_init_param ? *(const uint32_t *)b->custom_init_param : 42u;
    rc = bdgx_fill_uniform_xorshift(b->dev, dV0, S, S, m, S, seed);
    ts_block = bdg_wtime();
  } else if (!rc) {
    /* start block [u; v] on the host (zero-filled anonymous mapping, huge
     * pages), v half to the device */
    const size_t xbytes = n * m * es64;
    void *X = mmap(NULL, xbytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (X == MAP_FAILED) BDG_FAIL("out of host memory (%zu bytes)", xbytes);
    (void)madvise(X, xbytes, MADV_HUGEPAGE);
    /* a reuse block has u != v: both halves go up */
    const int two_halves = b->init_mode == BDG_INIT_REUSE && b->reuse_buf != NULL;
    bdg_fill_start_block(b, X);
    ts_block = bdg_wtime();
    if (two_halves) rc = bdgx_malloc(&dU0, S * m * es64);
    for (uint64_t j = 0; j < m && !rc; ++j) {
      rc = bdgx_memcpy_h2d(b->dev, (char *)dV0 + j * S * es64,
                           (const char *)X + (j * n + S) * es64, S * es64);
      if (!rc && dU0)
        rc = bdgx_memcpy_h2d(b->dev, (char *)dU0 + j * S * es64, (const char *)X + j * n * es64,
                             S * es64);
    }
    munmap(X, xbytes);
  }
  if (rc) {
    bdgx_free(dV0);
    bdgx_free(dU0);
    b->stats.status = rc;
    return rc;
  }
  if (b->formulation == BDG_FORM_FIRST_ORDER && !k_equals_m(b)) {
    /* the start halves (u = v unless a reuse block) pass to the solver,
     * which frees them as soon as it can */
    if (!dU0) dU0 = dV0;
    if (b->verbosity > 0)
      fprintf(stderr, "[bdg] setup: %.1f ms (fields %.1f, start block %.1f, upload %.1f)\n",
              1e3 * (bdg_wtime() - t_start), 1e3 * (ts_fields - t_start),
              1e3 * (ts_block - ts_fields), 1e3 * (bdg_wtime() - ts_block));
    return solve_first_order(b, &dU0, &dV0, learned, t_start);
  }
  /* the product form iterates on the v half only */
  bdgx_free(dU0);
  dU0 = NULL;
  lobpcg_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.S = S;
  cfg.m = (int)m;
  cfg.nev = (int)nev;
  cfg.cplx = c;
  cfg.tol = b->tol;
  cfg.zero_tol = 1e-3;
  cfg.verbosity = b->verbosity;
  cfg.gram_refresh = 25;
  cfg.maxiter = (int)b->maxIter;
  cfg.w_prec = b->precision == BDG_PREC_MIXED ? BDGX_FP32 : BDGX_FP64;
  cfg.img_tol = b->img_tol;
  cfg.refresh_every = (int)b->refresh_every;
  cfg.mixed_min_cols = (int)(m / 4 > 1 ? m / 4 : 1);
  cfg.k_equals_m = k_equals_m(b);
  cfg.learned = learned;
  cfg.lp_window = b->lp_window;
  cfg.lp_factor = b->lp_factor;
  cfg.graphs = b->graphs;
  cfg.freeze_tol = b->freeze_tol;
  lobpcg_result r;
  memset(&r, 0, sizeof(r));
  r.w2 = (double *)xcalloc(m, sizeof(double));
  r.zflag = (int *)xcalloc(m, sizeof(int));
  r.res = (double *)xcalloc(m, sizeof(double));
  void *dX = NULL, *dMX = NULL;
  result_prefault pf = {S * nev * es64, NULL, NULL, 0};
  pf.u = pin_take(pf.bytes);
  pf.v = pin_take(pf.bytes);
  pthread_t pf_thread;
  int pf_on = 0;
  if (pf.u && pf.v)
    pf.pinned = 1;
  else
    pf_on = pthread_create(&pf_thread, NULL, prefault_main, &pf) == 0;
  if (b->verbosity > 0)
    fprintf(stderr, "[bdg] setup: %.1f ms (fields %.1f, start block %.1f, upload %.1f)\n",
            1e3 * (bdg_wtime() - t_start), 1e3 * (ts_fields - t_start),
            1e3 * (ts_block - ts_fields), 1e3 * (bdg_wtime() - ts_block));
  rc = bdg_lobpcg_device(b->dev, &cfg, dV0, &r, &dX);
  b->stats.iterations = (uint64_t)r.iters;
  b->stats.column_applies = r.applies;
  b->stats.column_applies_fp32 = r.applies_fp32;
  b->stats.image_refreshes = r.refreshes;
  b->stats.learned_iterations = (uint64_t)r.learned_its;
  b->stats.learned_fallback_at = r.learned_fallback_at;
  b->stats.graph_launches = r.graph_launches;
  b->stats.frozen_skips = r.frozen_skips;
  b->stats.graph_captures = r.graph_captures;
  b->stats.restarts = (uint64_t)r.restarts;
  b->stats.seconds_rr = r.rr_seconds;
  if (b->verbosity > 0)
    fprintf(stderr,
            "[bdg] %s solve: %d iterations, %llu column applies (%llu fp32), %llu refreshes, "
            "%.3f s (Rayleigh-Ritz %.3f s, graph capture %.3f s), status %d\n",
            cfg.w_prec == BDGX_FP32 ? "mixed" : "fp64", r.iters, (unsigned long long)r.applies,
            (unsigned long long)r.applies_fp32, (unsigned long long)r.refreshes,
            bdg_wtime() - t_start, r.rr_seconds, r.capture_seconds, rc);
  bdgx_free(dV0);
  /* result extraction, timed by phase when verbose */
  const int tv = b->verbosity > 0;
  double tx[5] = {bdg_wtime(), 0, 0, 0, 0};
  /* M X for the u half */
  if (!rc) {
    rc = bdgx_malloc(&dMX, S * m * es64);
    if (!rc) rc = bdgx_apply(b->dev, BDGX_OP_M, BDGX_FP64, dX, dMX, m, S);
  }
  if (tv && !rc) rc = bdgx_sync(b->dev);
  tx[1] = bdg_wtime();
  if (pf_on) pthread_join(pf_thread, NULL);
  if (rc) {
    release_prefault(&pf);
    bdgx_free(dX);
    bdgx_free(dMX);
    free(r.w2);
    free(r.zflag);
    free(r.res);
    b->stats.status = rc;
    b->stats.seconds_total = bdg_wtime() - t_start;
    return rc;
  }
  /* results: w = signed sqrt(w2), v = X, u = M X / w, ||[u; v]|| = 1; the
   * columns are scaled on the device and copied straight into the result
   * arrays */
  b->eigvals = (double *)xcalloc(nev, sizeof(double));
  b->residuals = (double *)xcalloc(nev, sizeof(double));
  b->zero_modes = (int *)xcalloc(nev, sizeof(int));
  double scale = 1.0;
  for (uint64_t j = 0; j < nev; ++j) {
    const double w2 = r.w2[j];
    b->eigvals[j] = w2 >= 0 ? sqrt(w2) : -sqrt(-w2);
    scale = fmax(scale, fabs(b->eigvals[j]));
  }
  uint64_t conv = 0;
  for (uint64_t j = 0; j < nev; ++j) {
    b->residuals[j] = r.res[j];
    b->zero_modes[j] = r.zflag[j] || fabs(b->eigvals[j]) < 1e-3 * scale;
    /* the solver's own test: only M-null directions bypass the residual */
    if (r.zflag[j] || r.res[j] < b->tol) ++conv;
  }
  /* column norms ||X_j||^2 and ||M X_j||^2 (the residual kernel with w = 0) */
  double *dzero = NULL, *dn2 = NULL, *dcu = NULL, *dcv = NULL;
  void *dU = NULL;
  double *hn2 = (double *)xcalloc(3 * nev, sizeof(double));
  double *cu = (double *)xcalloc(nev * nev, sizeof(double));
  double *cv = (double *)xcalloc(nev * nev, sizeof(double));
  rc = bdgx_malloc((void **)&dzero, sizeof(double) * nev);
  if (!rc) rc = bdgx_malloc((void **)&dn2, sizeof(double) * 3 * nev);
  if (!rc) rc = bdgx_memset(b->dev, dzero, 0, sizeof(double) * nev);
  if (!rc)
    rc = bdgx_residual(b->dev, BDGX_FP64, dMX, dX, dMX, S, nev, S, dzero, 0, NULL, S, dn2);
  if (!rc) rc = bdgx_memcpy_d2h(b->dev, hn2, dn2, sizeof(double) * 3 * nev);
  tx[2] = bdg_wtime();
  if (!rc) {
    for (uint64_t j = 0; j < nev; ++j) {
      const double w = b->eigvals[j];
      const double inv = fabs(w) > 0 ? 1.0 / w : 1.0;
      const double nrm2 = hn2[3 * j + 2] * inv * inv + hn2[3 * j + 1];
      const double nrm = nrm2 > 0 ? 1.0 / sqrt(nrm2) : 1.0;
      cu[j + j * nev] = inv * nrm;
      cv[j + j * nev] = nrm;
    }
    rc = bdgx_malloc((void **)&dcu, sizeof(double) * nev * nev * (c ? 2 : 1));
  }
  if (!rc) rc = bdgx_malloc((void **)&dcv, sizeof(double) * nev * nev * (c ? 2 : 1));
  if (!rc && c) {
    /* complex combine coefficients: interleaved (re, im) */
    double *cz = (double *)xcalloc(2 * nev * nev, sizeof(double));
    for (uint64_t j = 0; j < nev * nev; ++j) cz[2 * j] = cu[j];
    rc = bdgx_memcpy_h2d(b->dev, dcu, cz, sizeof(double) * 2 * nev * nev);
    for (uint64_t j = 0; j < nev * nev; ++j) cz[2 * j] = cv[j];
    if (!rc) rc = bdgx_memcpy_h2d(b->dev, dcv, cz, sizeof(double) * 2 * nev * nev);
    free(cz);
  } else if (!rc) {
    rc = bdgx_memcpy_h2d(b->dev, dcu, cu, sizeof(double) * nev * nev);
    if (!rc) rc = bdgx_memcpy_h2d(b->dev, dcv, cv, sizeof(double) * nev * nev);
  }
  /* u = M X diag(inv nrm) into a fresh buffer, v = X diag(nrm) into M X */
  if (!rc) rc = bdgx_malloc(&dU, S * nev * es64);
  if (!rc) rc = bdgx_combine(b->dev, BDGX_FP64, dMX, S, nev, dcu, dU, S, nev, S);
  if (!rc) rc = bdgx_combine(b->dev, BDGX_FP64, dX, S, nev, dcv, dMX, S, nev, S);
  if (tv && !rc) rc = bdgx_sync(b->dev);
  tx[3] = bdg_wtime();
  if (!rc) {
    if (pf.u && pf.v) {
      b->modes_u = pf.u;
      b->modes_v = pf.v;
      b->modes_pinned = pf.pinned;
      b->modes_bytes = pf.bytes;
      pf.u = pf.v = NULL;
    } else {
      b->modes_u = xmalloc_big(S * nev * es64);
      b->modes_v = xmalloc_big(S * nev * es64);
    }
    rc = bdgx_memcpy_d2h(b->dev, b->modes_u, dU, S * nev * es64);
    if (!rc) rc = bdgx_memcpy_d2h(b->dev, b->modes_v, dMX, S * nev * es64);
  }
  tx[4] = bdg_wtime();
  if (tv)
    fprintf(stderr,
            "[bdg] results: %.1f ms (M X %.1f, norms %.1f, scaling %.1f, download %.1f)\n",
            1e3 * (tx[4] - tx[0]), 1e3 * (tx[1] - tx[0]), 1e3 * (tx[2] - tx[1]),
            1e3 * (tx[3] - tx[2]), 1e3 * (tx[4] - tx[3]));
  bdgx_free(dzero);
  bdgx_free(dn2);
  bdgx_free(dcu);
  bdgx_free(dcv);
  if (!rc && b->device_sweeps) {  /* u in dU, v in the head of dMX: kept for a sweep's reuse block */
    b->d_modes_u = dU;
    b->d_modes_v = dMX;
  } else {
    bdgx_free(dU);
    bdgx_free(dMX);
  }
  bdgx_free(dX);
  free(hn2);
  free(cu);
  free(cv);
  release_prefault(&pf);
  b->converged = conv;
  free(r.w2);
  free(r.zflag);
  free(r.res);
  b->stats.status = rc;
  b->stats.seconds_total = bdg_wtime() - t_start;
  return rc;
}

/* ---- results ---------------------------------------------------------------- */

uint64_t bdg_converged(const bdg_t *b) { return b->converged; }
const f64 *bdg_eigenvalues(const bdg_t *b) { return b->eigvals; }
const void *bdg_modes_u(const bdg_t *b) { return b->modes_u; }
const void *bdg_modes_v(const bdg_t *b) { return b->modes_v; }

/* ---- extensions (include/bdg/bdg_amd.h) ------------------------------------ */

void bdg_set_precision(bdg_t *b, bdg_precision_t p) { b->precision = (int)p; }
void bdg_set_formulation(bdg_t *b, bdg_formulation_t f) { b->formulation = (int)f; }
void bdg_set_locking(bdg_t *b, int max_active) { b->lock_max = max_active; }
void bdg_set_lock_guards(bdg_t *b, int guards) {
  if (guards < 0) {
    fprintf(stderr, "bdg_set_lock_guards: guards must be >= 0\n");
    exit(EXIT_FAILURE);
  }
  b->lock_guards = guards;
}
void bdg_set_fp32_phase(bdg_t *b, f64 fp32_tol, uint64_t stall_window) {
  b->fp32_tol = fp32_tol > 0 ? fp32_tol : 0.0;
  b->fp32_stall = (int)stall_window;
}
/* ---- batches of independent problems ---------------------------------- */

typedef struct {
  bdg_t **probs;
  int n;
  int *rc;
  int next; /* next problem to take, under lock */
  pthread_mutex_t lock;
} batch_state;

static void *batch_worker(void *arg) {
  batch_state *bs = (batch_state *)arg;
  for (;;) {
    pthread_mutex_lock(&bs->lock);
    const int i = bs->next < bs->n ? bs->next++ : -1;
    pthread_mutex_unlock(&bs->lock);
    if (i < 0) break;
    bs->rc[i] = bdg_solve(bs->probs[i]);
  }
  return NULL;
}

int bdg_solve_batch(bdg_t **probs, int n, int nthreads, int *rc) {
  if (!probs || n <= 0) return BDGX_EINVAL;
  int *codes = rc ? rc : (int *)xcalloc((size_t)n, sizeof(int));
  /* devices first, one at a time: context creation then happens before the
   * workers start */
  for (int i = 0; i < n; ++i) {
    codes[i] = probs[i] ? ensure_device(probs[i]) : BDGX_EINVAL;
  }
  int nt = nthreads > 0 && nthreads < n ? nthreads : n;
  batch_state bs;
  bs.probs = probs;
  bs.n = n;
  bs.rc = codes;
  bs.next = 0;
  pthread_mutex_init(&bs.lock, NULL);
  int first_err = 0;
  for (int i = 0; i < n; ++i)
    if (codes[i]) {
      first_err = codes[i];
      break;
    }
  if (!first_err) {
    pthread_t *th = (pthread_t *)xcalloc((size_t)nt, sizeof(pthread_t));
    int started = 0;
    for (int t = 0; t < nt; ++t)
      if (pthread_create(&th[t], NULL, batch_worker, &bs) == 0) ++started;
    if (started == 0) batch_worker(&bs); /* no threads: run them here */
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    free(th);
    for (int i = 0; i < n; ++i)
      if (codes[i]) {
        first_err = codes[i];
        break;
      }
  }
  pthread_mutex_destroy(&bs.lock);
  if (!rc) free(codes);
  return first_err;
}

void bdg_get_solve_stats(const bdg_t *b, bdg_solve_stats_t *out) {
  if (out) *out = b->stats;
}
const f64 *bdg_residuals(const bdg_t *b) { return b->residuals; }
const int *bdg_zero_modes(const bdg_t *b) { return b->zero_modes; }
bdgx_ctx *bdg_device_context(bdg_t *b) { return ensure_device(b) == BDGX_OK ? b->dev : NULL; }
const f64 *bdg_local_term_K(const bdg_t *b) { return b->ltK; }
const f64 *bdg_local_term_M(const bdg_t *b) { return b->ltM; }
const f64 *bdg_precond_weight_K(const bdg_t *b) { return b->pK; }
const f64 *bdg_precond_weight_M(const bdg_t *b) { return b->pM; }
f64 bdg_mu(const bdg_t *b) { return b->mu; }
void bdg_set_mixed_params(bdg_t *b, f64 img_tol, uint64_t refresh_every) {
  b->img_tol = img_tol;
  b->refresh_every = refresh_every;
}
void bdg_set_verbosity(bdg_t *b, int level) { b->verbosity = level; }
void bdg_set_graph_capture(bdg_t *b, int on) { b->graphs = on ? 1 : 0; }
void bdg_set_device_sweeps(bdg_t *b, int on) { b->device_sweeps = on ? 1 : 0; }
void bdg_set_fp64_sweep_only(bdg_t *b, int on) { b->sweep_only = on ? 1 : 0; }
void bdg_set_fp64_groups(bdg_t *b, int group) { b->fp64_groups = group > 0 ? group : 0; }
void bdg_set_freeze(bdg_t *b, f64 tol) { b->freeze_tol = tol > 0 ? tol : 0.0; }
void bdg_set_learned_fallback(bdg_t *b, int window, f64 factor) {
  b->lp_window = window > 0 ? window : 0;
  b->lp_factor = factor;
}
void bdg_set_learned_preconditioner(bdg_t *b, const float *theta) {
  freep((void **)&b->lp_theta);
  if (theta) {
    b->lp_theta = (float *)xcalloc(BDGX_LP_NPARAM, sizeof(float));
    memcpy(b->lp_theta, theta, sizeof(float) * BDGX_LP_NPARAM);
  }
}

f64 bdg_cb_harmonic_trap(uint64_t dim, const f64 *r, const void *param) {
  const f64 *om = (const f64 *)param;
  double v = 0.0;
  for (uint64_t d = 0; d < dim; ++d) v += 0.5 * om[d] * om[d] * r[d] * r[d];
  return v;
}

f64 bdg_cb_contact(const void *param, f64 density) { return *(const f64 *)param * density; }

void bdg_init_uniform_xorshift(const bdg_t *b, void *X, uint64_t n, uint64_t m, void *param) {
  uint32_t seed = param ? *(const uint32_t *)param : 42u;
  const uint64_t S = n / 2;
  for (uint64_t j = 0; j < m; ++j)
    for (uint64_t i = 0; i < S; ++i) {
      const double v = bdg_xrand(&seed) - 0.5;
      put(b, X, i, j, v);
      put(b, X, i + S, j, v);
    }
}
