// Device context and the extern "C" entry points of libbdg_amd's HIP layer.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "bdgx_internal.hpp"
#include "fft_core.hpp"
#include "plane.hpp"

thread_local hipError_t bdgx_last_hip_error = hipSuccess;

namespace bdgx {
int apply_dispatch(bdgx_ctx *c, int op, int prec, const void *dX, void *dY,
                   uint64_t ncols, uint64_t ld);
int gram_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                  const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b,
                  double *dG);
int gram_pair_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                       const void *dB, uint64_t ldb, const void *dB2, uint64_t ldb2,
                       uint64_t q, uint64_t nrows, double *dG);
int gram_noalloc_setup(int device);
int gram_noalloc_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                          const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b,
                          double *dG);
int combine2_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k1,
                      const void *dA2, uint64_t lda2, uint64_t k2, const double *dC, void *dY,
                      uint64_t ldy, uint64_t m, uint64_t nrows);
int combine_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                     const double *dC, void *dY, uint64_t ldy, uint64_t m, uint64_t nrows);
int combine_cols_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                          const double *dC, void *dY, uint64_t ldy, const int *cols,
                          uint64_t ncols, uint64_t nrows);
int residual_dispatch(bdgx_ctx *c, int prec, const void *dAX, const void *dX, const void *dMX,
                      uint64_t ld, uint64_t ncols, uint64_t nrows, const double *dlam,
                      int swap_x, void *dR, uint64_t ldr, double *dn2, const int *dcols);
int upload_cols(bdgx_ctx *c, const int *hidx, uint64_t ncols);
int convert_dispatch(bdgx_ctx *c, int ps, const void *src, uint64_t lds, int pd, void *dst,
                     uint64_t ldd, uint64_t nrows, uint64_t ncols);
int autotune_dispatch(bdgx_ctx *c, int prec, uint64_t nbatch);
int bench_pass_dispatch(bdgx_ctx *c, int pass, int prec, uint64_t nbatch, int iters, float *ms,
                        double *bytes);
int gather_dispatch(bdgx_ctx *c, int prec, const void *src, uint64_t lds, const int *hidx,
                    uint64_t ncols, void *dst, uint64_t ldd, uint64_t nrows);
template <typename R>
int meanfield_t(bdgx_ctx *c, const R *dpsi, R *dOut, double scale);
template <typename R>
int itp_kinetic_t(bdgx_ctx *c, R *dpsi, double dt);

int make_plan(int n, AxisPlan *p) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  int m = n;
  int f[32], nf = 0;
  while (m % 8 == 0 && m > 1) { f[nf++] = 8; m /= 8; }
  while (m % 4 == 0 && m > 1) { f[nf++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { f[nf++] = 2; m /= 2; }
  const int odd[] = {3, 5, 7};
  for (int t = 0; t < 3; ++t)
    while (m % odd[t] == 0 && m > 1) { f[nf++] = odd[t]; m /= odd[t]; }
  if (m != 1 || nf > 12) return BDGX_EINVAL;
  int Ns = 1;
  for (int s = 0; s < nf; ++s) {
    p->rad[s] = f[s];
    p->ns[s] = Ns;
    Ns *= f[s];
  }
  p->nst = nf;
  return BDGX_OK;
}

int ensure_spec(bdgx_ctx *c, size_t bytes) {
  if (bytes <= c->spec_bytes) return BDGX_OK;
  if (c->spec) {
    BDGX_HIP(hipStreamSynchronize(c->stream));
    BDGX_HIP(hipFree(c->spec));
    c->spec = nullptr;
    c->spec_bytes = 0;
  }
  BDGX_HIP(hipMalloc(&c->spec, bytes));
  c->spec_bytes = bytes;
  c->ws_gen++;
  return BDGX_OK;
}

int ensure_scratch(bdgx_ctx *c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return BDGX_OK;
  if (c->scratch) {
    BDGX_HIP(hipStreamSynchronize(c->stream));
    BDGX_HIP(hipFree(c->scratch));
    c->scratch = nullptr;
    c->scratch_bytes = 0;
  }
  BDGX_HIP(hipMalloc(&c->scratch, bytes));
  c->scratch_bytes = bytes;
  c->ws_gen++;
  return BDGX_OK;
}

// n reals: d[e] = (float)s[e]
__global__ void k_d2f(const double *s, float *d, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    d[e] = (float)s[e];
}

// n reals (a complex array counts its real and imaginary parts): the FP64
// copy uploaded as is, the FP32 copy converted from it on the device (same
// round-to-nearest conversion as on the host)
static int upload_both(bdgx_ctx *c, double **d64, float **d32, const double *src, size_t n) {
  if (!*d64) BDGX_HIP(hipMalloc((void **)d64, sizeof(double) * (n ? n : 1)));
  if (!*d32) BDGX_HIP(hipMalloc((void **)d32, sizeof(float) * (n ? n : 1)));
  if (n == 0) return BDGX_OK;
  BDGX_HIP(hipMemcpyAsync(*d64, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_d2f, dim3(1024), dim3(256), 0, c->stream, *d64, *d32, (long long)n);
  BDGX_HIP(hipGetLastError());
  BDGX_HIP(hipStreamSynchronize(c->stream));
  return BDGX_OK;
}

template <typename R>
static void free_fields(Fields<R> &f) {
  R **arr[] = {&f.ltK, &f.ltM, &f.pK, &f.pM, &f.wf, &f.U, &f.lpa[0], &f.lpa[1],
               &f.lpb[0], &f.lpb[1], &f.lpg[0], &f.lpg[1]};
  for (auto p : arr)
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  for (int d = 0; d < 3; ++d) {
    if (f.kx2[d]) (void)hipFree(f.kx2[d]);
    if (f.tw[d]) (void)hipFree(f.tw[d]);
    f.kx2[d] = nullptr;
    f.tw[d] = nullptr;
  }
}

// ---------------------------------------------------------------------------
// dipolar kernel U(k) on the device (reference src/dipolar.c:21-72)
__global__ void k_dipolar_kernel(double *U, int nx, int ny, int nz, int n0k, int cplx,
                                 double dkx, double dky, double dkz, double d0, double d1,
                                 double d2, double R) {
  const long long total = (long long)n0k * ny * nz;
  const double dd = d0 * d0 + d1 * d1 + d2 * d2;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(e % n0k);
    const long long rest = e / n0k;
    const int iy = (int)(rest % ny), iz = (int)(rest / ny);
    const double kz = dkz * (double)(iz <= nz / 2 ? iz : iz - nz);
    const double ky = dky * (double)(iy <= ny / 2 ? iy : iy - ny);
    const double kx = dkx * (double)((cplx && ix > nx / 2) ? ix - nx : ix);
    const double ks = kx * kx + ky * ky + kz * kz;
    if (ks < 1e-50) {
      U[e] = 0.0;
      continue;
    }
    const double kd = kx * d0 + ky * d1 + kz * d2;
    const double c2 = kd * kd / (ks * dd);
    const double kR = sqrt(ks) * R, kR2 = kR * kR;
    const double f = 1.0 + 3.0 * cos(kR) / kR2 - 3.0 * sin(kR) / (kR2 * kR);
    U[e] = (3.0 * c2 - 1.0) * f;
  }
}

// ---------------------------------------------------------------------------
// imaginary-time helpers
// psi *= exp(-0.5 dt (V + g|psi|^2 + phi)); accumulate sum |psi|^2
template <typename R, bool CPLX>
__global__ void k_itp_pot(R *psi, const double *V, const R *phi, double g, double dt,
                          long long S, double *norm) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S;
       i += (long long)gridDim.x * blockDim.x) {
    double pr, pi = 0.0;
    if (CPLX) {
      pr = (double)psi[2 * i];
      pi = (double)psi[2 * i + 1];
    } else {
      pr = (double)psi[i];
    }
    const double n = pr * pr + pi * pi;
    double ve = V[i] + g * n;
    if (phi) ve += (double)phi[i];
    const double f = exp(-0.5 * dt * ve);
    pr *= f;
    pi *= f;
    if (CPLX) {
      psi[2 * i] = (R)pr;
      psi[2 * i + 1] = (R)pi;
    } else {
      psi[i] = (R)pr;
    }
    s += pr * pr + pi * pi;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0 && norm) norm[blockIdx.x] = red[0];  // block partial
}

// Sum of the kItpBlocks block partials in a fixed order (every caller gets
// the same bits); blockDim.x == 256.
constexpr int kItpBlocks = 1024;

__device__ __forceinline__ double itp_sum_parts(const double *part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < kItpBlocks; k += 256) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double t = red[0];
  __syncthreads();
  return t;
}

template <typename R>
__global__ void __launch_bounds__(256) k_scale_by_norm(R *psi, long long n, const double *part,
                                                       double dV) {
  const double s = 1.0 / sqrt(itp_sum_parts(part) * dV);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    psi[i] = (R)((double)psi[i] * s);
}

__global__ void __launch_bounds__(256) k_sum_parts(const double *part, double *out) {
  const double t = itp_sum_parts(part);
  if (threadIdx.x == 0) out[0] = t;
}

// mu integrand: conj(psi) * (T psi + Veff psi), T psi supplied in tpsi
template <bool CPLX>
__global__ void k_itp_mu(const double *psi, const double *tpsi, const double *V,
                         const double *phi, double g, long long S, double *acc) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S;
       i += (long long)gridDim.x * blockDim.x) {
    double pr, pi = 0.0, tr, ti = 0.0;
    if (CPLX) {
      pr = psi[2 * i];
      pi = psi[2 * i + 1];
      tr = tpsi[2 * i];
      ti = tpsi[2 * i + 1];
    } else {
      pr = psi[i];
      tr = tpsi[i];
    }
    const double n = pr * pr + pi * pi;
    double ve = V[i] + g * n;
    if (phi) ve += phi[i];
    s += pr * (tr + ve * pr) + pi * (ti + ve * pi);
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[blockIdx.x] = red[0];  // block partial
}

}  // namespace bdgx

using namespace bdgx;

// ---- uniform start block: xorshift32 with jump-ahead -------------------
// One xorshift32 step is a linear map T on GF(2)^32; a 32x32 GF(2) matrix is
// stored as its 32 columns (column b = T applied to the unit vector b).
namespace {
struct Gf2 {
  uint32_t col[32];
};
uint32_t gf2_apply(const Gf2 &m, uint32_t x) {
  uint32_t r = 0;
  for (int b = 0; b < 32; ++b)
    if (x >> b & 1u) r ^= m.col[b];
  return r;
}
Gf2 gf2_mul(const Gf2 &a, const Gf2 &b) {  // a * b (apply b first)
  Gf2 r;
  for (int k = 0; k < 32; ++k) r.col[k] = gf2_apply(a, b.col[k]);
  return r;
}
uint32_t xs_step(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}
constexpr int kJumpBits = 40;  // jumps up to 2^40 steps
}  // namespace

// Elements [t * RUN, (t + 1) * RUN) of column blockIdx.y: the thread jumps
// its column's start state ahead by t * RUN steps with the powers T^(2^b),
// then steps sequentially.
template <bool CPLX>
__global__ void __launch_bounds__(256) k_fill_uniform(double *X, long long ld, long long nrows,
                                                      const uint32_t *col_state,
                                                      const uint32_t *pow2, int nbits) {
  constexpr int RUN = 64;
  const long long r0 = ((long long)blockIdx.x * 256 + threadIdx.x) * RUN;
  if (r0 >= nrows) return;
  uint32_t x = col_state[blockIdx.y];
  for (int b = 0; b < nbits; ++b) {
    if (((unsigned long long)r0 >> b) & 1ull) {
      const uint32_t *m = pow2 + 32 * b;
      uint32_t r = 0;
      for (int k = 0; k < 32; ++k)
        if (x >> k & 1u) r ^= m[k];
      x = r;
    }
  }
  double *col = X + (long long)blockIdx.y * ld * (CPLX ? 2 : 1);
  const long long r1 = r0 + RUN < nrows ? r0 + RUN : nrows;
  for (long long i = r0; i < r1; ++i) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    const double v = (double)(x & 0x7FFFFFFFu) / (double)0x7FFFFFFF - 0.5;
    if (CPLX) {
      col[2 * i] = v;
      col[2 * i + 1] = 0.0;
    } else {
      col[i] = v;
    }
  }
}

// ---- reuse block of a parameter sweep (reference src/bdg.c:112-207) ------
// Column j < nev: [U_j; V_j] plus sigma_j (z0, z1), sigma_j = noise *
// ||[U_j; V_j]||, (z0, z1) the Box-Muller pair of the xrand values number
// 2 (j S + i) + 1, + 2 of the xorshift32 stream; column j >= nev: both halves
// |z0| |psi0_i|.  Column norms: kReuseChunks fixed row chunks per column,
// summed in a fixed order (deterministic).
constexpr int kReuseChunks = 64;

template <bool CPLX>
__global__ void __launch_bounds__(256) k_colsq_partial(const double *U, const double *V,
                                                       long long ld, long long S,
                                                       double *part) {
  constexpr int W = CPLX ? 2 : 1;
  const int j = blockIdx.y;
  const long long len = (S + kReuseChunks - 1) / kReuseChunks;
  const long long r0 = (long long)blockIdx.x * len;
  const long long r1 = r0 + len < S ? r0 + len : S;
  const double *u = U + (long long)j * ld * W, *v = V + (long long)j * ld * W;
  double s = 0.0;
  for (long long i = r0 * W + threadIdx.x; i < r1 * W; i += 256) s += u[i] * u[i] + v[i] * v[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)j * kReuseChunks + blockIdx.x] = red[0];
}

template <bool CPLX>
__global__ void __launch_bounds__(256) k_reuse_fill(const double *U, const double *V,
                                                    long long ldm, int nev, const double *part,
                                                    double noise, const double *wf, double *OU,
                                                    double *OV, long long ldo, long long S,
                                                    const uint32_t *col_state,
                                                    const uint32_t *pow2, int nbits) {
  constexpr int RUN = 64;
  constexpr int W = CPLX ? 2 : 1;
  const int j = blockIdx.y;
  const long long r0 = ((long long)blockIdx.x * 256 + threadIdx.x) * RUN;
  if (r0 >= S) return;
  double sigma = 0.0;
  if (j < nev) {
    double s = 0.0;
    for (int q = 0; q < kReuseChunks; ++q) s += part[(long long)j * kReuseChunks + q];
    sigma = noise * sqrt(s);
  }
  // two stream values per row: jump 2 r0 steps from the column's start
  uint32_t x = col_state[j];
  const unsigned long long jump = 2ull * (unsigned long long)r0;
  for (int b = 0; b < nbits; ++b) {
    if ((jump >> b) & 1ull) {
      const uint32_t *m = pow2 + 32 * b;
      uint32_t r = 0;
      for (int k = 0; k < 32; ++k)
        if (x >> k & 1u) r ^= m[k];
      x = r;
    }
  }
  const long long r1 = r0 + RUN < S ? r0 + RUN : S;
  double *ou = OU + (long long)j * ldo * W, *ov = OV + (long long)j * ldo * W;
  for (long long i = r0; i < r1; ++i) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    double u1 = (double)(x & 0x7FFFFFFFu) / (double)0x7FFFFFFF;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    const double u2 = (double)(x & 0x7FFFFFFFu) / (double)0x7FFFFFFF;
    u1 = fmax(u1, 1e-10);
    const double rr = sqrt(-2.0 * log(u1));
    const double ang = 2.0 * M_PI * u2;
    if (j < nev) {
      const double *u = U + (long long)j * ldm * W, *v = V + (long long)j * ldm * W;
      ou[W * i] = u[W * i] + sigma * (rr * cos(ang));
      ov[W * i] = v[W * i] + sigma * (rr * sin(ang));
      if (CPLX) {
        ou[W * i + 1] = u[W * i + 1];
        ov[W * i + 1] = v[W * i + 1];
      }
    } else {
      double a = 1.0;
      if (wf) a = CPLX ? sqrt(wf[2 * i] * wf[2 * i] + wf[2 * i + 1] * wf[2 * i + 1]) : fabs(wf[i]);
      const double val = fabs(rr * cos(ang)) * a;
      ou[W * i] = val;
      ov[W * i] = val;
      if (CPLX) {
        ou[W * i + 1] = 0.0;
        ov[W * i + 1] = 0.0;
      }
    }
  }
}

// device blocks kept between bdgx_free and bdgx_malloc (below)
namespace {
constexpr size_t kDevCacheMax = (size_t)64 << 30;
struct DevCache {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void *>> freeb;  // (device, bytes) -> blocks
  std::map<void *, std::pair<int, size_t>> live;                // blocks handed out
  size_t kept[64] = {0};                                         // bytes kept per device
};
DevCache &dev_cache() {
  static DevCache c;
  return c;
}
void release_kept(int dev) {
  DevCache &c = dev_cache();
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> g(c.mu);
    for (auto it = c.freeb.begin(); it != c.freeb.end();) {
      if (it->first.first == dev) {
        for (void *q : it->second) drop.push_back(q);
        it = c.freeb.erase(it);
      } else {
        ++it;
      }
    }
    if (dev >= 0 && dev < 64) c.kept[dev] = 0;
  }
  for (void *q : drop) (void)hipFree(q);
}
}  // namespace

extern "C" {

int bdgx_reuse_block(bdgx_ctx *c, const void *dU, const void *dV, uint64_t ldm, uint64_t nev,
                     void *dOU, void *dOV, uint64_t ldo, uint64_t ncols, uint64_t nrows,
                     double noise_frac, uint32_t seed) {
  if (!c || !dOU || !dOV || nrows == 0 || ncols == 0 || ncols > 65535 || nev > ncols ||
      ldo < nrows || (nev > 0 && (!dU || !dV || ldm < nrows)) ||
      2 * nrows * ncols > (1ull << kJumpBits))
    return BDGX_EINVAL;
  std::vector<uint32_t> pw(32 * kJumpBits);
  Gf2 t;
  for (int b = 0; b < 32; ++b) t.col[b] = xs_step(1u << b);
  for (int b = 0; b < kJumpBits; ++b) {
    memcpy(&pw[32 * b], t.col, sizeof(t.col));
    t = gf2_mul(t, t);
  }
  // column start states: 2 nrows stream values per column
  std::vector<uint32_t> st(ncols);
  uint32_t s0 = seed;
  const uint64_t per_col = 2 * nrows;
  for (uint64_t j = 0; j < ncols; ++j) {
    st[j] = s0;
    for (int b = 0; b < kJumpBits; ++b)
      if (per_col >> b & 1ull) {
        Gf2 m;
        memcpy(m.col, &pw[32 * b], sizeof(m.col));
        s0 = gf2_apply(m, s0);
      }
  }
  uint32_t *dtab = nullptr;
  double *dpart = nullptr;
  const size_t tb = sizeof(uint32_t) * (pw.size() + st.size());
  BDGX_HIP(hipMalloc((void **)&dtab, tb));
  int rc = BDGX_OK;
  if (hipMalloc((void **)&dpart, sizeof(double) * kReuseChunks * (nev ? nev : 1)) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc && (hipMemcpyAsync(dtab, pw.data(), sizeof(uint32_t) * pw.size(),
                             hipMemcpyHostToDevice, c->stream) != hipSuccess ||
              hipMemcpyAsync(dtab + pw.size(), st.data(), sizeof(uint32_t) * st.size(),
                             hipMemcpyHostToDevice, c->stream) != hipSuccess))
    rc = BDGX_EHIP;
  const double *wf = c->has_wf ? c->f64.wf : nullptr;
  const double *U = (const double *)dU, *V = (const double *)dV;
  if (!rc && nev > 0) {
    const dim3 grid(kReuseChunks, (unsigned)nev);
    if (c->cplx)
      hipLaunchKernelGGL(k_colsq_partial<true>, grid, dim3(256), 0, c->stream, U, V,
                         (long long)ldm, (long long)nrows, dpart);
    else
      hipLaunchKernelGGL(k_colsq_partial<false>, grid, dim3(256), 0, c->stream, U, V,
                         (long long)ldm, (long long)nrows, dpart);
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  if (!rc) {
    const unsigned gx = (unsigned)((nrows + 256 * 64 - 1) / (256 * 64));
    const dim3 grid(gx, (unsigned)ncols);
    if (c->cplx)
      hipLaunchKernelGGL(k_reuse_fill<true>, grid, dim3(256), 0, c->stream, U, V, (long long)ldm,
                         (int)nev, dpart, noise_frac, wf, (double *)dOU, (double *)dOV,
                         (long long)ldo, (long long)nrows, dtab + pw.size(), dtab, kJumpBits);
    else
      hipLaunchKernelGGL(k_reuse_fill<false>, grid, dim3(256), 0, c->stream, U, V,
                         (long long)ldm, (int)nev, dpart, noise_frac, wf, (double *)dOU,
                         (double *)dOV, (long long)ldo, (long long)nrows, dtab + pw.size(), dtab,
                         kJumpBits);
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  // the tables and partials must outlive the kernels: wait, then free
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = BDGX_EHIP;
  (void)hipFree(dtab);
  if (dpart) (void)hipFree(dpart);
  return rc;
}

const char *bdgx_error_string(int code) {
  switch (code) {
    case BDGX_OK: return "ok";
    case BDGX_EINVAL: return "invalid argument or unsupported shape";
    case BDGX_EHIP: return hipGetErrorString(bdgx_last_hip_error);
    case BDGX_ESTATE: return "required field not uploaded";
    case BDGX_ENOMEM: return "out of memory";
    case BDGX_ETOOBIG:
      return "problem exceeds the fixed scratch of the allocation-free entry point";
    default: return "unknown error";
  }
}

int bdgx_ctx_create(bdgx_ctx **out, uint64_t dim, const uint64_t *N, const double *L,
                    int complex_psi0) {
  if (!out || dim < 1 || dim > 3 || !N || !L) return BDGX_EINVAL;
  *out = nullptr;
  bdgx_ctx *c = new bdgx_ctx();
  memset(c, 0, sizeof(*c));
  c->dim = (int)dim;
  c->cplx = complex_psi0 ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    c->N[d] = d < (int)dim ? N[d] : 1;
    c->L[d] = d < (int)dim ? L[d] : 1.0;
    if (c->N[d] < 1 || c->N[d] > 4096) {
      delete c;
      return BDGX_EINVAL;
    }
  }
  Geo &g = c->geo;
  g.nx = (int)c->N[0];
  g.ny = (int)c->N[1];
  g.nz = (int)c->N[2];
  g.n0k = c->cplx ? g.nx : g.nx / 2 + 1;
  g.S = (long long)g.nx * g.ny * g.nz;
  g.Kk = (long long)g.n0k * g.ny * g.nz;
  tune_defaults(c);
  for (int d = 0; d < 3; ++d)
    if (make_plan((int)c->N[d], &c->plan[d])) {
      delete c;
      return BDGX_EINVAL;
    }
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return BDGX_EHIP;
  }
  // device address of the allocation-free Gram's fixed scratch (once per
  // device, with this device current)
  if (gram_noalloc_setup(c->device) != BDGX_OK) {
    bdgx_ctx_destroy(c);
    return BDGX_EHIP;
  }
  // k^2 tables (reference src/matmul_ctx.c:102-113) and twiddles
  for (int d = 0; d < 3; ++d) {
    const int n = (int)c->N[d];
    std::vector<double> k2(n), twr(2 * n);
    const double dk = 2.0 * M_PI / c->L[d];
    for (int j = 0; j < n; ++j) {
      const double kj = (d < (int)dim) ? ((j <= n / 2) ? dk * j : dk * (j - n)) : 0.0;
      k2[j] = kj * kj;
      const double a = -2.0 * M_PI * (double)j / (double)n;
      twr[2 * j] = cos(a);
      twr[2 * j + 1] = sin(a);
    }
    int rc = upload_both(c, &c->f64.kx2[d], &c->f32.kx2[d], k2.data(), n);
    if (!rc)
      rc = upload_both(c, reinterpret_cast<double **>(&c->f64.tw[d]),
                       reinterpret_cast<float **>(&c->f32.tw[d]), twr.data(), 2 * n);
    if (rc) {
      bdgx_ctx_destroy(c);
      return rc;
    }
  }
  *out = c;
  return BDGX_OK;
}

void bdgx_ctx_destroy(bdgx_ctx *c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  free_fields(c->f64);
  free_fields(c->f32);
  if (c->spec) (void)hipFree(c->spec);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->idx_buf) (void)hipFree(c->idx_buf);
  if (c->lp_conv) (void)hipFree(c->lp_conv);
  if (c->lp_rho) (void)hipFree(c->lp_rho);
  if (c->lp_tmp) (void)hipFree(c->lp_tmp);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

uint64_t bdgx_ctx_size(const bdgx_ctx *c) { return (uint64_t)c->geo.S; }
uint64_t bdgx_ctx_ksize(const bdgx_ctx *c) { return (uint64_t)c->geo.Kk; }

int bdgx_set_fields(bdgx_ctx *c, const double *ltK, const double *ltM, const void *wf,
                    const double *pK, const double *pM, double mu) {
  if (!c) return BDGX_EINVAL;
  BDGX_HIP(hipStreamSynchronize(c->stream));  // kernels in flight may read the fields
  const size_t S = (size_t)c->geo.S;
  int rc = 0;
  if (ltK) rc = rc ? rc : upload_both(c, &c->f64.ltK, &c->f32.ltK, ltK, S);
  if (ltM) rc = rc ? rc : upload_both(c, &c->f64.ltM, &c->f32.ltM, ltM, S);
  if (pK) rc = rc ? rc : upload_both(c, &c->f64.pK, &c->f32.pK, pK, S);
  if (pM) rc = rc ? rc : upload_both(c, &c->f64.pM, &c->f32.pM, pM, S);
  if (wf)
    rc = rc ? rc
            : upload_both(c, &c->f64.wf, &c->f32.wf, (const double *)wf, S * (c->cplx ? 2 : 1));
  if (rc) return rc;
  if (ltK && ltM) c->has_lt = 1;
  if (pK && pM) c->has_p = 1;
  if (wf) c->has_wf = 1;
  c->mu = mu;
  return BDGX_OK;
}

// pK = 1 / sqrt(max(floor, ltK + mu)), pM from ltM; FP32 copies converted
__global__ void k_derive_precond(const double *ltK, const double *ltM, double mu, double floor,
                                 double *pK, double *pM, float *pK32, float *pM32, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const double a = 1.0 / sqrt(fmax(floor, ltK[e] + mu));
    const double b = 1.0 / sqrt(fmax(floor, ltM[e] + mu));
    pK[e] = a;
    pM[e] = b;
    pK32[e] = (float)a;
    pM32[e] = (float)b;
  }
}

int bdgx_derive_precond(bdgx_ctx *c, double floor) {
  if (!c || !c->has_lt) return BDGX_EINVAL;
  const size_t S = (size_t)c->geo.S;
  if (!c->f64.pK) BDGX_HIP(hipMalloc((void **)&c->f64.pK, sizeof(double) * (S ? S : 1)));
  if (!c->f64.pM) BDGX_HIP(hipMalloc((void **)&c->f64.pM, sizeof(double) * (S ? S : 1)));
  if (!c->f32.pK) BDGX_HIP(hipMalloc((void **)&c->f32.pK, sizeof(float) * (S ? S : 1)));
  if (!c->f32.pM) BDGX_HIP(hipMalloc((void **)&c->f32.pM, sizeof(float) * (S ? S : 1)));
  hipLaunchKernelGGL(k_derive_precond, dim3(1024), dim3(256), 0, c->stream, c->f64.ltK,
                     c->f64.ltM, c->mu, floor, c->f64.pK, c->f64.pM, c->f32.pK, c->f32.pM,
                     (long long)S);
  BDGX_HIP(hipGetLastError());
  BDGX_HIP(hipStreamSynchronize(c->stream));
  c->has_p = 1;
  return BDGX_OK;
}

int bdgx_reset_fields(bdgx_ctx *c) {
  if (!c) return BDGX_EINVAL;
  BDGX_HIP(hipStreamSynchronize(c->stream));
  c->has_wf = c->has_U = c->has_lt = c->has_p = 0;
  c->mu = c->g_ddi = 0.0;
  return BDGX_OK;
}

int bdgx_dipolar_set_kernel(bdgx_ctx *c, double g_ddi, const double *dir, double R) {
  if (!c || !dir || c->dim != 3) return BDGX_EINVAL;
  BDGX_HIP(hipStreamSynchronize(c->stream));
  const size_t K = (size_t)c->geo.Kk;
  if (!c->f64.U) BDGX_HIP(hipMalloc((void **)&c->f64.U, sizeof(double) * K));
  if (!c->f32.U) BDGX_HIP(hipMalloc((void **)&c->f32.U, sizeof(float) * K));
  const Geo &g = c->geo;
  hipLaunchKernelGGL(k_dipolar_kernel, dim3(1024), dim3(256), 0, c->stream, c->f64.U, g.nx,
                     g.ny, g.nz, g.n0k, c->cplx, 2.0 * M_PI / c->L[0], 2.0 * M_PI / c->L[1],
                     2.0 * M_PI / c->L[2], dir[0], dir[1], dir[2], R);
  BDGX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_d2f, dim3(1024), dim3(256), 0, c->stream, c->f64.U, c->f32.U,
                     (long long)K);
  BDGX_HIP(hipGetLastError());
  BDGX_HIP(hipStreamSynchronize(c->stream));
  c->g_ddi = g_ddi;
  c->has_U = 1;
  return BDGX_OK;
}

int bdgx_dipolar_get_kernel(bdgx_ctx *c, double *h) {
  if (!c || !c->has_U) return BDGX_ESTATE;
  BDGX_HIP(hipMemcpyAsync(h, c->f64.U, sizeof(double) * c->geo.Kk, hipMemcpyDeviceToHost,
                          c->stream));
  BDGX_HIP(hipStreamSynchronize(c->stream));
  return BDGX_OK;
}

int bdgx_dipolar_meanfield(bdgx_ctx *c, double *h) {
  if (!c || !c->has_U || !c->has_wf) return BDGX_ESTATE;
  double *d = nullptr;
  BDGX_HIP(hipMalloc((void **)&d, sizeof(double) * c->geo.S));
  int rc = meanfield_t<double>(c, c->f64.wf, d, c->g_ddi);
  if (!rc) {
    if (hipMemcpyAsync(h, d, sizeof(double) * c->geo.S, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = BDGX_EHIP;
  }
  (void)hipFree(d);
  return rc;
}

int bdgx_apply(bdgx_ctx *c, int op, int prec, const void *dX, void *dY, uint64_t ncols,
               uint64_t ld) {
  if (!c || !dX || !dY) return BDGX_EINVAL;
  if (ncols == 0) return BDGX_OK;
  const uint64_t need = (op == BDGX_OP_LREP || op == BDGX_OP_SWAP || op == BDGX_OP_PRECOND_LREP)
                            ? 2 * (uint64_t)c->geo.S
                            : (uint64_t)c->geo.S;
  if (ld < need) return BDGX_EINVAL;
  return apply_dispatch(c, op, prec, dX, dY, ncols, ld);
}

int bdgx_apply_host(bdgx_ctx *c, int op, const void *hX, void *hY, uint64_t ncols,
                    uint64_t ld) {
  if (!c || !hX || !hY) return BDGX_EINVAL;
  const size_t bytes = sizeof(double) * (c->cplx ? 2 : 1) * (size_t)ld * ncols;
  void *dX = nullptr, *dY = nullptr;
  BDGX_HIP(hipMalloc(&dX, bytes));
  if (hipMalloc(&dY, bytes) != hipSuccess) {
    (void)hipFree(dX);
    return BDGX_EHIP;
  }
  int rc = BDGX_OK;
  if (hipMemcpyAsync(dX, hX, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc && hipMemsetAsync(dY, 0, bytes, c->stream) != hipSuccess) rc = BDGX_EHIP;
  if (!rc) rc = bdgx_apply(c, op, BDGX_FP64, dX, dY, ncols, ld);
  if (!rc && hipMemcpyAsync(hY, dY, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = BDGX_EHIP;
  (void)hipFree(dX);
  (void)hipFree(dY);
  return rc;
}

int bdgx_gram(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p, const void *dB,
              uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b, double *dG) {
  if (!c || !dA || !dB || !dG) return BDGX_EINVAL;
  if (swap_b && (nrows & 1)) return BDGX_EINVAL;
  return gram_dispatch(c, prec, dA, lda, p, dB, ldb, q, nrows, swap_b, dG);
}

int bdgx_gram_noalloc(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                      const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b,
                      double *dG) {
  if (!c || !dA || !dB || !dG) return BDGX_EINVAL;
  if (swap_b && (nrows & 1)) return BDGX_EINVAL;
  return gram_noalloc_dispatch(c, prec, dA, lda, p, dB, ldb, q, nrows, swap_b, dG);
}

int bdgx_gram_pair(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                   const void *dB1, uint64_t ldb1, const void *dB2, uint64_t ldb2, uint64_t q,
                   uint64_t nrows, double *dG) {
  if (!c || !dA || !dB1 || !dB2 || !dG) return BDGX_EINVAL;
  return gram_pair_dispatch(c, prec, dA, lda, p, dB1, ldb1, dB2, ldb2, q, nrows, dG);
}

int bdgx_combine(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                 const double *dC, void *dY, uint64_t ldy, uint64_t m, uint64_t nrows) {
  if (!c || !dA || !dC || !dY) return BDGX_EINVAL;
  return combine_dispatch(c, prec, dA, lda, k, dC, dY, ldy, m, nrows);
}

int bdgx_combine2(bdgx_ctx *c, int prec, const void *dA1, uint64_t lda1, uint64_t k1,
                  const void *dA2, uint64_t lda2, uint64_t k2, const double *dC, void *dY,
                  uint64_t ldy, uint64_t m, uint64_t nrows) {
  if (!c || !dA1 || !dC || !dY || (k2 && !dA2)) return BDGX_EINVAL;
  return combine2_dispatch(c, prec, dA1, lda1, k1, dA2, lda2, k2, dC, dY, ldy, m, nrows);
}

int bdgx_combine_cols(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                      const double *dC, void *dY, uint64_t ldy, const int *cols, uint64_t ncols,
                      uint64_t nrows) {
  if (!c || !dA || !dC || !dY || (!cols && ncols)) return BDGX_EINVAL;
  return combine_cols_dispatch(c, prec, dA, lda, k, dC, dY, ldy, cols, ncols, nrows);
}

int bdgx_residual(bdgx_ctx *c, int prec, const void *dAX, const void *dX, const void *dMX,
                  uint64_t ld, uint64_t ncols, uint64_t nrows, const double *dlam, int swap_x,
                  void *dR, uint64_t ldr, double *dn2) {
  if (!c || !dAX || !dX || !dlam || !dn2) return BDGX_EINVAL;
  if (swap_x && (nrows & 1)) return BDGX_EINVAL;
  return residual_dispatch(c, prec, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2,
                           nullptr);
}

int bdgx_residual_cols(bdgx_ctx *c, int prec, const void *dAX, const void *dX, const void *dMX,
                       uint64_t ld, const int *cols, uint64_t ncols, uint64_t nrows,
                       const double *dlam, int swap_x, void *dR, uint64_t ldr, double *dn2) {
  if (!c || !dAX || !dX || !dlam || !dn2 || (ncols && !cols)) return BDGX_EINVAL;
  if (swap_x && (nrows & 1)) return BDGX_EINVAL;
  if (ncols == 0) return BDGX_OK;
  int rc = upload_cols(c, cols, ncols);
  if (rc) return rc;
  return residual_dispatch(c, prec, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2,
                           c->idx_buf);
}

int bdgx_gather_cols(bdgx_ctx *c, int prec, const void *src, uint64_t lds, const int *idx,
                     uint64_t ncols, void *dst, uint64_t ldd, uint64_t nrows) {
  if (!c || !src || !dst || (!idx && ncols)) return BDGX_EINVAL;
  return gather_dispatch(c, prec, src, lds, idx, ncols, dst, ldd, nrows);
}

int bdgx_convert(bdgx_ctx *c, int ps, const void *src, uint64_t lds, int pd, void *dst,
                 uint64_t ldd, uint64_t nrows, uint64_t ncols) {
  if (!c) return BDGX_EINVAL;
  return convert_dispatch(c, ps, src, lds, pd, dst, ldd, nrows, ncols);
}

int bdgx_itp_groundstate(bdgx_ctx *c, const double *omega, double g, double g_dd,
                         const double *dir, double cutoff, double dt, uint64_t nsteps,
                         void *h_psi, double *mu_out) {
  if (!c || !omega || !h_psi) return BDGX_EINVAL;
  if (g_dd != 0.0 && (c->dim != 3 || !dir)) return BDGX_EINVAL;
  const Geo &G = c->geo;
  const long long S = G.S;
  const int W = c->cplx ? 2 : 1;
  double dV = 1.0;
  for (int d = 0; d < c->dim; ++d) dV *= c->L[d] / (double)c->N[d];
  // harmonic trap on the grid of reference src/setup.c:60-81
  std::vector<double> V((size_t)S);
  for (int iz = 0; iz < G.nz; ++iz)
    for (int iy = 0; iy < G.ny; ++iy)
      for (int ix = 0; ix < G.nx; ++ix) {
        const int idx[3] = {ix, iy, iz};
        double v = 0.0;
        for (int d = 0; d < c->dim; ++d) {
          const double r = ((double)idx[d] - 0.5 * (double)c->N[d]) * c->L[d] / (double)c->N[d];
          v += 0.5 * omega[d] * omega[d] * r * r;
        }
        V[((size_t)iz * G.ny + iy) * G.nx + ix] = v;
      }
  double *dV_ = nullptr, *dpsi = nullptr, *dphi = nullptr, *dnorm = nullptr, *dT = nullptr;
  int rc = BDGX_OK;
  const double saved_g = c->g_ddi;
  const int saved_hasU = c->has_U;
  auto cleanup = [&]() {
    if (dV_) (void)hipFree(dV_);
    if (dpsi) (void)hipFree(dpsi);
    if (dphi) (void)hipFree(dphi);
    if (dnorm) (void)hipFree(dnorm);
    if (dT) (void)hipFree(dT);
    c->g_ddi = saved_g;
    c->has_U = saved_hasU;
  };
#define ITP_HIP(x)          \
  do {                      \
    if ((x) != hipSuccess) { \
      cleanup();            \
      return BDGX_EHIP;     \
    }                       \
  } while (0)
  ITP_HIP(hipMalloc((void **)&dV_, sizeof(double) * S));
  ITP_HIP(hipMalloc((void **)&dpsi, sizeof(double) * S * W));
  ITP_HIP(hipMalloc((void **)&dphi, sizeof(double) * S));
  // dnorm: block partials of the norm (kItpBlocks), of mu (kItpBlocks), mu
  ITP_HIP(hipMalloc((void **)&dnorm, sizeof(double) * (2 * kItpBlocks + 1)));
  ITP_HIP(hipMalloc((void **)&dT, sizeof(double) * S * W));
  ITP_HIP(hipMemcpyAsync(dV_, V.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
  ITP_HIP(hipMemcpyAsync(dpsi, h_psi, sizeof(double) * S * W, hipMemcpyHostToDevice, c->stream));
  const bool dip = g_dd != 0.0;
  if (dip) {
    // kernel for the ITP's own dipole parameters (restored on exit)
    double *save64 = c->f64.U;
    float *save32 = c->f32.U;
    c->f64.U = nullptr;
    c->f32.U = nullptr;
    rc = bdgx_dipolar_set_kernel(c, g_dd, dir, cutoff);
    double *itpU64 = c->f64.U;
    float *itpU32 = c->f32.U;
    if (rc) {
      c->f64.U = save64;
      c->f32.U = save32;
      cleanup();
      return rc;
    }
    // run the loop with the ITP kernel, then restore the context's own kernel
    auto potential = [&](double *nrm) -> int {
      int r = meanfield_t<double>(c, dpsi, dphi, g_dd);
      if (r) return r;
      if (c->cplx)
        hipLaunchKernelGGL((k_itp_pot<double, true>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                           dV_, dphi, g, dt, S, nrm);
      else
        hipLaunchKernelGGL((k_itp_pot<double, false>), dim3(kItpBlocks), dim3(256), 0, c->stream,
                           dpsi, dV_, dphi, g, dt, S, nrm);
      return hipGetLastError() == hipSuccess ? BDGX_OK : BDGX_EHIP;
    };
    {
      // normalise the start vector
      double *nrm = dnorm;
      if (c->cplx)
        hipLaunchKernelGGL((k_itp_pot<double, true>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                           dV_, (const double *)nullptr, 0.0, 0.0, S, nrm);
      else
        hipLaunchKernelGGL((k_itp_pot<double, false>), dim3(kItpBlocks), dim3(256), 0, c->stream,
                           dpsi, dV_, (const double *)nullptr, 0.0, 0.0, S, nrm);
      hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                         S * W, (const double *)nrm, dV);
    }
    for (uint64_t st = 0; st < nsteps && !rc; ++st) {
      rc = potential(nullptr);
      if (!rc) rc = itp_kinetic_t<double>(c, dpsi, dt);
      if (!rc) rc = potential(dnorm);
      if (!rc)
        hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                           S * W, (const double *)dnorm, dV);
    }
    if (!rc) rc = meanfield_t<double>(c, dpsi, dphi, g_dd);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(itpU64);
    (void)hipFree(itpU32);
    c->f64.U = save64;
    c->f32.U = save32;
  } else {
    if (c->cplx)
      hipLaunchKernelGGL((k_itp_pot<double, true>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                         dV_, (const double *)nullptr, 0.0, 0.0, S, dnorm);
    else
      hipLaunchKernelGGL((k_itp_pot<double, false>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                         dV_, (const double *)nullptr, 0.0, 0.0, S, dnorm);
    hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                       S * W, (const double *)dnorm, dV);
    for (uint64_t st = 0; st < nsteps && !rc; ++st) {
      if (c->cplx)
        hipLaunchKernelGGL((k_itp_pot<double, true>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                           dV_, (const double *)nullptr, g, dt, S, (double *)nullptr);
      else
        hipLaunchKernelGGL((k_itp_pot<double, false>), dim3(kItpBlocks), dim3(256), 0, c->stream,
                           dpsi, dV_, (const double *)nullptr, g, dt, S, (double *)nullptr);
      rc = itp_kinetic_t<double>(c, dpsi, dt);
      if (rc) break;
      if (c->cplx)
        hipLaunchKernelGGL((k_itp_pot<double, true>), dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                           dV_, (const double *)nullptr, g, dt, S, dnorm);
      else
        hipLaunchKernelGGL((k_itp_pot<double, false>), dim3(kItpBlocks), dim3(256), 0, c->stream,
                           dpsi, dV_, (const double *)nullptr, g, dt, S, dnorm);
      hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi,
                         S * W, (const double *)dnorm, dV);
    }
    ITP_HIP(hipMemsetAsync(dphi, 0, sizeof(double) * S, c->stream));
  }
  if (rc) {
    cleanup();
    return rc;
  }
  // mu = sum conj(psi) (T psi + Veff psi) dV
  ITP_HIP(hipMemcpyAsync(dT, dpsi, sizeof(double) * S * W, hipMemcpyDeviceToDevice, c->stream));
  rc = itp_kinetic_t<double>(c, dT, -1.0);  // dt < 0 selects the plain k^2/2 multiplier
  if (rc) {
    cleanup();
    return rc;
  }
  if (c->cplx)
    hipLaunchKernelGGL(k_itp_mu<true>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi, dT, dV_,
                       dip ? dphi : nullptr, g, S, dnorm + kItpBlocks);
  else
    hipLaunchKernelGGL(k_itp_mu<false>, dim3(kItpBlocks), dim3(256), 0, c->stream, dpsi, dT, dV_,
                       dip ? dphi : nullptr, g, S, dnorm + kItpBlocks);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream,
                     (const double *)(dnorm + kItpBlocks), dnorm + 2 * kItpBlocks);
  double acc = 0.0;
  ITP_HIP(hipMemcpyAsync(&acc, dnorm + 2 * kItpBlocks, sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  ITP_HIP(hipMemcpyAsync(h_psi, dpsi, sizeof(double) * S * W, hipMemcpyDeviceToHost, c->stream));
  ITP_HIP(hipStreamSynchronize(c->stream));
#undef ITP_HIP
  if (mu_out) *mu_out = acc * dV;
  cleanup();
  return BDGX_OK;
}

// Device blocks released by bdgx_free are kept for the next bdgx_malloc of
// the same size on the same device (a solve allocates the same sizes at
// the same points every time: ~0.3 ms per hipMalloc / hipFree of a large
// block, at locking and between precision phases).  A release waits for
// the device first, as hipFree does (work queued on any stream may still
// use the block).  At most kDevCacheMax bytes are kept; when hipMalloc runs
// out of memory the kept blocks are released and it is tried again, and
// bdgx_mem_info counts kept blocks as free.

int bdgx_malloc(void **p, size_t bytes) {
  if (!p) return BDGX_EINVAL;
  const size_t b = bytes ? bytes : 1;
  int dev = 0;
  BDGX_HIP(hipGetDevice(&dev));
  DevCache &c = dev_cache();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.freeb.find({dev, b});
    if (it != c.freeb.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      if (dev >= 0 && dev < 64) c.kept[dev] -= b;
      c.live[*p] = {dev, b};
      return BDGX_OK;
    }
  }
  hipError_t e = hipMalloc(p, b);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    release_kept(dev);
    e = hipMalloc(p, b);
  }
  BDGX_HIP(e);
  std::lock_guard<std::mutex> g(c.mu);
  c.live[*p] = {dev, b};
  return BDGX_OK;
}
int bdgx_mem_info(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return BDGX_EINVAL;
  BDGX_HIP(hipMemGetInfo(free_bytes, total_bytes));
  int dev = 0;
  BDGX_HIP(hipGetDevice(&dev));
  DevCache &c = dev_cache();
  std::lock_guard<std::mutex> g(c.mu);
  if (dev >= 0 && dev < 64) *free_bytes += c.kept[dev];
  return BDGX_OK;
}
int bdgx_free(void *p) {
  if (!p) return BDGX_OK;
  DevCache &c = dev_cache();
  std::pair<int, size_t> key{-1, 0};
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it != c.live.end()) {
      key = it->second;
      c.live.erase(it);
    }
  }
  if (key.first < 0 || key.first >= 64) {  // not from bdgx_malloc
    BDGX_HIP(hipFree(p));
    return BDGX_OK;
  }
  BDGX_HIP(hipDeviceSynchronize());
  {
    std::lock_guard<std::mutex> g(c.mu);
    if (c.kept[key.first] + key.second <= kDevCacheMax) {
      c.freeb[key].push_back(p);
      c.kept[key.first] += key.second;
      return BDGX_OK;
    }
  }
  BDGX_HIP(hipFree(p));
  return BDGX_OK;
}
int bdgx_host_alloc(void **p, size_t bytes) {
  if (!p) return BDGX_EINVAL;
  *p = nullptr;
  BDGX_HIP(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
  return BDGX_OK;
}
int bdgx_host_free(void *p) {
  if (p) BDGX_HIP(hipHostFree(p));
  return BDGX_OK;
}
int bdgx_memcpy_h2d(bdgx_ctx *c, void *d, const void *s, size_t n) {
  BDGX_HIP(hipMemcpyAsync(d, s, n, hipMemcpyHostToDevice, c->stream));
  BDGX_HIP(hipStreamSynchronize(c->stream));
  return BDGX_OK;
}
int bdgx_memcpy_h2d_async(bdgx_ctx *c, void *d, const void *s, size_t n) {
  if (!c || (n && (!d || !s))) return BDGX_EINVAL;
  BDGX_HIP(hipMemcpyAsync(d, s, n, hipMemcpyHostToDevice, c->stream));
  return BDGX_OK;
}
int bdgx_memcpy_d2h(bdgx_ctx *c, void *d, const void *s, size_t n) {
  BDGX_HIP(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToHost, c->stream));
  BDGX_HIP(hipStreamSynchronize(c->stream));
  return BDGX_OK;
}
int bdgx_memcpy_d2h_pair(bdgx_ctx *c, void *d1, const void *s1, void *d2, const void *s2,
                         size_t n) {
  if (!c) return BDGX_EINVAL;
  if (!c->aux) BDGX_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  // the second copy waits for the work queued so far on the context stream
  hipEvent_t ready;
  BDGX_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ready, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->aux, ready, 0);
  if (e == hipSuccess) e = hipMemcpyAsync(d1, s1, n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d2, s2, n, hipMemcpyDeviceToHost, c->aux);
  const hipError_t e1 = hipStreamSynchronize(c->stream), e2 = hipStreamSynchronize(c->aux);
  (void)hipEventDestroy(ready);
  BDGX_HIP(e);
  BDGX_HIP(e1);
  BDGX_HIP(e2);
  return BDGX_OK;
}
int bdgx_fill_uniform_xorshift(bdgx_ctx *c, void *dX, uint64_t ld, uint64_t nrows,
                               uint64_t ncols, uint64_t steps_per_col, uint32_t seed) {
  if (!c || !dX || ld < nrows || nrows == 0 || ncols == 0 || ncols > 65535 ||
      nrows > (1ull << kJumpBits) || steps_per_col > (1ull << kJumpBits))
    return BDGX_EINVAL;
  // T^(2^b), b < kJumpBits
  std::vector<uint32_t> pw(32 * kJumpBits);
  Gf2 t;
  for (int b = 0; b < 32; ++b) t.col[b] = xs_step(1u << b);
  for (int b = 0; b < kJumpBits; ++b) {
    memcpy(&pw[32 * b], t.col, sizeof(t.col));
    t = gf2_mul(t, t);
  }
  // column start states: s_0 = seed, s_{j+1} = T^steps_per_col s_j
  std::vector<uint32_t> st(ncols);
  uint32_t s0 = seed;
  for (uint64_t j = 0; j < ncols; ++j) {
    st[j] = s0;
    for (int b = 0; b < kJumpBits; ++b)
      if (steps_per_col >> b & 1ull) {
        Gf2 m;
        memcpy(m.col, &pw[32 * b], sizeof(m.col));
        s0 = gf2_apply(m, s0);
      }
  }
  uint32_t *dtab = nullptr;
  const size_t tb = sizeof(uint32_t) * (pw.size() + st.size());
  BDGX_HIP(hipMalloc((void **)&dtab, tb));
  int rc = BDGX_OK;
  if (hipMemcpyAsync(dtab, pw.data(), sizeof(uint32_t) * pw.size(), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess ||
      hipMemcpyAsync(dtab + pw.size(), st.data(), sizeof(uint32_t) * st.size(),
                     hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc) {
    const unsigned gx = (unsigned)((nrows + 256 * 64 - 1) / (256 * 64));
    const dim3 grid(gx, (unsigned)ncols);
    if (c->cplx)
      hipLaunchKernelGGL(k_fill_uniform<true>, grid, dim3(256), 0, c->stream, (double *)dX,
                         (long long)ld, (long long)nrows, dtab + pw.size(), dtab, kJumpBits);
    else
      hipLaunchKernelGGL(k_fill_uniform<false>, grid, dim3(256), 0, c->stream, (double *)dX,
                         (long long)ld, (long long)nrows, dtab + pw.size(), dtab, kJumpBits);
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  // the tables must outlive the kernel: wait, then free
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = BDGX_EHIP;
  (void)hipFree(dtab);
  return rc;
}

int bdgx_memset(bdgx_ctx *c, void *d, int v, size_t n) {
  BDGX_HIP(hipMemsetAsync(d, v, n, c->stream));
  return BDGX_OK;
}
int bdgx_sync(bdgx_ctx *c) {
  BDGX_HIP(hipStreamSynchronize(c->stream));
  return BDGX_OK;
}
void *bdgx_stream(bdgx_ctx *c) { return (void *)c->stream; }
int bdgx_timer_begin(bdgx_ctx *c) {
  BDGX_HIP(hipEventRecord(c->ev0, c->stream));
  return BDGX_OK;
}
int bdgx_timer_end(bdgx_ctx *c, float *ms) {
  BDGX_HIP(hipEventRecord(c->ev1, c->stream));
  BDGX_HIP(hipEventSynchronize(c->ev1));
  BDGX_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return BDGX_OK;
}
struct bdgx_graph {
  hipGraphExec_t exec;
  unsigned gen;
};

int bdgx_graph_begin(bdgx_ctx *c) {
  if (!c) return BDGX_EINVAL;
  BDGX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  return BDGX_OK;
}

int bdgx_graph_end(bdgx_ctx *c, bdgx_graph **out) {
  if (!c) return BDGX_EINVAL;
  if (out) *out = nullptr;
  hipGraph_t g = nullptr;
  BDGX_HIP(hipStreamEndCapture(c->stream, &g));
  if (!out) {  // abandon the capture
    if (g) (void)hipGraphDestroy(g);
    return BDGX_OK;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return BDGX_EHIP;
  *out = new bdgx_graph{exec, c->ws_gen};
  return BDGX_OK;
}

int bdgx_graph_launch(bdgx_ctx *c, const bdgx_graph *g) {
  if (!c || !g) return BDGX_EINVAL;
  if (g->gen != c->ws_gen) return BDGX_ESTATE;  // workspace moved since capture
  BDGX_HIP(hipGraphLaunch(g->exec, c->stream));
  return BDGX_OK;
}

void bdgx_graph_destroy(bdgx_graph *g) {
  if (!g) return;
  (void)hipGraphExecDestroy(g->exec);
  delete g;
}

int bdgx_set_chunk_bytes(bdgx_ctx *c, uint64_t bytes) {
  if (!c) return BDGX_EINVAL;
  c->chunk_bytes = (size_t)bytes;
  return BDGX_OK;
}

int bdgx_reserve(bdgx_ctx *c, uint64_t spec_bytes, uint64_t scratch_bytes) {
  if (!c) return BDGX_EINVAL;
  int rc = ensure_spec(c, (size_t)spec_bytes);
  if (!rc) rc = ensure_scratch(c, (size_t)scratch_bytes);
  return rc;
}

// Process-wide record of tuned launch variants, keyed by device, grid, psi0
// kind, precision and batch width (like planner wisdom): a new context for a
// shape already tuned in this process takes the recorded choice instead of
// timing again.  Guarded by a mutex (contexts may tune from several threads).
namespace {
struct TuneRecord {
  int device, cplx, prec;
  int N[3];
  uint64_t nbatch;
  int var[5];
};
std::mutex g_tune_lock;
std::vector<TuneRecord> g_tune_records;

bool tune_key_eq(const TuneRecord &r, const bdgx_ctx *c, int prec, uint64_t nbatch) {
  return r.device == c->device && r.cplx == c->cplx && r.prec == prec && r.nbatch == nbatch &&
         r.N[0] == c->geo.nx && r.N[1] == c->geo.ny && r.N[2] == c->geo.nz;
}
}  // namespace

int bdgx_autotune(bdgx_ctx *c, int prec, uint64_t nbatch, int force) {
  if (!c || (prec != BDGX_FP64 && prec != BDGX_FP32)) return BDGX_EINVAL;
  const int pi = prec == BDGX_FP64 ? 0 : 1;
  if (c->tuned[pi] && !force) return BDGX_OK;
  if (!force) {
    std::lock_guard<std::mutex> g(g_tune_lock);
    for (const TuneRecord &r : g_tune_records)
      if (tune_key_eq(r, c, prec, nbatch)) {
        for (int k = 0; k < 5; ++k) c->tune_var[pi][k] = r.var[k];
        c->tuned[pi] = 1;
        return BDGX_OK;
      }
  }
  const int rc = autotune_dispatch(c, prec, nbatch);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_tune_lock);
  TuneRecord rec;
  rec.device = c->device;
  rec.cplx = c->cplx;
  rec.prec = prec;
  rec.N[0] = c->geo.nx;
  rec.N[1] = c->geo.ny;
  rec.N[2] = c->geo.nz;
  rec.nbatch = nbatch;
  for (int k = 0; k < 5; ++k) rec.var[k] = c->tune_var[pi][k];
  for (TuneRecord &r : g_tune_records)
    if (tune_key_eq(r, c, prec, nbatch)) {
      r = rec;
      return BDGX_OK;
    }
  g_tune_records.push_back(rec);
  return BDGX_OK;
}

int bdgx_get_launch_variant(const bdgx_ctx *c, int kind, int prec, int *variant) {
  if (!c || !variant || kind < 0 || kind > 4 || (prec != BDGX_FP64 && prec != BDGX_FP32))
    return BDGX_EINVAL;
  *variant = c->tune_var[prec == BDGX_FP64 ? 0 : 1][kind];
  return BDGX_OK;
}

int bdgx_set_launch_variant(bdgx_ctx *c, int kind, int prec, int variant) {
  if (!c || kind < 0 || kind > 4 || (prec != BDGX_FP64 && prec != BDGX_FP32) || variant < 0 ||
      variant > (kind == BDGX_TUNE_Z ? 3 : kind <= BDGX_TUNE_PLANE_INV ? 7 : 3))
    return BDGX_EINVAL;
  c->tune_var[prec == BDGX_FP64 ? 0 : 1][kind] = variant;
  return BDGX_OK;
}

int bdgx_bench_pass(bdgx_ctx *c, int pass, int prec, uint64_t nbatch, int iters, float *ms,
                    double *bytes) {
  if (!c || !ms || !bytes || nbatch == 0 || iters < 0) return BDGX_EINVAL;
  return bench_pass_dispatch(c, pass, prec, nbatch, iters, ms, bytes);
}

uint64_t bdgx_workspace_bytes(const bdgx_ctx *c) {
  return (uint64_t)(c->spec_bytes + c->scratch_bytes);
}

}  // extern "C"
