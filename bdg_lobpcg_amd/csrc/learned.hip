// Learned preconditioner for the iLOBPCG search directions (FP32, real psi0).
//
//   N(r) = T'_M(V),  V = U + s0 sig sum_h v_h tanh(sum_nb w_{h,nb} U(x+nb)/(s0 sig) + d_h),
//   U = T'_K(r),  T'_X v = a_X * F^-1[ g_X * F(b_X * v) ],  X in {K, M}
//   a_X = p_X exp(alpha_X . phi),  b_X = p_X exp(beta_X . phi),  phi = (rho, rho^2)
//   g_X(k) = exp(interp(kappa_X, |k|)) / (S * safe(mu + k^2/2))
//   sig = rms(r) per column,  s0 = mean(p_K^2) / safe(mu)
//
// A spectral layer, a 3x3x3 stencil layer with HID hidden channels and a
// residual connection, and a second spectral layer.  The spectral layers
// generalise the analytic preconditioner T_M T_K (reference
// src/precond_impl.inc:21-111; N is exactly T_M T_K at theta = 0 with v = 0)
// and run through the same FFT pipeline as every operator (pre-weight,
// k-space table, post-weight).  The stencil acts on the already smoothed U and
// its output is smoothed again by T'_M, so it cannot put unfiltered high-k
// content into a search direction; N is positively homogeneous in r.  The
// stencil kernels read the 27 neighbours of a point once per hidden channel
// and never store the hidden activations (the adjoint stores one channel at a
// time).  The parameter gradient recomputes the forward and runs the adjoint
// (the filters are self-adjoint because g is real and even in k): two more
// table filters, two pairs of forward spectra for the kappa gradients, the
// transposed stencil, and reductions accumulated in FP64, all on the device.
#include <math.h>
#include <string.h>

#include <vector>

#include "bdgx.h"
#include "bdgx_internal.hpp"
#include "fft_core.hpp"

namespace bdgx {

int lp_filter_f32(bdgx_ctx *c, int w, int mode, const float *X, float *Y, long long ncols,
                  long long ld);
int lp_spectra_f32(bdgx_ctx *c, int w, const float *X2, long long ncols, long long ld2);

constexpr int LPK = BDGX_LP_NKNOT;
constexpr int LPH = BDGX_LP_HIDDEN;
constexpr int LP_OFF_W = 8 + 2 * LPK;
constexpr int LP_NCONV = 29 * LPH;

struct LpCoef {
  float aK[2], bK[2], aM[2], bM[2];
  float kK[LPK], kM[LPK];
};

// ---------------------------------------------------------------------------
// setup kernels
__global__ void k_lp_rho(const float *wf, long long S, float *rho) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S;
       i += (long long)gridDim.x * blockDim.x)
    rho[i] = wf[i] * wf[i];
}

// max of a non-negative array into *out (float bit patterns order like ints)
__global__ void k_lp_max(const float *x, long long n, unsigned int *out) {
  __shared__ float red[256];
  float m = 0.f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, x[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// out[0] += sum x^2
__global__ void k_lp_sumsq(const float *x, long long n, double *out) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    s += (double)x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

__global__ void k_lp_weights(LpCoef cf, const float *pK, const float *pM, float *rho,
                             const unsigned int *rmax, long long S, float *aK, float *bK,
                             float *aM, float *bM) {
  const float inv = 1.0f / fmaxf(__uint_as_float(*rmax), 1e-30f);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S;
       i += (long long)gridDim.x * blockDim.x) {
    const float r = rho[i] * inv;
    rho[i] = r;
    const float r2 = r * r;
    aK[i] = pK[i] * expf(cf.aK[0] * r + cf.aK[1] * r2);
    bK[i] = pK[i] * expf(cf.bK[0] * r + cf.bK[1] * r2);
    aM[i] = pM[i] * expf(cf.aM[0] * r + cf.aM[1] * r2);
    bM[i] = pM[i] * expf(cf.bM[0] * r + cf.bM[1] * r2);
  }
}

__device__ __forceinline__ void lp_knot(float kabs, float kmax, int &i, float &f) {
  const float t = kmax > 0.f ? kabs / kmax * (float)(LPK - 1) : 0.f;
  i = (int)floorf(t);
  if (i > LPK - 2) i = LPK - 2;
  if (i < 0) i = 0;
  f = t - (float)i;
}

__global__ void k_lp_tables(LpCoef cf, Geo g, const float *kx2, const float *ky2,
                            const float *kz2, double mu, float kmax, float *gK, float *gM) {
  const long long total = g.Kk;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int kx = (int)(e % g.n0k);
    const long long r = e / g.n0k;
    const int y = (int)(r % g.ny), z = (int)(r / g.ny);
    const double k2 = (double)kx2[kx] + (double)ky2[y] + (double)kz2[z];
    double d = mu + 0.5 * k2;
    if (fabs(d) < 1e-8) d = 1e-8;
    const double base = 1.0 / ((double)g.S * d);
    int i;
    float f;
    lp_knot((float)sqrt(k2), kmax, i, f);
    gK[e] = (float)(base * exp((double)((1.f - f) * cf.kK[i] + f * cf.kK[i + 1])));
    gM[e] = (float)(base * exp((double)((1.f - f) * cf.kM[i] + f * cf.kM[i + 1])));
  }
}

// Column rms in two launches: part[col * LP_NCHUNK + chunk] = sum of squares
// over one of LP_NCHUNK contiguous chunks (grid LP_NCHUNK x ncols), then
// sig[col] = sqrt(sum_chunk part / S) in a fixed order (deterministic).
constexpr int LP_NCHUNK = 64;

__global__ void __launch_bounds__(256) k_lp_colsq(const float *R, long long ld, long long S,
                                                  double *part) {
  const float *r = R + (long long)blockIdx.y * ld;
  const long long len = (S + LP_NCHUNK - 1) / LP_NCHUNK;
  const long long i0 = (long long)blockIdx.x * len;
  const long long i1 = i0 + len < S ? i0 + len : S;
  double s = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += (double)r[i] * r[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * LP_NCHUNK + blockIdx.x] = red[0];
}

__global__ void k_lp_colfinal(const double *part, int ncols, long long S, float *sig) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  double s = 0.0;
  for (int k = 0; k < LP_NCHUNK; ++k) s += part[(long long)col * LP_NCHUNK + k];
  sig[col] = (float)fmax(sqrt(s / (double)S), 1e-30);
}

// ---------------------------------------------------------------------------
// stencil layer
// The 27 periodic neighbours of point i: index of offset (dz, dy, dx) is
// row[(dz + 1) * 3 + dy + 1] + col[dx + 1]; neighbour n = (dz+1)*9 + (dy+1)*3
// + (dx+1) and the opposite offset is 26 - n.  One division chain per point.
struct LpNb {
  long long row[9];
  int col[3];
  __device__ __forceinline__ LpNb(long long i, const Geo &g) {
    const int x = (int)(i % g.nx);
    const long long r = i / g.nx;
    const int y = (int)(r % g.ny), z = (int)(r / g.ny);
    const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
    const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
    const int zs[3] = {z == 0 ? g.nz - 1 : z - 1, z, z == g.nz - 1 ? 0 : z + 1};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      col[a] = xs[a];
#pragma unroll
      for (int b = 0; b < 3; ++b) row[a * 3 + b] = ((long long)zs[a] * g.ny + ys[b]) * g.nx;
    }
  }
  __device__ __forceinline__ long long at(int n) const { return row[n / 3] + col[n % 3]; }
};

// tanh from one exponential: 1 - 2 / (exp(2z) + 1), exact limits +-1
__device__ __forceinline__ float lp_tanh(float z) {
  const float e = __expf(2.f * z);
  return 1.f - 2.f / (e + 1.f);
}

// V = U + s0 sig_col sum_h v_h tanh(z_h),  z_h = sum_nb w_{h,nb} U(x+nb)/(s0 sig_col) + d_h
__global__ void __launch_bounds__(256) k_lp_stencil_fwd(const float *U, int ncols, Geo g,
                                                        const float *prm, float s0,
                                                        const float *sig, float *V) {
  __shared__ float sp[LP_NCONV];
  for (int i = threadIdx.x; i < LP_NCONV; i += blockDim.x) sp[i] = prm[i];
  __syncthreads();
  const float *w = sp, *d = sp + 27 * LPH, *v = sp + 28 * LPH;
  const long long total = g.S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / g.S);
    const long long i = e - (long long)col * g.S;
    const float *u = U + (long long)col * g.S;
    const float sc = s0 * sig[col], isc = 1.f / sc;
    const LpNb nbi(i, g);
    float nb[27];
#pragma unroll
    for (int n = 0; n < 27; ++n) nb[n] = u[nbi.at(n)] * isc;
    float s = 0.f;
#pragma unroll 4
    for (int h = 0; h < LPH; ++h) {
      float z = d[h];
#pragma unroll
      for (int q = 0; q < 27; ++q) z += w[h * 27 + q] * nb[q];
      s += v[h] * lp_tanh(z);
    }
    V[e] = u[i] + sc * s;
  }
}

// Parameter gradient of the stencil layer for hidden channel h = blockIdx.y,
// and Q = dV v_h (1 - tanh(z_h)^2) for channel `qh` (when Q != null), the
// input of the transposed stencil.
__global__ void __launch_bounds__(256) k_lp_stencil_bwd(const float *U, const float *dV,
                                                        int ncols, Geo g, const float *prm,
                                                        float s0, const float *sig,
                                                        double *grad, int qh, float *Q) {
  const int h = Q ? qh : blockIdx.y;
  const float *w = prm + h * 27;
  const float dh = prm[27 * LPH + h], vh = prm[28 * LPH + h];
  float gw[27], gd = 0.f, gv = 0.f;
#pragma unroll
  for (int q = 0; q < 27; ++q) gw[q] = 0.f;
  const long long total = g.S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / g.S);
    const long long i = e - (long long)col * g.S;
    const float *u = U + (long long)col * g.S;
    const float sc = s0 * sig[col], isc = 1.f / sc;
    const LpNb nbi(i, g);
    float nb[27];
#pragma unroll
    for (int n = 0; n < 27; ++n) nb[n] = u[nbi.at(n)] * isc;
    float z = dh;
#pragma unroll
    for (int q = 0; q < 27; ++q) z += w[q] * nb[q];
    const float t = lp_tanh(z);
    const float dvv = dV[e];
    const float q = dvv * vh * (1.f - t * t);
    if (Q) {
      Q[e] = q;
      continue;
    }
    gv += sc * dvv * t;
    gd += sc * q;
#pragma unroll
    for (int k = 0; k < 27; ++k) gw[k] += sc * q * nb[k];
  }
  if (Q) return;
  // block sum of the 29 partials through LDS
  __shared__ float red[256];
  for (int k = 0; k < 29; ++k) {
    const float val = k < 27 ? gw[k] : (k == 27 ? gd : gv);
    red[threadIdx.x] = val;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const int o = k < 27 ? LP_OFF_W + h * 27 + k
                           : (k == 27 ? LP_OFF_W + 27 * LPH + h : LP_OFF_W + 28 * LPH + h);
      atomicAdd(grad + o, (double)red[0]);
    }
    __syncthreads();
  }
}

// dU(y) += sum_nb w_{h,nb} Q(y - nb)   (transposed stencil of channel h)
__global__ void __launch_bounds__(256) k_lp_stencil_t(const float *Q, int ncols, Geo g,
                                                      const float *prm, int h, float *dU) {
  __shared__ float ws[27];
  if (threadIdx.x < 27) ws[threadIdx.x] = prm[h * 27 + threadIdx.x];
  __syncthreads();
  const long long total = g.S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / g.S);
    const long long i = e - (long long)col * g.S;
    const float *q = Q + (long long)col * g.S;
    const LpNb nbi(i, g);
    float acc = 0.f;
#pragma unroll
    for (int n = 0; n < 27; ++n) acc += ws[n] * q[nbi.at(26 - n)];
    dU[e] += acc;
  }
}

// ---------------------------------------------------------------------------
// reductions and elementwise helpers
// out[0..1] += sum_{x, col} P Q wgt (rho, rho^2)   (wgt may be null)
__global__ void __launch_bounds__(256) k_lp_pointgrad(const float *P, const float *Q,
                                                      const float *wgt, const float *rho,
                                                      long long S, int ncols, long long ld,
                                                      double *out) {
  double s0 = 0.0, s1 = 0.0;
  const long long total = S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / S);
    const long long i = e - (long long)col * S;
    double t = (double)P[i + col * ld] * (double)Q[i + col * ld];
    if (wgt) t *= (double)wgt[i];
    const double r = rho[i];
    s0 += t * r;
    s1 += t * r * r;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) atomicAdd(out + threadIdx.x, red[threadIdx.x][0]);
}

__global__ void k_lp_mulw(float *X, const float *w, long long S, int ncols, long long ld) {
  const long long total = S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / S);
    const long long i = e - (long long)col * S;
    X[i + col * ld] *= w[i];
  }
}

// X2[:, col] = [A[:, col]; B[:, col]] (ld2 = 2 S)
__global__ void k_lp_pack2(const float *A, const float *B, long long S, int ncols, long long ld,
                           float *X2) {
  const long long total = S * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int col = (int)(e / S);
    const long long i = e - (long long)col * S;
    X2[2 * S * col + i] = A[i + col * ld];
    X2[2 * S * col + S + i] = B[i + col * ld];
  }
}

// kappa gradient: out[knot] += sum_{k, col} wt(kx) Re(conj(A_k) B_k) g_k * basis
__global__ void __launch_bounds__(256) k_lp_knotgrad(const cx<float> *spec, int ncols, Geo g,
                                                     const float *kx2, const float *ky2,
                                                     const float *kz2, float kmax,
                                                     const float *gt, double *out) {
  __shared__ double hist[LPK];
  if (threadIdx.x < LPK) hist[threadIdx.x] = 0.0;
  __syncthreads();
  const long long total = g.Kk * ncols;
  const long long chan = g.Kk * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long kidx = e % g.Kk;
    const int kx = (int)(kidx % g.n0k);
    const long long r = kidx / g.n0k;
    const int y = (int)(r % g.ny), z = (int)(r / g.ny);
    const cx<float> A = spec[e], B = spec[chan + e];
    const bool paired = kx > 0 && !((g.nx & 1) == 0 && kx == g.nx / 2);
    const double q = (paired ? 2.0 : 1.0) * ((double)A.re * B.re + (double)A.im * B.im) *
                     (double)gt[kidx];
    int i;
    float f;
    lp_knot(sqrtf(kx2[kx] + ky2[y] + kz2[z]), kmax, i, f);
    atomicAdd(&hist[i], (1.0 - f) * q);
    atomicAdd(&hist[i + 1], (double)f * q);
  }
  __syncthreads();
  if (threadIdx.x < LPK) atomicAdd(out + threadIdx.x, hist[threadIdx.x]);
}

// ---------------------------------------------------------------------------
static int lp_alloc(float **p, size_t n) {
  if (!*p) BDGX_HIP(hipMalloc((void **)p, sizeof(float) * (n ? n : 1)));
  return BDGX_OK;
}

static int ensure_lp_tmp(bdgx_ctx *c, size_t bytes) {
  if (bytes <= c->lp_tmp_bytes) return BDGX_OK;
  if (c->lp_tmp) {
    BDGX_HIP(hipStreamSynchronize(c->stream));
    BDGX_HIP(hipFree(c->lp_tmp));
    c->lp_tmp = nullptr;
    c->lp_tmp_bytes = 0;
  }
  BDGX_HIP(hipMalloc(&c->lp_tmp, bytes));
  c->lp_tmp_bytes = bytes;
  c->ws_gen++;
  return BDGX_OK;
}

static float lp_kmax(const bdgx_ctx *c) {
  double s = 0.0;
  for (int d = 0; d < c->dim; ++d) {
    const double kn = M_PI * (double)c->N[d] / c->L[d];  // largest |k_d| on the grid
    s += kn * kn;
  }
  return (float)sqrt(s);
}

static unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace bdgx

using namespace bdgx;

// sig[col] = rms of column col of R (S x n, ld S)
static void lp_colnorm(const float *R, long long S, long long n, double *part, float *sig,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_lp_colsq, dim3(LP_NCHUNK, (unsigned)n), dim3(256), 0, st, R, S, S, part);
  hipLaunchKernelGGL(k_lp_colfinal, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st,
                     (const double *)part, (int)n, S, sig);
}

extern "C" {

int bdgx_lp_set_params(bdgx_ctx *c, const float *theta) {
  if (!c || !theta) return BDGX_EINVAL;
  if (c->cplx) return BDGX_EINVAL;
  if (!c->has_wf || !c->has_p) return BDGX_ESTATE;
  const long long S = c->geo.S;
  Fields<float> &F = c->f32;
  int rc = 0;
  for (int w = 0; w < 2 && !rc; ++w) {
    rc = lp_alloc(&F.lpa[w], (size_t)S);
    if (!rc) rc = lp_alloc(&F.lpb[w], (size_t)S);
    if (!rc) rc = lp_alloc(&F.lpg[w], (size_t)c->geo.Kk);
  }
  if (!rc) rc = lp_alloc(&c->lp_rho, (size_t)S);
  if (!rc) rc = lp_alloc(&c->lp_conv, (size_t)LP_NCONV);
  if (rc) return rc;
  LpCoef cf;
  memcpy(cf.aK, theta + 0, 2 * sizeof(float));
  memcpy(cf.bK, theta + 2, 2 * sizeof(float));
  memcpy(cf.aM, theta + 4, 2 * sizeof(float));
  memcpy(cf.bM, theta + 6, 2 * sizeof(float));
  memcpy(cf.kK, theta + 8, LPK * sizeof(float));
  memcpy(cf.kM, theta + 8 + LPK, LPK * sizeof(float));
  BDGX_HIP(hipStreamSynchronize(c->stream));  // earlier applies may read the tables
  BDGX_HIP(hipMemcpyAsync(c->lp_conv, theta + LP_OFF_W, sizeof(float) * LP_NCONV,
                          hipMemcpyHostToDevice, c->stream));
  rc = ensure_scratch(c, 2 * sizeof(double));
  if (rc) return rc;
  unsigned int *dmax = reinterpret_cast<unsigned int *>(c->scratch);
  double *dsum = reinterpret_cast<double *>(c->scratch) + 1;
  BDGX_HIP(hipMemsetAsync(c->scratch, 0, 2 * sizeof(double), c->stream));
  hipLaunchKernelGGL(k_lp_rho, dim3(grid_for(S)), dim3(256), 0, c->stream, F.wf, S, c->lp_rho);
  hipLaunchKernelGGL(k_lp_max, dim3(grid_for(S)), dim3(256), 0, c->stream, c->lp_rho, S, dmax);
  hipLaunchKernelGGL(k_lp_sumsq, dim3(grid_for(S)), dim3(256), 0, c->stream, F.pK, S, dsum);
  hipLaunchKernelGGL(k_lp_weights, dim3(grid_for(S)), dim3(256), 0, c->stream, cf, F.pK, F.pM,
                     c->lp_rho, dmax, S, F.lpa[0], F.lpb[0], F.lpa[1], F.lpb[1]);
  hipLaunchKernelGGL(k_lp_tables, dim3(grid_for(c->geo.Kk)), dim3(256), 0, c->stream, cf, c->geo,
                     F.kx2[0], F.kx2[1], F.kx2[2], c->mu, lp_kmax(c), F.lpg[0], F.lpg[1]);
  BDGX_HIP(hipGetLastError());
  double hsum = 0.0;
  BDGX_HIP(hipMemcpyAsync(&hsum, dsum, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BDGX_HIP(hipStreamSynchronize(c->stream));
  const double smu = fabs(c->mu) < 1e-8 ? 1e-8 : c->mu;
  c->lp_s0 = hsum / (double)S / smu;
  c->has_lp = 1;
  return BDGX_OK;
}

int bdgx_lp_apply(bdgx_ctx *c, int prec, const void *dR, void *dW, uint64_t ncols, uint64_t ld) {
  if (!c || prec != BDGX_FP32 || c->cplx || dR == dW || ld != (uint64_t)c->geo.S)
    return BDGX_EINVAL;
  if (!c->has_lp) return BDGX_ESTATE;
  if (ncols == 0) return BDGX_OK;
  const long long S = c->geo.S, n = (long long)ncols;
  // column partial sums (n x LP_NCHUNK doubles), U, V (S x n each), sig (n)
  int rc = ensure_lp_tmp(c, sizeof(double) * (size_t)n * LP_NCHUNK +
                                sizeof(float) * ((size_t)S * n * 2 + n));
  if (rc) return rc;
  double *part = reinterpret_cast<double *>(c->lp_tmp);
  float *U = reinterpret_cast<float *>(part + n * LP_NCHUNK);
  float *V = U + S * n, *sig = V + S * n;
  const float *R = reinterpret_cast<const float *>(dR);
  float *W = reinterpret_cast<float *>(dW);
  lp_colnorm(R, S, n, part, sig, c->stream);
  rc = lp_filter_f32(c, 0, 0, R, U, n, S);
  if (rc) return rc;
  hipLaunchKernelGGL(k_lp_stencil_fwd, dim3(grid_for(S * n)), dim3(256), 0, c->stream,
                     (const float *)U, (int)n, c->geo, c->lp_conv, (float)c->lp_s0,
                     (const float *)sig, V);
  BDGX_HIP(hipGetLastError());
  return lp_filter_f32(c, 1, 0, V, W, n, S);
}

int bdgx_lp_backward(bdgx_ctx *c, const float *dR, const float *dGW, uint64_t ncols, uint64_t ld,
                     float *h_grad) {
  if (!c || !dR || !dGW || !h_grad || c->cplx || ld != (uint64_t)c->geo.S) return BDGX_EINVAL;
  if (!c->has_lp) return BDGX_ESTATE;
  if (ncols == 0) return BDGX_OK;
  const long long S = c->geo.S, n = (long long)ncols;
  Fields<float> &F = c->f32;
  // column partial sums (n x LP_NCHUNK doubles), U, V, Wsp, dV, dU, T, Q
  // (S x n each), X2 (2S x n), sig (n)
  int rc = ensure_lp_tmp(c, sizeof(double) * (size_t)n * LP_NCHUNK +
                                sizeof(float) * ((size_t)S * n * 9 + n));
  if (rc) return rc;
  double *part = reinterpret_cast<double *>(c->lp_tmp);
  float *U = reinterpret_cast<float *>(part + n * LP_NCHUNK);
  float *V = U + S * n, *Wsp = V + S * n, *dV = Wsp + S * n, *dU = dV + S * n;
  float *T = dU + S * n, *Q = T + S * n, *X2 = Q + S * n, *sig = X2 + 2 * S * n;
  rc = ensure_scratch(c, sizeof(double) * BDGX_LP_NPARAM);
  if (rc) return rc;
  double *dg = reinterpret_cast<double *>(c->scratch);
  BDGX_HIP(hipMemsetAsync(dg, 0, sizeof(double) * BDGX_LP_NPARAM, c->stream));
  const unsigned gb = grid_for(S * n);
  const float kmax = lp_kmax(c), s0 = (float)c->lp_s0;
  const hipStream_t st = c->stream;
  // forward intermediates
  lp_colnorm(dR, S, n, part, sig, st);
  rc = lp_filter_f32(c, 0, 0, dR, U, n, S);
  if (!rc) {
    hipLaunchKernelGGL(k_lp_stencil_fwd, dim3(gb), dim3(256), 0, st, (const float *)U, (int)n,
                       c->geo, c->lp_conv, s0, (const float *)sig, V);
    rc = lp_filter_f32(c, 1, 0, V, Wsp, n, S);
  }
  // M layer: alpha_M, beta_M, kappa_M; dV = b_M * C_M(a_M dW)
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, dGW, (const float *)Wsp,
                       (const float *)nullptr, c->lp_rho, S, (int)n, S, dg + 4);
    rc = lp_filter_f32(c, 1, 1, dGW, dV, n, S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, (const float *)V,
                       (const float *)dV, (const float *)F.lpb[1], c->lp_rho, S, (int)n, S, dg + 6);
    hipLaunchKernelGGL(k_lp_pack2, dim3(gb), dim3(256), 0, st, dGW, (const float *)V, S, (int)n, S,
                       X2);
    rc = lp_spectra_f32(c, 1, X2, n, 2 * S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_knotgrad, dim3(grid_for(c->geo.Kk * n)), dim3(256), 0, st,
                       reinterpret_cast<const cx<float> *>(c->spec), (int)n, c->geo, F.kx2[0],
                       F.kx2[1], F.kx2[2], kmax, (const float *)F.lpg[1], dg + 8 + LPK);
    hipLaunchKernelGGL(k_lp_mulw, dim3(gb), dim3(256), 0, st, dV, (const float *)F.lpb[1], S,
                       (int)n, S);
    // stencil layer: parameters, then dU = dV + transposed stencil of each channel
    hipLaunchKernelGGL(k_lp_stencil_bwd, dim3(gb > 1024 ? 1024 : gb, LPH), dim3(256), 0, st,
                       (const float *)U, (const float *)dV, (int)n, c->geo, c->lp_conv, s0,
                       (const float *)sig, dg, 0, (float *)nullptr);
    if (hipMemcpyAsync(dU, dV, sizeof(float) * S * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
      rc = BDGX_EHIP;
    for (int h = 0; h < LPH && !rc; ++h) {
      hipLaunchKernelGGL(k_lp_stencil_bwd, dim3(gb), dim3(256), 0, st, (const float *)U,
                         (const float *)dV, (int)n, c->geo, c->lp_conv, s0, (const float *)sig, dg,
                         h, Q);
      hipLaunchKernelGGL(k_lp_stencil_t, dim3(gb), dim3(256), 0, st, (const float *)Q, (int)n,
                         c->geo, c->lp_conv, h, dU);
    }
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  // K layer with upstream dU
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, (const float *)dU,
                       (const float *)U, (const float *)nullptr, c->lp_rho, S, (int)n, S, dg + 0);
    rc = lp_filter_f32(c, 0, 1, dU, T, n, S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, dR, (const float *)T,
                       (const float *)F.lpb[0], c->lp_rho, S, (int)n, S, dg + 2);
    hipLaunchKernelGGL(k_lp_pack2, dim3(gb), dim3(256), 0, st, (const float *)dU, dR, S, (int)n, S,
                       X2);
    rc = lp_spectra_f32(c, 0, X2, n, 2 * S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_knotgrad, dim3(grid_for(c->geo.Kk * n)), dim3(256), 0, st,
                       reinterpret_cast<const cx<float> *>(c->spec), (int)n, c->geo, F.kx2[0],
                       F.kx2[1], F.kx2[2], kmax, (const float *)F.lpg[0], dg + 8);
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  std::vector<double> hg(BDGX_LP_NPARAM, 0.0);
  if (!rc && hipMemcpyAsync(hg.data(), dg, sizeof(double) * BDGX_LP_NPARAM, hipMemcpyDeviceToHost,
                            st) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = BDGX_EHIP;
  if (rc) return rc;
  for (int i = 0; i < BDGX_LP_NPARAM; ++i) h_grad[i] = (float)hg[i];
  return BDGX_OK;
}

// One spectral layer on its own (layer 0: T'_K, 1: T'_M): the first-order
// solver's per-half preconditioner W_U = T'_K R_U, W_V = T'_M R_V.
int bdgx_lp_apply_layer(bdgx_ctx *c, int layer, const void *dR, void *dW, uint64_t ncols,
                        uint64_t ld) {
  if (!c || (layer != 0 && layer != 1) || c->cplx || !dR || !dW || dR == dW ||
      ld != (uint64_t)c->geo.S)
    return BDGX_EINVAL;
  if (!c->has_lp) return BDGX_ESTATE;
  if (ncols == 0) return BDGX_OK;
  return lp_filter_f32(c, layer, 0, reinterpret_cast<const float *>(dR),
                       reinterpret_cast<float *>(dW), (long long)ncols, c->geo.S);
}

// Parameter gradient of sum <dGW, T'_X R> for layer X alone (the K-layer
// part of bdgx_lp_backward with the layer's own offsets): alpha_X from
// dGW . W, beta_X from R . b_X C_X(a_X dGW), kappa_X from the spectra of
// (dGW, R).  Entries of the other layer and of the stencil are zero.
int bdgx_lp_backward_layer(bdgx_ctx *c, int layer, const float *dR, const float *dGW,
                           uint64_t ncols, uint64_t ld, float *h_grad) {
  if (!c || (layer != 0 && layer != 1) || !dR || !dGW || !h_grad || c->cplx ||
      ld != (uint64_t)c->geo.S)
    return BDGX_EINVAL;
  if (!c->has_lp) return BDGX_ESTATE;
  for (int i = 0; i < BDGX_LP_NPARAM; ++i) h_grad[i] = 0.f;
  if (ncols == 0) return BDGX_OK;
  const long long S = c->geo.S, n = (long long)ncols;
  Fields<float> &F = c->f32;
  // W, T (S x n each), X2 (2S x n)
  int rc = ensure_lp_tmp(c, sizeof(float) * (size_t)S * n * 4);
  if (rc) return rc;
  float *W = reinterpret_cast<float *>(c->lp_tmp);
  float *T = W + S * n, *X2 = T + S * n;
  rc = ensure_scratch(c, sizeof(double) * BDGX_LP_NPARAM);
  if (rc) return rc;
  double *dg = reinterpret_cast<double *>(c->scratch);
  BDGX_HIP(hipMemsetAsync(dg, 0, sizeof(double) * BDGX_LP_NPARAM, c->stream));
  const unsigned gb = grid_for(S * n);
  const hipStream_t st = c->stream;
  const int oa = layer == 0 ? 0 : 4, ob = oa + 2, ok = 8 + layer * LPK;
  rc = lp_filter_f32(c, layer, 0, dR, W, n, S);
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, dGW, (const float *)W,
                       (const float *)nullptr, c->lp_rho, S, (int)n, S, dg + oa);
    rc = lp_filter_f32(c, layer, 1, dGW, T, n, S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_pointgrad, dim3(gb), dim3(256), 0, st, dR, (const float *)T,
                       (const float *)F.lpb[layer], c->lp_rho, S, (int)n, S, dg + ob);
    hipLaunchKernelGGL(k_lp_pack2, dim3(gb), dim3(256), 0, st, dGW, dR, S, (int)n, S, X2);
    rc = lp_spectra_f32(c, layer, X2, n, 2 * S);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lp_knotgrad, dim3(grid_for(c->geo.Kk * n)), dim3(256), 0, st,
                       reinterpret_cast<const cx<float> *>(c->spec), (int)n, c->geo, F.kx2[0],
                       F.kx2[1], F.kx2[2], lp_kmax(c), (const float *)F.lpg[layer], dg + ok);
    if (hipGetLastError() != hipSuccess) rc = BDGX_EHIP;
  }
  std::vector<double> hg(BDGX_LP_NPARAM, 0.0);
  if (!rc && hipMemcpyAsync(hg.data(), dg, sizeof(double) * BDGX_LP_NPARAM, hipMemcpyDeviceToHost,
                            st) != hipSuccess)
    rc = BDGX_EHIP;
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = BDGX_EHIP;
  if (rc) return rc;
  for (int i = 0; i < BDGX_LP_NPARAM; ++i) h_grad[i] = (float)hg[i];
  return BDGX_OK;
}

}  // extern "C"
