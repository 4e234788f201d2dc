// Tall-skinny block kernels for the iLOBPCG step on gfx950: Gram products
// A^H B (optionally with B's halves swapped, giving the swap-operator Gram
// without materialising Swap*B), block combinations A*C, residuals and
// precision conversion.  Storage is FP64 or FP32 (real or interleaved
// complex); all reductions accumulate in FP64.
#include <string.h>

#include "bdgx_internal.hpp"
#include "fft_core.hpp"

namespace bdgx {

// ---------------------------------------------------------------------------
// Gram: partial[chunk][i + j*p] = sum_{r in chunk} conj(A[r,i]) B[r',j]
constexpr int GT = 64;   // output tile edge
constexpr int GK = 32;   // rows per LDS stage

template <typename T>
struct Elem;
template <>
struct Elem<double> {
  static constexpr bool cplx = false;
};
template <>
struct Elem<float> {
  static constexpr bool cplx = false;
};

template <typename R, bool CPLX>
struct GramArgs {
  const R *A, *B;
  long long lda, ldb, nrows, half;
  int p, q, swap_b, nchunk;
  long long rows_per_chunk;
  double *part;  // nchunk * p * q (complex: 2x)
};

template <typename R, bool CPLX>
__global__ void __launch_bounds__(256) k_gram(GramArgs<R, CPLX> a) {
  constexpr int W = CPLX ? 2 : 1;
  __shared__ double As[GK][GT * W + 1];
  __shared__ double Bs[GK][GT * W + 1];
  const int tilesP = (a.p + GT - 1) / GT;
  const int tilesQ = (a.q + GT - 1) / GT;
  const int tiles = tilesP * tilesQ;
  const int chunk = blockIdx.x / tiles;
  const int tt = blockIdx.x - chunk * tiles;
  const int tp = tt % tilesP, tq = tt / tilesP;
  const int i0 = tp * GT, j0 = tq * GT;
  const long long rs = (long long)chunk * a.rows_per_chunk;
  long long re = rs + a.rows_per_chunk;
  if (re > a.nrows) re = a.nrows;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4][W];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int w = 0; w < W; ++w) acc[u][v][w] = 0.0;
  for (long long r0 = rs; r0 < re; r0 += GK) {
    // stage GK rows x GT columns of A and B
    for (int e = threadIdx.x; e < GK * GT; e += blockDim.x) {
      const int rr = e % GK, cc = e / GK;
      const long long r = r0 + rr;
      const bool okr = r < re;
      const int ia = i0 + cc, jb = j0 + cc;
      long long rb = r;
      if (a.swap_b) rb = (r < a.half) ? r + a.half : r - a.half;
      if (CPLX) {
        double ar = 0, ai = 0, br = 0, bi = 0;
        if (okr && ia < a.p) {
          ar = (double)a.A[2 * (r + ia * a.lda)];
          ai = (double)a.A[2 * (r + ia * a.lda) + 1];
        }
        if (okr && jb < a.q) {
          br = (double)a.B[2 * (rb + jb * a.ldb)];
          bi = (double)a.B[2 * (rb + jb * a.ldb) + 1];
        }
        As[rr][2 * cc] = ar;
        As[rr][2 * cc + 1] = ai;
        Bs[rr][2 * cc] = br;
        Bs[rr][2 * cc + 1] = bi;
      } else {
        As[rr][cc] = (okr && ia < a.p) ? (double)a.A[r + ia * a.lda] : 0.0;
        Bs[rr][cc] = (okr && jb < a.q) ? (double)a.B[rb + jb * a.ldb] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      double av[4][W], bv[4][W];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < W; ++w) {
          av[u][w] = As[kk][(ty + 16 * u) * W + w];
          bv[u][w] = Bs[kk][(tx + 16 * u) * W + w];
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (CPLX) {
            // conj(a) * b
            acc[u][v][0] += av[u][0] * bv[v][0] + av[u][W - 1] * bv[v][W - 1];
            acc[u][v][W - 1] += av[u][0] * bv[v][W - 1] - av[u][W - 1] * bv[v][0];
          } else {
            acc[u][v][0] += av[u][0] * bv[v][0];
          }
        }
    }
    __syncthreads();
  }
  double *out = a.part + (long long)chunk * a.p * a.q * W;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
      if (i < a.p && j < a.q) {
#pragma unroll
        for (int w = 0; w < W; ++w) out[(i + (long long)j * a.p) * W + w] = acc[u][v][w];
      }
    }
}

// Sum of nchunk partial Grams: 64 outputs x 4 chunk lanes per workgroup,
// lanes combined in LDS (deterministic order).
__global__ void __launch_bounds__(256) k_reduce_chunks(const double *part, int nchunk,
                                                       long long len, double *out) {
  __shared__ double red[4][64];
  const int el = threadIdx.x & 63, lane = threadIdx.x >> 6;
  const long long e = blockIdx.x * 64LL + el;
  double s = 0.0;
  if (e < len)
    for (int c = lane; c < nchunk; c += 4) s += part[(long long)c * len + e];
  red[lane][el] = s;
  __syncthreads();
  if (lane == 0 && e < len) out[e] = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
}

template <typename R, bool CPLX>
static int gram_t(bdgx_ctx *c, const void *dA, uint64_t lda, uint64_t p,
                  const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows,
                  int swap_b, double *dG) {
  if (p == 0 || q == 0) return BDGX_OK;
  const int W = CPLX ? 2 : 1;
  const int tiles = (int)(((p + GT - 1) / GT) * ((q + GT - 1) / GT));
  int nchunk = 1024 / tiles;
  if (nchunk > 256) nchunk = 256;
  if (nchunk < 1) nchunk = 1;
  long long rpc = ((long long)nrows + nchunk - 1) / nchunk;
  rpc = ((rpc + GK - 1) / GK) * GK;
  if (rpc < GK) rpc = GK;
  nchunk = (int)(((long long)nrows + rpc - 1) / rpc);
  if (nchunk < 1) nchunk = 1;
  const size_t need = sizeof(double) * (size_t)nchunk * p * q * W;
  int rc = ensure_scratch(c, need);
  if (rc) return rc;
  GramArgs<R, CPLX> a;
  a.A = reinterpret_cast<const R *>(dA);
  a.B = reinterpret_cast<const R *>(dB);
  a.lda = (long long)lda;
  a.ldb = (long long)ldb;
  a.nrows = (long long)nrows;
  a.half = (long long)nrows / 2;
  a.p = (int)p;
  a.q = (int)q;
  a.swap_b = swap_b;
  a.nchunk = nchunk;
  a.rows_per_chunk = rpc;
  a.part = reinterpret_cast<double *>(c->scratch);
  hipLaunchKernelGGL((k_gram<R, CPLX>), dim3(tiles * nchunk), dim3(256), 0, c->stream, a);
  BDGX_HIP(hipGetLastError());
  const long long len = (long long)p * q * W;
  const int nb = (int)((len + 63) / 64);
  hipLaunchKernelGGL(k_reduce_chunks, dim3(nb), dim3(256), 0, c->stream,
                     (const double *)c->scratch, nchunk, len, dG);
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

int gram_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                  const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows,
                  int swap_b, double *dG) {
  if (prec == BDGX_FP64)
    return c->cplx ? gram_t<double, true>(c, dA, lda, p, dB, ldb, q, nrows, swap_b, dG)
                   : gram_t<double, false>(c, dA, lda, p, dB, ldb, q, nrows, swap_b, dG);
  if (prec == BDGX_FP32)
    return c->cplx ? gram_t<float, true>(c, dA, lda, p, dB, ldb, q, nrows, swap_b, dG)
                   : gram_t<float, false>(c, dA, lda, p, dB, ldb, q, nrows, swap_b, dG);
  return BDGX_EINVAL;
}

// ---------------------------------------------------------------------------
// combine: Y[r, j] = sum_i A[r, i] C[i, j]
constexpr int CJ = 16;  // output columns per pass

template <typename R, bool CPLX>
__global__ void __launch_bounds__(256) k_combine(const R *A, long long lda, int k,
                                                 const double *C, R *Y, long long ldy,
                                                 int m, long long nrows) {
  constexpr int W = CPLX ? 2 : 1;
  extern __shared__ __align__(16) char smem_raw[];
  double *Cs = reinterpret_cast<double *>(smem_raw);  // k x CJ (x W)
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (int j0 = 0; j0 < m; j0 += CJ) {
    const int nj = (m - j0 < CJ) ? m - j0 : CJ;
    __syncthreads();
    for (int e = threadIdx.x; e < k * CJ; e += blockDim.x) {
      const int i = e % k, jj = e / k;
#pragma unroll
      for (int w = 0; w < W; ++w)
        Cs[(i + jj * k) * W + w] = jj < nj ? C[(i + (long long)(j0 + jj) * k) * W + w] : 0.0;
    }
    __syncthreads();
    if (r < nrows) {
      double acc[CJ][W];
#pragma unroll
      for (int jj = 0; jj < CJ; ++jj)
#pragma unroll
        for (int w = 0; w < W; ++w) acc[jj][w] = 0.0;
      for (int i = 0; i < k; ++i) {
        double ar, ai = 0.0;
        if (CPLX) {
          ar = (double)A[2 * (r + i * lda)];
          ai = (double)A[2 * (r + i * lda) + 1];
        } else {
          ar = (double)A[r + i * lda];
        }
#pragma unroll
        for (int jj = 0; jj < CJ; ++jj) {
          if (CPLX) {
            const double cr = Cs[(i + jj * k) * 2], ci = Cs[(i + jj * k) * 2 + 1];
            acc[jj][0] += ar * cr - ai * ci;
            acc[jj][W - 1] += ar * ci + ai * cr;
          } else {
            acc[jj][0] += ar * Cs[i + jj * k];
          }
        }
      }
#pragma unroll
      for (int jj = 0; jj < CJ; ++jj) {
        if (jj < nj) {
          const long long o = r + (long long)(j0 + jj) * ldy;
          if (CPLX) {
            Y[2 * o] = (R)acc[jj][0];
            Y[2 * o + 1] = (R)acc[jj][W - 1];
          } else {
            Y[o] = (R)acc[jj][0];
          }
        }
      }
    }
  }
}

template <typename R, bool CPLX>
static int combine_t(bdgx_ctx *c, const void *dA, uint64_t lda, uint64_t k,
                     const double *dC, void *dY, uint64_t ldy, uint64_t m,
                     uint64_t nrows) {
  if (m == 0 || nrows == 0) return BDGX_OK;
  const int W = CPLX ? 2 : 1;
  const size_t sm = sizeof(double) * (size_t)k * CJ * W;
  if (sm > 64 * 1024) return BDGX_EINVAL;
  const long long nb = ((long long)nrows + 255) / 256;
  hipLaunchKernelGGL((k_combine<R, CPLX>), dim3((unsigned)nb), dim3(256), sm, c->stream,
                     reinterpret_cast<const R *>(dA), (long long)lda, (int)k, dC,
                     reinterpret_cast<R *>(dY), (long long)ldy, (int)m, (long long)nrows);
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

int combine_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                     const double *dC, void *dY, uint64_t ldy, uint64_t m,
                     uint64_t nrows) {
  if (prec == BDGX_FP64)
    return c->cplx ? combine_t<double, true>(c, dA, lda, k, dC, dY, ldy, m, nrows)
                   : combine_t<double, false>(c, dA, lda, k, dC, dY, ldy, m, nrows);
  if (prec == BDGX_FP32)
    return c->cplx ? combine_t<float, true>(c, dA, lda, k, dC, dY, ldy, m, nrows)
                   : combine_t<float, false>(c, dA, lda, k, dC, dY, ldy, m, nrows);
  return BDGX_EINVAL;
}

// ---------------------------------------------------------------------------
// residual R = AX - lambda * X' (X' = Swap(X) or X) and squared column norms
// ||R||^2, ||X||^2, ||MX||^2 accumulated in FP64.
template <typename R, bool CPLX>
__global__ void __launch_bounds__(256) k_residual(const R *AX, const R *X, const R *MX,
                                                  long long ld, long long nrows,
                                                  const double *lam, int swap_x, R *Rr,
                                                  long long ldr, double *n2) {
  constexpr int W = CPLX ? 2 : 1;
  const int col = blockIdx.y;
  const long long half = nrows / 2;
  const double l = lam[col];
  double sr = 0.0, sx = 0.0, sm = 0.0;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nrows;
       r += (long long)gridDim.x * blockDim.x) {
    const long long rs = swap_x ? (r < half ? r + half : r - half) : r;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const double ax = (double)AX[(r + col * ld) * W + w];
      const double xs = (double)X[(rs + col * ld) * W + w];
      const double xv = (double)X[(r + col * ld) * W + w];
      const double rv = ax - l * xs;
      if (Rr) Rr[(r + col * ldr) * W + w] = (R)rv;
      sr += rv * rv;
      sx += xv * xv;
      if (MX) {
        const double mv = (double)MX[(r + col * ld) * W + w];
        sm += mv * mv;
      }
    }
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = sx;
  red[2][threadIdx.x] = sm;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicAdd(n2 + 3 * col + threadIdx.x, red[threadIdx.x][0]);
}

template <typename R, bool CPLX>
static int residual_t(bdgx_ctx *c, const void *dAX, const void *dX, const void *dMX,
                      uint64_t ld, uint64_t ncols, uint64_t nrows, const double *dlam,
                      int swap_x, void *dR, uint64_t ldr, double *dn2) {
  if (ncols == 0) return BDGX_OK;
  BDGX_HIP(hipMemsetAsync(dn2, 0, sizeof(double) * 3 * ncols, c->stream));
  long long nbx = ((long long)nrows + 255) / 256;
  if (nbx > 512) nbx = 512;
  hipLaunchKernelGGL((k_residual<R, CPLX>), dim3((unsigned)nbx, (unsigned)ncols), dim3(256),
                     0, c->stream, reinterpret_cast<const R *>(dAX),
                     reinterpret_cast<const R *>(dX), reinterpret_cast<const R *>(dMX),
                     (long long)ld, (long long)nrows, dlam, swap_x, reinterpret_cast<R *>(dR),
                     (long long)ldr, dn2);
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

int residual_dispatch(bdgx_ctx *c, int prec, const void *dAX, const void *dX, const void *dMX,
                      uint64_t ld, uint64_t ncols, uint64_t nrows, const double *dlam,
                      int swap_x, void *dR, uint64_t ldr, double *dn2) {
  if (prec == BDGX_FP64)
    return c->cplx ? residual_t<double, true>(c, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2)
                   : residual_t<double, false>(c, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2);
  if (prec == BDGX_FP32)
    return c->cplx ? residual_t<float, true>(c, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2)
                   : residual_t<float, false>(c, dAX, dX, dMX, ld, ncols, nrows, dlam, swap_x, dR, ldr, dn2);
  return BDGX_EINVAL;
}

// ---------------------------------------------------------------------------
// column gather: dst[:, j] = src[:, idx[j]]
template <typename T>
__global__ void k_gather(const T *src, long long lds, const int *idx, int ncols, T *dst,
                         long long ldd, long long nrows) {
  const int j = blockIdx.y;
  const long long base = (long long)idx[j] * lds;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nrows;
       r += (long long)gridDim.x * blockDim.x)
    dst[r + j * ldd] = src[base + r];
}

int gather_dispatch(bdgx_ctx *c, int prec, const void *src, uint64_t lds, const int *hidx,
                    uint64_t ncols, void *dst, uint64_t ldd, uint64_t nrows) {
  if (ncols == 0) return BDGX_OK;
  int rc = ensure_scratch(c, 0);
  (void)rc;
  // indices go through a small device buffer owned by the context
  if (!c->idx_buf || c->idx_cap < ncols) {
    if (c->idx_buf) {
      BDGX_HIP(hipStreamSynchronize(c->stream));
      BDGX_HIP(hipFree(c->idx_buf));
    }
    c->idx_cap = ncols < 1024 ? 1024 : ncols;
    BDGX_HIP(hipMalloc((void **)&c->idx_buf, sizeof(int) * c->idx_cap));
  }
  BDGX_HIP(hipMemcpyAsync(c->idx_buf, hidx, sizeof(int) * ncols, hipMemcpyHostToDevice,
                          c->stream));
  const int W = c->cplx ? 2 : 1;
  long long nbx = ((long long)nrows + 255) / 256;
  if (nbx > 1024) nbx = 1024;
  if (prec == BDGX_FP64) {
    if (W == 2)
      hipLaunchKernelGGL(k_gather<double2>, dim3((unsigned)nbx, (unsigned)ncols), dim3(256), 0,
                         c->stream, (const double2 *)src, (long long)lds, c->idx_buf, (int)ncols,
                         (double2 *)dst, (long long)ldd, (long long)nrows);
    else
      hipLaunchKernelGGL(k_gather<double>, dim3((unsigned)nbx, (unsigned)ncols), dim3(256), 0,
                         c->stream, (const double *)src, (long long)lds, c->idx_buf, (int)ncols,
                         (double *)dst, (long long)ldd, (long long)nrows);
  } else {
    if (W == 2)
      hipLaunchKernelGGL(k_gather<float2>, dim3((unsigned)nbx, (unsigned)ncols), dim3(256), 0,
                         c->stream, (const float2 *)src, (long long)lds, c->idx_buf, (int)ncols,
                         (float2 *)dst, (long long)ldd, (long long)nrows);
    else
      hipLaunchKernelGGL(k_gather<float>, dim3((unsigned)nbx, (unsigned)ncols), dim3(256), 0,
                         c->stream, (const float *)src, (long long)lds, c->idx_buf, (int)ncols,
                         (float *)dst, (long long)ldd, (long long)nrows);
  }
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

template <typename S, typename D>
__global__ void k_convert(const S *src, long long lds, D *dst, long long ldd,
                          long long nrows, int ncols) {
  const long long total = nrows * ncols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long col = e / nrows, r = e - col * nrows;
    dst[r + col * ldd] = (D)src[r + col * lds];
  }
}

int convert_dispatch(bdgx_ctx *c, int ps, const void *src, uint64_t lds, int pd,
                     void *dst, uint64_t ldd, uint64_t nrows, uint64_t ncols) {
  const int W = c->cplx ? 2 : 1;
  const long long nr = (long long)nrows * W;
  const long long ls = (long long)lds * W, ldd2 = (long long)ldd * W;
  if (ps == BDGX_FP64 && pd == BDGX_FP32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(2048), dim3(256), 0, c->stream,
                       (const double *)src, ls, (float *)dst, ldd2, nr, (int)ncols);
  else if (ps == BDGX_FP32 && pd == BDGX_FP64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0, c->stream,
                       (const float *)src, ls, (double *)dst, ldd2, nr, (int)ncols);
  else if (ps == BDGX_FP64 && pd == BDGX_FP64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(2048), dim3(256), 0, c->stream,
                       (const double *)src, ls, (double *)dst, ldd2, nr, (int)ncols);
  else if (ps == BDGX_FP32 && pd == BDGX_FP32)
    hipLaunchKernelGGL((k_convert<float, float>), dim3(2048), dim3(256), 0, c->stream,
                       (const float *)src, ls, (float *)dst, ldd2, nr, (int)ncols);
  else
    return BDGX_EINVAL;
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

}  // namespace bdgx
