// Operator pipelines of the BdG hot path on gfx950.
//
// Every operator of the reference (kinetic, K, M, Lrep, precond, dipolar
// convolution, dipolar mean field) is "pointwise pre-weight -> 3D FFT ->
// k-space multiplier -> inverse 3D FFT -> pointwise combine".  Here one block
// apply runs as five kinds of launches over all columns at once:
//
//   xfwd   x-axis forward transform; the pre-weights (p_K, p_M, conj(psi0),
//          |psi0|^2) are applied while the rows are staged in LDS.  Real
//          vectors use the two-for-one trick (rows 2q, 2q+1 packed as the real
//          and imaginary part of one complex transform, untangled on store),
//          writing the Nx/2+1 half spectrum.
//   ypass  y-axis transform on the spectrum (pencils of P adjacent kx).
//   zpass  z-axis forward transform, multiply by the channel's k-space factor
//          (k^2/2S, U(k)/S, 1/(S*safe(mu+k^2/2)), ...), z-axis inverse -- one
//          launch, the pencil never leaves LDS.
//   ypass  inverse.
//   xinv   x-axis inverse transform plus the output combine
//          y = sum_c coef_c * chan_c + lt * x (two channels of one output row
//          share one complex transform in the real case).
//
// Spectrum layout in HBM: spec[((chan * ncols + col) * Nz + iz) * Ny + iy][kx],
// i.e. the reference's r2c layout (src/matmul_ctx.c:116-131) per column.
#include <math.h>
#include <string.h>

#include "bdgx_internal.hpp"
#include "fft_core.hpp"
#include "plane.hpp"

namespace bdgx {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// x-axis forward, real input (two-for-one)
template <typename R>
__global__ void __launch_bounds__(kThreads) k_xfwd_real(XArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = a.g.nx, n0k = a.g.n0k;
  const int rows = a.g.ny * a.g.nz;
  const int RB = a.rb, hp = RB / 2;
  const int nblk = (rows + RB - 1) / RB;
  const int col = blockIdx.x / nblk;
  const int r0 = (blockIdx.x - col * nblk) * RB;
  const int npen = a.nch * hp;
  const int ps = nx;
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + nx;
  cx<R> *bufB = bufA + npen * ps;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = a.tw[i];
  const R *xcol = a.X + (long long)col * a.ld + (long long)a.in_half * a.g.S;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int rr = e / nx, i = e - rr * nx;
    const int r = r0 + rr;
    const long long gi = (long long)r * nx + i;
    const bool ok = r < rows;
    const R v = ok ? xcol[gi] : R(0);
    for (int c = 0; c < a.nch; ++c) {
      R w = v;
      if (ok) {
        if (a.pre[c] == PRE_REAL || a.pre[c] == PRE_CONJ_PSI)
          w = v * a.prew[c][gi];
        else if (a.pre[c] == PRE_ABS2)
          w = v * v;
      }
      cx<R> *pp = bufA + (c * hp + (rr >> 1)) * ps + i;
      if (rr & 1)
        pp->im = w;
      else
        pp->re = w;
    }
  }
  __syncthreads();
  const cx<R> *res = lds_fft<R, false>(bufA, bufB, tw, a.pl, npen, ps);
  const int total = npen * n0k;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int p = e / n0k, k = e - p * n0k;
    const int c = p / hp, q = p - c * hp;
    const cx<R> zk = res[p * ps + k];
    const cx<R> zn = res[p * ps + (k == 0 ? 0 : nx - k)];
    const cx<R> A = cmk<R>(R(0.5) * (zk.re + zn.re), R(0.5) * (zk.im - zn.im));
    const cx<R> B = cmk<R>(R(0.5) * (zk.im + zn.im), R(-0.5) * (zk.re - zn.re));
    const int ra = r0 + 2 * q, rb = ra + 1;
    cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                (long long)col * a.g.Kk;
    if (ra < rows) sp[(long long)ra * n0k + k] = A;
    if (rb < rows) sp[(long long)rb * n0k + k] = B;
  }
}

// ---------------------------------------------------------------------------
// x-axis inverse to real output, plus the combine.
template <typename R>
__global__ void __launch_bounds__(kThreads) k_xinv_real(XArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = a.g.nx, n0k = a.g.n0k;
  const int rows = a.g.ny * a.g.nz;
  const int RB = a.rb;
  const int nblk = (rows + RB - 1) / RB;
  const int col = blockIdx.x / nblk;
  const int r0 = (blockIdx.x - col * nblk) * RB;
  const bool two = a.nch == 2;
  const int npen = two ? RB : RB / 2;
  const int ps = nx;
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + nx;
  cx<R> *bufB = bufA + npen * ps;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = a.tw[i];
  const long long colK = (long long)col * a.g.Kk;
  const cx<R> *s0 = a.spec + (long long)a.chan_base * a.chan_stride + colK;
  const cx<R> *s1 = two ? s0 + a.chan_stride : s0;
  const bool even = (nx & 1) == 0;
  for (int e = threadIdx.x; e < npen * nx; e += blockDim.x) {
    const int p = e / nx, k = e - p * nx;
    const bool cj = k >= n0k;
    const int kk = cj ? nx - k : k;
    int ra, rb;
    if (two) {
      ra = r0 + p;
      rb = ra;
    } else {
      ra = r0 + 2 * p;
      rb = ra + 1;
    }
    cx<R> A0 = cmk<R>(0, 0), A1 = cmk<R>(0, 0);
    if (ra < rows) A0 = s0[(long long)ra * n0k + kk];
    if (rb < rows) A1 = s1[(long long)rb * n0k + kk];
    if (cj) {
      A0.im = -A0.im;
      A1.im = -A1.im;
    }
    if (kk == 0 || (even && kk == nx / 2)) {
      A0.im = 0;
      A1.im = 0;
    }
    bufA[p * ps + k] = cmk<R>(A0.re - A1.im, A0.im + A1.re);
  }
  __syncthreads();
  const cx<R> *res = lds_fft<R, true>(bufA, bufB, tw, a.pl, npen, ps);
  const R *xcol = a.X + (long long)col * a.ld + (long long)a.in_half * a.g.S;
  R *ycol = a.Y + (long long)col * a.ld + (long long)a.out_half * a.g.S;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int rr = e / nx, i = e - rr * nx;
    const int r = r0 + rr;
    if (r >= rows) continue;
    const long long gi = (long long)r * nx + i;
    R c0, c1 = 0;
    if (two) {
      const cx<R> z = res[rr * ps + i];
      c0 = z.re;
      c1 = z.im;
    } else {
      const cx<R> z = res[(rr >> 1) * ps + i];
      c0 = (rr & 1) ? z.im : z.re;
    }
    R out = a.posts[0] * c0;
    if (a.post[0] != POST_ONE) out *= a.postw[0][gi];
    if (two) {
      R t = a.posts[1] * c1;
      if (a.post[1] != POST_ONE) t *= a.postw[1][gi];
      out += t;
    }
    if (a.lt) out += a.lt[gi] * xcol[gi];
    ycol[gi] = out;
  }
}

// ---------------------------------------------------------------------------
// x-axis forward, complex input
template <typename R>
__global__ void __launch_bounds__(kThreads) k_xfwd_cplx(XArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = a.g.nx;
  const int rows = a.g.ny * a.g.nz;
  const int RB = a.rb;
  const int nblk = (rows + RB - 1) / RB;
  const int col = blockIdx.x / nblk;
  const int r0 = (blockIdx.x - col * nblk) * RB;
  const int npen = a.nch * RB;
  const int ps = nx;
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + nx;
  cx<R> *bufB = bufA + npen * ps;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = a.tw[i];
  const cx<R> *xcol = reinterpret_cast<const cx<R> *>(a.X) +
                      (long long)col * a.ld + (long long)a.in_half * a.g.S;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int rr = e / nx, i = e - rr * nx;
    const int r = r0 + rr;
    const long long gi = (long long)r * nx + i;
    const bool ok = r < rows;
    const cx<R> v = ok ? xcol[gi] : cmk<R>(0, 0);
    for (int c = 0; c < a.nch; ++c) {
      cx<R> w = v;
      if (ok) {
        if (a.pre[c] == PRE_REAL) {
          w = cscale(v, a.prew[c][gi]);
        } else if (a.pre[c] == PRE_CONJ_PSI) {
          const cx<R> ps0 = reinterpret_cast<const cx<R> *>(a.prew[c])[gi];
          w = cmul(cconj(ps0), v);
        } else if (a.pre[c] == PRE_ABS2) {
          w = cmk<R>(v.re * v.re + v.im * v.im, 0);
        }
      }
      bufA[(c * RB + rr) * ps + i] = w;
    }
  }
  __syncthreads();
  const cx<R> *res = lds_fft<R, false>(bufA, bufB, tw, a.pl, npen, ps);
  for (int e = threadIdx.x; e < npen * nx; e += blockDim.x) {
    const int p = e / nx, k = e - p * nx;
    const int c = p / RB, rr = p - c * RB;
    const int r = r0 + rr;
    if (r >= rows) continue;
    cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                (long long)col * a.g.Kk;
    sp[(long long)r * nx + k] = res[p * ps + k];
  }
}

// ---------------------------------------------------------------------------
// x-axis inverse, complex, plus combine.
template <typename R>
__global__ void __launch_bounds__(kThreads) k_xinv_cplx(XArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = a.g.nx;
  const int rows = a.g.ny * a.g.nz;
  const int RB = a.rb;
  const int nblk = (rows + RB - 1) / RB;
  const int col = blockIdx.x / nblk;
  const int r0 = (blockIdx.x - col * nblk) * RB;
  const int npen = a.nch * RB;
  const int ps = nx;
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + nx;
  cx<R> *bufB = bufA + npen * ps;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) tw[i] = a.tw[i];
  for (int e = threadIdx.x; e < npen * nx; e += blockDim.x) {
    const int p = e / nx, k = e - p * nx;
    const int c = p / RB, rr = p - c * RB;
    const int r = r0 + rr;
    const cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                      (long long)col * a.g.Kk;
    bufA[p * ps + k] = r < rows ? sp[(long long)r * nx + k] : cmk<R>(0, 0);
  }
  __syncthreads();
  const cx<R> *res = lds_fft<R, true>(bufA, bufB, tw, a.pl, npen, ps);
  const cx<R> *xcol = reinterpret_cast<const cx<R> *>(a.X) +
                      (long long)col * a.ld + (long long)a.in_half * a.g.S;
  for (int e = threadIdx.x; e < RB * nx; e += blockDim.x) {
    const int rr = e / nx, i = e - rr * nx;
    const int r = r0 + rr;
    if (r >= rows) continue;
    const long long gi = (long long)r * nx + i;
    cx<R> out = cmk<R>(0, 0);
    for (int c = 0; c < a.nch; ++c) {
      cx<R> t = cscale(res[(c * RB + rr) * ps + i], a.posts[c]);
      if (a.post[c] == POST_REAL)
        t = cscale(t, a.postw[c][gi]);
      else if (a.post[c] == POST_PSI)
        t = cmul(t, reinterpret_cast<const cx<R> *>(a.postw[c])[gi]);
      out = out + t;
    }
    if (a.lt) out = out + cscale(xcol[gi], a.lt[gi]);
    if (a.out_real) {
      a.Y[(long long)col * a.ld + (long long)a.out_half * a.g.S + gi] = out.re;
    } else {
      reinterpret_cast<cx<R> *>(a.Y)[(long long)col * a.ld +
                                     (long long)a.out_half * a.g.S + gi] = out;
    }
  }
}

// ---------------------------------------------------------------------------
// y-axis transform in place on the spectrum.
template <typename R, bool INV>
__global__ void __launch_bounds__(kThreads) k_ypass(YZArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int ny = a.g.ny, nz = a.g.nz, n0k = a.g.n0k, P = a.P;
  const int ntile = (n0k + P - 1) / P;
  const int t = blockIdx.x % ntile;
  const int rest = blockIdx.x / ntile;
  const int z = rest % nz;
  const int b = rest / nz;
  const int kx0 = t * P;
  const int ps = ny + 1;
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + ny;
  cx<R> *bufB = bufA + P * ps;
  for (int i = threadIdx.x; i < ny; i += blockDim.x) tw[i] = a.tw[i];
  cx<R> *base = a.spec + (long long)b * a.g.Kk + (long long)z * ny * n0k;
  for (int e = threadIdx.x; e < P * ny; e += blockDim.x) {
    const int y = e / P, p = e - y * P;
    const int kx = kx0 + p;
    bufA[p * ps + y] = kx < n0k ? base[(long long)y * n0k + kx] : cmk<R>(0, 0);
  }
  __syncthreads();
  const cx<R> *res = lds_fft<R, INV>(bufA, bufB, tw, a.pl, P, ps);
  for (int e = threadIdx.x; e < P * ny; e += blockDim.x) {
    const int y = e / P, p = e - y * P;
    const int kx = kx0 + p;
    if (kx < n0k) base[(long long)y * n0k + kx] = res[p * ps + y];
  }
}

template <typename R>
__device__ __forceinline__ R kmul(const YZArgs<R> &a, int code, int kx, int y,
                                  int z) {
  const R k2 = a.kx2[kx] + a.ky2[y] + a.kz2[z];
  switch (code) {
    case MUL_KIN:
      return k2 * (R(0.5) * a.invS);
    case MUL_DIP:
      return a.U[((long long)z * a.g.ny + y) * a.g.n0k + kx] * a.invS;
    case MUL_PRE: {
      R d = a.mu + R(0.5) * k2;
      if (fabs((double)d) < 1e-8) d = R(1e-8);
      return a.invS / d;
    }
    case MUL_ITPKIN:
      return R(exp(-(double)a.dt * 0.5 * (double)k2)) * a.invS;
    case MUL_TABK:
    case MUL_TABM:
      return a.tab[code - MUL_TABK][((long long)z * a.g.ny + y) * a.g.n0k + kx];
    default:
      return a.invS;
  }
}

// z-axis forward, k-space multiply, z-axis inverse -- fused.
template <typename R>
__global__ void __launch_bounds__(kThreads) k_zpass(YZArgs<R> a) {
  extern __shared__ __align__(16) char smem_raw[];
  const int ny = a.g.ny, nz = a.g.nz, n0k = a.g.n0k, P = a.P;
  const int ntile = (n0k + P - 1) / P;
  const int t = blockIdx.x % ntile;
  const int rest = blockIdx.x / ntile;
  const int y = rest % ny;
  const int b = rest / ny;
  const int kx0 = t * P;
  const int ps = nz + 1;
  const long long zs = (long long)n0k * ny;
  const int code = a.mul[b / a.ncols];
  cx<R> *tw = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *bufA = tw + nz;
  cx<R> *bufB = bufA + P * ps;
  for (int i = threadIdx.x; i < nz; i += blockDim.x) tw[i] = a.tw[i];
  cx<R> *base = a.spec + (long long)b * a.g.Kk + (long long)y * n0k;
  for (int e = threadIdx.x; e < P * nz; e += blockDim.x) {
    const int z = e / P, p = e - z * P;
    const int kx = kx0 + p;
    bufA[p * ps + z] = kx < n0k ? base[z * zs + kx] : cmk<R>(0, 0);
  }
  __syncthreads();
  cx<R> *res = lds_fft<R, false>(bufA, bufB, tw, a.pl, P, ps);
  if (a.fwd_only) {
    for (int e = threadIdx.x; e < P * nz; e += blockDim.x) {
      const int z = e / P, p = e - z * P;
      const int kx = kx0 + p;
      if (kx < n0k) base[z * zs + kx] = res[p * ps + z];
    }
    return;
  }
  cx<R> *oth = res == bufA ? bufB : bufA;
  for (int e = threadIdx.x; e < P * nz; e += blockDim.x) {
    const int z = e / P, p = e - z * P;
    const int kx = kx0 + p;
    if (kx < n0k) res[p * ps + z] = cscale(res[p * ps + z], kmul(a, code, kx, y, z));
  }
  __syncthreads();
  const cx<R> *fin = lds_fft<R, true>(res, oth, tw, a.pl, P, ps);
  for (int e = threadIdx.x; e < P * nz; e += blockDim.x) {
    const int z = e / P, p = e - z * P;
    const int kx = kx0 + p;
    if (kx < n0k) base[z * zs + kx] = fin[p * ps + z];
  }
}

// Swap: y = [x_v; x_u] (reference src/matmulSwap_impl.inc:14-26).
template <typename T>
__global__ void k_swap(const T *X, T *Y, long long S, long long ld, int ncols) {
  const long long total = (long long)ncols * S;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long col = e / S, i = e - col * S;
    const T *x = X + col * ld;
    T *y = Y + col * ld;
    const T lo = x[i], hi = x[S + i];
    y[i] = hi;
    y[S + i] = lo;
  }
}

// ---------------------------------------------------------------------------
// host side

struct HalfJob {
  int in_half, out_half, nch;
  int pre[kMaxChan];
  int mul[kMaxChan];
  int post[kMaxChan];
  double posts[kMaxChan];
  int prew_sel[kMaxChan];   // field selector, see field_ptr()
  int postw_sel[kMaxChan];
  int lt_sel;
  int out_real;
};

enum FieldSel { F_NONE = 0, F_LTK, F_LTM, F_PK, F_PM, F_WF, F_LAK, F_LBK, F_LAM, F_LBM };

template <typename R>
static const R *field_ptr(const Fields<R> &f, int sel) {
  switch (sel) {
    case F_LTK: return f.ltK;
    case F_LTM: return f.ltM;
    case F_PK: return f.pK;
    case F_PM: return f.pM;
    case F_WF: return f.wf;
    case F_LAK: return f.lpa[0];
    case F_LBK: return f.lpb[0];
    case F_LAM: return f.lpa[1];
    case F_LBM: return f.lpb[1];
    default: return nullptr;
  }
}

template <typename R>
static Fields<R> &fields(bdgx_ctx *c);
template <>
Fields<double> &fields<double>(bdgx_ctx *c) {
  return c->f64;
}
template <>
Fields<float> &fields<float>(bdgx_ctx *c) {
  return c->f32;
}

static int rows_per_wg(int nx, int npen_per_2rows, bool real) {
  // aim for ~1024 complex elements of pencils per workgroup
  int rb = 1024 / (nx > 0 ? nx : 1);
  if (real) rb = (2 * rb) / (npen_per_2rows > 0 ? npen_per_2rows : 1);
  else rb = rb / (npen_per_2rows > 0 ? npen_per_2rows : 1);
  if (rb < 2) rb = 2;
  rb &= ~1;
  return rb;
}

template <typename R>
static XArgs<R> make_xargs(bdgx_ctx *c, const HalfJob &J, int base, int nc, const R *Xc, R *Yc,
                           long long ld, cx<R> *spec, long long cstride, const R *lt_override) {
  Fields<R> &F = fields<R>(c);
  XArgs<R> a;
  memset(&a, 0, sizeof(a));
  a.g = c->geo;
  a.ncols = nc;
  a.ld = ld;
  a.X = Xc;
  a.Y = Yc;
  a.in_half = J.in_half;
  a.out_half = J.out_half;
  a.nch = J.nch;
  a.chan_base = base;
  for (int q = 0; q < J.nch; ++q) {
    a.pre[q] = J.pre[q];
    a.prew[q] = field_ptr(F, J.prew_sel[q]);
    a.post[q] = J.post[q];
    a.postw[q] = field_ptr(F, J.postw_sel[q]);
    a.posts[q] = (R)J.posts[q];
  }
  a.lt = lt_override ? lt_override : field_ptr(F, J.lt_sel);
  a.out_real = J.out_real;
  a.spec = spec;
  a.chan_stride = cstride;
  a.tw = F.tw[0];
  a.tw_y = F.tw[1];
  a.pl = c->plan[0];
  return a;
}

// x-axis pass (forward: pre-weights + transform into the spectrum; inverse:
// transform back + combine into Y) for one half-job over nc columns.
template <typename R>
static int launch_x(bdgx_ctx *c, XArgs<R> &a, bool inverse) {
  const Geo &g = c->geo;
  const int rows = g.ny * g.nz;
  const int nc = a.ncols;
  if (!c->cplx) {
    int npen;
    if (!inverse) {
      a.rb = rows_per_wg(g.nx, a.nch, true);
      npen = a.nch * (a.rb / 2);
    } else if (a.nch == 2) {
      int rb = 1024 / g.nx;  // one pencil per row (both channels in one transform)
      if (rb < 2) rb = 2;
      a.rb = rb & ~1;
      npen = a.rb;
    } else {
      a.rb = rows_per_wg(g.nx, 1, true);
      npen = a.rb / 2;
    }
    const size_t sm = sizeof(cx<R>) * ((size_t)g.nx + 2 * (size_t)npen * g.nx);
    const int nblk = (rows + a.rb - 1) / a.rb;
    if (inverse)
      hipLaunchKernelGGL(k_xinv_real<R>, dim3(nblk * nc), dim3(kThreads), sm, c->stream, a);
    else
      hipLaunchKernelGGL(k_xfwd_real<R>, dim3(nblk * nc), dim3(kThreads), sm, c->stream, a);
  } else {
    a.rb = rows_per_wg(g.nx, a.nch, false);
    const int npen = a.nch * a.rb;
    const size_t sm = sizeof(cx<R>) * ((size_t)g.nx + 2 * (size_t)npen * g.nx);
    const int nblk = (rows + a.rb - 1) / a.rb;
    if (inverse)
      hipLaunchKernelGGL(k_xinv_cplx<R>, dim3(nblk * nc), dim3(kThreads), sm, c->stream, a);
    else
      hipLaunchKernelGGL(k_xfwd_cplx<R>, dim3(nblk * nc), dim3(kThreads), sm, c->stream, a);
  }
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

template <typename R>
static YZArgs<R> make_yz(bdgx_ctx *c, const int *mul, int C, int nc, cx<R> *spec) {
  const Geo &g = c->geo;
  Fields<R> &F = fields<R>(c);
  YZArgs<R> yz;
  memset(&yz, 0, sizeof(yz));
  yz.g = g;
  yz.ncols = nc;
  yz.nb = C * nc;
  yz.P = sizeof(R) == 8 ? 8 : 16;
  if (yz.P > g.n0k) yz.P = g.n0k;
  yz.spec = spec;
  for (int ch = 0; ch < C; ++ch) yz.mul[ch] = mul[ch];
  yz.kx2 = F.kx2[0];
  yz.ky2 = F.kx2[1];
  yz.kz2 = F.kx2[2];
  yz.U = F.U;
  yz.tab[0] = F.lpg[0];
  yz.tab[1] = F.lpg[1];
  yz.mu = (R)c->mu;
  yz.invS = (R)(1.0 / (double)g.S);
  yz.dt = (R)c->itp_dt;
  return yz;
}

// y-axis transform over all batches (skipped when Ny == 1)
template <typename R>
static int launch_y(bdgx_ctx *c, YZArgs<R> yz, bool inverse) {
  const Geo &g = c->geo;
  if (g.ny <= 1) return BDGX_OK;
  Fields<R> &F = fields<R>(c);
  yz.tw = F.tw[1];
  yz.pl = c->plan[1];
  const int ntile = (g.n0k + yz.P - 1) / yz.P;
  const size_t sm = sizeof(cx<R>) * ((size_t)g.ny + 2 * (size_t)yz.P * (g.ny + 1));
  if (inverse)
    hipLaunchKernelGGL((k_ypass<R, true>), dim3(ntile * g.nz * yz.nb), dim3(kThreads), sm,
                       c->stream, yz);
  else
    hipLaunchKernelGGL((k_ypass<R, false>), dim3(ntile * g.nz * yz.nb), dim3(kThreads), sm,
                       c->stream, yz);
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

// fused z-axis forward, k-space multiply, z-axis inverse
template <typename R>
static int launch_z(bdgx_ctx *c, YZArgs<R> z) {
  const Geo &g = c->geo;
  Fields<R> &F = fields<R>(c);
  z.tw = F.tw[2];
  z.pl = c->plan[2];
  if (zpass_fixed_supported<R>(c)) return launch_zpass_fixed<R>(c, z);
  const int ntile = (g.n0k + z.P - 1) / z.P;
  const size_t sm = sizeof(cx<R>) * ((size_t)g.nz + 2 * (size_t)z.P * (g.nz + 1));
  hipLaunchKernelGGL(k_zpass<R>, dim3(ntile * g.ny * z.nb), dim3(kThreads), sm, c->stream, z);
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

template <typename R>
static int run_jobs(bdgx_ctx *c, const HalfJob *jobs, int njobs, const R *X,
                    R *Y, long long ncols_total, long long ld, const R *lt_override,
                    bool fwd_only = false) {
  const Geo &g = c->geo;
  int C = 0;
  int mul[kMaxChan];
  for (int j = 0; j < njobs; ++j)
    for (int q = 0; q < jobs[j].nch; ++q) {
      if (C >= kMaxChan) return BDGX_EINVAL;
      mul[C++] = jobs[j].mul[q];
    }
  // chunk columns so the spectrum workspace stays bounded
  const size_t per_col = (size_t)C * (size_t)g.Kk * sizeof(cx<R>);
  const size_t budget = c->chunk_bytes ? c->chunk_bytes : (size_t)6 << 30;
  long long chunk = (long long)(budget / (per_col ? per_col : 1));
  if (chunk < 1) chunk = 1;
  if (chunk > ncols_total) chunk = ncols_total;
  int rc = ensure_spec(c, per_col * (size_t)chunk);
  if (rc) return rc;
  cx<R> *spec = reinterpret_cast<cx<R> *>(c->spec);
  const int cw = c->cplx ? 2 : 1;  // R per element
  for (long long c0 = 0; c0 < ncols_total; c0 += chunk) {
    const int nc = (int)((ncols_total - c0 < chunk) ? ncols_total - c0 : chunk);
    const R *Xc = X + c0 * ld * cw;
    R *Yc = Y ? Y + c0 * ld * cw : nullptr;
    const long long cstride = (long long)nc * g.Kk;
    // whole-plane x/y kernels where the spectrum plane fits in LDS, else
    // separate x and y pencil passes
    const bool plane = plane_supported<R>(c);
    int base = 0;
    for (int j = 0; j < njobs; ++j) {
      XArgs<R> a = make_xargs<R>(c, jobs[j], base, nc, Xc, Yc, ld, spec, cstride, lt_override);
      if ((rc = plane ? launch_plane_x<R>(c, a, false) : launch_x<R>(c, a, false))) return rc;
      base += jobs[j].nch;
    }
    YZArgs<R> yz = make_yz<R>(c, mul, C, nc, spec);
    yz.fwd_only = fwd_only ? 1 : 0;
    if (!plane && (rc = launch_y<R>(c, yz, false))) return rc;
    if ((rc = launch_z<R>(c, yz))) return rc;
    if (fwd_only) {
      // the caller reads the spectrum; a single chunk is required
      if (chunk < ncols_total) return BDGX_EINVAL;
      return BDGX_OK;
    }
    if (!plane && (rc = launch_y<R>(c, yz, true))) return rc;
    base = 0;
    for (int j = 0; j < njobs; ++j) {
      XArgs<R> a = make_xargs<R>(c, jobs[j], base, nc, Xc, Yc, ld, spec, cstride, lt_override);
      if ((rc = plane ? launch_plane_x<R>(c, a, true) : launch_x<R>(c, a, true))) return rc;
      base += jobs[j].nch;
    }
  }
  return BDGX_OK;
}

static HalfJob job_kin(int half, int lt_sel) {
  HalfJob J;
  memset(&J, 0, sizeof(J));
  J.in_half = J.out_half = half;
  J.nch = 1;
  J.pre[0] = PRE_NONE;
  J.mul[0] = MUL_KIN;
  J.post[0] = POST_ONE;
  J.posts[0] = 1.0;
  J.lt_sel = lt_sel;
  return J;
}

static HalfJob job_M(const bdgx_ctx *c, int half) {
  if (!(c->has_U && c->dim == 3)) return job_kin(half, F_LTM);
  HalfJob J = job_kin(half, F_LTM);
  J.nch = 2;
  J.pre[1] = PRE_CONJ_PSI;
  J.prew_sel[1] = F_WF;
  J.mul[1] = MUL_DIP;
  J.post[1] = POST_PSI;
  J.postw_sel[1] = F_WF;
  J.posts[1] = 2.0 * c->g_ddi;
  return J;
}

static HalfJob job_pre(int half, int psel) {
  HalfJob J;
  memset(&J, 0, sizeof(J));
  J.in_half = J.out_half = half;
  J.nch = 1;
  J.pre[0] = PRE_REAL;
  J.prew_sel[0] = psel;
  J.mul[0] = MUL_PRE;
  J.post[0] = POST_REAL;
  J.postw_sel[0] = psel;
  J.posts[0] = 1.0;
  return J;
}

static HalfJob job_dipconv(const bdgx_ctx *c) {
  HalfJob J;
  memset(&J, 0, sizeof(J));
  J.nch = 1;
  J.pre[0] = PRE_CONJ_PSI;
  J.prew_sel[0] = F_WF;
  J.mul[0] = MUL_DIP;
  J.post[0] = POST_PSI;
  J.postw_sel[0] = F_WF;
  J.posts[0] = c->g_ddi;
  return J;
}

template <typename R>
static int apply_t(bdgx_ctx *c, int op, const void *dX, void *dY, uint64_t ncols,
                   uint64_t ld) {
  const R *X = reinterpret_cast<const R *>(dX);
  R *Y = reinterpret_cast<R *>(dY);
  HalfJob jobs[2];
  int nj = 0;
  const bool needs_wf = (op == BDGX_OP_DIPOLAR_CONV) ||
                        ((op == BDGX_OP_M || op == BDGX_OP_LREP) && c->has_U && c->dim == 3);
  if (needs_wf && !c->has_wf) return BDGX_ESTATE;
  switch (op) {
    case BDGX_OP_KINETIC: jobs[nj++] = job_kin(0, F_NONE); break;
    case BDGX_OP_K:
      if (!c->has_lt) return BDGX_ESTATE;
      jobs[nj++] = job_kin(0, F_LTK);
      break;
    case BDGX_OP_M:
      if (!c->has_lt) return BDGX_ESTATE;
      jobs[nj++] = job_M(c, 0);
      break;
    case BDGX_OP_LREP:
      if (!c->has_lt) return BDGX_ESTATE;
      jobs[nj++] = job_kin(0, F_LTK);
      jobs[nj++] = job_M(c, 1);
      break;
    case BDGX_OP_PRECOND_K:
      if (!c->has_p) return BDGX_ESTATE;
      jobs[nj++] = job_pre(0, F_PK);
      break;
    case BDGX_OP_PRECOND_M:
      if (!c->has_p) return BDGX_ESTATE;
      jobs[nj++] = job_pre(0, F_PM);
      break;
    case BDGX_OP_PRECOND_LREP:
      if (!c->has_p) return BDGX_ESTATE;
      jobs[nj++] = job_pre(0, F_PK);
      jobs[nj++] = job_pre(1, F_PM);
      break;
    case BDGX_OP_DIPOLAR_CONV:
      if (!c->has_U) return BDGX_ESTATE;
      jobs[nj++] = job_dipconv(c);
      break;
    case BDGX_OP_SWAP: {
      const long long S = c->geo.S;
      if (c->cplx) {
        hipLaunchKernelGGL(k_swap<cx<R>>, dim3(1024), dim3(256), 0, c->stream,
                           reinterpret_cast<const cx<R> *>(X), reinterpret_cast<cx<R> *>(Y),
                           S, (long long)ld, (int)ncols);
      } else {
        hipLaunchKernelGGL(k_swap<R>, dim3(1024), dim3(256), 0, c->stream, X, Y, S,
                           (long long)ld, (int)ncols);
      }
      BDGX_HIP(hipGetLastError());
      return BDGX_OK;
    }
    default:
      return BDGX_EINVAL;
  }
  // In-place (dX == dY) is safe: every forward x pass finishes before the first
  // inverse x pass, and each output element is written by the thread that reads
  // the matching input element.
  return run_jobs<R>(c, jobs, nj, X, Y, (long long)ncols, (long long)ld, nullptr);
}

// scale * Re IFFT(U/S FFT(|psi|^2)) into dOut (real, size) -- the dipolar mean
// field of reference src/dipolar_meanfield_impl.inc:20-55 with scale = g_ddi.
template <typename R>
int meanfield_t(bdgx_ctx *c, const R *dpsi, R *dOut, double scale) {
  HalfJob J;
  memset(&J, 0, sizeof(J));
  J.nch = 1;
  J.pre[0] = PRE_ABS2;
  J.mul[0] = MUL_DIP;
  J.post[0] = POST_ONE;
  J.posts[0] = scale;
  J.out_real = 1;
  return run_jobs<R>(c, &J, 1, dpsi, dOut, 1, c->geo.S, nullptr);
}

// In place: psi = IFFT(exp(-dt k^2/2)/S * FFT psi) for dt >= 0, or the kinetic
// operator IFFT(k^2/(2S) * FFT psi) for dt < 0.
template <typename R>
int itp_kinetic_t(bdgx_ctx *c, R *dpsi, double dt) {
  HalfJob J = job_kin(0, F_NONE);
  if (dt >= 0.0) J.mul[0] = MUL_ITPKIN;
  c->itp_dt = dt >= 0.0 ? dt : 0.0;
  return run_jobs<R>(c, &J, 1, dpsi, dpsi, 1, c->geo.S, nullptr);
}

// Learned-preconditioner building blocks (FP32; see learned.hip).
// lp_filter: Y = a_w * C_w(b_w * X) (mode 0) or Y = C_w(a_w * X) (mode 1),
// C_w the k-space table filter of stage w (0: K, 1: M).
int lp_filter_f32(bdgx_ctx *c, int w, int mode, const float *X, float *Y, long long ncols,
                  long long ld) {
  HalfJob J;
  memset(&J, 0, sizeof(J));
  J.nch = 1;
  J.pre[0] = PRE_REAL;
  J.prew_sel[0] = mode == 0 ? (w ? F_LBM : F_LBK) : (w ? F_LAM : F_LAK);
  J.mul[0] = w ? MUL_TABM : MUL_TABK;
  J.post[0] = mode == 0 ? POST_REAL : POST_ONE;
  J.postw_sel[0] = mode == 0 ? (w ? F_LAM : F_LAK) : F_NONE;
  J.posts[0] = 1.0;
  return run_jobs<float>(c, &J, 1, X, Y, ncols, ld, nullptr);
}

// Forward 3D spectra of a_w * P (channel 0) and b_w * Q (channel 1) for the
// columns of X2 = [P; Q] (2*size rows, ld2), left in the context's spectrum
// workspace: spec[(chan * ncols + col) * Kk + kidx].
int lp_spectra_f32(bdgx_ctx *c, int w, const float *X2, long long ncols, long long ld2) {
  HalfJob J[2];
  memset(J, 0, sizeof(J));
  for (int q = 0; q < 2; ++q) {
    J[q].in_half = q;
    J[q].nch = 1;
    J[q].pre[0] = PRE_REAL;
    J[q].prew_sel[0] = q == 0 ? (w ? F_LAM : F_LAK) : (w ? F_LBM : F_LBK);
    J[q].mul[0] = MUL_ONE;
    J[q].posts[0] = 1.0;
  }
  return run_jobs<float>(c, J, 2, X2, nullptr, ncols, ld2, nullptr, true);
}

// Times one pass kind over nbatch single-channel (kinetic) batches.
template <typename R>
static int bench_pass_t(bdgx_ctx *c, int pass, uint64_t nbatch, int iters, float *ms,
                        double *bytes) {
  const Geo &g = c->geo;
  const int nc = (int)nbatch;
  const int cw = c->cplx ? 2 : 1;
  const size_t es = sizeof(R) * cw;
  int rc = ensure_spec(c, sizeof(cx<R>) * (size_t)g.Kk * nbatch);
  if (rc) return rc;
  cx<R> *spec = reinterpret_cast<cx<R> *>(c->spec);
  R *X = nullptr, *Y = nullptr;
  BDGX_HIP(hipMalloc((void **)&X, es * (size_t)g.S * nbatch));
  if (hipMalloc((void **)&Y, es * (size_t)g.S * nbatch) != hipSuccess) {
    (void)hipFree(X);
    return BDGX_EHIP;
  }
  (void)hipMemsetAsync(X, 0, es * (size_t)g.S * nbatch, c->stream);
  (void)hipMemsetAsync(spec, 0, sizeof(cx<R>) * (size_t)g.Kk * nbatch, c->stream);
  HalfJob J = job_kin(0, F_NONE);
  int mul[kMaxChan] = {MUL_KIN, MUL_KIN, MUL_KIN};
  const long long cstride = (long long)nc * g.Kk;
  XArgs<R> a = make_xargs<R>(c, J, 0, nc, X, Y, (long long)g.S, spec, cstride, nullptr);
  const YZArgs<R> yz = make_yz<R>(c, mul, 1, nc, spec);
  const double specb = (double)sizeof(cx<R>) * (double)g.Kk * (double)nbatch;
  const double realb = (double)es * (double)g.S * (double)nbatch;
  switch (pass) {
    case BDGX_PASS_XFWD:
    case BDGX_PASS_PLANE_FWD: *bytes = realb + specb; break;
    case BDGX_PASS_XINV:
    case BDGX_PASS_PLANE_INV: *bytes = specb + realb; break;
    default: *bytes = 2.0 * specb; break;
  }
  if ((pass == BDGX_PASS_PLANE_FWD || pass == BDGX_PASS_PLANE_INV) && !plane_supported<R>(c)) {
    (void)hipFree(X);
    (void)hipFree(Y);
    return BDGX_EINVAL;
  }
  auto one = [&]() -> int {
    switch (pass) {
      case BDGX_PASS_XFWD: return launch_x<R>(c, a, false);
      case BDGX_PASS_YFWD: return launch_y<R>(c, yz, false);
      case BDGX_PASS_Z: return launch_z<R>(c, yz);
      case BDGX_PASS_YINV: return launch_y<R>(c, yz, true);
      case BDGX_PASS_XINV: return launch_x<R>(c, a, true);
      case BDGX_PASS_PLANE_FWD: return launch_plane_x<R>(c, a, false);
      case BDGX_PASS_PLANE_INV: return launch_plane_x<R>(c, a, true);
      default: return BDGX_EINVAL;
    }
  };
  rc = one();  // warm-up
  if (!rc && hipEventRecord(c->ev0, c->stream) != hipSuccess) rc = BDGX_EHIP;
  for (int i = 0; i < iters && !rc; ++i) rc = one();
  if (!rc && hipEventRecord(c->ev1, c->stream) != hipSuccess) rc = BDGX_EHIP;
  if (!rc && hipEventSynchronize(c->ev1) != hipSuccess) rc = BDGX_EHIP;
  float t = 0.f;
  if (!rc && hipEventElapsedTime(&t, c->ev0, c->ev1) != hipSuccess) rc = BDGX_EHIP;
  *ms = iters > 0 ? t / (float)iters : 0.f;
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(X);
  (void)hipFree(Y);
  return rc;
}

int combine_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t k,
                     const double *dC, void *dY, uint64_t ldy, uint64_t m, uint64_t nrows);
int gram_dispatch(bdgx_ctx *c, int prec, const void *dA, uint64_t lda, uint64_t p,
                  const void *dB, uint64_t ldb, uint64_t q, uint64_t nrows, int swap_b,
                  double *dG);

// Launch-shape autotuner: every variant of every kernel kind with variants
// that this grid uses (plane and fixed z passes on nbatch kinetic batches,
// the real block combine nbatch -> nbatch), 1 warm-up + 3 timed launches,
// HIP events on the ctx stream; the fastest variant is kept per kind.
template <typename R>
static int autotune_t(bdgx_ctx *c, uint64_t nbatch) {
  const int pi = sizeof(R) == 8 ? 0 : 1;
  const bool plane = plane_supported<R>(c), zfix = zpass_fixed_supported<R>(c);
  const bool comb = !c->cplx;
  if ((!plane && !zfix && !comb) || nbatch == 0) {
    c->tuned[pi] = 1;
    return BDGX_OK;
  }
  const Geo &g = c->geo;
  const int nc = (int)nbatch;
  const size_t es = sizeof(R) * (c->cplx ? 2 : 1);
  int rc = ensure_spec(c, sizeof(cx<R>) * (size_t)g.Kk * nbatch);
  if (rc) return rc;
  cx<R> *spec = reinterpret_cast<cx<R> *>(c->spec);
  // X holds 3 nbatch columns and Y 2 nbatch: the combine is timed at the
  // solver's main shape [X | P | W] -> [X | P] (3 nbatch -> 2 nbatch)
  R *X = nullptr, *Y = nullptr;
  BDGX_HIP(hipMalloc((void **)&X, es * (size_t)g.S * nbatch * 3));
  if (hipMalloc((void **)&Y, es * (size_t)g.S * nbatch * 2) != hipSuccess) {
    (void)hipFree(X);
    return BDGX_EHIP;
  }
  double *dC = nullptr;
  if (hipMalloc((void **)&dC, sizeof(double) * nbatch * nbatch * 6) != hipSuccess) {
    (void)hipFree(X);
    (void)hipFree(Y);
    return BDGX_EHIP;
  }
  (void)hipMemsetAsync(dC, 0, sizeof(double) * nbatch * nbatch * 6, c->stream);
  (void)hipMemsetAsync(X, 0, es * (size_t)g.S * nbatch * 3, c->stream);
  (void)hipMemsetAsync(spec, 0, sizeof(cx<R>) * (size_t)g.Kk * nbatch, c->stream);
  HalfJob J = job_kin(0, F_NONE);
  int mul[kMaxChan] = {MUL_KIN, MUL_KIN, MUL_KIN};
  XArgs<R> a = make_xargs<R>(c, J, 0, nc, X, Y, (long long)g.S, spec, (long long)nc * g.Kk,
                             nullptr);
  const YZArgs<R> yz = make_yz<R>(c, mul, 1, nc, spec);
  const int kinds[5] = {BDGX_TUNE_PLANE_FWD, BDGX_TUNE_Z, BDGX_TUNE_PLANE_INV,
                        BDGX_TUNE_COMBINE, BDGX_TUNE_GRAM};
  const int prec = sizeof(R) == 8 ? BDGX_FP64 : BDGX_FP32;
  // Gram output (3 nbatch x nbatch doubles): the head of dC
  for (int q = 0; q < 5 && !rc; ++q) {
    const int kind = kinds[q];
    const bool on = kind == BDGX_TUNE_Z ? zfix
                    : (kind == BDGX_TUNE_COMBINE || kind == BDGX_TUNE_GRAM) ? comb
                                                                             : plane;
    if (!on) continue;
    auto one = [&]() -> int {
      // launch_z fills in the z twiddles and plan before the fixed z pass
      if (kind == BDGX_TUNE_Z) return launch_z<R>(c, yz);
      if (kind == BDGX_TUNE_COMBINE)
        return combine_dispatch(c, prec, X, (uint64_t)g.S, 3 * nbatch, dC, Y, (uint64_t)g.S,
                                2 * nbatch, (uint64_t)g.S);
      if (kind == BDGX_TUNE_GRAM)
        return gram_dispatch(c, prec, X, (uint64_t)g.S, 3 * nbatch, Y, (uint64_t)g.S, nbatch,
                             (uint64_t)g.S, 0, dC);
      return launch_plane_x<R>(c, a, kind == BDGX_TUNE_PLANE_INV);
    };
    int best = c->tune_var[pi][kind];
    float best_ms = 0.f;
    const int nvar = kind == BDGX_TUNE_Z ? zpass_variants(c)
                     : (kind == BDGX_TUNE_PLANE_FWD || kind == BDGX_TUNE_PLANE_INV)
                         ? plane_variants(c)
                         : 4;
    for (int v = 0; v < nvar && !rc; ++v) {
      c->tune_var[pi][kind] = v;
      rc = one();
      if (!rc && hipEventRecord(c->ev0, c->stream) != hipSuccess) rc = BDGX_EHIP;
      for (int i = 0; i < 3 && !rc; ++i) rc = one();
      if (!rc && hipEventRecord(c->ev1, c->stream) != hipSuccess) rc = BDGX_EHIP;
      if (!rc && hipEventSynchronize(c->ev1) != hipSuccess) rc = BDGX_EHIP;
      float t = 0.f;
      if (!rc && hipEventElapsedTime(&t, c->ev0, c->ev1) != hipSuccess) rc = BDGX_EHIP;
      if (!rc && (v == 0 || t < best_ms)) {
        best = v;
        best_ms = t;
      }
    }
    c->tune_var[pi][kind] = best;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(X);
  (void)hipFree(Y);
  (void)hipFree(dC);
  if (!rc) c->tuned[pi] = 1;
  return rc;
}

int autotune_dispatch(bdgx_ctx *c, int prec, uint64_t nbatch) {
  if (prec == BDGX_FP64) return autotune_t<double>(c, nbatch);
  if (prec == BDGX_FP32) return autotune_t<float>(c, nbatch);
  return BDGX_EINVAL;
}

int bench_pass_dispatch(bdgx_ctx *c, int pass, int prec, uint64_t nbatch, int iters, float *ms,
                        double *bytes) {
  if (prec == BDGX_FP64) return bench_pass_t<double>(c, pass, nbatch, iters, ms, bytes);
  if (prec == BDGX_FP32) return bench_pass_t<float>(c, pass, nbatch, iters, ms, bytes);
  return BDGX_EINVAL;
}

int apply_dispatch(bdgx_ctx *c, int op, int prec, const void *dX, void *dY,
                   uint64_t ncols, uint64_t ld) {
  if (prec == BDGX_FP64) return apply_t<double>(c, op, dX, dY, ncols, ld);
  if (prec == BDGX_FP32) return apply_t<float>(c, op, dX, dY, ncols, ld);
  return BDGX_EINVAL;
}

template int meanfield_t<double>(bdgx_ctx *, const double *, double *, double);
template int meanfield_t<float>(bdgx_ctx *, const float *, float *, double);
template int itp_kinetic_t<double>(bdgx_ctx *, double *, double);
template int itp_kinetic_t<float>(bdgx_ctx *, float *, double);

}  // namespace bdgx
