// Batched mixed-radix Stockham FFT on pencils staged in LDS (gfx950).
//
// A workgroup holds `npen` pencils of length n in LDS (pencil p at offset
// p*ps) and runs the radix stages of an AxisPlan over all of them at once:
// every thread owns whole radix-R butterflies, reads its R inputs from one
// LDS buffer, applies the stage twiddles from an LDS table, runs the
// butterfly in registers and writes the R outputs (auto-sorted) into the
// other buffer.  Radices 2, 3, 4, 5, 7 and 8 are unrolled at compile time so
// that the per-butterfly values stay in VGPRs.
#pragma once

#include "bdgx_internal.hpp"

namespace bdgx {

template <typename R>
__device__ __forceinline__ cx<R> cmk(R a, R b) {
  cx<R> r;
  r.re = a;
  r.im = b;
  return r;
}
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
  return cmk<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
  return cmk<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return cmk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) {
  return cmk<R>(a.re * s, a.im * s);
}
template <typename R>
__device__ __forceinline__ cx<R> cconj(cx<R> a) {
  return cmk<R>(a.re, -a.im);
}
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV>
__device__ __forceinline__ cx<R> rot90(cx<R> a) {
  return INV ? cmk<R>(-a.im, a.re) : cmk<R>(a.im, -a.re);
}

// In-register DFT of length RAD; forward uses exp(-2 pi i qr/RAD).
template <typename R, bool INV, int RAD>
struct Bfly;

template <typename R, bool INV>
struct Bfly<R, INV, 2> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const cx<R> a = v[0];
    v[0] = a + v[1];
    v[1] = a - v[1];
  }
};

template <typename R, bool INV>
struct Bfly<R, INV, 3> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const R c = R(-0.5), s = R(0.86602540378443864676);
    const cx<R> t1 = v[1] + v[2];
    const cx<R> t2 = v[1] - v[2];
    const cx<R> m = v[0] + cscale(t1, c);
    const cx<R> r = rot90<R, INV>(cscale(t2, s));
    v[0] = v[0] + t1;
    v[1] = m + r;
    v[2] = m - r;
  }
};

template <typename R, bool INV>
struct Bfly<R, INV, 4> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const cx<R> a0 = v[0] + v[2], a1 = v[0] - v[2];
    const cx<R> b0 = v[1] + v[3];
    const cx<R> b1 = rot90<R, INV>(v[1] - v[3]);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + b1;
    v[3] = a1 - b1;
  }
};

template <typename R, bool INV>
struct Bfly<R, INV, 5> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    const cx<R> p14 = v[1] + v[4], m14 = v[1] - v[4];
    const cx<R> p23 = v[2] + v[3], m23 = v[2] - v[3];
    const cx<R> a1 = v[0] + cscale(p14, c1) + cscale(p23, c2);
    const cx<R> a2 = v[0] + cscale(p14, c2) + cscale(p23, c1);
    const cx<R> b1 = rot90<R, INV>(cscale(m14, s1) + cscale(m23, s2));
    const cx<R> b2 = rot90<R, INV>(cscale(m14, s2) - cscale(m23, s1));
    v[0] = v[0] + p14 + p23;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
  }
};

template <typename R, bool INV>
struct Bfly<R, INV, 8> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const R h = R(0.70710678118654752440);
    cx<R> e[4] = {v[0], v[2], v[4], v[6]};
    cx<R> o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<R, INV, 4>::run(e, nullptr, 0);
    Bfly<R, INV, 4>::run(o, nullptr, 0);
    // o[q] *= w8^q, w8 = exp(-+ 2 pi i / 8)
    const cx<R> o1 = o[1], o3 = o[3];
    o[1] = INV ? cmk<R>(h * (o1.re - o1.im), h * (o1.re + o1.im))
               : cmk<R>(h * (o1.re + o1.im), h * (o1.im - o1.re));
    o[2] = rot90<R, INV>(o[2]);
    o[3] = INV ? cmk<R>(-h * (o3.re + o3.im), h * (o3.re - o3.im))
               : cmk<R>(h * (o3.im - o3.re), -h * (o3.re + o3.im));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = e[q] + o[q];
      v[q + 4] = e[q] - o[q];
    }
  }
};

// 16 = 4 x 4: x[4 n1 + n2] -> DFT4 over n1, twiddle w16^(n2 k1), DFT4 over
// n2 -> X[k1 + 4 k2] (natural order out)
template <typename R, bool INV>
__device__ __forceinline__ cx<R> tw16(cx<R> a, R c, R s) {  // a * exp(-+ i theta)
  return cmul(a, cmk<R>(c, INV ? s : -s));
}
template <typename R, bool INV>
struct Bfly<R, INV, 16> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *, int) {
    const R c1 = R(0.92387953251128675613), s1 = R(0.38268343236508977173);
    const R h = R(0.70710678118654752440);
    cx<R> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cx<R> t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Bfly<R, INV, 4>::run(t, nullptr, 0);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // w16^j for j = n2 k1 in {1, 2, 3, 4, 6, 9}: (cos, sin) of 2 pi j / 16
    a[1][1] = tw16<R, INV>(a[1][1], c1, s1);
    a[1][2] = tw16<R, INV>(a[1][2], h, h);
    a[1][3] = tw16<R, INV>(a[1][3], s1, c1);
    a[2][1] = tw16<R, INV>(a[2][1], h, h);
    a[2][2] = rot90<R, INV>(a[2][2]);
    a[2][3] = tw16<R, INV>(a[2][3], -h, h);
    a[3][1] = tw16<R, INV>(a[3][1], s1, c1);
    a[3][2] = tw16<R, INV>(a[3][2], -h, h);
    a[3][3] = tw16<R, INV>(a[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<R> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Bfly<R, INV, 4>::run(t, nullptr, 0);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// generic odd prime radix from the n-point twiddle table
template <typename R, bool INV>
struct Bfly<R, INV, 7> {
  static __device__ __forceinline__ void run(cx<R> *v, const cx<R> *tw, int n) {
    const int unit = n / 7;
    cx<R> out[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      cx<R> acc = v[0];
#pragma unroll
      for (int r = 1; r < 7; ++r) {
        cx<R> w = tw[((q * r) % 7) * unit];
        if (INV) w.im = -w.im;
        acc = acc + cmul(v[r], w);
      }
      out[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) v[q] = out[q];
  }
};

template <typename R, bool INV, int RAD>
__device__ __forceinline__ void stockham_stage(const cx<R> *__restrict__ a,
                                               cx<R> *__restrict__ b,
                                               const cx<R> *tw, int n, int Ns,
                                               int npen, int ps) {
  const int nr = n / RAD;
  const int step = n / (Ns * RAD);
  const int total = npen * nr;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int p = t / nr;
    const int j = t - p * nr;
    const int k = j % Ns;
    const cx<R> *src = a + p * ps + j;
    cx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[r] = src[r * nr];
    if (k) {
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        cx<R> w = tw[r * k * step];
        if (INV) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
    }
    Bfly<R, INV, RAD>::run(v, tw, n);
    cx<R> *dst = b + p * ps + (j / Ns) * Ns * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) dst[r * Ns] = v[r];
  }
}

// Runs all stages; returns the buffer (a or b) that holds the result.
// Caller must __syncthreads() after filling `a`; this returns after a barrier.
template <typename R, bool INV>
__device__ cx<R> *lds_fft(cx<R> *a, cx<R> *b, const cx<R> *tw,
                          const AxisPlan &pl, int npen, int ps) {
  const int n = pl.n;
  for (int s = 0; s < pl.nst; ++s) {
    const int Ns = pl.ns[s];
    switch (pl.rad[s]) {
      case 2: stockham_stage<R, INV, 2>(a, b, tw, n, Ns, npen, ps); break;
      case 3: stockham_stage<R, INV, 3>(a, b, tw, n, Ns, npen, ps); break;
      case 4: stockham_stage<R, INV, 4>(a, b, tw, n, Ns, npen, ps); break;
      case 5: stockham_stage<R, INV, 5>(a, b, tw, n, Ns, npen, ps); break;
      case 7: stockham_stage<R, INV, 7>(a, b, tw, n, Ns, npen, ps); break;
      default: stockham_stage<R, INV, 8>(a, b, tw, n, Ns, npen, ps); break;
    }
    __syncthreads();
    cx<R> *t = a;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace bdgx
