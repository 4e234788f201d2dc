// Compile-time-sized Stockham FFT for P pencils held element-major in LDS
// (element i of pencil p at i*P + p).  Consecutive threads own consecutive
// pencils, so LDS accesses are conflict-free and, for pencils along a strided
// axis of the spectrum (consecutive p = consecutive kx), the first stage can
// read its inputs straight from HBM and the last stage write its outputs
// straight back, coalesced.  All index arithmetic is on constants.
#pragma once

#include "fft_core.hpp"

namespace bdgx {

// radix plan: as many 8s as possible, then 4, 2, 3, 5; with r16 a leftover
// 2 is merged with the first 8 into a 16 (128 = 16 x 8: one LDS round trip
// fewer per transform, at 16 complex values per thread in the butterfly)
constexpr bool kFixedRadix16 = true;

struct FixedPlan {
  int nst;
  int rad[12];
  int ns[12];
};

constexpr FixedPlan make_fixed_plan(int n, bool r16 = kFixedRadix16) {
  FixedPlan p{};
  int m = n, nf = 0;
  int f[12] = {};
  while (m % 8 == 0 && m > 1) { f[nf++] = 8; m /= 8; }
  if (r16 && nf > 0 && m == 2) {
    f[0] = 16;
    m = 1;
  }
  while (m % 4 == 0 && m > 1) { f[nf++] = 4; m /= 4; }
  while (m % 2 == 0 && m > 1) { f[nf++] = 2; m /= 2; }
  while (m % 3 == 0 && m > 1) { f[nf++] = 3; m /= 3; }
  while (m % 5 == 0 && m > 1) { f[nf++] = 5; m /= 5; }
  int Ns = 1;
  for (int s = 0; s < nf; ++s) {
    p.rad[s] = f[s];
    p.ns[s] = Ns;
    Ns *= f[s];
  }
  p.nst = nf;
  return p;
}

// Sources / sinks of a stage: LDS (element-major) or a strided HBM pencil.
template <typename R>
struct LdsIO {
  cx<R> *buf;
  template <int P>
  __device__ __forceinline__ cx<R> ld(int i, int p) const { return buf[i * P + p]; }
  template <int P>
  __device__ __forceinline__ void st(int i, int p, cx<R> v) const { buf[i * P + p] = v; }
};

template <typename R>
struct GlobIO {  // element i of pencil p at base[i * stride + p * pstride]; p < valid
  cx<R> *base;
  long long stride;
  int valid;
  long long pstride = 1;
  template <int P>
  __device__ __forceinline__ cx<R> ld(int i, int p) const {
    return p < valid ? base[i * stride + p * pstride] : cmk<R>(0, 0);
  }
  template <int P>
  __device__ __forceinline__ void st(int i, int p, cx<R> v) const {
    if (p < valid) base[i * stride + p * pstride] = v;
  }
};

// optional per-output multiplier applied after the butterfly of the last
// forward stage (identity by default)
struct NoMul {
  template <typename R>
  __device__ __forceinline__ cx<R> operator()(cx<R> v, int, int) const { return v; }
};

template <typename R, bool INV, int N, int RAD, int NS, int P, int NT, class SRC, class DST,
          class MUL>
__device__ __forceinline__ void fixed_stage(const SRC &src, const DST &dst, const cx<R> *tw,
                                            const MUL &mul) {
  constexpr int NR = N / RAD;
  constexpr int STEP = N / (NS * RAD);
  constexpr int TOTAL = P * NR;
  for (int t = threadIdx.x; t < TOTAL; t += NT) {
    const int p = t % P;   // P is a power of two: a mask
    const int j = t / P;
    const int k = j % NS;
    cx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[r] = src.template ld<P>(j + r * NR, p);
    if (NS > 1 && k) {
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        cx<R> w = tw[r * k * STEP];
        if (INV) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
    }
    Bfly<R, INV, RAD>::run(v, tw, N);
    const int base = (j / NS) * NS * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) dst.template st<P>(base + r * NS, p, mul(v[r], base + r * NS, p));
  }
}

// Stage dispatcher over the compile-time plan.  Stage 0 reads `first` and
// writes X, stage 1 reads X and writes Y, and so on alternating; the last
// stage writes `last` (through `mul`) instead.  Barriers follow every stage.
// With nst stages the last LDS buffer written before `last` is X when nst is
// even, Y when odd (for nst == 1 none is).
template <typename R, bool INV, int N, int P, int NT, int S, bool R16 = kFixedRadix16,
          class FIRST, class LAST, class MUL>
__device__ __forceinline__ void fixed_stages(const FIRST &first, const LAST &last, cx<R> *X,
                                             cx<R> *Y, const cx<R> *tw, const MUL &mul) {
  constexpr FixedPlan pl = make_fixed_plan(N, R16);
  if constexpr (S < pl.nst) {
    constexpr int RAD = pl.rad[S];
    constexpr int NS = pl.ns[S];
    constexpr bool is_first = S == 0, is_last = S == pl.nst - 1;
    LdsIO<R> in{(S % 2 == 1) ? X : Y};
    LdsIO<R> out{(S % 2 == 0) ? X : Y};
    if constexpr (is_first && is_last)
      fixed_stage<R, INV, N, RAD, NS, P, NT>(first, last, tw, mul);
    else if constexpr (is_first)
      fixed_stage<R, INV, N, RAD, NS, P, NT>(first, out, tw, NoMul{});
    else if constexpr (is_last)
      fixed_stage<R, INV, N, RAD, NS, P, NT>(in, last, tw, mul);
    else
      fixed_stage<R, INV, N, RAD, NS, P, NT>(in, out, tw, NoMul{});
    __syncthreads();
    fixed_stages<R, INV, N, P, NT, S + 1, R16>(first, last, X, Y, tw, mul);
  }
}

}  // namespace bdgx
