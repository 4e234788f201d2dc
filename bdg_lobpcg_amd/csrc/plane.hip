// Whole-plane x/y transforms for real vectors on gfx950.
//
// For grids whose (y, kx) spectrum plane fits in LDS (Nx = Ny = 32, 64, 128
// in FP32 and FP64: 137 KB at 128^2 FP64 of the 160 KB per CU), one
// workgroup owns one z-plane of one column:
//
//   forward  read the real plane (contiguous, coalesced) once into registers;
//            per channel: pre-weight, x transform of the row pairs packed as
//            complex rows (two-for-one), untangle into the Nx/2+1 half
//            spectrum, y transform, write the spectrum plane (contiguous);
//   inverse  per channel: read the spectrum plane, y inverse, repack row
//            pairs, x inverse, fold the real rows into the output kept in
//            registers; then add lt * x and write the output plane.
//
// This replaces the x pass and the y pass of each direction (five launches
// per transform pair become three with the fused z pass), and every HBM
// access of these two kernels is a contiguous plane.  The transforms run
// in place (decimation in frequency forward, in time inverse), so a single
// LDS plane suffices.  All sizes are compile-time constants.
#include <string.h>

#include "bdgx_internal.hpp"
#include "fft_core.hpp"
#include "fft_fixed.hpp"
#include "plane.hpp"

namespace bdgx {

// In-place mixed-radix transforms over NP pencils, element i of pencil p at
// buf[p*PST + i*EST].  Decimation in frequency (DIF) takes natural order and
// leaves X[k] at position dif_pos<N, R16>(k) (digits of k in the plan's radices,
// first radix least significant, reversed); decimation in time (DIT) takes
// that order and returns natural order.  Every butterfly reads and writes the
// same RAD positions, so a stage needs no staging buffer and only one barrier;
// each thread keeps one butterfly in registers at a time.  PFAST maps
// consecutive threads to consecutive pencils (pencils adjacent in memory),
// otherwise to consecutive butterflies of one pencil.
template <int N, bool R16>
__device__ __forceinline__ int dif_pos(int k) {
  constexpr FixedPlan pl = make_fixed_plan(N, R16);
  int p = 0, span = N;
#pragma unroll
  for (int s = 0; s < pl.nst; ++s) {
    span /= pl.rad[s];
    p += (k % pl.rad[s]) * span;
    k /= pl.rad[s];
  }
  return p;
}

// LDS index of element e of the row-major x layout (rows of N complex
// values), for ES-byte elements (0 = no swizzle).  A pass of 16 FP64 (32
// FP32) complex values covers the 64 LDS banks once; within each such chunk
// the low bits of e are XORed with g(h), h = chunk index mod 8:
//   FP64: g(h) = 2h + (h & 1)            (chunks of 16)
//   FP32: g(h) = (5h + 4 (h & 1)) mod 32 (chunks of 32)
// With these the lanes of a pass hit distinct bank groups in every x stage
// (radices 8, 8, 2 at N = 128), in the digit-reversed untangle and repack
// and in the row loads (FP32: at most 2-way), where the plain layout has up
// to 8-way conflicts.  The permutation stays inside its chunk, so a row keeps
// its LDS footprint.
template <int ES>
__device__ __forceinline__ int swz(int e) {
  if constexpr (ES == 16) {
    const int h = (e >> 4) & 7;
    return e ^ (2 * h + (h & 1));
  } else if constexpr (ES == 8) {
    const int h = (e >> 5) & 7;
    return e ^ ((5 * h + 4 * (h & 1)) & 31);
  } else {
    return e;
  }
}

// Swizzle of the x layout of an N-point plane: the permutation stays inside
// its chunk (16 FP64 / 32 FP32 values), so a row must be whole chunks for
// the in-place untangle and repack rounds (which rewrite rows in place) to
// touch disjoint LDS ranges; other row lengths (N = 48, 80 in FP32) run
// unswizzled.
template <typename R, int N>
constexpr int plane_swz() {
  return N % (sizeof(R) == 8 ? 16 : 32) == 0 ? (int)sizeof(cx<R>) : 0;
}

template <typename R, bool INV, bool DIT, int N, int RAD, int L, int NP, int PST, int EST,
          int T, bool PFAST, int SWZ>
__device__ __forceinline__ void ip_stage(cx<R> *buf, const cx<R> *tw, int tid) {
  constexpr int LR = L / RAD;
  constexpr int NBF = N / RAD;
  constexpr int TOTAL = NP * NBF;
  constexpr int TSTEP = N / L;
  for (int t = tid; t < TOTAL; t += T) {
    const int p = PFAST ? t % NP : t / NBF;
    const int u = PFAST ? t / NP : t % NBF;
    const int blk = u / LR, m = u % LR;
    const int e0 = p * PST + (blk * L + m) * EST;
    cx<R> v[RAD];
#pragma unroll
    for (int j = 0; j < RAD; ++j) v[j] = buf[swz<SWZ>(e0 + j * LR * EST)];
    if (DIT && LR > 1 && m) {
#pragma unroll
      for (int j = 1; j < RAD; ++j) {
        cx<R> w = tw[j * m * TSTEP];
        if (INV) w.im = -w.im;
        v[j] = cmul(v[j], w);
      }
    }
    Bfly<R, INV, RAD>::run(v, tw, N);
    if (!DIT && LR > 1 && m) {
#pragma unroll
      for (int q = 1; q < RAD; ++q) {
        cx<R> w = tw[q * m * TSTEP];
        if (INV) w.im = -w.im;
        v[q] = cmul(v[q], w);
      }
    }
#pragma unroll
    for (int q = 0; q < RAD; ++q) buf[swz<SWZ>(e0 + q * LR * EST)] = v[q];
  }
  __syncthreads();
}

// DIF: stages S = 0 .. nst-1 with block length L = N / ns[S]
template <typename R, bool INV, int N, int NP, int PST, int EST, int T, bool PFAST, int SWZ,
          int S, bool R16>
__device__ __forceinline__ void dif_fft(cx<R> *buf, const cx<R> *tw, int tid) {
  constexpr FixedPlan pl = make_fixed_plan(N, R16);
  if constexpr (S < pl.nst) {
    ip_stage<R, INV, false, N, pl.rad[S], N / pl.ns[S], NP, PST, EST, T, PFAST, SWZ>(buf, tw, tid);
    dif_fft<R, INV, N, NP, PST, EST, T, PFAST, SWZ, S + 1, R16>(buf, tw, tid);
  }
}

// DIT: the same stages in reverse order (S = nst-1 .. 0)
template <typename R, bool INV, int N, int NP, int PST, int EST, int T, bool PFAST, int SWZ,
          int S, bool R16>
__device__ __forceinline__ void dit_fft(cx<R> *buf, const cx<R> *tw, int tid) {
  constexpr FixedPlan pl = make_fixed_plan(N, R16);
  if constexpr (S >= 0) {
    ip_stage<R, INV, true, N, pl.rad[S], N / pl.ns[S], NP, PST, EST, T, PFAST, SWZ>(buf, tw, tid);
    dit_fft<R, INV, N, NP, PST, EST, T, PFAST, SWZ, S - 1, R16>(buf, tw, tid);
  }
}

template <int N, bool R16>
constexpr int last_stage() {
  return make_fixed_plan(N, R16).nst - 1;
}

// rows handled per round of the in-place untangle / repack: the largest
// power of two dividing N/2 that keeps at most `per` staged values per
// thread (the rounds cover the N/2 row pairs exactly)
template <int N, int T, int per>
constexpr int group_rows(int row_len) {
  int p = 1;
  while ((N / 2) % (2 * p) == 0 && 2 * p * row_len <= per * T) p *= 2;
  return p;
}

// the largest group size <= g that divides n (loads issued in whole groups)
constexpr int div_group(int n, int g) {
  while (g > 1 && n % g) --g;
  return g;
}

// x transform, untangle, y transform of the weighted plane of channel c held
// in LDS (complex rows of row pairs, swizzled), then the spectrum plane out
template <typename R, int N, int T, bool R16 = kFixedRadix16>
__device__ __forceinline__ void plane_fwd_transform(const XArgs<R> &a, cx<R> *buf,
                                                    const cx<R> *twx, const cx<R> *twy, int c,
                                                    int col, int z, int tid) {
  constexpr int NK = N / 2 + 1;
  constexpr int SW = plane_swz<R, N>();
  dif_fft<R, false, N, N / 2, N, 1, T, false, SW, 0, R16>(buf, twx, tid);
  // untangle Z_q (digit-reversed in k) into the half spectra of rows 2q,
  // 2q+1 (natural kx, row stride NK); rounds of G rows from the top down,
  // so that no row is overwritten before it is read
  constexpr int G = group_rows<N, T, 5>(NK);
  constexpr int NU = G * NK;
  constexpr int BU = (NU + T - 1) / T;
  for (int q0 = N / 2 - G; q0 >= 0; q0 -= G) {
    cx<R> ua[BU], ub[BU];
#pragma unroll
    for (int b = 0; b < BU; ++b) {
      const int t = tid + b * T;
      if (NU % T == 0 || t < NU) {
        const int q = q0 + t / NK, k = t % NK;
        const cx<R> zk = buf[swz<SW>(q * N + dif_pos<N, R16>(k))];
        const cx<R> zn = buf[swz<SW>(q * N + dif_pos<N, R16>(k == 0 ? 0 : N - k))];
        ua[b] = cmk<R>(R(0.5) * (zk.re + zn.re), R(0.5) * (zk.im - zn.im));
        ub[b] = cmk<R>(R(0.5) * (zk.im + zn.im), R(-0.5) * (zk.re - zn.re));
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BU; ++b) {
      const int t = tid + b * T;
      if (NU % T == 0 || t < NU) {
        const int q = q0 + t / NK, k = t % NK;
        buf[(2 * q) * NK + k] = ua[b];
        buf[(2 * q + 1) * NK + k] = ub[b];
      }
    }
    __syncthreads();
  }
  // y transform over the NK adjacent pencils; rows come out digit-reversed
  dif_fft<R, false, N, NK, 1, NK, T, true, 0, 0, R16>(buf, twy, tid);
  cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride + (long long)col * a.g.Kk +
              (long long)z * N * NK;
  if constexpr (sizeof(R) == 8 || T > 4 * N) {  // (no registers to spare)
    for (int e = tid; e < N * NK; e += T) {
      const int y = e / NK, k = e % NK;
      sp[e] = buf[dif_pos<N, R16>(y) * NK + k];
    }
    __syncthreads();
    return;
  }
  // LDS reads in groups of SG per thread ahead of their stores
  constexpr int ES = (N * NK + T - 1) / T;
  constexpr int SG = 4;
#pragma unroll
  for (int b0 = 0; b0 < ES; b0 += SG) {
    cx<R> v[SG];
#pragma unroll
    for (int i = 0; i < SG; ++i) {
      const int b = b0 + i, e = tid + b * T;
      if (b < ES && (N * NK % T == 0 || b < ES - 1 || e < N * NK)) {
        const int y = e / NK, k = e % NK;
        v[i] = buf[dif_pos<N, R16>(y) * NK + k];
      }
    }
#pragma unroll
    for (int i = 0; i < SG; ++i) {
      const int b = b0 + i, e = tid + b * T;
      if (b < ES && (N * NK % T == 0 || b < ES - 1 || e < N * NK)) sp[e] = v[i];
    }
  }
  __syncthreads();
}

// Register budget of the one-plane kernels: FP32 planes at N = 128 (67 KB
// of LDS) run two inverse workgroups per CU when each 4N-thread workgroup
// keeps to 128 VGPRs (4 waves per SIMD), with loads grouped by 2 instead of
// 4.
template <typename R, int N, int T>
constexpr bool plane_two_per_cu() {
  return sizeof(R) == 4 && N == 128 && T == 4 * N;
}
template <typename R, int N, int T>
constexpr int plane_waves_per_simd() {
  return plane_two_per_cu<R, N, T>() ? 4 : 1;
}
template <typename R, int N, int T>
constexpr int plane_load_group() {
  return plane_two_per_cu<R, N, T>() ? 2 : 4;
}
// Radix-16 stages in the inverse kernels except the FP64 persistent one,
// which holds the next plane's spectrum and the output rows: there radix 16
// spills about 100 VGPRs and runs 16% slower (the FP32 kernel held to 128
// VGPRs spills 25 and still gains 3%).  The radix plan only orders values
// inside one kernel's LDS; HBM planes are natural order either way.
template <typename R, int N, int T, bool PERSIST>
constexpr bool plane_inv_r16() {
  return kFixedRadix16 && !(PERSIST && sizeof(R) == 8);
}

// One plane per workgroup, nothing held across the transform (fewer
// registers: more workgroups per CU where the LDS allows it).  Each channel
// reads its input values in groups of LG per thread, a group's loads all in
// flight together.
template <typename R, int N, int T, bool R16>
__global__ void __launch_bounds__(T) k_plane_fwd_real(XArgs<R> a) {
  constexpr int SW = plane_swz<R, N>();
  constexpr int EP = N * N / 2 / T;  // row-pair values per thread
  static_assert(EP * T == N * N / 2, "whole row pairs per thread");
  constexpr int LG = div_group(EP, plane_load_group<R, N, T>());
  static_assert(EP % LG == 0, "load groups");
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * (N / 2 + 1)];
  const int nz = a.g.nz;
  const int z = blockIdx.x % nz;
  const int col = blockIdx.x / nz;
  for (int i = threadIdx.x; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const long long plane0 = (long long)z * N * N;
  const R *xp = a.X + (long long)col * a.ld + (long long)a.in_half * a.g.S + plane0;
  for (int c = 0; c < a.nch; ++c) {
    const int tid = threadIdx.x;
    const int pre = a.pre[c];
    const bool weighted = pre == PRE_REAL || pre == PRE_CONJ_PSI;
    const R *pw = a.prew[c] + plane0;
#pragma unroll
    for (int b0 = 0; b0 < EP; b0 += LG) {
      R x0[LG], x1[LG], q0[LG], q1[LG];
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        const int g = tid + (b0 + i) * T;
        const int g0 = 2 * (g / N) * N + g % N;
        x0[i] = xp[g0];
        x1[i] = xp[g0 + N];
      }
      if (weighted) {
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          const int g = tid + (b0 + i) * T;
          const int g0 = 2 * (g / N) * N + g % N;
          q0[i] = pw[g0];
          q1[i] = pw[g0 + N];
        }
      } else if (pre == PRE_ABS2) {
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          q0[i] = x0[i];
          q1[i] = x1[i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          q0[i] = R(1);
          q1[i] = R(1);
        }
      }
#pragma unroll
      for (int i = 0; i < LG; ++i)
        buf[swz<SW>(tid + (b0 + i) * T)] = cmk<R>(x0[i] * q0[i], x1[i] * q1[i]);
    }
    __syncthreads();
    plane_fwd_transform<R, N, T, R16>(a, buf, twx, twy, c, col, z, tid);
  }
}

// Persistent forward kernel: each workgroup takes planes blockIdx.x,
// + gridDim.x, ... and loads the next plane's values into registers while the
// current plane is transformed in LDS, so that its HBM reads overlap the LDS
// work.
template <typename R, int N, int T, bool R16>
__global__ void __launch_bounds__(T) k_plane_fwd_persist(XArgs<R> a, int nplanes) {
  constexpr int NK = N / 2 + 1;
  constexpr int SW = plane_swz<R, N>();
  constexpr int EP = N * N / 2 / T;  // row-pair values per thread
  static_assert(EP * T == N * N / 2, "whole row pairs per thread");
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * NK];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const int nz = a.g.nz;
  R x0[EP], x1[EP];
  auto load = [&](int p) {
    const int z = p % nz, col = p / nz;
    const R *xp = a.X + (long long)col * a.ld + (long long)a.in_half * a.g.S +
                  (long long)z * N * N;
#pragma unroll
    for (int b = 0; b < EP; ++b) {
      const int g = tid + b * T;
      const int g0 = 2 * (g / N) * N + g % N;
      x0[b] = xp[g0];
      x1[b] = xp[g0 + N];
    }
  };
  int p = blockIdx.x;
  if (p < nplanes) load(p);
  for (; p < nplanes; p += gridDim.x) {
    const int z = p % nz, col = p / nz;
    const long long plane0 = (long long)z * N * N;
    for (int c = 0; c < a.nch; ++c) {
      const int pre = a.pre[c];
      if (pre == PRE_REAL || pre == PRE_CONJ_PSI) {
        // all weights requested before the first is used: one HBM round
        // trip per channel, not one per value
        const R *pw = a.prew[c] + plane0;
        R q0[EP], q1[EP];
#pragma unroll
        for (int b = 0; b < EP; ++b) {
          const int g = tid + b * T;
          const int g0 = 2 * (g / N) * N + g % N;
          q0[b] = pw[g0];
          q1[b] = pw[g0 + N];
        }
#pragma unroll
        for (int b = 0; b < EP; ++b) buf[swz<SW>(tid + b * T)] = cmk<R>(x0[b] * q0[b], x1[b] * q1[b]);
      } else if (pre == PRE_ABS2) {
#pragma unroll
        for (int b = 0; b < EP; ++b) buf[swz<SW>(tid + b * T)] = cmk<R>(x0[b] * x0[b], x1[b] * x1[b]);
      } else {
#pragma unroll
        for (int b = 0; b < EP; ++b) buf[swz<SW>(tid + b * T)] = cmk<R>(x0[b], x1[b]);
      }
      __syncthreads();
      // the registers are free after the last channel's LDS write: fetch
      // the next plane now, its loads in flight during the transform
      if (c == a.nch - 1 && p + (int)gridDim.x < nplanes) load(p + gridDim.x);
      plane_fwd_transform<R, N, T, R16>(a, buf, twx, twy, c, col, z, tid);
    }
  }
}

// y inverse, repack, x inverse of channel c's spectrum plane held in LDS
// (rows digit-reversed), then acc += post-weighted real rows
template <typename R, int N, int T, int EPT, int PGMAX, bool R16>
__device__ __forceinline__ void plane_inv_transform(const XArgs<R> &a, cx<R> *buf,
                                                    const cx<R> *twx, const cx<R> *twy, int c,
                                                    long long plane0, R (&acc0)[EPT],
                                                    R (&acc1)[EPT], int tid) {
  constexpr int NK = N / 2 + 1;
  constexpr int SW = plane_swz<R, N>();
  dit_fft<R, true, N, NK, 1, NK, T, true, 0, last_stage<N, R16>(), R16>(buf, twy, tid);
  // repack: complex row q = spectrum(row 2q) + i spectrum(row 2q+1),
  // extended to all kx by Hermitian symmetry, digit-reversed in x; rounds
  // of G rows from the bottom up
  constexpr int G = group_rows<N, T, 8>(N);
  constexpr int NQ = G * N;
  constexpr int BQ = (NQ + T - 1) / T;
  for (int q0 = 0; q0 < N / 2; q0 += G) {
    cx<R> zq[BQ];
#pragma unroll
    for (int b = 0; b < BQ; ++b) {
      const int t = tid + b * T;
      if (NQ % T == 0 || t < NQ) {
        const int q = q0 + t / N, k = t % N;
        const bool cj = k >= NK;
        const int kk = cj ? N - k : k;
        cx<R> A0 = buf[(2 * q) * NK + kk], A1 = buf[(2 * q + 1) * NK + kk];
        if (cj) {
          A0.im = -A0.im;
          A1.im = -A1.im;
        }
        if (kk == 0 || kk == N / 2) {
          A0.im = 0;
          A1.im = 0;
        }
        zq[b] = cmk<R>(A0.re - A1.im, A0.im + A1.re);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BQ; ++b) {
      const int t = tid + b * T;
      if (NQ % T == 0 || t < NQ) {
        const int q = q0 + t / N, k = t % N;
        buf[swz<SW>(q * N + dif_pos<N, R16>(k))] = zq[b];
      }
    }
    __syncthreads();
  }
  dit_fft<R, true, N, N / 2, N, 1, T, false, SW, last_stage<N, R16>(), R16>(buf, twx, tid);
  const bool weighted = a.post[c] != POST_ONE;
  const R ps = a.posts[c];
  const R *pw = a.postw[c] + plane0;
  if constexpr (PGMAX > 1) {
    // weights requested in groups of PG values per thread, a group's loads
    // all in flight before the first is used
    constexpr int PG = div_group(EPT, PGMAX);
    static_assert(EPT % PG == 0, "weight groups");
#pragma unroll
    for (int b0 = 0; b0 < EPT; b0 += PG) {
      R q0[PG], q1[PG];
      if (weighted) {
#pragma unroll
        for (int i = 0; i < PG; ++i) {
          const int g = tid + (b0 + i) * T;
          const int g0 = 2 * (g / N) * N + g % N;
          const bool in = N * N / 2 % T == 0 || g < N * N / 2;
          q0[i] = ps * pw[in ? g0 : 0];
          q1[i] = ps * pw[in ? g0 + N : 0];
        }
      } else {
#pragma unroll
        for (int i = 0; i < PG; ++i) {
          q0[i] = ps;
          q1[i] = ps;
        }
      }
#pragma unroll
      for (int i = 0; i < PG; ++i) {
        const int g = tid + (b0 + i) * T;
        if (N * N / 2 % T == 0 || g < N * N / 2) {
          const cx<R> zv = buf[swz<SW>(g)];
          acc0[b0 + i] += zv.re * q0[i];
          acc1[b0 + i] += zv.im * q1[i];
        }
      }
    }
  } else {  // one weight at a time
#pragma unroll
    for (int b = 0; b < EPT; ++b) {
      const int g = tid + b * T;
      if (N * N / 2 % T == 0 || g < N * N / 2) {
        const int g0 = 2 * (g / N) * N + g % N;
        const cx<R> zv = buf[swz<SW>(g)];
        R v0 = ps * zv.re, v1 = ps * zv.im;
        if (weighted) {
          v0 *= pw[g0];
          v1 *= pw[g0 + N];
        }
        acc0[b] += v0;
        acc1[b] += v1;
      }
    }
  }
  __syncthreads();
}

// Y = acc + lt * X for one plane; lt and X are read in groups of SG values
// per thread, all loads of a group in flight together (Y may alias X, so
// the loads of a group are issued before its stores)
template <typename R, int N, int T, int EPT, int SGMAX>
__device__ __forceinline__ void plane_inv_store(const XArgs<R> &a, int col, long long plane0,
                                                const R (&acc0)[EPT], const R (&acc1)[EPT],
                                                int tid) {
  constexpr int SG = div_group(EPT, SGMAX);
  static_assert(EPT % SG == 0, "store groups");
  const R *xp = a.X + (long long)col * a.ld + (long long)a.in_half * a.g.S + plane0;
  R *yp = a.Y + (long long)col * a.ld + (long long)a.out_half * a.g.S + plane0;
  const R *lt = a.lt ? a.lt + plane0 : nullptr;
#pragma unroll
  for (int b0 = 0; b0 < EPT; b0 += SG) {
    R o0[SG], o1[SG];
    if (lt) {
      R l0[SG], l1[SG], u0[SG], u1[SG];
#pragma unroll
      for (int i = 0; i < SG; ++i) {
        const int g = tid + (b0 + i) * T;
        if (N * N / 2 % T == 0 || g < N * N / 2) {
          const int g0 = 2 * (g / N) * N + g % N;
          l0[i] = lt[g0];
          l1[i] = lt[g0 + N];
          u0[i] = xp[g0];
          u1[i] = xp[g0 + N];
        }
      }
#pragma unroll
      for (int i = 0; i < SG; ++i) {
        o0[i] = acc0[b0 + i] + l0[i] * u0[i];
        o1[i] = acc1[b0 + i] + l1[i] * u1[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < SG; ++i) {
        o0[i] = acc0[b0 + i];
        o1[i] = acc1[b0 + i];
      }
    }
#pragma unroll
    for (int i = 0; i < SG; ++i) {
      const int g = tid + (b0 + i) * T;
      if (N * N / 2 % T == 0 || g < N * N / 2) {
        const int g0 = 2 * (g / N) * N + g % N;
        yp[g0] = o0[i];
        yp[g0 + N] = o1[i];
      }
    }
  }
}

// One plane per workgroup, nothing held across the transform but the
// output accumulators; each channel's spectrum plane is read in groups of
// LG values per thread, a group's loads all in flight together.
template <typename R, int N, int T, bool R16>
__global__ void __launch_bounds__(T, (plane_waves_per_simd<R, N, T>())) k_plane_inv_real(XArgs<R> a) {
  constexpr int NK = N / 2 + 1;
  constexpr int EPT = (N * N / 2 + T - 1) / T;
  constexpr int ES = (N * NK + T - 1) / T;  // spectrum values per thread
  constexpr int LG = plane_load_group<R, N, T>();
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * NK];
  const int nz = a.g.nz;
  const int z = blockIdx.x % nz;
  const int col = blockIdx.x / nz;
  for (int i = threadIdx.x; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const long long plane0 = (long long)z * N * N;
  R acc0[EPT], acc1[EPT];
#pragma unroll
  for (int b = 0; b < EPT; ++b) {
    acc0[b] = R(0);
    acc1[b] = R(0);
  }
  for (int c = 0; c < a.nch; ++c) {
    // thread index opaque to the compiler inside the channel loop: keeps it
    // from hoisting every swizzled LDS address of the body out of a loop that
    // usually runs once or twice
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                      (long long)col * a.g.Kk + (long long)z * N * NK;
    // rows into digit-reversed order for the DIT y inverse
#pragma unroll
    for (int b0 = 0; b0 < ES; b0 += LG) {
      cx<R> v[LG];
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        const int e = tid + (b0 + i) * T;
        if (b0 + i < ES && (N * NK % T == 0 || b0 + i < ES - 1 || e < N * NK)) v[i] = sp[e];
      }
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        const int e = tid + (b0 + i) * T;
        if (b0 + i < ES && (N * NK % T == 0 || b0 + i < ES - 1 || e < N * NK)) {
          const int y = e / NK, k = e % NK;
          buf[dif_pos<N, R16>(y) * NK + k] = v[i];
        }
      }
    }
    __syncthreads();
    plane_inv_transform<R, N, T, EPT, LG, R16>(a, buf, twx, twy, c,
                                                                           plane0, acc0, acc1, tid);
  }
  plane_inv_store<R, N, T, EPT, LG>(a, col, plane0, acc0, acc1, threadIdx.x);
}

// Persistent inverse kernel: workgroups take planes blockIdx.x, + gridDim.x,
// ...  A channel's spectrum plane passes through registers on its way into
// LDS (all of a thread's loads in flight at once); the next plane's first
// channel is loaded while the last channel of the current plane is
// transformed, so those HBM reads overlap the LDS work.
template <typename R, int N, int T, bool R16>
__global__ void __launch_bounds__(T) k_plane_inv_persist(XArgs<R> a, int nplanes) {
  constexpr int NK = N / 2 + 1;
  constexpr int EPT = (N * N / 2 + T - 1) / T;
  constexpr int ES = (N * NK + T - 1) / T;  // spectrum values per thread
  // post-weight and store loads in groups of WG / SG per thread (FP64: no
  // registers to spare for weight groups next to the accumulators and the
  // next plane's spectrum)
  constexpr int WG = sizeof(R) == 8 ? 1 : 4;
  constexpr int SG = 4;
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * NK];
  for (int i = threadIdx.x; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const int nz = a.g.nz;
  // next plane's first-channel spectrum, real and imaginary parts apart (a
  // plain register array: no aggregate to keep in scratch)
  R sre[ES], sim[ES];
  int p = blockIdx.x;
  if (p < nplanes) {
    const cx<R> *sp = a.spec + (long long)a.chan_base * a.chan_stride +
                      (long long)(p / nz) * a.g.Kk + (long long)(p % nz) * N * NK;
#pragma unroll
    for (int b = 0; b < ES; ++b) {
      const int e = threadIdx.x + b * T;
      if (N * NK % T == 0 || b < ES - 1 || e < N * NK) {
        const cx<R> v = sp[e];
        sre[b] = v.re;
        sim[b] = v.im;
      }
    }
  }
  for (; p < nplanes; p += gridDim.x) {
    const int z = p % nz, col = p / nz;
    const long long plane0 = (long long)z * N * N;
    R acc0[EPT], acc1[EPT];
#pragma unroll
    for (int b = 0; b < EPT; ++b) {
      acc0[b] = R(0);
      acc1[b] = R(0);
    }
    for (int c = 0; c < a.nch; ++c) {
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      if (c > 0) {  // this channel's spectrum plane into the (free) registers
        const cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                          (long long)col * a.g.Kk + (long long)z * N * NK;
#pragma unroll
        for (int b = 0; b < ES; ++b) {
          const int e = tid + b * T;
          if (N * NK % T == 0 || b < ES - 1 || e < N * NK) {
            const cx<R> v = sp[e];
            sre[b] = v.re;
            sim[b] = v.im;
          }
        }
      }
#pragma unroll
      for (int b = 0; b < ES; ++b) {
        const int e = tid + b * T;
        if (N * NK % T == 0 || b < ES - 1 || e < N * NK) {
          const int y = e / NK, k = e % NK;
          buf[dif_pos<N, R16>(y) * NK + k] = cmk<R>(sre[b], sim[b]);
        }
      }
      __syncthreads();
      // the registers are free after the last channel's LDS write
      const int pn = p + (int)gridDim.x;
      if (c == a.nch - 1 && pn < nplanes) {
        const cx<R> *sp = a.spec + (long long)a.chan_base * a.chan_stride +
                          (long long)(pn / nz) * a.g.Kk + (long long)(pn % nz) * N * NK;
#pragma unroll
        for (int b = 0; b < ES; ++b) {
          const int e = threadIdx.x + b * T;
          if (N * NK % T == 0 || b < ES - 1 || e < N * NK) {
            const cx<R> v = sp[e];
            sre[b] = v.re;
            sim[b] = v.im;
          }
        }
      }
      plane_inv_transform<R, N, T, EPT, WG, R16>(a, buf, twx, twy, c,
                                                                          plane0, acc0, acc1, tid);
    }
    plane_inv_store<R, N, T, EPT, SG>(a, col, plane0, acc0, acc1, threadIdx.x);
  }
}

// ---------------------------------------------------------------------------
// Fused z pass with compile-time Nz: P consecutive (y, kx) pencils per
// workgroup held element-major in LDS (element z of pencil p at z*P + p).  The
// first forward stage reads its butterfly inputs straight from the spectrum in
// HBM (one contiguous segment per z), the channel's k-space factor is
// applied as the last forward stage stores, and the last inverse stage writes
// straight back to HBM.
template <typename R>
struct ZMul {
  const YZArgs<R> *a;
  int code, valid;
  const R *kxy2;  // kx^2 + ky^2 of pencil p (LDS)
  const R *kz2;   // kz^2 (LDS)
  __device__ __forceinline__ cx<R> operator()(cx<R> v, int z, int p) const {
    if (p >= valid) return v;
    const R k2 = kxy2[p] + kz2[z];
    R f;
    switch (code) {
      case MUL_KIN:
        f = k2 * (R(0.5) * a->invS);
        break;
      case MUL_PRE: {
        R d = a->mu + R(0.5) * k2;
        if (fabs((double)d) < 1e-8) d = R(1e-8);
        f = a->invS / d;
        break;
      }
      case MUL_ITPKIN:
        f = R(exp(-(double)a->dt * 0.5 * (double)k2)) * a->invS;
        break;
      case MUL_DIP:
      case MUL_TABK:
      case MUL_TABM:
        return v;  // tabulated factors: applied after the forward transform
      default:
        f = a->invS;
    }
    return cscale(v, f);
  }
};

// channels whose k-space factor is a table in HBM (the dipolar kernel, the
// learned tables)
__host__ __device__ __forceinline__ bool mul_is_table(int code) {
  return code == MUL_DIP || code == MUL_TABK || code == MUL_TABM;
}

// Workgroups of a batch: the ny * n0k (y, kx) pencils of one spectrum, taken
// in memory order (flat index f = y * n0k + kx), P consecutive pencils per
// workgroup.  A z row of a workgroup is then P consecutive elements starting
// at a multiple of P: one whole aligned 128-byte line when ny * n0k is a
// multiple of P, so no line is split between two workgroups (which may run
// on different XCDs, each fetching the line into its own L2).
template <int P>
__host__ __device__ __forceinline__ int z_tiles_per_batch(int n0k, int ny) {
  return (n0k * ny + P - 1) / P;
}

template <typename R, int NZ, int P, int T, bool R16>
__global__ void __launch_bounds__(T) k_zpass_fixed(YZArgs<R> a) {
  __shared__ cx<R> tw[NZ];
  __shared__ cx<R> bufA[NZ * P];
  __shared__ cx<R> bufB[NZ * P];
  __shared__ R kxy2[P];
  __shared__ R kz2[NZ];
  constexpr int EU = NZ * P / T;  // tabulated factors per thread
  static_assert(EU * T == NZ * P, "whole tile per thread set");
  const int ny = a.g.ny, n0k = a.g.n0k;
  const int npen = ny * n0k;
  const int per_b = z_tiles_per_batch<P>(n0k, ny);
  const int b = blockIdx.x / per_b;
  const int f0 = (blockIdx.x % per_b) * P;
  const int valid = npen - f0 < P ? npen - f0 : P;
  const int code = a.fwd_only ? MUL_ONE : a.mul[b / a.ncols];
  // a table factor of element e = z * P + p of the tile, e = tid + q * T,
  // requested now and applied once the forward transform is done
  R uf[EU];
  if (mul_is_table(code)) {
    const R *src = code == MUL_DIP ? a.U : a.tab[code - MUL_TABK];
    const R sc = code == MUL_DIP ? a.invS : R(1);
#pragma unroll
    for (int q = 0; q < EU; ++q) {
      const int e = threadIdx.x + q * T;
      const int z = e / P, pp = e % P;
      const R u = src[(long long)z * npen + f0 + (pp < valid ? pp : 0)];
      uf[q] = pp < valid ? u * sc : R(0);
    }
  }
  for (int i = threadIdx.x; i < NZ; i += T) {
    tw[i] = a.tw[i];
    kz2[i] = a.kz2[i];
  }
  if (threadIdx.x < P) {
    const int f = f0 + (int)threadIdx.x < npen ? f0 + (int)threadIdx.x : npen - 1;
    const int y = f / n0k;
    kxy2[threadIdx.x] = a.kx2[f - y * n0k] + a.ky2[y];
  }
  __syncthreads();
  cx<R> *base = a.spec + (long long)b * a.g.Kk + f0;
  GlobIO<R> g{base, (long long)npen, valid, 1};
  ZMul<R> mul{&a, code, valid, kxy2, kz2};
  constexpr int nst = make_fixed_plan(NZ, R16).nst;
  cx<R> *F = (nst % 2 == 1) ? bufA : bufB;  // last LDS buffer of the forward
  cx<R> *O = (F == bufA) ? bufB : bufA;
  if (a.fwd_only) {
    fixed_stages<R, false, NZ, P, T, 0, R16>(g, g, bufA, bufB, tw, NoMul{});
    return;
  }
  // forward: global -> A -> B ... ; the last stage stores into the buffer
  // it does not read (F)
  if constexpr (nst == 1)
    fixed_stages<R, false, NZ, P, T, 0, R16>(g, LdsIO<R>{bufA}, bufA, bufB, tw, mul);
  else
    fixed_stages<R, false, NZ, P, T, 0, R16>(g, LdsIO<R>{F}, bufA, bufB, tw, mul);
  if (mul_is_table(code)) {
#pragma unroll
    for (int q = 0; q < EU; ++q) {
      const int e = threadIdx.x + q * T;
      F[e] = cscale(F[e], uf[q]);
    }
    __syncthreads();
  }
  // inverse: F -> O -> F ... -> global
  if constexpr (nst == 1)
    fixed_stages<R, true, NZ, P, T, 0, R16>(LdsIO<R>{bufA}, g, bufB, bufA, tw, NoMul{});
  else
    fixed_stages<R, true, NZ, P, T, 0, R16>(LdsIO<R>{F}, g, O, F, tw, NoMul{});
}

template <typename R>
bool zpass_fixed_supported(const bdgx_ctx *c) {
  const int nz = c->geo.nz;
  return nz == 32 || nz == 48 || nz == 64 || nz == 80 || nz == 96 || nz == 128 || nz == 192;
}

// Radix plans of the fixed kernels: a size whose plan ends in 8 x ... x 2
// (N = 128) also has the plan with the 2 merged into a radix-16 first stage
// (one LDS round trip fewer, 16 values per butterfly in registers).  Both are
// compiled where they differ and the autotuner times both.
constexpr bool has_alt_plan(int n) {
  return make_fixed_plan(n, true).nst != make_fixed_plan(n, false).nst;
}

// Launch shapes of the fixed z pass: variant bit 0 halves T (T = 256 at
// NZ >= 128, 128 below: two butterflies per thread in the radix-8 stages),
// bit 1 takes the other radix plan (where NZ has one).  P = 8 (FP64) or 16
// (FP32) adjacent kx pencils: one 128-byte segment per z row.
template <typename R, int NZ, int T, bool R16>
static void launch_z_nt(const YZArgs<R> &z, dim3 grid, hipStream_t s) {
  constexpr int P = sizeof(R) == 8 ? 8 : 16;
  hipLaunchKernelGGL((k_zpass_fixed<R, NZ, P, T, R16>), grid, dim3(T), 0, s, z);
}

template <typename R, int NZ, int TB>
static void launch_z_v(const YZArgs<R> &z, dim3 grid, int v, hipStream_t s) {
  constexpr bool D = kFixedRadix16;
  if constexpr (has_alt_plan(NZ)) {
    if (v & 2) {
      if (v & 1) launch_z_nt<R, NZ, TB / 2, !D>(z, grid, s);
      else launch_z_nt<R, NZ, TB, !D>(z, grid, s);
      return;
    }
  }
  if (v & 1) launch_z_nt<R, NZ, TB / 2, D>(z, grid, s);
  else launch_z_nt<R, NZ, TB, D>(z, grid, s);
}

template <typename R>
int launch_zpass_fixed(bdgx_ctx *c, const YZArgs<R> &z) {
  constexpr int P = sizeof(R) == 8 ? 8 : 16;
  const Geo &g = c->geo;
  if (!z.tw || !z.spec || !z.kx2 || !z.ky2 || !z.kz2 || z.nb < 1) return BDGX_EINVAL;
  const dim3 grid((unsigned)(z_tiles_per_batch<P>(g.n0k, g.ny) * z.nb));
  const int v = c->tune_var[sizeof(R) == 8 ? 0 : 1][BDGX_TUNE_Z];
  switch (g.nz) {
    case 32: launch_z_v<R, 32, 128>(z, grid, v, c->stream); break;
    case 48: launch_z_v<R, 48, 128>(z, grid, v, c->stream); break;
    case 64: launch_z_v<R, 64, 128>(z, grid, v, c->stream); break;
    case 80: launch_z_v<R, 80, 128>(z, grid, v, c->stream); break;
    case 96: launch_z_v<R, 96, 128>(z, grid, v, c->stream); break;
    case 128: launch_z_v<R, 128, 256>(z, grid, v, c->stream); break;
    case 192: launch_z_v<R, 192, 256>(z, grid, v, c->stream); break;
    default: return BDGX_EINVAL;
  }
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

template bool zpass_fixed_supported<float>(const bdgx_ctx *);
template bool zpass_fixed_supported<double>(const bdgx_ctx *);
template int launch_zpass_fixed<float>(bdgx_ctx *, const YZArgs<float> &);
template int launch_zpass_fixed<double>(bdgx_ctx *, const YZArgs<double> &);

// ---------------------------------------------------------------------------
// Whole-plane x/y transforms for complex vectors (c2c, Nx = Ny = N).  One
// workgroup owns one z-plane of one column; the N x N complex plane sits in
// LDS in the swizzled row-major x layout (147 KB at N = 96 FP64).
//
//   forward  read the complex plane into registers once; per channel:
//            pre-weight (real weight, conj(psi0), |x|^2), x transform of the
//            N rows, y transform of the N columns (both DIF, in place), write
//            the spectrum plane spec[y][kx] in natural order;
//   inverse  per channel: read the spectrum plane into digit-reversed
//            positions, y inverse of the columns, x inverse of the rows (DIT,
//            in place), accumulate the post-weighted values (posts, real
//            weight, psi0) in registers; then add lt * x and write the output
//            plane (complex, or its real part).
//
// The formulas per channel are those of the pencil kernels k_xfwd_cplx /
// k_xinv_cplx; only the order of the FFT work differs.
template <typename R, int N>
constexpr bool cplane_fits() {
  return (size_t)N * N * sizeof(cx<R>) + 2 * N * sizeof(cx<R>) <= 160 * 1024;
}

template <typename R, int N, int T>
__global__ void __launch_bounds__(T) k_cplane_fwd(XArgs<R> a) {
  constexpr int SW = plane_swz<R, N>();
  constexpr bool R16 = kFixedRadix16;
  constexpr int EP = N * N / T;  // plane values per thread
  static_assert(EP * T == N * N, "whole plane per thread set");
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * N];
  const int tid = threadIdx.x;
  const int nz = a.g.nz;
  const int z = blockIdx.x % nz;
  const int col = blockIdx.x / nz;
  for (int i = tid; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const long long plane0 = (long long)z * N * N;
  const cx<R> *xp = reinterpret_cast<const cx<R> *>(a.X) + (long long)col * a.ld +
                    (long long)a.in_half * a.g.S + plane0;
  cx<R> xv[EP];
#pragma unroll
  for (int b = 0; b < EP; ++b) xv[b] = xp[tid + b * T];
  for (int c = 0; c < a.nch; ++c) {
    const int pre = a.pre[c];
    if (pre == PRE_REAL) {
      const R *pw = a.prew[c] + plane0;
      R q[EP];
#pragma unroll
      for (int b = 0; b < EP; ++b) q[b] = pw[tid + b * T];
#pragma unroll
      for (int b = 0; b < EP; ++b) buf[swz<SW>(tid + b * T)] = cscale(xv[b], q[b]);
    } else if (pre == PRE_CONJ_PSI) {
      const cx<R> *pw = reinterpret_cast<const cx<R> *>(a.prew[c]) + plane0;
      constexpr int LG = div_group(EP, 4);
#pragma unroll
      for (int b0 = 0; b0 < EP; b0 += LG) {
        cx<R> q[LG];
#pragma unroll
        for (int i = 0; i < LG; ++i) q[i] = pw[tid + (b0 + i) * T];
#pragma unroll
        for (int i = 0; i < LG; ++i)
          buf[swz<SW>(tid + (b0 + i) * T)] = cmul(cconj(q[i]), xv[b0 + i]);
      }
    } else if (pre == PRE_ABS2) {
#pragma unroll
      for (int b = 0; b < EP; ++b)
        buf[swz<SW>(tid + b * T)] = cmk<R>(xv[b].re * xv[b].re + xv[b].im * xv[b].im, R(0));
    } else {
#pragma unroll
      for (int b = 0; b < EP; ++b) buf[swz<SW>(tid + b * T)] = xv[b];
    }
    __syncthreads();
    // x transform of the N rows, then y transform of the N columns (pencil
    // p = x position, adjacent pencils adjacent in LDS)
    dif_fft<R, false, N, N, N, 1, T, false, SW, 0, R16>(buf, twx, tid);
    dif_fft<R, false, N, N, 1, N, T, true, SW, 0, R16>(buf, twy, tid);
    cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride + (long long)col * a.g.Kk +
                plane0;
    constexpr int SG = div_group(EP, 4);
#pragma unroll
    for (int b0 = 0; b0 < EP; b0 += SG) {
      cx<R> v[SG];
#pragma unroll
      for (int i = 0; i < SG; ++i) {
        const int e = tid + (b0 + i) * T;
        const int y = e / N, k = e % N;
        v[i] = buf[swz<SW>(dif_pos<N, R16>(y) * N + dif_pos<N, R16>(k))];
      }
#pragma unroll
      for (int i = 0; i < SG; ++i) sp[tid + (b0 + i) * T] = v[i];
    }
    __syncthreads();
  }
}

template <typename R, int N, int T>
__global__ void __launch_bounds__(T) k_cplane_inv(XArgs<R> a) {
  constexpr int SW = plane_swz<R, N>();
  constexpr bool R16 = kFixedRadix16;
  constexpr int EP = N * N / T;
  static_assert(EP * T == N * N, "whole plane per thread set");
  constexpr int LG = div_group(EP, 4);
  __shared__ cx<R> twx[N];
  __shared__ cx<R> twy[N];
  __shared__ cx<R> buf[N * N];
  const int nz = a.g.nz;
  const int z = blockIdx.x % nz;
  const int col = blockIdx.x / nz;
  for (int i = threadIdx.x; i < N; i += T) {
    twx[i] = a.tw[i];
    twy[i] = a.tw_y[i];
  }
  const long long plane0 = (long long)z * N * N;
  cx<R> acc[EP];
#pragma unroll
  for (int b = 0; b < EP; ++b) acc[b] = cmk<R>(0, 0);
  for (int c = 0; c < a.nch; ++c) {
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const cx<R> *sp = a.spec + (long long)(a.chan_base + c) * a.chan_stride +
                      (long long)col * a.g.Kk + plane0;
#pragma unroll
    for (int b0 = 0; b0 < EP; b0 += LG) {
      cx<R> v[LG];
#pragma unroll
      for (int i = 0; i < LG; ++i) v[i] = sp[tid + (b0 + i) * T];
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        const int e = tid + (b0 + i) * T;
        const int y = e / N, k = e % N;
        buf[swz<SW>(dif_pos<N, R16>(y) * N + dif_pos<N, R16>(k))] = v[i];
      }
    }
    __syncthreads();
    dit_fft<R, true, N, N, 1, N, T, true, SW, last_stage<N, R16>(), R16>(buf, twy, tid);
    dit_fft<R, true, N, N, N, 1, T, false, SW, last_stage<N, R16>(), R16>(buf, twx, tid);
    const int post = a.post[c];
    const R ps = a.posts[c];
#pragma unroll
    for (int b0 = 0; b0 < EP; b0 += LG) {
      cx<R> w[LG];
      if (post == POST_REAL) {
        const R *pw = a.postw[c] + plane0;
#pragma unroll
        for (int i = 0; i < LG; ++i) w[i] = cmk<R>(ps * pw[tid + (b0 + i) * T], R(0));
      } else if (post == POST_PSI) {
        const cx<R> *pw = reinterpret_cast<const cx<R> *>(a.postw[c]) + plane0;
#pragma unroll
        for (int i = 0; i < LG; ++i) w[i] = cscale(pw[tid + (b0 + i) * T], ps);
      } else {
#pragma unroll
        for (int i = 0; i < LG; ++i) w[i] = cmk<R>(ps, R(0));
      }
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        const cx<R> zv = buf[swz<SW>(tid + (b0 + i) * T)];
        acc[b0 + i] = acc[b0 + i] + (post == POST_PSI ? cmul(zv, w[i]) : cscale(zv, w[i].re));
      }
    }
    __syncthreads();
  }
  const int tid = threadIdx.x;
  const long long obase = (long long)col * a.ld + (long long)a.out_half * a.g.S + plane0;
  if (a.lt) {
    const cx<R> *xp = reinterpret_cast<const cx<R> *>(a.X) + (long long)col * a.ld +
                      (long long)a.in_half * a.g.S + plane0;
    const R *lt = a.lt + plane0;
#pragma unroll
    for (int b0 = 0; b0 < EP; b0 += LG) {
      cx<R> u[LG];
      R l[LG];
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        u[i] = xp[tid + (b0 + i) * T];
        l[i] = lt[tid + (b0 + i) * T];
      }
#pragma unroll
      for (int i = 0; i < LG; ++i) acc[b0 + i] = acc[b0 + i] + cscale(u[i], l[i]);
    }
  }
  if (a.out_real) {
#pragma unroll
    for (int b = 0; b < EP; ++b) a.Y[obase + tid + b * T] = acc[b].re;
  } else {
    cx<R> *yp = reinterpret_cast<cx<R> *>(a.Y) + obase;
#pragma unroll
    for (int b = 0; b < EP; ++b) yp[tid + b * T] = acc[b];
  }
}

template <typename R, int N, int T>
static void launch_cplane(const XArgs<R> &a, bool inverse, hipStream_t s) {
  const unsigned grid = (unsigned)(a.g.nz * a.ncols);
  if (inverse)
    hipLaunchKernelGGL((k_cplane_inv<R, N, T>), dim3(grid), dim3(T), 0, s, a);
  else
    hipLaunchKernelGGL((k_cplane_fwd<R, N, T>), dim3(grid), dim3(T), 0, s, a);
}

// complex variants: even = 4N threads, odd = 8N threads (no persistent
// forms; variants 2 / 3 repeat 0 / 1)
template <typename R, int N>
static void launch_cplane_v(const XArgs<R> &a, bool inverse, int v, hipStream_t s) {
  if constexpr (cplane_fits<R, N>()) {
    if (v & 1)
      launch_cplane<R, N, 8 * N>(a, inverse, s);
    else
      launch_cplane<R, N, 4 * N>(a, inverse, s);
  }
}

template <typename R, int N, int T, bool R16>
static void launch_plane(const XArgs<R> &a, bool inverse, hipStream_t s) {
  const unsigned grid = (unsigned)(a.g.nz * a.ncols);
  if (inverse)
    hipLaunchKernelGGL((k_plane_inv_real<R, N, T, R16>), dim3(grid), dim3(T), 0, s, a);
  else
    hipLaunchKernelGGL((k_plane_fwd_real<R, N, T, R16>), dim3(grid), dim3(T), 0, s, a);
}

template <typename R>
bool plane_supported(const bdgx_ctx *c) {
  const Geo &g = c->geo;
  if (c->dim != 3 || g.nx != g.ny) return false;
  if (c->cplx)  // the complex plane in LDS: N = 128 only in FP32
    return g.nx == 32 || g.nx == 48 || g.nx == 64 || g.nx == 80 || g.nx == 96 ||
           (g.nx == 128 && sizeof(R) == 4);
  return g.nx == 32 || g.nx == 48 || g.nx == 64 || g.nx == 80 || g.nx == 96 || g.nx == 128;
}

// Launch shapes of the whole-plane kernels: variant bits 0-1 = the one-plane
// kernels with 4N / 8N threads (0 / 1) or the persistent kernels with as
// many workgroups as fit on the device at once, each then walking several
// planes, 4N / 8N threads (2 / 3); bit 2 = the other radix plan (where N has
// one; the default plan is radix 16 first except in the FP64 persistent
// inverse, which spills with it).
template <typename R, int N, int T, bool INV, bool R16>
static void launch_plane_persist(const bdgx_ctx *c, const XArgs<R> &a, hipStream_t s) {
  static const int per_cu = [] {
    int nb = 0;
    const void *fn = INV ? reinterpret_cast<const void *>(&k_plane_inv_persist<R, N, T, R16>)
                         : reinterpret_cast<const void *>(&k_plane_fwd_persist<R, N, T, R16>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, T, 0) != hipSuccess || nb < 1)
      nb = 1;
    return nb;
  }();
  const int nplanes = a.g.nz * a.ncols;
  int grid = c->ncu * per_cu;
  if (grid > nplanes) grid = nplanes;
  if (grid < 1) grid = 1;
  if (INV)
    hipLaunchKernelGGL((k_plane_inv_persist<R, N, T, R16>), dim3(grid), dim3(T), 0, s, a,
                       nplanes);
  else
    hipLaunchKernelGGL((k_plane_fwd_persist<R, N, T, R16>), dim3(grid), dim3(T), 0, s, a,
                       nplanes);
}

// default radix plan of each plane kernel, and the launch of shape v & 3
// with the default plan (ALT = false) or the other one
template <typename R, int N, int T, bool INV, bool PERSIST>
constexpr bool plane_default_r16() {
  return INV ? plane_inv_r16<R, N, T, PERSIST>() : kFixedRadix16;
}
template <typename R, int N, bool INV, bool ALT>
static void launch_plane_shape(const bdgx_ctx *c, const XArgs<R> &a, int shape, hipStream_t s) {
  switch (shape) {
    case 2:
      launch_plane_persist<R, N, 4 * N, INV, plane_default_r16<R, N, 4 * N, INV, true>() != ALT>(
          c, a, s);
      break;
    case 3:
      launch_plane_persist<R, N, 8 * N, INV, plane_default_r16<R, N, 8 * N, INV, true>() != ALT>(
          c, a, s);
      break;
    case 1:
      launch_plane<R, N, 8 * N, plane_default_r16<R, N, 8 * N, INV, false>() != ALT>(a, INV, s);
      break;
    default:
      launch_plane<R, N, 4 * N, plane_default_r16<R, N, 4 * N, INV, false>() != ALT>(a, INV, s);
  }
}

template <typename R, int N>
static void launch_plane_v(const bdgx_ctx *c, const XArgs<R> &a, bool inverse, int v,
                           hipStream_t s) {
  const int shape = v & 3;
  if constexpr (has_alt_plan(N)) {
    if (v & 4) {
      if (inverse) launch_plane_shape<R, N, true, true>(c, a, shape, s);
      else launch_plane_shape<R, N, false, true>(c, a, shape, s);
      return;
    }
  }
  if (inverse) launch_plane_shape<R, N, true, false>(c, a, shape, s);
  else launch_plane_shape<R, N, false, false>(c, a, shape, s);
}

template <typename R>
int launch_plane_x(bdgx_ctx *c, const XArgs<R> &a, bool inverse) {
  if (!a.tw || !a.tw_y || !a.spec || !a.X || (inverse && !a.Y) || a.ncols < 1)
    return BDGX_EINVAL;
  const int v = c->tune_var[sizeof(R) == 8 ? 0 : 1][inverse ? BDGX_TUNE_PLANE_INV
                                                              : BDGX_TUNE_PLANE_FWD];
  if (c->cplx) {
    if (!plane_supported<R>(c)) return BDGX_EINVAL;
    switch (c->geo.nx) {
      case 32: launch_cplane_v<R, 32>(a, inverse, v, c->stream); break;
      case 48: launch_cplane_v<R, 48>(a, inverse, v, c->stream); break;
      case 64: launch_cplane_v<R, 64>(a, inverse, v, c->stream); break;
      case 80: launch_cplane_v<R, 80>(a, inverse, v, c->stream); break;
      case 96: launch_cplane_v<R, 96>(a, inverse, v, c->stream); break;
      case 128: launch_cplane_v<R, 128>(a, inverse, v, c->stream); break;
      default: return BDGX_EINVAL;
    }
    BDGX_HIP(hipGetLastError());
    return BDGX_OK;
  }
  switch (c->geo.nx) {
    case 32: launch_plane_v<R, 32>(c, a, inverse, v, c->stream); break;
    case 48: launch_plane_v<R, 48>(c, a, inverse, v, c->stream); break;
    case 64: launch_plane_v<R, 64>(c, a, inverse, v, c->stream); break;
    case 80: launch_plane_v<R, 80>(c, a, inverse, v, c->stream); break;
    case 96: launch_plane_v<R, 96>(c, a, inverse, v, c->stream); break;
    case 128: launch_plane_v<R, 128>(c, a, inverse, v, c->stream); break;
    default: return BDGX_EINVAL;
  }
  BDGX_HIP(hipGetLastError());
  return BDGX_OK;
}

// Measured defaults (MI355X, 32 columns): at N = 128 FP64 (one 133 KB
// plane per CU) the persistent kernels with 4N threads; FP32 (two 67 KB
// planes per CU) the persistent forward with 8N threads and the inverse with
// one plane per 4N-thread workgroup; 8N threads, one plane per workgroup, at
// N = 32, 64.  The z pass default is variant 0.
void tune_defaults(bdgx_ctx *c) {
  for (int p = 0; p < 2; ++p) {
    const bool n128 = c->geo.nx == 128;
    c->tune_var[p][BDGX_TUNE_PLANE_FWD] = n128 ? (p == 0 ? 2 : 3) : 1;
    c->tune_var[p][BDGX_TUNE_PLANE_INV] = n128 ? (p == 0 ? 2 : 0) : 1;
    c->tune_var[p][BDGX_TUNE_Z] = 0;
    c->tune_var[p][BDGX_TUNE_COMBINE] = 0;
    c->tune_var[p][BDGX_TUNE_GRAM] = 0;
    c->tuned[p] = 0;
  }
}

// number of launch variants the autotuner times: shapes x radix plans
int plane_variants(const bdgx_ctx *c) {
  if (c->cplx) return 2;
  return has_alt_plan(c->geo.nx) ? 8 : 4;
}
int zpass_variants(const bdgx_ctx *c) { return has_alt_plan(c->geo.nz) ? 4 : 2; }

template bool plane_supported<float>(const bdgx_ctx *);
template bool plane_supported<double>(const bdgx_ctx *);
template int launch_plane_x<float>(bdgx_ctx *, const XArgs<float> &, bool);
template int launch_plane_x<double>(bdgx_ctx *, const XArgs<double> &, bool);

}  // namespace bdgx
