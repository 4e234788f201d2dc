// Launch arguments of the FFT pass kernels, shared by the pencil passes
// (ops.hip) and the whole-plane x/y kernels (plane.hip).
#pragma once

#include "bdgx_internal.hpp"

namespace bdgx {

template <typename R>
struct XArgs {
  Geo g;
  int ncols;
  long long ld;
  const R *X;
  R *Y;
  int in_half, out_half;
  int nch, chan_base;
  int pre[kMaxChan];
  const R *prew[kMaxChan];
  int post[kMaxChan];
  const R *postw[kMaxChan];
  R posts[kMaxChan];
  const R *lt;
  int out_real;
  cx<R> *spec;
  long long chan_stride;
  const cx<R> *tw;
  const cx<R> *tw_y;  // y-axis twiddles (whole-plane kernels)
  AxisPlan pl;
  int rb;  // rows per workgroup
};

template <typename R>
struct YZArgs {
  Geo g;
  int ncols, nb, P;
  cx<R> *spec;
  const cx<R> *tw;
  AxisPlan pl;
  // z pass only
  int mul[kMaxChan];
  int fwd_only;  // z forward transform only: the spectrum stays in HBM
  const R *kx2, *ky2, *kz2, *U;
  const R *tab[2];  // learned k-space tables (MUL_TABK, MUL_TABM)
  R mu, invS, dt;
};

// Whole-plane x/y transforms (real vectors, Nx = Ny in {32, 64, 128}).
template <typename R>
bool plane_supported(const bdgx_ctx *c);
template <typename R>
int launch_plane_x(bdgx_ctx *c, const XArgs<R> &a, bool inverse);
// Fused z pass with compile-time Nz in {32, 64, 128}.
template <typename R>
bool zpass_fixed_supported(const bdgx_ctx *c);
template <typename R>
int launch_zpass_fixed(bdgx_ctx *c, const YZArgs<R> &z);
// measured default launch variants of the plane and fixed z kernels
void tune_defaults(bdgx_ctx *c);
// launch variants (shapes x radix plans) of the plane / fixed z kernels
int plane_variants(const bdgx_ctx *c);
int zpass_variants(const bdgx_ctx *c);

}  // namespace bdgx
