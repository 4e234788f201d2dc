// Internal definitions shared by the HIP translation units of libbdg_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bdgx.h"

#define BDGX_HIP(call)                                   \
  do {                                                   \
    hipError_t e__ = (call);                             \
    if (e__ != hipSuccess) {                             \
      bdgx_last_hip_error = e__;                         \
      return BDGX_EHIP;                                  \
    }                                                    \
  } while (0)

extern thread_local hipError_t bdgx_last_hip_error;

namespace bdgx {

// complex value of precision R
template <typename R>
struct cx {
  R re, im;
};

// One axis FFT plan: mixed-radix Stockham stages.
struct AxisPlan {
  int n;
  int nst;
  int rad[12];
  int ns[12];
};

constexpr int kMaxChan = 3;

// weights applied when a channel pencil is formed (load side)
enum PreKind { PRE_NONE = 0, PRE_REAL = 1, PRE_CONJ_PSI = 2, PRE_ABS2 = 3 };
// coefficient applied when a channel is folded into the output (store side)
enum PostKind { POST_ONE = 0, POST_REAL = 1, POST_PSI = 2, POST_RE = 3 };
// k-space multipliers applied between the forward and inverse z transform
enum MulKind { MUL_ONE = 0, MUL_KIN = 1, MUL_DIP = 2, MUL_PRE = 3, MUL_ITPKIN = 4,
               MUL_TABK = 5, MUL_TABM = 6 };  // TAB*: learned k-space tables

struct Geo {
  int nx, ny, nz, n0k;
  long long S, Kk;
};

// Device-side copies of the physics fields in one precision.
template <typename R>
struct Fields {
  R *ltK = nullptr, *ltM = nullptr, *pK = nullptr, *pM = nullptr;
  R *wf = nullptr;    // size (real) or 2*size (complex, interleaved)
  R *U = nullptr;     // k_size dipolar kernel
  R *kx2[3] = {nullptr, nullptr, nullptr};
  cx<R> *tw[3] = {nullptr, nullptr, nullptr};  // per-axis twiddles exp(-2 pi i m/n)
  // learned preconditioner (FP32 only): pointwise weights a, b and k-space
  // tables g (1/S folded in) of its two spectral filters, [0] K and [1] M
  R *lpa[2] = {nullptr, nullptr}, *lpb[2] = {nullptr, nullptr}, *lpg[2] = {nullptr, nullptr};
};

}  // namespace bdgx

struct bdgx_ctx {
  int dim, cplx, device;
  uint64_t N[3];
  double L[3];
  bdgx::Geo geo;
  bdgx::AxisPlan plan[3];
  hipStream_t stream;
  hipStream_t aux;  // second copy stream (created on first use)
  hipEvent_t ev0, ev1;
  bdgx::Fields<double> f64;
  bdgx::Fields<float> f32;
  double mu, g_ddi;
  double itp_dt;  // time step of the imaginary-time kinetic multiplier
  int has_wf, has_U, has_lt, has_p;
  // spectrum workspace (grown on demand)
  void *spec;
  size_t spec_bytes;
  // small device scratch for gram partials etc.
  void *scratch;
  size_t scratch_bytes;
  // device copy of column indices for gathers
  int *idx_buf;
  uint64_t idx_cap;
  // learned preconditioner: stencil parameters (w, d, v), features rho,
  // workspace
  float *lp_conv;
  float *lp_rho;
  void *lp_tmp;
  size_t lp_tmp_bytes;
  double lp_s0;  // stencil-layer scale mean(p_K^2) / safe(mu)
  int has_lp;
  // launch-shape variants of the fixed-size FFT kernels, per precision
  // (index 0 FP64, 1 FP32): [BDGX_TUNE_PLANE_FWD, _PLANE_INV, _Z, _COMBINE,
  // _GRAM]; set to measured defaults at creation, re-chosen by bdgx_autotune
  int tune_var[2][5];
  int tuned[2];
  // bumped whenever spec / scratch / lp_tmp / idx_buf move: graphs captured
  // before a move hold stale pointers and must not be launched
  unsigned ws_gen;
  // spectrum bytes per column chunk of an operator apply (0: 6 GiB)
  size_t chunk_bytes;
  int ncu;  // compute units of the device (persistent launches)
};

namespace bdgx {
int ensure_spec(bdgx_ctx *c, size_t bytes);
int ensure_scratch(bdgx_ctx *c, size_t bytes);
int make_plan(int n, AxisPlan *p);
}  // namespace bdgx
