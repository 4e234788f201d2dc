/*
 * bdg_oracle.h -- FP64 CPU oracle for the BdG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * operator algorithms (pstuermer/BDG_LOBPCG src/kinetic_impl.inc and the other operator
 * .inc files, src/dipolar.c,
 * src/setup.c, src/matmul_ctx.c), used by the test-suite, the smoke check and
 * the CPU-baseline leg of the benchmark as the thing results are checked
 * against.  The product library never links or calls it.
 *
 * All arrays are host memory.  Real-space vectors are x-fastest,
 * i = ix + Nx*(iy + Ny*iz); the r2c spectrum is ix + N0k*(iy + Ny*iz) with
 * N0k = Nx/2+1 (real psi0) or Nx (complex psi0).  FFTs are unnormalised
 * (forward sign -1), as with FFTW.
 */
#ifndef BDG_ORACLE_H
#define BDG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* ---- FFT (own mixed-radix Stockham; any N >= 1) ---------------------- */
/* In-place c2c of a dim-D array (N[0] fastest).  data: interleaved re,im. */
int orc_fft_c2c(int dim, const int64_t *N, double *data, int sign);
/* r2c: in real[size] -> out complex[(N0/2+1)*N1*N2]. */
int orc_fft_r2c(int dim, const int64_t *N, const double *in, double *out);
/* c2r: in complex[(N0/2+1)*N1*N2] (destroyed) -> out real[size]. */
int orc_fft_c2r(int dim, const int64_t *N, double *in, double *out);

/* ---- context (reference src/matmul_ctx.c:21-172) ---------------------- */
orc_ctx *orc_ctx_new(int dim, const int64_t *N, const double *L, int complex_psi0);
void orc_ctx_free(orc_ctx *ctx);
int64_t orc_ctx_size(const orc_ctx *ctx);
int64_t orc_ctx_ksize(const orc_ctx *ctx);
void orc_ctx_get_k2(const orc_ctx *ctx, double *out);           /* k_size */
void orc_ctx_get_kx2(const orc_ctx *ctx, int d, double *out);   /* N[d]   */

/* Copy physics fields into the context (NULL leaves a field untouched). */
void orc_ctx_set_fields(orc_ctx *ctx, const double *ltK, const double *ltM,
                        const double *wf, const double *pK, const double *pM,
                        double mu);
/* Dipolar: build the kernel (reference src/dipolar.c:21-72) and mark active. */
void orc_dipolar_set_kernel(orc_ctx *ctx, double g_ddi, const double *dir,
                            double cutoff_radius);
void orc_dipolar_get_kernel(const orc_ctx *ctx, double *out);   /* k_size */
/* g_ddi * Phi_dd (reference src/dipolar_meanfield_impl.inc:20-55), size. */
void orc_dipolar_meanfield(orc_ctx *ctx, double *gphi_out);

/* ---- operators: one vector, like the reference ------------------------ */
/* x, y are f64[len] for real psi0, interleaved c64[len] for complex psi0.  */
void orc_kinetic(orc_ctx *ctx, const double *x, double *y);
void orc_matmulK(orc_ctx *ctx, const double *x, double *y);
void orc_matmulM(orc_ctx *ctx, const double *x, double *y);
void orc_matmulLrep(orc_ctx *ctx, const double *x, double *y);
void orc_matmulSwap(orc_ctx *ctx, const double *x, double *y);
void orc_precondK(orc_ctx *ctx, const double *x, double *y);
void orc_precondM(orc_ctx *ctx, const double *x, double *y);
void orc_precondLrep(orc_ctx *ctx, const double *x, double *y);
void orc_dipolar_conv(orc_ctx *ctx, const double *x, double *y);

/* Block forms: ncols columns, leading dimension ld (elements). op codes:
 * 0 kinetic, 1 K, 2 M, 3 Lrep, 4 Swap, 5 precondK, 6 precondM,
 * 7 precondLrep, 8 dipolar_conv. */
void orc_apply_block(orc_ctx *ctx, int op, const double *X, double *Y,
                     int64_t ncols, int64_t ld);

/* ---- setup (reference src/setup.c:17-177) ------------------------------ */
/* Harmonic trap 0.5*sum_d (w_d r_d)^2 added into ltK, ltM. */
void orc_setup_trap_harmonic(const orc_ctx *ctx, const double *omega,
                             double *ltK, double *ltM);
/* Contact interactions: K += gK*n, M += gK*n + 2*gM*n with n = |psi|^2. */
void orc_setup_contact(const orc_ctx *ctx, const double *wf, double gK,
                       double gM, double *ltK, double *ltM);
/* p = 1/sqrt(max(1e-8, lt)) before, then lt -= mu. */
void orc_setup_mu(int64_t size, double mu, double *ltK, double *ltM,
                  double *pK, double *pM);

/* ---- psi0 by imaginary-time split-step -------------------------------- */
/* See orc_itp_groundstate in bdg_oracle.c for the exact step definition.  */
int orc_itp_groundstate(const orc_ctx *ctx, const double *omega, double g,
                        double g_dd, const double *dir, double cutoff,
                        double dt, int64_t nsteps, double *psi /*in/out*/,
                        double *mu_out);

/* ---- start blocks ------------------------------------------------------ */
/* xorshift32 uniform (0,1] (reference src/bdg_internal.h:53-65). */
uint32_t orc_xorshift32(uint32_t *state);
double orc_xrand(uint32_t *state);
/* reuse block of a sweep (reference src/bdg.c:112-207) */
void orc_reuse_block(int64_t S, int64_t nev, int64_t m, int cplx, const double *U,
                     const double *V, const double *wf, double noise, uint32_t seed,
                     double *X);
/* Column-major n x ncols block, u = v = xrand-0.5 (seed given). */
void orc_start_block_uniform(const orc_ctx *ctx, uint32_t seed, int64_t ncols,
                             double *X);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
