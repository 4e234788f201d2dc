/*
 * bdg.h -- public C API of the BdG eigensolver (MI355X build).
 *
 * Drop-in for the reference's public header (reference include/bdg/bdg.h:1-246):
 * the same opaque handle, enums, callback types and 22 functions with the same
 * signatures, argument meaning, ownership and error behaviour.  A program
 * written against the reference header compiles and links against this one
 * unchanged; the operator applications and the iLOBPCG iteration then run on
 * the GPU (one HIP device per process, selected by HIP_VISIBLE_DEVICES /
 * the current device at bdg_set_system time).
 *
 * Problem: A x = lambda B x with A = diag(K, M), B = [[0, I], [I, 0]] (swap),
 * x = [u; v] of length n = 2*size.  The returned eigenvalues are the BdG
 * frequencies omega (not omega^2), sorted ascending.
 */
#ifndef BDG_H
#define BDG_H

#include "lobpcg/types.h"
#include <stdarg.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque handle (reference include/bdg/bdg.h:23). */
typedef struct bdg_t bdg_t;

/* Start-block strategies (reference include/bdg/bdg.h:29-34); values fixed. */
typedef enum {
  BDG_INIT_PLANEWAVE = 0,  /* lowest-|k| cos/sin columns weighted by |psi0| (default) */
  BDG_INIT_WF_WEIGHTED,    /* |gaussian noise| * |psi0|, u = v */
  BDG_INIT_REUSE,          /* previous modes + noise (set by bdg_reuse_modes) */
  BDG_INIT_CUSTOM          /* user callback fills the block */
} bdg_init_mode_t;

/* Planewave geometry hints (reference include/bdg/bdg.h:36-40). */
typedef enum {
  BDG_GEOM_AUTO = 0,       /* lowest |k|^2 over all axes */
  BDG_GEOM_ELONGATED,      /* k only along the longest box axis */
  BDG_GEOM_RING            /* (kx, ky) pairs, kz = 0 */
} bdg_geom_hint_t;

/*
 * Custom start-block callback (reference include/bdg/bdg.h:53-54).
 * X is column-major n x sizeSub of f64 (real psi0) or c64 (complex psi0);
 * rows [0, size) are the u half, [size, n) the v half.  The block must be
 * B-positive (u = v per column is the simple choice).
 */
typedef void (*bdg_init_fn)(const bdg_t *bdg, void *X,
                            uint64_t n, uint64_t sizeSub, void *param);

/* ---- lifecycle ------------------------------------------------------- */

/* Allocate a problem on a dim-D grid N[0..dim) (x fastest) of box lengths L. */
bdg_t *bdg_alloc(uint64_t dim, const uint64_t *N, const f64 *L, int complex_psi0);

/* Release everything (host and device) and set *bdg = NULL. */
void bdg_free(bdg_t **bdg);

/*
 * Clear the physics (trap, psi0, interactions, dipolar, mu, results) while
 * keeping the grid, the device context and solver parameters.  After the call
 * only the "system" state remains; result pointers become invalid.
 */
void bdg_reset(bdg_t *bdg);

/* ---- setup, in order ------------------------------------------------- */

/* Build the k-space description and the device context. */
void bdg_set_system(bdg_t *bdg);

/* Add V_trap(r) to the K and M local terms (additive; needs the system). */
void bdg_set_trap(bdg_t *bdg,
                  f64 (*V_trap)(uint64_t dim, const f64 *r, const void *param),
                  const void *param);

/* Copy the condensate wavefunction (f64* or c64*, wf_size == grid size). */
void bdg_set_wavefunction(bdg_t *bdg, const void *wf, uint64_t wf_size);

/*
 * Add local interaction terms (once; needs psi0):
 *   K += U_intK(|psi|^2),   M += U_intK(|psi|^2) + 2 U_intM(|psi|^2).
 */
void bdg_set_local_interactions(bdg_t *bdg,
                                f64 (*U_intK)(const void *param, f64 density),
                                f64 (*U_intM)(const void *param, f64 density),
                                const void *param);

/* Subtract mu from both local terms; build the preconditioner weights first. */
void bdg_set_mu(bdg_t *bdg, f64 mu);

/*
 * Dipolar interaction (3D only, before bdg_set_mu): builds the cutoff kernel
 * in k-space and adds g_ddi * Phi_dd to both local terms.
 */
void bdg_set_dipolar(bdg_t *bdg, f64 g_ddi, const f64 *dipole_dir, f64 cutoff_radius);

/* ---- solver ---------------------------------------------------------- */

void bdg_set_solver_params(bdg_t *bdg, uint64_t nev, uint64_t sizeSub,
                           uint64_t maxIter, f64 tol);

/*
 * Start-block strategy.  For BDG_INIT_PLANEWAVE pass fn = NULL and
 * param = (void*)(intptr_t)hint (NULL = BDG_GEOM_AUTO).  Survives bdg_reset.
 */
void bdg_set_init_mode(bdg_t *bdg, bdg_init_mode_t mode,
                       bdg_init_fn fn, void *param);

/*
 * Keep the current modes (+ gaussian noise of stddev noise_frac*||column||)
 * as the next start block and switch the init mode to BDG_INIT_REUSE.
 * Returns 0, or -3 when no results exist.
 */
int bdg_reuse_modes(bdg_t *bdg, f64 noise_frac);

/* Solve.  Returns 0 on success, nonzero on a device or numerical failure. */
int bdg_solve(bdg_t *bdg);

/* ---- results (borrowed pointers, valid until reset/free/next solve) --- */

uint64_t bdg_converged(const bdg_t *bdg);
const f64 *bdg_eigenvalues(const bdg_t *bdg);
const void *bdg_modes_u(const bdg_t *bdg);   /* size*nev column-major */
const void *bdg_modes_v(const bdg_t *bdg);

/* ---- file I/O -------------------------------------------------------- */

/* Text loader, one "re im" pair per line.  0, -1 (file) or -2 (format). */
int bdg_load_wavefunction(bdg_t *bdg, const char *filename);
int bdg_load_wavefunction_fmt(bdg_t *bdg, const char *fmt, ...);

/* Append one tab-separated %.8e row of eigenvalues.  0, -1 or -3. */
int bdg_write_eigenvalues(const bdg_t *bdg, const char *filename);

/* Write mode mode_idx to <basename>_u_<idx>.dat / _v_.  0, -1, -2 or -3. */
int bdg_write_mode_u(const bdg_t *bdg, uint64_t mode_idx, const char *basename);
int bdg_write_mode_v(const bdg_t *bdg, uint64_t mode_idx, const char *basename);

#ifdef __cplusplus
}
#endif

#endif /* BDG_H */
