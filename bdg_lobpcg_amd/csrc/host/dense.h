/* Small dense FP64 linear algebra for the Rayleigh-Ritz step (host side).
 * Matrices are column-major; complex matrices are interleaved (re, im). */
#ifndef BDG_AMD_DENSE_H
#define BDG_AMD_DENSE_H

#include <stdint.h>

/* Real symmetric eigendecomposition: on exit A holds the orthonormal
 * eigenvectors (columns) and w the eigenvalues in ascending order.
 * Returns 0, or -1 if the QL iteration did not converge. */
int dn_syev(int n, double *A, double *w);

/* Complex Hermitian eigendecomposition (interleaved A, n x n).  Same output
 * convention; A is overwritten by the eigenvectors. */
int dn_heev(int n, double *A, double *w);

/* The m lowest eigenpairs of the real symmetric tridiagonal matrix with
 * diagonal d[n] and off-diagonal e[n-1]: w[m] ascending and Y (n x m)
 * orthonormal, by implicit QR without vectors and inverse iteration
 * (Gram-Schmidt inside clusters closer than 1e-3 ||T||).  Returns 0 or -1. */
int dn_trid_low(int n, const double *d, const double *e, int m, double *w, double *Y);

/* The m lowest eigenpairs of Hermitian G (n x n, complex interleaved when
 * cplx; not modified): w[m] ascending, X (n x m) orthonormal eigenvectors.
 * Returns 0 or < 0. */
int dn_eig_low(int n, int m, int cplx, const double *G, double *w, double *X);

/* Rayleigh-Ritz for the Hermitian-definite pencil (G_A, G_M) on a basis
 * with Euclidean Gram G_V (all k x k, complex interleaved when cplx):
 *   w2      out: the m smallest Ritz values (w^2), ascending
 *   Y       out: k x m coefficients (complex when cplx)
 *   zflags  out (optional): 1 for directions in the null space of G_M (zero
 *           modes of a singular M), else 0
 * Returns the number of pairs produced (<= m), or < 0 on failure. */
int dn_rayleigh_ritz(int k, int m, int cplx, const double *GV, const double *GM,
                     const double *GA, double *w2, double *Y, int *zflags);

/* out (q x q) = Q^H G Q, G k x k Hermitian, Q k x q (complex interleaved
 * when cplx): the Gram set of a combined basis V Q from the Gram of V. */
void dn_qhgq(int k, int q, int cplx, const double *G, const double *Q, double *out);

/* out (nu x nv) = Qu^H G Qv for G (ku x kv), Qu (ku x nu), Qv (kv x nv)
 * (complex interleaved when cplx): the Gram between two combined bases. */
void dn_project(int ku, int kv, int nu, int nv, int cplx, const double *G, const double *Qu,
                const double *Qv, double *out);

/* 1 when the diagonally scaled G (k x k, Hermitian) has a Cholesky factor
 * whose every pivot is >= tau, else 0 (also when a diagonal entry is not
 * positive). */
int dn_chol_ok(int k, int cplx, const double *G, double tau);

/* First-order pencil Rayleigh-Ritz (lrep.c): Ritz pairs of
 * G_K a = w G_W b, G_M b = w G_W^H a on (nearly) orthonormal bases of ku and
 * kv columns (G_K ku x ku, G_M kv x kv, G_W ku x kv).  w2[m] ascending (a
 * negative value: a pair with imaginary w, then K u = -|w| v, M v = |w| u),
 * A (ku x m), B (kv x m), ||a||^2 + ||b||^2 = 1.  Returns the number of
 * pairs, or < 0. */
int dn_lrep_rr(int ku, int kv, int m, int cplx, const double *GK, const double *GM,
               const double *GW, double *w2, double *A, double *B);

/* Orthonormal update coefficients CQ (k x (m + p)) = [Q1 | Q2], A = Q1 R11,
 * Q2 spanning the part of A[m:, act] (first m rows zeroed) outside span(A);
 * parts below drop x their norm left out.  Returns m + p or < 0. */
int dn_lrep_update(int k, int m, int cplx, const double *A, const int *act, int nact,
                   double drop, double *CQ, double *R11);

/* C ((nq + nw) x r) with [Q | W] C orthonormal and orthogonal to Q, from
 * GQW = Q^H W and GWW = W^H W (Q orthonormal); directions that keep less
 * than drop of their norm are left out.  Returns r. */
int dn_orth_coef(int nq, int nw, int cplx, const double *GQW, const double *GWW, double drop,
                 double *C);

#endif
