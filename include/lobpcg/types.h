/*
 * Scalar typedefs used by the public BdG header.
 *
 * The reference's public header includes "lobpcg/types.h" from an external,
 * un-vendored eigensolver library (reference include/bdg/bdg.h:4).  Only the
 * two scalar names it relies on are needed here: f64 (real double) and c64
 * (double-precision complex, laid out as {re, im}).
 */
#ifndef BDG_AMD_LOBPCG_TYPES_H
#define BDG_AMD_LOBPCG_TYPES_H

#ifdef __cplusplus
#include <complex>
typedef double f64;
typedef std::complex<double> c64;
#else
#include <complex.h>
typedef double f64;
typedef double _Complex c64;
#endif

#endif /* BDG_AMD_LOBPCG_TYPES_H */
