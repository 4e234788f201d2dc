/*
 * constants.h -- physical constants used by programs written against the
 * reference API (reference include/bdg/constants.h; its examples include it
 * for unit conversion).  CODATA 2018 values in SI units.
 */
#ifndef BDG_CONSTANTS_H
#define BDG_CONSTANTS_H

#define HBAR 1.054571817e-34        /* J s */
#define BOHR_RADIUS 5.29177210903e-11 /* m */
#define MASS_UNIT 1.6605390666e-27  /* kg, atomic mass unit */

#endif /* BDG_CONSTANTS_H */
