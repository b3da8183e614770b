/* Oracle BLAS/LAPACK subset with the Fortran ABI (sgemv_, dsymv_, ...).
 *
 * TEST INFRASTRUCTURE ONLY.  SPAMS calls its arithmetic through an external
 * Fortran BLAS/LAPACK that it does not vendor (linalg/cblas_alt_template.h:
 * 291-370 declares the symbols).  This file provides those symbols so that the
 * reference's own headers can be compiled in place into oracle/_ref, and so
 * that the C restatement in spx_oracle.c shares one set of BLAS semantics with
 * it.  The routines follow the published BLAS reference specification; see
 * blas_body.inc.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define REAL float
#define F(x) s##x##_
#define I(x) is##x##_
#include "blas_body.inc"
#undef REAL
#undef F
#undef I

#define LSAME LSAME_d
#define REAL double
#define F(x) d##x##_
#define I(x) id##x##_
#include "blas_body.inc"
#undef REAL
#undef F
#undef I
#undef LSAME
