/*
 * HODLR x vector.  Replaces include/hmat_lib/vector_algebra.h:3-8 of the
 * reference (same signature, src/vector_algebra.c:131-197).
 */
#ifndef HMAT_LIB_VECTOR_ALGEBRA_H
#define HMAT_LIB_VECTOR_ALGEBRA_H

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;

double *multiply_vector(const struct TreeHODLR *hodlr, const double *vector,
                        double *out, int *ierr);

#ifdef __cplusplus
}
#endif

#endif
