/*
 * Dense -> HODLR compression.  Replaces include/hmat_lib/constructors.h:3-14
 * of the reference (same signature).  The off-diagonal blocks are
 * compressed on the GPU; the host tree receives malloc'd copies of every
 * factor and diagonal block, and the device copy stays cached for later
 * products.  Unlike the reference, `matrix` is never overwritten.
 */
#ifndef HMAT_LIB_CONSTRUCTORS_H
#define HMAT_LIB_CONSTRUCTORS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;

/* Reference: src/constructors.c:629-704. */
int dense_to_tree_hodlr(struct TreeHODLR *hodlr, const int m, const int *ms,
                        double *matrix, const double svd_threshold, int *ierr
#ifdef _TEST_HODLR
                        , void *(*permanent_allocator)(size_t size),
                        void (*free)(void *ptr)
#endif
);

#ifdef __cplusplus
}
#endif

#endif
