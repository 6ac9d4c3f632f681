/*
 * HODLR x dense products.  Replaces include/hmat_lib/dense_algebra.h:4-79 of
 * the reference; every signature is the reference's.
 */
#ifndef HMAT_LIB_DENSE_ALGEBRA_H
#define HMAT_LIB_DENSE_ALGEBRA_H

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;
struct HODLRInternalNode;

/* max rank over the tree x matrix_a.  Reference src/dense_algebra.c:39-44. */
int compute_multiply_hodlr_dense_workspace(const struct TreeHODLR *hodlr,
                                           const int matrix_a);

/* out = H * matrix.  Reference src/dense_algebra.c:361-429. */
double *multiply_hodlr_dense(const struct TreeHODLR *hodlr,
                             const double *matrix, const int matrix_n,
                             const int matrix_ld, double *out,
                             const int out_ld, int *ierr);

/* out = H^T * matrix.  Reference src/dense_algebra.c:617-648. */
double *multiply_hodlr_transpose_dense(const struct TreeHODLR *hodlr,
                                       const double *matrix,
                                       const int matrix_n,
                                       const int matrix_ld, double *out,
                                       const int out_ld, int *ierr);

/* out = matrix * H.  Reference src/dense_algebra.c:957-1026. */
double *multiply_dense_hodlr(const struct TreeHODLR *hodlr,
                             const double *matrix, const int matrix_m,
                             const int matrix_ld, double *out,
                             const int out_ld, int *ierr);

/* Subtree forms.  Reference src/dense_algebra.c:492-552, 710-771,
 * 1087-1151.  `queue` and `workspace` are accepted for signature
 * compatibility; the GPU path keeps its own scratch. */
void multiply_internal_node_dense(const struct HODLRInternalNode *internal,
                                  const int height, const double *matrix,
                                  const int matrix_n, const int matrix_ld,
                                  const struct HODLRInternalNode **queue,
                                  double *workspace, double *out,
                                  const int out_ld);

void multiply_internal_node_transpose_dense(
    const struct HODLRInternalNode *internal, const int height,
    const double *matrix, const int matrix_n, const int matrix_ld,
    const struct HODLRInternalNode **queue, double *workspace, double *out,
    const int out_ld);

void multiply_dense_internal_node(const struct HODLRInternalNode *internal,
                                  const int height, const double *matrix,
                                  const int matrix_m, const int matrix_ld,
                                  const struct HODLRInternalNode **queue,
                                  double *workspace, double *out,
                                  const int out_ld);

#ifdef __cplusplus
}
#endif

#endif
