/*
 * HODLR x HODLR product and recompression of a low-rank off-diagonal node.
 *
 * multiply_hodlr_hodlr is the reference's public entry point
 * (include/hmat_lib/hodlr_algebra.h:4-10 of hmat_lib).  The reference keeps
 * its two recompression routines static (src/hodlr_algebra.c:97-211
 * `recompress`, QR based, used when s < min(m, n); src/hodlr_algebra.c:
 * 280-343 `recompress_large_s`, dense SVD of U*V^T, used otherwise) and
 * reaches them only through multiply_hodlr_hodlr (src/hodlr_algebra.c:
 * 835-845, 1116-1126).  This header exports them directly so that callers
 * can recompress on the GPU.  Both replace u, v and s of the node in place
 * (u and v are freed and re-malloc'd) and use the reference's absolute
 * truncation rule: keep sigma_i while sigma_i > svd_threshold, at least one
 * column.
 *
 * Return value: 0, or a negative LAPACK-style argument index on bad input.
 */
#ifndef HMAT_LIB_HODLR_ALGEBRA_H
#define HMAT_LIB_HODLR_ALGEBRA_H

#ifdef __cplusplus
extern "C" {
#endif

struct NodeOffDiagonal;
struct TreeHODLR;

/* out = hodlr1 @ hodlr2 (same height and block sizes).  `out` must be an
 * allocated tree of that height holding no data; its diagonal and
 * off-diagonal arrays are malloc'd (free with free_tree_data /
 * free_tree_hodlr).  Every off-diagonal block is recompressed with the
 * absolute rule above.  Replaces src/hodlr_algebra.c:1452-1553. */
int multiply_hodlr_hodlr(const struct TreeHODLR *hodlr1,
                         const struct TreeHODLR *hodlr2,
                         struct TreeHODLR *out, const double svd_threshold,
                         int *ierr);

/* Workspace sizes of the reference's multiply_hodlr_hodlr: *size1 =
 * s1 * s2, *size2 = s1 * (largest diagonal leaf of hodlr1), with s1, s2 the
 * largest ranks of hodlr1 and hodlr2.  Replaces src/hodlr_algebra.c:
 * 1356-1374 (exported there without a header declaration). */
void compute_workspace_multiply_hodlr_hodlr(const struct TreeHODLR *const hodlr1,
                                            const struct TreeHODLR *const hodlr2,
                                            unsigned int *const size1,
                                            unsigned int *const size2);

/* QR method: U = Q_U R_U, V = Q_V R_V, SVD of R_U R_V^T (s x s).
 * Requires node->s < min(node->m, node->n).
 * Replaces src/hodlr_algebra.c:97-211. */
int recompress(struct NodeOffDiagonal *node, const int m_smaller,
               const double svd_threshold, int *ierr);

/* Dense method: SVD of the m x n product U V^T.
 * Replaces src/hodlr_algebra.c:280-343. */
int recompress_large_s(struct NodeOffDiagonal *node, const int m_smaller,
                       const double svd_threshold, int *ierr);

#ifdef __cplusplus
}
#endif

#endif
