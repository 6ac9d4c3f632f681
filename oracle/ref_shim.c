/*
 * ORACLE RECIPE - TEST INFRASTRUCTURE ONLY.
 *
 * Builds the reference hmat_lib from its own, unmodified sources where they
 * lie (the include path is set by oracle/Makefile to the reference's src/
 * directory) as one translation unit, the same way the reference's unit
 * tests reach its static functions (tests/unit_tests/test_constructors.c:17,
 * test_hodlr_hodlr_algebra.c:19).  The result, oracle/_ref/libhmat_ref.so,
 * exports the reference's public API unchanged plus three ref_* wrappers of
 * its static hot-path routines.  It is a checker; the product never loads it.
 */
#include "allocators.c"
#include "constructors.c"
#include "dense_algebra.c"
#include "hodlr_algebra.c"
#include "lapack_wrapper.c"
#include "utils.c"
#include "vector_algebra.c"

/* src/constructors.c:308-367 */
int ref_compress_off_diagonal(struct NodeOffDiagonal *node, int m_smaller,
                              int matrix_ld, double *matrix, double *s,
                              double *u, double *vt, double svd_threshold,
                              int *ierr) {
  return compress_off_diagonal(node, m_smaller, matrix_ld, matrix, s, u, vt,
                               svd_threshold, ierr);
}

/* src/hodlr_algebra.c:97-211 */
int ref_recompress(struct NodeOffDiagonal *node, int m_smaller,
                   double svd_threshold, int *ierr) {
  return recompress(node, m_smaller, svd_threshold, ierr);
}

/* src/hodlr_algebra.c:280-343 */
int ref_recompress_large_s(struct NodeOffDiagonal *node, int m_smaller,
                           double svd_threshold, int *ierr) {
  return recompress_large_s(node, m_smaller, svd_threshold, ierr);
}
