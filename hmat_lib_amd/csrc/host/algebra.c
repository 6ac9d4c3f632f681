/* Products and recompression: host side of the drop-in API.
 *
 * Same signatures, argument checks and output semantics as the reference
 * (src/vector_algebra.c:131-197, src/dense_algebra.c:39-1151,
 * src/utils.c:5-26, src/hodlr_algebra.c:97-343); the arithmetic runs on
 * the device through hm_hip_*.  Host arrays are staged through
 * process-wide device buffers; the tree itself comes from the side table. */
#include <stdlib.h>
#include <string.h>

#include "hm_host.h"
#include "hmat_lib/dense_algebra.h"
#include "hmat_lib/hodlr_algebra.h"
#include "hmat_lib/utils.h"
#include "hmat_lib/vector_algebra.h"

static int max_rank_sub(const struct HODLRInternalNode *root, int height) {
  struct HODLRInternalNode **all = hm_levels(root, height);
  if (!all) return 0;
  int smax = 0;
  for (long k = 0; k < (1L << height) - 1; k++)
    for (int c = 1; c < 3; c++) {
      int s = all[k]->children[c].leaf->data.off_diagonal.s;
      if (s > smax) smax = s;
    }
  free(all);
  return smax;
}

/* Reference: src/utils.c:5-26. */
HM_EXPORT int get_highest_s(const struct TreeHODLR *const hodlr) {
  return max_rank_sub(hodlr->root, hodlr->height);
}

/* Reference: src/dense_algebra.c:39-44. */
HM_EXPORT int compute_multiply_hodlr_dense_workspace(const struct TreeHODLR *hodlr,
                                                     const int matrix_a) {
  return get_highest_s(hodlr) * matrix_a;
}

/* Y(rows_y x cols) = op(H) X on host arrays.  mode 0: H X; 1: H^T X;
 * 2: X H (X is `other` x N).  N is the order of the (sub)tree. */
static int host_product(const struct HODLRInternalNode *node, int height,
                        int mode, const double *x, int other, int ldx,
                        double *y, int ldy) {
  int rc = SUCCESS;
  hm_dtree *dt = hm_side_get(node, height, &rc);
  if (!dt) return rc;
  const int n = node->m;
  if (n <= 0 || other <= 0) return SUCCESS;
  const size_t elems = (size_t)n * (size_t)other;
  hm_staging_lock();
  double *dx = NULL, *dy = NULL;
  rc = hm_staging(elems * sizeof(double), elems * sizeof(double), &dx, &dy);
  if (rc == SUCCESS) {
    if (mode == 2) {
      /* X is other x n (ld ldx); device copy ld = other */
      rc = hm_hip_copy2d_h2d(dx, other, x, ldx, other, n, NULL);
      if (rc == SUCCESS)
        rc = hm_hip_matmat(dt, 1, other, dx, other, 1, dy, other, 1, NULL);
      if (rc == SUCCESS) rc = hm_hip_copy2d_d2h(y, ldy, dy, other, other, n, NULL);
    } else {
      rc = hm_hip_copy2d_h2d(dx, n, x, ldx, n, other, NULL);
      if (rc == SUCCESS)
        rc = hm_hip_matmat(dt, mode, other, dx, 1, n, dy, 1, n, NULL);
      if (rc == SUCCESS) rc = hm_hip_copy2d_d2h(y, ldy, dy, n, n, other, NULL);
    }
  }
  hm_staging_unlock();
  return hm_report(rc, "HODLR product");
}

/* Reference: src/vector_algebra.c:131-197. */
HM_EXPORT double *multiply_vector(const struct TreeHODLR *hodlr,
                                  const double *vector, double *out, int *ierr) {
  if (hodlr == NULL || vector == NULL || vector == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  const int alloc = out == NULL;
  if (alloc) {
    out = malloc((size_t)(hodlr->root->m > 0 ? hodlr->root->m : 1) * sizeof(double));
    if (out == NULL) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  int rc = SUCCESS;
  hm_dtree *dt = hm_side_get(hodlr->root, hodlr->height, &rc);
  const int n = hodlr->root->m;
  if (dt && n > 0) {
    hm_staging_lock();
    double *dx = NULL, *dy = NULL;
    rc = hm_staging((size_t)n * sizeof(double), (size_t)n * sizeof(double), &dx, &dy);
    if (rc == SUCCESS) rc = hm_hip_memcpy_h2d(dx, vector, (size_t)n * sizeof(double), NULL);
    if (rc == SUCCESS) rc = hm_hip_matvec(dt, dx, dy, NULL);
    if (rc == SUCCESS) rc = hm_hip_memcpy_d2h(out, dy, (size_t)n * sizeof(double), NULL);
    hm_staging_unlock();
    hm_report(rc, "multiply_vector");
  }
  if (rc != SUCCESS) {
    *ierr = rc;
    if (alloc) free(out);
    return NULL;
  }
  return out;
}

static double *dense_entry(const struct TreeHODLR *hodlr, const double *matrix,
                           int other, int matrix_ld, double *out, int out_ld,
                           int *ierr, int mode) {
  if (hodlr == NULL || matrix == NULL || matrix == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  const int alloc = out == NULL;
  if (alloc) {
    size_t cols = mode == 2 ? (size_t)hodlr->root->m : (size_t)other;
    out = malloc((size_t)out_ld * cols * sizeof(double) + sizeof(double));
    if (out == NULL) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  int rc = host_product(hodlr->root, hodlr->height, mode, matrix, other,
                        matrix_ld, out, out_ld);
  if (rc != SUCCESS) {
    *ierr = rc;
    if (alloc) free(out);
    return NULL;
  }
  return out;
}

/* Reference: src/dense_algebra.c:361-429. */
HM_EXPORT double *multiply_hodlr_dense(const struct TreeHODLR *hodlr,
                                       const double *matrix, const int matrix_n,
                                       const int matrix_ld, double *out,
                                       const int out_ld, int *ierr) {
  return dense_entry(hodlr, matrix, matrix_n, matrix_ld, out, out_ld, ierr, 0);
}

/* Reference: src/dense_algebra.c:617-648. */
HM_EXPORT double *multiply_hodlr_transpose_dense(const struct TreeHODLR *hodlr,
                                                 const double *matrix,
                                                 const int matrix_n,
                                                 const int matrix_ld,
                                                 double *out, const int out_ld,
                                                 int *ierr) {
  return dense_entry(hodlr, matrix, matrix_n, matrix_ld, out, out_ld, ierr, 1);
}

/* Reference: src/dense_algebra.c:957-1026. */
HM_EXPORT double *multiply_dense_hodlr(const struct TreeHODLR *hodlr,
                                       const double *matrix, const int matrix_m,
                                       const int matrix_ld, double *out,
                                       const int out_ld, int *ierr) {
  return dense_entry(hodlr, matrix, matrix_m, matrix_ld, out, out_ld, ierr, 2);
}

/* Reference: src/dense_algebra.c:492-552.  No error channel in the
 * signature: failures are reported on stderr. */
HM_EXPORT void multiply_internal_node_dense(const struct HODLRInternalNode *internal,
                                            const int height, const double *matrix,
                                            const int matrix_n, const int matrix_ld,
                                            const struct HODLRInternalNode **queue,
                                            double *workspace, double *out,
                                            const int out_ld) {
  (void)queue;
  (void)workspace;
  host_product(internal, height, 0, matrix, matrix_n, matrix_ld, out, out_ld);
}

/* Reference: src/dense_algebra.c:710-771. */
HM_EXPORT void multiply_internal_node_transpose_dense(
    const struct HODLRInternalNode *internal, const int height,
    const double *matrix, const int matrix_n, const int matrix_ld,
    const struct HODLRInternalNode **queue, double *workspace, double *out,
    const int out_ld) {
  (void)queue;
  (void)workspace;
  host_product(internal, height, 1, matrix, matrix_n, matrix_ld, out, out_ld);
}

/* Reference: src/dense_algebra.c:1087-1151. */
HM_EXPORT void multiply_dense_internal_node(const struct HODLRInternalNode *internal,
                                            const int height, const double *matrix,
                                            const int matrix_m, const int matrix_ld,
                                            const struct HODLRInternalNode **queue,
                                            double *workspace, double *out,
                                            const int out_ld) {
  (void)queue;
  (void)workspace;
  host_product(internal, height, 2, matrix, matrix_m, matrix_ld, out, out_ld);
}

/* Recompression of one host node through the batched device routine. */
static int recompress_host(struct NodeOffDiagonal *node, double thr, int method,
                           int *ierr) {
  if (!node || node->s <= 0 || node->m <= 0 || node->n <= 0 || !node->u || !node->v) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  const int m = node->m, n = node->n, s = node->s;
  double *du = NULL, *dv = NULL;
  int rc = hm_hip_malloc((void **)&du, (size_t)m * s * sizeof(double));
  if (rc == SUCCESS) rc = hm_hip_malloc((void **)&dv, (size_t)n * s * sizeof(double));
  if (rc == SUCCESS) rc = hm_hip_memcpy_h2d(du, node->u, (size_t)m * s * sizeof(double), NULL);
  if (rc == SUCCESS) rc = hm_hip_memcpy_h2d(dv, node->v, (size_t)n * s * sizeof(double), NULL);
  int k = 0;
  if (rc == SUCCESS) rc = hm_hip_recompress_batch(1, &m, &n, &s, &du, &dv, thr, method, &k, NULL);
  double *un = NULL, *vn = NULL;
  if (rc == SUCCESS) {
    un = malloc((size_t)m * k * sizeof(double));
    vn = malloc((size_t)n * k * sizeof(double));
    if (!un || !vn) rc = ALLOCATION_FAILURE;
  }
  if (rc == SUCCESS) rc = hm_hip_memcpy_d2h(un, du, (size_t)m * k * sizeof(double), NULL);
  if (rc == SUCCESS) rc = hm_hip_memcpy_d2h(vn, dv, (size_t)n * k * sizeof(double), NULL);
  hm_hip_free(du);
  hm_hip_free(dv);
  if (rc != SUCCESS) {
    free(un);
    free(vn);
    *ierr = hm_report(rc, "recompress");
    return 0;
  }
  free(node->u);
  free(node->v);
  node->u = un;
  node->v = vn;
  node->s = k;
  return 0;
}

/* Reference: src/hodlr_algebra.c:97-211. */
HM_EXPORT int recompress(struct NodeOffDiagonal *node, const int m_smaller,
                         const double svd_threshold, int *ierr) {
  (void)m_smaller;
  return recompress_host(node, svd_threshold, 0, ierr);
}

/* Reference: src/hodlr_algebra.c:280-343. */
HM_EXPORT int recompress_large_s(struct NodeOffDiagonal *node, const int m_smaller,
                                 const double svd_threshold, int *ierr) {
  (void)m_smaller;
  return recompress_host(node, svd_threshold, 1, ierr);
}

/* Reference: src/hodlr_algebra.c:1356-1374 (non-static there, documented in
 * docs/source/api/hodlr_algebra.rst:7, no header declaration).  The sizes of
 * the two workspaces the reference's multiply_hodlr_hodlr allocates:
 * size1 = s1 * s2 and size2 = s1 * (largest diagonal leaf of hodlr1), s1, s2
 * the largest ranks of the two trees.  The device product needs no caller
 * workspace; the query is kept so callers that size buffers with it link. */
HM_EXPORT void compute_workspace_multiply_hodlr_hodlr(
    const struct TreeHODLR *const hodlr1, const struct TreeHODLR *const hodlr2,
    unsigned int *restrict const size1, unsigned int *restrict const size2) {
  const int s1 = get_highest_s(hodlr1);
  const int s2 = get_highest_s(hodlr2);
  *size1 = (unsigned int)(s1 * s2);
  int largest_m = 0;
  for (long c = 0; c < 2 * hodlr1->len_work_queue; c++) {
    const int m = hodlr1->innermost_leaves[c]->data.diagonal.m;
    if (m > largest_m) largest_m = m;
  }
  *size2 = (unsigned int)(s1 * largest_m);
}

/* Reference: src/hodlr_algebra.c:1452-1553.  Both inputs come from the side
 * table (uploaded on first use), the product is formed and recompressed on
 * the device, then copied into malloc'd arrays of `out` (which must hold no
 * data, as in the reference) and kept in the side table for later products.
 * Trees of different heights or leaf sizes give INPUT_ERROR (the reference
 * leaves that undefined).  Unlike the reference, which never writes them,
 * the internal-node sizes of `out` are set. */
HM_EXPORT int multiply_hodlr_hodlr(const struct TreeHODLR *hodlr1,
                                   const struct TreeHODLR *hodlr2,
                                   struct TreeHODLR *out,
                                   const double svd_threshold, int *ierr) {
  if (hodlr1 == NULL || hodlr2 == NULL || out == NULL) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  *ierr = SUCCESS;
  if (hodlr1->height != hodlr2->height || out->height != hodlr1->height) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  int rc = SUCCESS;
  hm_dtree *da = hm_side_get(hodlr1->root, hodlr1->height, &rc);
  hm_dtree *db = da ? hm_side_get(hodlr2->root, hodlr2->height, &rc) : NULL;
  hm_dtree *dc = NULL;
  if (da && db) rc = hm_hip_multiply_hodlr_hodlr(da, db, svd_threshold, &dc, NULL);
  if (rc == SUCCESS && dc) rc = hm_hip_tree_download(dc, out, malloc, NULL);
  if (rc != SUCCESS) {
    if (dc) hm_hip_tree_destroy(dc);
    *ierr = hm_report(rc, "multiply_hodlr_hodlr");
    return 0;
  }
  hm_side_put(out->root, out->height, dc);
  return 0;
}
