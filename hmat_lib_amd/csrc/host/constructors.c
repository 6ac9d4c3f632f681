/* dense_to_tree_hodlr: host side of the drop-in compression.
 *
 * Reference: src/constructors.c:629-704.  Block sizes are set on the host
 * tree exactly as the reference does (halves: top-left child gets
 * m - m/2, src/constructors.c:44-95; custom leaf sizes summed upwards,
 * 133-176, INPUT_ERROR when they do not add up to m, 648-656).  The dense
 * matrix is copied to the device once, every off-diagonal block is
 * compressed there (hm_hip_compress_dense), and the result is copied back
 * into malloc'd u/v/data arrays of the host tree; the device tree stays
 * cached for later products.  The input matrix is not modified. */
#include <stdlib.h>

#include "hm_host.h"
#include "hmat_lib/constructors.h"

static void set_offdiag_dims(struct HODLRInternalNode *p, int m0, int m3) {
  p->children[1].leaf->data.off_diagonal.m = m0;
  p->children[1].leaf->data.off_diagonal.n = m3;
  p->children[2].leaf->data.off_diagonal.m = m3;
  p->children[2].leaf->data.off_diagonal.n = m0;
}

static int block_sizes(struct TreeHODLR *t, int m, const int *ms) {
  struct HODLRInternalNode **all = hm_levels(t->root, t->height);
  if (!all) return ALLOCATION_FAILURE;
  const int h = t->height;
  if (!ms) {
    t->root->m = m;
    for (int d = 0; d < h; d++) {
      struct HODLRInternalNode **lev = all + ((1L << d) - 1);
      for (long i = 0; i < (1L << d); i++) {
        struct HODLRInternalNode *p = lev[i];
        const int small = p->m / 2, large = p->m - small;
        set_offdiag_dims(p, large, small);
        if (d + 1 < h) {
          p->children[0].internal->m = large;
          p->children[3].internal->m = small;
        } else {
          p->children[0].leaf->data.diagonal.m = large;
          p->children[3].leaf->data.diagonal.m = small;
        }
      }
    }
  } else {
    for (int d = h - 1; d >= 0; d--) {
      struct HODLRInternalNode **lev = all + ((1L << d) - 1);
      for (long i = 0; i < (1L << d); i++) {
        struct HODLRInternalNode *p = lev[i];
        int m0, m3;
        if (d == h - 1) {
          m0 = ms[2 * i];
          m3 = ms[2 * i + 1];
          p->children[0].leaf->data.diagonal.m = m0;
          p->children[3].leaf->data.diagonal.m = m3;
        } else {
          m0 = p->children[0].internal->m;
          m3 = p->children[3].internal->m;
        }
        p->m = m0 + m3;
        set_offdiag_dims(p, m0, m3);
      }
    }
  }
  free(all);
  return SUCCESS;
}

static int dense_to_tree_impl(struct TreeHODLR *hodlr, const int m,
                              const int *ms, const double *matrix,
                              const double thr, int *ierr,
                              void *(*alloc)(size_t)) {
  if (hodlr == NULL || matrix == NULL) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  *ierr = SUCCESS;
  int rc = block_sizes(hodlr, m, ms);
  if (rc != SUCCESS) {
    *ierr = rc;
    return 0;
  }
  if (ms != NULL && hodlr->root->m != m) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  const long nleaf = 1L << hodlr->height;
  int *leaves = malloc((size_t)nleaf * sizeof(int));
  if (!leaves) {
    *ierr = ALLOCATION_FAILURE;
    return 0;
  }
  for (long i = 0; i < nleaf; i++)
    leaves[i] = hodlr->innermost_leaves[i]->data.diagonal.m;
  double *d_a = NULL;
  size_t bytes = (size_t)m * (size_t)m * sizeof(double);
  rc = hm_hip_malloc((void **)&d_a, bytes);
  if (rc == SUCCESS) rc = hm_hip_memcpy_h2d(d_a, matrix, bytes, NULL);
  hm_dtree *dt = NULL;
  if (rc == SUCCESS)
    rc = hm_hip_compress_dense(d_a, m > 0 ? m : 1, m, hodlr->height, leaves,
                               thr, NULL, &dt, NULL, NULL);
  hm_hip_free(d_a);
  free(leaves);
  if (rc == SUCCESS) rc = hm_hip_tree_download(dt, hodlr, alloc, NULL);
  if (rc != SUCCESS) {
    if (dt) hm_hip_tree_destroy(dt);
    *ierr = hm_report(rc, "dense_to_tree_hodlr");
    return 0;
  }
  hm_side_put(hodlr->root, hodlr->height, dt);
  return 0;
}

#ifndef _TEST_HODLR
HM_EXPORT int dense_to_tree_hodlr(struct TreeHODLR *hodlr, const int m,
                                  const int *ms, double *matrix,
                                  const double svd_threshold, int *ierr) {
  return dense_to_tree_impl(hodlr, m, ms, matrix, svd_threshold, ierr, malloc);
}
#else
HM_EXPORT int dense_to_tree_hodlr(struct TreeHODLR *hodlr, const int m,
                                  const int *ms, double *matrix,
                                  const double svd_threshold, int *ierr,
                                  void *(*permanent_allocator)(size_t),
                                  void (*fr)(void *)) {
  (void)fr;
  return dense_to_tree_impl(hodlr, m, ms, matrix, svd_threshold, ierr,
                            permanent_allocator);
}
#endif
