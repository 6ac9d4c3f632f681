/* Tree skeleton allocation and freeing (host only).
 *
 * Same API and observable structure as the reference (src/allocators.c):
 * height >= 1, 2^h - 1 internal nodes, 3*2^h - 2 leaves, children order
 * [0] top-left, [1] top-right, [2] bottom-left, [3] bottom-right,
 * innermost_leaves and work_queue left to right, zeroed sizes and NULL
 * data.  The node arrays of the monolithic form are laid out in heap order
 * (root at index 0, children of k at 2k+1 and 2k+2), with the off-diagonal
 * leaves of internal node k at leaf indices 2k, 2k+1 followed by the
 * diagonal leaves.  Freeing a tree also drops its cached device copy. */
#include <stdlib.h>
#include <string.h>

#include "hm_host.h"
#include "hmat_lib/allocators.h"

typedef void *(*hm_malloc_fn)(size_t);
typedef void (*hm_free_fn)(void *);

static void init_offdiag(struct HODLRLeafNode *leaf,
                         struct HODLRInternalNode *parent) {
  leaf->type = OFFDIAGONAL;
  leaf->parent = parent;
  leaf->data.off_diagonal.u = NULL;
  leaf->data.off_diagonal.v = NULL;
  leaf->data.off_diagonal.m = 0;
  leaf->data.off_diagonal.s = 0;
  leaf->data.off_diagonal.n = 0;
}

static void init_diag(struct HODLRLeafNode *leaf,
                      struct HODLRInternalNode *parent) {
  leaf->type = DIAGONAL;
  leaf->parent = parent;
  leaf->data.diagonal.data = NULL;
  leaf->data.diagonal.m = 0;
}

static void init_internal(struct HODLRInternalNode *node,
                          struct HODLRInternalNode *parent) {
  node->parent = parent;
  node->m = 0;
  for (int c = 0; c < 4; c++) node->children[c].leaf = NULL;
}

/* Link pre-allocated nodes into a tree.  internal[k] in heap order; leaves
 * off-diagonal first (2 per internal node), then diagonal (2 per bottom
 * internal node). */
static void link_tree(int height, struct TreeHODLR *t,
                      struct HODLRInternalNode **internal,
                      struct HODLRLeafNode **leaves,
                      struct HODLRInternalNode **work_queue,
                      struct HODLRLeafNode **innermost) {
  const long nleaf = 1L << height, nint = nleaf - 1;
  for (long k = 0; k < nint; k++)
    init_internal(internal[k], k == 0 ? NULL : internal[(k - 1) / 2]);
  for (long k = 0; k < nint; k++) {
    struct HODLRInternalNode *p = internal[k];
    p->children[1].leaf = leaves[2 * k];
    p->children[2].leaf = leaves[2 * k + 1];
    init_offdiag(leaves[2 * k], p);
    init_offdiag(leaves[2 * k + 1], p);
    if (2 * k + 2 < nint) {
      p->children[0].internal = internal[2 * k + 1];
      p->children[3].internal = internal[2 * k + 2];
    }
  }
  const long first_bottom = nleaf / 2 - 1;
  for (long q = 0; q < nleaf / 2; q++) {
    struct HODLRInternalNode *p = internal[first_bottom + q];
    struct HODLRLeafNode *d0 = leaves[2 * nint + 2 * q];
    struct HODLRLeafNode *d3 = leaves[2 * nint + 2 * q + 1];
    init_diag(d0, p);
    init_diag(d3, p);
    p->children[0].leaf = d0;
    p->children[3].leaf = d3;
    work_queue[q] = p;
    innermost[2 * q] = d0;
    innermost[2 * q + 1] = d3;
  }
  t->height = height;
  t->root = internal[0];
  t->innermost_leaves = innermost;
  t->len_work_queue = nleaf / 2;
  t->work_queue = work_queue;
}

/* Reference: src/allocators.c:229-344.  Every node is its own malloc. */
HM_EXPORT struct TreeHODLR *allocate_tree(const int height, int *ierr) {
  if (height < 1 || height > 30) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  const long nleaf = 1L << height, nint = nleaf - 1;
  const long nleaves_all = 2 * nint + nleaf;
  struct TreeHODLR *t = malloc(sizeof(*t));
  struct HODLRInternalNode **internal = calloc(nint, sizeof(*internal));
  struct HODLRLeafNode **leaves = calloc(nleaves_all, sizeof(*leaves));
  struct HODLRInternalNode **wq = malloc((nleaf / 2) * sizeof(*wq));
  struct HODLRLeafNode **inner = malloc(nleaf * sizeof(*inner));
  int ok = t && internal && leaves && wq && inner;
  for (long k = 0; ok && k < nint; k++) ok = (internal[k] = malloc(sizeof(**internal))) != NULL;
  for (long k = 0; ok && k < nleaves_all; k++) ok = (leaves[k] = malloc(sizeof(**leaves))) != NULL;
  if (!ok) {
    if (internal) for (long k = 0; k < nint; k++) free(internal[k]);
    if (leaves) for (long k = 0; k < nleaves_all; k++) free(leaves[k]);
    free(internal); free(leaves); free(wq); free(inner); free(t);
    *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  link_tree(height, t, internal, leaves, wq, inner);
  t->memory_internal_ptr = NULL;
  t->memory_leaf_ptr = NULL;
  free(internal);
  free(leaves);
  *ierr = SUCCESS;
  return t;
}

/* Reference: src/allocators.c:499-512 (sizes in bytes). */
HM_EXPORT void compute_construct_tree_array_sizes(
    const int height, size_t *size_internal_nodes, size_t *size_leaf_nodes,
    size_t *size_work_queue, size_t *size_innermost_leaves) {
  const size_t half = (size_t)1 << (height - 1);
  *size_work_queue = half * sizeof(struct HODLRInternalNode *);
  *size_internal_nodes = (2 * half - 1) * sizeof(struct HODLRInternalNode);
  *size_innermost_leaves = 2 * half * sizeof(struct HODLRLeafNode *);
  *size_leaf_nodes = (6 * half - 2) * sizeof(struct HODLRLeafNode);
}

/* Reference: src/allocators.c:561-635. */
HM_EXPORT void construct_tree(const int height, struct TreeHODLR *hodlr,
                              struct HODLRInternalNode *internal_nodes,
                              struct HODLRLeafNode *leaf_nodes,
                              struct HODLRInternalNode **work_queue,
                              struct HODLRLeafNode **innermost_leaves,
                              int *ierr) {
  if (height < 1 || height > 30) {
    *ierr = INPUT_ERROR;
    return;
  }
  const long nleaf = 1L << height, nint = nleaf - 1;
  const long nleaves_all = 2 * nint + nleaf;
  struct HODLRInternalNode **ip = malloc(nint * sizeof(*ip));
  struct HODLRLeafNode **lp = malloc(nleaves_all * sizeof(*lp));
  if (!ip || !lp) {
    free(ip); free(lp);
    *ierr = ALLOCATION_FAILURE;
    return;
  }
  for (long k = 0; k < nint; k++) ip[k] = internal_nodes + k;
  for (long k = 0; k < nleaves_all; k++) lp[k] = leaf_nodes + k;
  link_tree(height, hodlr, ip, lp, work_queue, innermost_leaves);
  hodlr->memory_internal_ptr = internal_nodes;
  hodlr->memory_leaf_ptr = leaf_nodes;
  free(ip);
  free(lp);
  *ierr = SUCCESS;
}

static struct TreeHODLR *monolithic_impl(int height, int *ierr, hm_malloc_fn mal,
                                         hm_free_fn fr) {
  if (height < 1 || height > 30) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  size_t si, sl, sw, sil;
  compute_construct_tree_array_sizes(height, &si, &sl, &sw, &sil);
  struct TreeHODLR *t = mal(sizeof(*t));
  struct HODLRInternalNode *in = mal(si);
  struct HODLRLeafNode *lv = mal(sl);
  struct HODLRInternalNode **wq = mal(sw);
  struct HODLRLeafNode **il = mal(sil);
  if (!t || !in || !lv || !wq || !il) {
    if (t) fr(t);
    if (in) fr(in);
    if (lv) fr(lv);
    if (wq) fr(wq);
    if (il) fr(il);
    *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  construct_tree(height, t, in, lv, wq, il, ierr);
  if (*ierr != SUCCESS) {
    fr(t); fr(in); fr(lv); fr(wq); fr(il);
    return NULL;
  }
  return t;
}

static void free_data_impl(struct TreeHODLR *t, hm_free_fn fr) {
  if (!t || !t->root) return;
  hm_side_drop_tree(t);
  struct HODLRInternalNode **all = hm_levels(t->root, t->height);
  if (!all) return;
  const long nint = (1L << t->height) - 1;
  for (long k = 0; k < nint; k++) {
    for (int c = 1; c < 3; c++) {
      struct NodeOffDiagonal *o = &all[k]->children[c].leaf->data.off_diagonal;
      if (o->u) fr(o->u);
      if (o->v) fr(o->v);
      o->u = NULL;
      o->v = NULL;
    }
  }
  for (long i = 0; i < (1L << t->height); i++) {
    struct NodeDiagonal *d = &t->innermost_leaves[i]->data.diagonal;
    if (d->data) fr(d->data);
    d->data = NULL;
  }
  free(all);
}

static void free_tree_impl(struct TreeHODLR **pt, hm_free_fn fr) {
  if (!pt || !*pt) return;
  struct TreeHODLR *t = *pt;
  free_data_impl(t, fr);
  if (t->memory_leaf_ptr != NULL) {
    fr(t->memory_leaf_ptr);
    fr(t->memory_internal_ptr);
    fr(t->work_queue);
    fr(t->innermost_leaves);
    fr(t);
    *pt = NULL;
    return;
  }
  struct HODLRInternalNode **all = hm_levels(t->root, t->height);
  const long nint = (1L << t->height) - 1;
  if (all) {
    for (long k = 0; k < nint; k++) {
      fr(all[k]->children[1].leaf);
      fr(all[k]->children[2].leaf);
      if (k >= nint / 2) {
        fr(all[k]->children[0].leaf);
        fr(all[k]->children[3].leaf);
      }
    }
    for (long k = 0; k < nint; k++) fr(all[k]);
    free(all);
  }
  fr(t->work_queue);
  fr(t->innermost_leaves);
  fr(t);
  *pt = NULL;
}

#ifndef _TEST_HODLR
/* Reference: src/allocators.c:404-464. */
HM_EXPORT struct TreeHODLR *allocate_tree_monolithic(const int height, int *ierr) {
  return monolithic_impl(height, ierr, malloc, free);
}
/* Reference: src/allocators.c:659-708. */
HM_EXPORT void free_tree_data(struct TreeHODLR *hodlr) { free_data_impl(hodlr, free); }
/* Reference: src/allocators.c:747-839. */
HM_EXPORT void free_tree_hodlr(struct TreeHODLR **hodlr_ptr) {
  free_tree_impl(hodlr_ptr, free);
}
#else
HM_EXPORT struct TreeHODLR *allocate_tree_monolithic(const int height, int *ierr,
                                                     void *(*mal)(size_t),
                                                     void (*fr)(void *)) {
  return monolithic_impl(height, ierr, mal, fr);
}
HM_EXPORT void free_tree_data(struct TreeHODLR *hodlr, void (*fr)(void *)) {
  free_data_impl(hodlr, fr);
}
HM_EXPORT void free_tree_hodlr(struct TreeHODLR **hodlr_ptr, void (*fr)(void *)) {
  free_tree_impl(hodlr_ptr, fr);
}
#endif
