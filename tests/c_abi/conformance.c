/*
 * A C caller of the hmat_lib API, compiled against the headers of include/hmat_lib and
 * linked to libhmat_lib.so, used the way the reference's own callers use it
 * (dev/test.c:145-240, tests/test_real_data.c:115-236 of hmat_lib): build
 * trees, compress a dense matrix, walk the tree reading u / v / s / data
 * directly, and call every product entry point.
 *
 *   conformance host   tree skeletons, sizes and the rank / workspace
 *                      queries only (no device needed)
 *   conformance full   everything, checked against dense products formed
 *                      here from the walked tree (exact algebra)
 *
 * Prints one line per check and exits non-zero on the first failure.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hmat_lib/allocators.h"
#include "hmat_lib/constructors.h"
#include "hmat_lib/dense_algebra.h"
#include "hmat_lib/error.h"
#include "hmat_lib/hodlr.h"
#include "hmat_lib/hodlr_algebra.h"
#include "hmat_lib/utils.h"
#include "hmat_lib/vector_algebra.h"

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      printf("FAIL %s:%d ", __FILE__, __LINE__);           \
      printf(__VA_ARGS__);                                 \
      printf("\n");                                        \
      exit(1);                                             \
    }                                                      \
  } while (0)

/* Internal nodes of depth d (left to right) by walking children[0] / [3]. */
static void nodes_at_depth(struct HODLRInternalNode *root, int depth,
                           struct HODLRInternalNode **out, int *count) {
  if (depth == 0) {
    out[(*count)++] = root;
    return;
  }
  nodes_at_depth(root->children[0].internal, depth - 1, out, count);
  nodes_at_depth(root->children[3].internal, depth - 1, out, count);
}

/* Dense matrix of a (sub)tree of `height` internal-node layers below and
 * including `node`, from its public fields only; returns the order. */
static int walk_dense(const struct HODLRInternalNode *node, int height,
                      double *a, int lda) {
  int m0, m1;
  if (height == 1) {
    const struct NodeDiagonal *d0 = &node->children[0].leaf->data.diagonal;
    const struct NodeDiagonal *d3 = &node->children[3].leaf->data.diagonal;
    CHECK(node->children[0].leaf->type == DIAGONAL, "leaf type");
    m0 = d0->m;
    m1 = d3->m;
    for (int j = 0; j < m0; j++)
      for (int i = 0; i < m0; i++) a[i + (long)j * lda] = d0->data[i + (long)j * m0];
    for (int j = 0; j < m1; j++)
      for (int i = 0; i < m1; i++)
        a[m0 + i + (long)(m0 + j) * lda] = d3->data[i + (long)j * m1];
  } else {
    m0 = walk_dense(node->children[0].internal, height - 1, a, lda);
    m1 = walk_dense(node->children[3].internal, height - 1, a + m0 + (long)m0 * lda, lda);
  }
  for (int c = 1; c <= 2; c++) {
    const struct HODLRLeafNode *leaf = node->children[c].leaf;
    const struct NodeOffDiagonal *o = &leaf->data.off_diagonal;
    CHECK(leaf->type == OFFDIAGONAL, "off-diagonal type");
    const int r0 = c == 1 ? 0 : m0, c0 = c == 1 ? m0 : 0;
    CHECK(o->m == (c == 1 ? m0 : m1) && o->n == (c == 1 ? m1 : m0), "block dims");
    CHECK(o->s >= 1, "rank");
    for (int j = 0; j < o->n; j++)
      for (int i = 0; i < o->m; i++) {
        double v = 0.0;
        for (int k = 0; k < o->s; k++) v += o->u[i + (long)k * o->m] * o->v[j + (long)k * o->n];
        a[r0 + i + (long)(c0 + j) * lda] = v;
      }
  }
  return m0 + m1;
}

/* C = op(A) op(B), column-major, n x n times n x r etc. */
static void gemm(int ta, int tb, int m, int n, int k, const double *a, int lda,
                 const double *b, int ldb, double *c, int ldc) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int p = 0; p < k; p++)
        s += (ta ? a[p + (long)i * lda] : a[i + (long)p * lda]) *
             (tb ? b[j + (long)p * ldb] : b[p + (long)j * ldb]);
      c[i + (long)j * ldc] = s;
    }
}

static double rel(const double *x, const double *y, long count) {
  double num = 0.0, den = 0.0;
  for (long i = 0; i < count; i++) {
    num += (x[i] - y[i]) * (x[i] - y[i]);
    den += y[i] * y[i];
  }
  return sqrt(num / (den > 0 ? den : 1.0));
}

static void host_checks(void) {
  int ierr = -1;
  for (int h = 1; h <= 5; h++) {
    size_t si, sl, sw, sil;
    compute_construct_tree_array_sizes(h, &si, &sl, &sw, &sil);
    const size_t ni = ((size_t)1 << h) - 1, nl = 3 * ((size_t)1 << h) - 2;
    CHECK(si == ni * sizeof(struct HODLRInternalNode), "internal array size h=%d", h);
    CHECK(sl == nl * sizeof(struct HODLRLeafNode), "leaf array size h=%d", h);
    CHECK(sw == ((size_t)1 << (h - 1)) * sizeof(struct HODLRInternalNode *), "queue size");
    CHECK(sil == ((size_t)1 << h) * sizeof(struct HODLRLeafNode *), "leaf list size");
    struct TreeHODLR tree;
    struct HODLRInternalNode *in = malloc(si);
    struct HODLRLeafNode *lf = malloc(sl);
    struct HODLRInternalNode **wq = malloc(sw);
    struct HODLRLeafNode **il = malloc(sil);
    construct_tree(h, &tree, in, lf, wq, il, &ierr);
    CHECK(ierr == SUCCESS, "construct_tree h=%d", h);
    CHECK(tree.height == h && tree.len_work_queue == (1L << (h - 1)), "tree header");
    CHECK(tree.root->parent == NULL, "root parent");
    /* innermost leaves: the diagonal children of the deepest internal
     * nodes, top-left to bottom-right */
    struct HODLRInternalNode **deep = malloc(sizeof(*deep) << h);
    int cnt = 0;
    nodes_at_depth(tree.root, h - 1, deep, &cnt);
    CHECK(cnt == 1 << (h - 1), "deepest count");
    for (int q = 0; q < cnt; q++) {
      CHECK(tree.innermost_leaves[2 * q] == deep[q]->children[0].leaf, "leaf order");
      CHECK(tree.innermost_leaves[2 * q + 1] == deep[q]->children[3].leaf, "leaf order");
      CHECK(deep[q]->children[1].leaf->parent == deep[q], "parent link");
    }
    /* sizes and ranks set by hand: rank of block (depth d, q, child c) */
    for (int d = 0; d < h; d++) {
      cnt = 0;
      nodes_at_depth(tree.root, d, deep, &cnt);
      for (int q = 0; q < cnt; q++)
        for (int c = 1; c <= 2; c++)
          deep[q]->children[c].leaf->data.off_diagonal.s = 1 + ((3 * d + q + c) % 7);
    }
    for (long c = 0; c < 2 * tree.len_work_queue; c++)
      tree.innermost_leaves[c]->data.diagonal.m = 10 + (int)c;
    int smax = 0;
    for (int d = 0; d < h; d++) {
      cnt = 0;
      nodes_at_depth(tree.root, d, deep, &cnt);
      for (int q = 0; q < cnt; q++)
        for (int c = 1; c <= 2; c++)
          if (deep[q]->children[c].leaf->data.off_diagonal.s > smax)
            smax = deep[q]->children[c].leaf->data.off_diagonal.s;
    }
    CHECK(get_highest_s(&tree) == smax, "get_highest_s h=%d", h);
    CHECK(compute_multiply_hodlr_dense_workspace(&tree, 13) == 13 * smax, "workspace");
    unsigned int w1 = 0, w2 = 0;
    compute_workspace_multiply_hodlr_hodlr(&tree, &tree, &w1, &w2);
    const int largest = 10 + (1 << h) - 1;
    CHECK(w1 == (unsigned)(smax * smax) && w2 == (unsigned)(smax * largest),
          "hodlr_hodlr workspace h=%d: %u %u", h, w1, w2);
    free(deep);
    free(in);
    free(lf);
    free(wq);
    free(il);
  }
  struct TreeHODLR *t = allocate_tree(3, &ierr);
  CHECK(t && ierr == SUCCESS, "allocate_tree");
  free_tree_hodlr(&t);
  CHECK(t == NULL, "free_tree_hodlr clears the pointer");
  t = allocate_tree_monolithic(4, &ierr);
  CHECK(t && ierr == SUCCESS && t->height == 4, "allocate_tree_monolithic");
  free_tree_hodlr(&t);
  t = allocate_tree(0, &ierr);
  CHECK(t == NULL && ierr == INPUT_ERROR, "height 0 is an input error");
  printf("ok host: skeletons, sizes, rank and workspace queries\n");
}

static void full_checks(void) {
  const int h = 3, n = 203;
  const int ms[8] = {30, 25, 1, 47, 27, 33, 9, 31};
  int ierr = -1;
  double *x = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) x[i] = (double)i / n + 0.37 * sin(0.9 * i);
  double *a = malloc(sizeof(double) * n * n);
  double *a_in = malloc(sizeof(double) * n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      a[i + (long)j * n] = exp(-fabs(x[i] - x[j]) / 0.3) + (i == j ? 0.5 : 0.0);
  memcpy(a_in, a, sizeof(double) * n * n);

  struct TreeHODLR *t = allocate_tree_monolithic(h, &ierr);
  CHECK(ierr == SUCCESS, "allocate");
  dense_to_tree_hodlr(t, n, ms, a, 1e-12, &ierr);
  CHECK(ierr == SUCCESS, "dense_to_tree_hodlr: %d", ierr);
  CHECK(t->root->m == n, "root m");
  for (int c = 0; c < 8; c++)
    CHECK(t->innermost_leaves[c]->data.diagonal.m == ms[c], "leaf %d size", c);

  /* the tree from its public fields */
  double *hd = calloc((size_t)n * n, sizeof(double));
  CHECK(walk_dense(t->root, h, hd, n) == n, "walk order");
  const double err_c = rel(hd, a_in, (long)n * n);
  printf("compression: |walked tree - A| / |A| = %.2e, highest rank %d\n", err_c,
         get_highest_s(t));
  CHECK(err_c < 1e-10, "compression error %.2e", err_c);

  /* every product entry point vs the walked dense matrix */
  const int r = 7;
  double *X = malloc(sizeof(double) * n * r), *Xt = malloc(sizeof(double) * r * n);
  for (long i = 0; i < (long)n * r; i++) X[i] = cos(0.31 * i) + 0.1 * (i % 5);
  for (long i = 0; i < (long)n * r; i++) Xt[i] = sin(0.17 * i) - 0.2 * (i % 3);
  double *ref = malloc(sizeof(double) * n * r), *out = malloc(sizeof(double) * n * r);

  double *y = multiply_vector(t, x, NULL, &ierr);
  CHECK(y && ierr == SUCCESS, "multiply_vector");
  gemm(0, 0, n, 1, n, hd, n, x, n, ref, n);
  CHECK(rel(y, ref, n) < 1e-12, "multiply_vector %.2e", rel(y, ref, n));
  free(y);

  double *o = multiply_hodlr_dense(t, X, r, n, out, n, &ierr);
  gemm(0, 0, n, r, n, hd, n, X, n, ref, n);
  CHECK(o == out && ierr == SUCCESS && rel(out, ref, (long)n * r) < 1e-12,
        "multiply_hodlr_dense %.2e", rel(out, ref, (long)n * r));
  o = multiply_hodlr_dense(t, X, r, n, NULL, n, &ierr);
  CHECK(o && rel(o, ref, (long)n * r) < 1e-12, "multiply_hodlr_dense, library output");
  free(o);

  o = multiply_hodlr_transpose_dense(t, X, r, n, out, n, &ierr);
  gemm(1, 0, n, r, n, hd, n, X, n, ref, n);
  CHECK(ierr == SUCCESS && rel(out, ref, (long)n * r) < 1e-12,
        "multiply_hodlr_transpose_dense %.2e", rel(out, ref, (long)n * r));

  o = multiply_dense_hodlr(t, Xt, r, r, out, r, &ierr);
  gemm(0, 0, r, n, n, Xt, r, hd, n, ref, r);
  CHECK(ierr == SUCCESS && rel(out, ref, (long)n * r) < 1e-12,
        "multiply_dense_hodlr %.2e", rel(out, ref, (long)n * r));

  /* internal-node forms on the bottom-right subtree, with the caller's
   * queue and workspace sized as the reference documents */
  struct HODLRInternalNode *sub = t->root->children[3].internal;
  const int ms3 = sub->m, off = n - ms3;
  const int wsz = compute_multiply_hodlr_dense_workspace(t, r);
  double *ws = malloc(sizeof(double) * (wsz > 0 ? wsz : 1));
  const struct HODLRInternalNode **queue = malloc(sizeof(*queue) * t->len_work_queue);
  double *hsub = malloc(sizeof(double) * ms3 * ms3);
  for (int j = 0; j < ms3; j++)
    for (int i = 0; i < ms3; i++) hsub[i + (long)j * ms3] = hd[off + i + (long)(off + j) * n];
  multiply_internal_node_dense(sub, h - 1, X, r, n, queue, ws, out, ms3);
  gemm(0, 0, ms3, r, ms3, hsub, ms3, X, n, ref, ms3);
  CHECK(rel(out, ref, (long)ms3 * r) < 1e-12, "multiply_internal_node_dense");
  multiply_internal_node_transpose_dense(sub, h - 1, X, r, n, queue, ws, out, ms3);
  gemm(1, 0, ms3, r, ms3, hsub, ms3, X, n, ref, ms3);
  CHECK(rel(out, ref, (long)ms3 * r) < 1e-12, "multiply_internal_node_transpose_dense");
  multiply_dense_internal_node(sub, h - 1, Xt, r, r, queue, ws, out, r);
  gemm(0, 0, r, ms3, ms3, Xt, r, hsub, ms3, ref, r);
  CHECK(rel(out, ref, (long)ms3 * r) < 1e-12, "multiply_dense_internal_node");

  /* H x H: the product tree, walked, vs the dense square */
  unsigned int w1 = 0, w2 = 0;
  compute_workspace_multiply_hodlr_hodlr(t, t, &w1, &w2);
  const int s1 = get_highest_s(t);
  CHECK(w1 == (unsigned)(s1 * s1) && w2 == (unsigned)(s1 * 47), "hh workspace");
  struct TreeHODLR *p = allocate_tree_monolithic(h, &ierr);
  multiply_hodlr_hodlr(t, t, p, 1e-13, &ierr);
  CHECK(ierr == SUCCESS, "multiply_hodlr_hodlr: %d", ierr);
  double *pd = calloc((size_t)n * n, sizeof(double)), *hh = malloc(sizeof(double) * n * n);
  walk_dense(p->root, h, pd, n);
  gemm(0, 0, n, n, n, hd, n, hd, n, hh, n);
  const double err_hh = rel(pd, hh, (long)n * n);
  printf("multiply_hodlr_hodlr: |walked product - H H| / |H H| = %.2e\n", err_hh);
  CHECK(err_hh < 1e-11, "multiply_hodlr_hodlr error %.2e", err_hh);

  free_tree_hodlr(&p);
  free_tree_hodlr(&t);
  free(pd), free(hh), free(hsub), free(queue), free(ws), free(ref), free(out);
  free(X), free(Xt), free(hd), free(a), free(a_in), free(x);
  printf("ok full: compression, tree walk, 7 product entry points, H x H\n");
}

int main(int argc, char **argv) {
  host_checks();
  if (argc > 1 && strcmp(argv[1], "full") == 0) full_checks();
  return 0;
}
