/*
 * ORACLE - TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch CPU restatement of the reference hmat_lib HODLR hot path,
 * used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * as the checker.  Nothing in the product (libhmat_lib, libhmat_hip) links,
 * loads or calls this file.
 *
 * It follows the reference's algorithm and its BLAS/LAPACK call sequence
 * (so that, linked against the same LAPACK, it reproduces the reference's
 * results bit for bit), with 64-bit offsets so it also runs past the
 * reference's int32 limits (N >= 32768).  Each function cites the
 * reference source it restates.  Arithmetic is FP64, column-major.
 *
 * LAPACK/BLAS come from the host library this file is linked against
 * (Fortran LP64 symbols dgesdd_, dgeqrt_, dgemqrt_, dtrmm_, dgemm_, dgemv_).
 *
 * How each part is pinned (checked by tests/test_oracle.py):
 *   block sizes          reference tables test_constructors.c:149-254,
 *                        342-436 (tests/golden/block_sizes.json)
 *   tree allocation      closed-form counts (test_allocators.c:82-99)
 *   compression          exact Laplacian/identity/zero known answers
 *                        (test_constructors.c:725-873), closed-form rank 1
 *                        of the sorted exponential kernel, bitwise
 *                        agreement with oracle/_ref (the reference built
 *                        from its own sources) where that builds
 *   products             known answers of test_vector_algebra.c:54-154,
 *                        239-360 and test_hodlr_dense_algebra.c:163-276;
 *                        the dense sum of the tree's blocks (exact algebra)
 *   recompression        12-digit golden values of
 *                        test_hodlr_hodlr_algebra.c:1643-1669 and the zero
 *                        cases 1575-1630; reconstruction invariants
 *   H x H product        the known answers of test_hodlr_hodlr_algebra.c:
 *                        100-339, the dense product of random trees (exact
 *                        algebra), bitwise agreement with oracle/_ref
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/hmat_lib/error.h"
#include "../include/hmat_lib/hodlr.h"

/* Fortran LP64 BLAS/LAPACK prototypes (the reference declares the same
 * seven in include/internal/blas_wrapper.h:1-23 and lapack_wrapper.h). */
void dgemm_(const char *, const char *, const int *, const int *, const int *,
            const double *, const double *, const int *, const double *,
            const int *, const double *, double *, const int *);
void dgemv_(const char *, const int *, const int *, const double *,
            const double *, const int *, const double *, const int *,
            const double *, double *, const int *);
void dgeqrt_(const int *, const int *, const int *, double *, const int *,
             double *, const int *, double *, int *);
void dtrmm_(const char *, const char *, const char *, const char *,
            const int *, const int *, const double *, const double *,
            const int *, double *, const int *);
void dgemqrt_(const char *, const char *, const int *, const int *,
              const int *, const int *, const double *, const int *,
              const double *, const int *, double *, const int *, double *,
              int *);
void dgesdd_(const char *, const int *, const int *, double *, const int *,
             double *, double *, const int *, double *, const int *, double *,
             const int *, int *, int *);

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------
 * Tree skeleton.  Restates allocate_tree_monolithic + construct_tree
 * (src/allocators.c:404-464, 561-635): 2^h - 1 internal nodes, 3*2^h - 2
 * leaves, innermost_leaves and work_queue left to right.  Built here by a
 * breadth-first walk instead of the reference's strided queue.
 * ---------------------------------------------------------------------- */
static void orc_init_offdiag(struct HODLRLeafNode *leaf,
                             struct HODLRInternalNode *parent) {
  leaf->type = OFFDIAGONAL;
  leaf->parent = parent;
  leaf->data.off_diagonal.u = NULL;
  leaf->data.off_diagonal.v = NULL;
  leaf->data.off_diagonal.m = 0;
  leaf->data.off_diagonal.s = 0;
  leaf->data.off_diagonal.n = 0;
}

static void orc_init_diag(struct HODLRLeafNode *leaf,
                          struct HODLRInternalNode *parent) {
  leaf->type = DIAGONAL;
  leaf->parent = parent;
  leaf->data.diagonal.data = NULL;
  leaf->data.diagonal.m = 0;
}

ORC_API struct TreeHODLR *orc_allocate_tree(int height, int *ierr) {
  if (height < 1) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  const long nleaf = 1L << height;
  const long ninternal = nleaf - 1;
  const long noffleaf = 2 * ninternal;
  struct TreeHODLR *t = calloc(1, sizeof(*t));
  struct HODLRInternalNode *in = calloc(ninternal, sizeof(*in));
  struct HODLRLeafNode *lv = calloc(noffleaf + nleaf, sizeof(*lv));
  struct HODLRInternalNode **wq = calloc(nleaf / 2, sizeof(*wq));
  struct HODLRLeafNode **il = calloc(nleaf, sizeof(*il));
  if (!t || !in || !lv || !wq || !il) {
    free(t); free(in); free(lv); free(wq); free(il);
    *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  /* internal node k (heap order: children of k are 2k+1, 2k+2) */
  long next_leaf = 0;
  for (long k = 0; k < ninternal; k++) {
    in[k].parent = (k == 0) ? NULL : &in[(k - 1) / 2];
    in[k].m = 0;
    long c0 = 2 * k + 1, c3 = 2 * k + 2;
    for (int c = 1; c < 3; c++) {
      in[k].children[c].leaf = &lv[next_leaf++];
      orc_init_offdiag(in[k].children[c].leaf, &in[k]);
    }
    if (c0 < ninternal) {
      in[k].children[0].internal = &in[c0];
      in[k].children[3].internal = &in[c3];
    }
  }
  /* the bottom internal layer (heap indices nleaf/2-1 .. nleaf-2) */
  for (long q = 0; q < nleaf / 2; q++) {
    struct HODLRInternalNode *p = &in[nleaf / 2 - 1 + q];
    wq[q] = p;
    p->children[0].leaf = &lv[next_leaf++];
    p->children[3].leaf = &lv[next_leaf++];
    orc_init_diag(p->children[0].leaf, p);
    orc_init_diag(p->children[3].leaf, p);
    il[2 * q] = p->children[0].leaf;
    il[2 * q + 1] = p->children[3].leaf;
  }
  t->height = height;
  t->root = &in[0];
  t->innermost_leaves = il;
  t->len_work_queue = nleaf / 2;
  t->work_queue = wq;
  t->memory_internal_ptr = in;
  t->memory_leaf_ptr = lv;
  *ierr = SUCCESS;
  return t;
}

/* Internal nodes of a (sub)tree per depth, left to right.  levels[d] holds
 * 2^d pointers, d = 0 .. height-1.  Returned as one malloc'd block. */
static struct HODLRInternalNode **orc_levels(
    const struct HODLRInternalNode *root, int height) {
  long total = (1L << height) - 1;
  struct HODLRInternalNode **all = malloc(total * sizeof(*all));
  all[0] = (struct HODLRInternalNode *)root;
  long base = 0;
  for (int d = 0; d + 1 < height; d++) {
    long cnt = 1L << d, nbase = base + cnt;
    for (long i = 0; i < cnt; i++) {
      all[nbase + 2 * i] = all[base + i]->children[0].internal;
      all[nbase + 2 * i + 1] = all[base + i]->children[3].internal;
    }
    base = nbase;
  }
  return all;
}
#define ORC_LEVEL(all, d) ((all) + ((1L << (d)) - 1))

ORC_API void orc_free_tree_data(struct TreeHODLR *t) {
  if (!t) return;
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  long total = (1L << t->height) - 1;
  for (long k = 0; k < total; k++) {
    for (int c = 1; c < 3; c++) {
      struct NodeOffDiagonal *o = &all[k]->children[c].leaf->data.off_diagonal;
      free(o->u); free(o->v);
      o->u = NULL; o->v = NULL;
    }
  }
  for (long i = 0; i < (1L << t->height); i++) {
    free(t->innermost_leaves[i]->data.diagonal.data);
    t->innermost_leaves[i]->data.diagonal.data = NULL;
  }
  free(all);
}

ORC_API void orc_free_tree(struct TreeHODLR *t) {
  if (!t) return;
  orc_free_tree_data(t);
  free(t->memory_internal_ptr);
  free(t->memory_leaf_ptr);
  free(t->work_queue);
  free(t->innermost_leaves);
  free(t);
}

/* ------------------------------------------------------------------------
 * Block sizes.  Halves: src/constructors.c:44-95 (top-left child gets
 * m - m/2).  Custom: src/constructors.c:133-176 (sums leaves upward).
 * ---------------------------------------------------------------------- */
static void orc_set_offdiag_dims(struct HODLRInternalNode *p, int m0, int m3) {
  p->children[1].leaf->data.off_diagonal.m = m0;
  p->children[1].leaf->data.off_diagonal.n = m3;
  p->children[2].leaf->data.off_diagonal.m = m3;
  p->children[2].leaf->data.off_diagonal.n = m0;
}

ORC_API void orc_block_sizes_halves(struct TreeHODLR *t, int m) {
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  t->root->m = m;
  for (int d = 0; d < t->height; d++) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    for (long i = 0; i < (1L << d); i++) {
      struct HODLRInternalNode *p = lev[i];
      int small = p->m / 2, large = p->m - small;
      orc_set_offdiag_dims(p, large, small);
      if (d + 1 < t->height) {
        p->children[0].internal->m = large;
        p->children[3].internal->m = small;
      } else {
        p->children[0].leaf->data.diagonal.m = large;
        p->children[3].leaf->data.diagonal.m = small;
      }
    }
  }
  free(all);
}

ORC_API void orc_block_sizes_custom(struct TreeHODLR *t, const int *ms) {
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  for (int d = t->height - 1; d >= 0; d--) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    for (long i = 0; i < (1L << d); i++) {
      struct HODLRInternalNode *p = lev[i];
      int m0, m3;
      if (d == t->height - 1) {
        m0 = ms[2 * i];
        m3 = ms[2 * i + 1];
        p->children[0].leaf->data.diagonal.m = m0;
        p->children[3].leaf->data.diagonal.m = m3;
      } else {
        m0 = p->children[0].internal->m;
        m3 = p->children[3].internal->m;
      }
      p->m = m0 + m3;
      orc_set_offdiag_dims(p, m0, m3);
    }
  }
  free(all);
}

/* ------------------------------------------------------------------------
 * svd_double (src/lapack_wrapper.c:8-59): dgesdd JOBZ='S' with a workspace
 * query; only info < 0 is an error.
 * ---------------------------------------------------------------------- */
static int orc_svd(int m, int n, int k, int lda, double *a, double *s,
                   double *u, double *vt, int *ierr) {
  double wq;
  int lwork = -1, info = 0;
  int *iwork = malloc(8 * (size_t)(k > 0 ? k : 1) * sizeof(int));
  if (!iwork) {
    *ierr = SVD_ALLOCATION_FAILURE;
    return 0;
  }
  dgesdd_("S", &m, &n, a, &lda, s, u, &m, vt, &k, &wq, &lwork, iwork, &info);
  if (info < 0) {
    free(iwork);
    *ierr = SVD_FAILURE;
    return info;
  }
  lwork = (int)wq;
  double *work = malloc((size_t)lwork * sizeof(double));
  if (!work) {
    free(iwork);
    *ierr = SVD_ALLOCATION_FAILURE;
    return info;
  }
  dgesdd_("S", &m, &n, a, &lda, s, u, &m, vt, &k, work, &lwork, iwork, &info);
  if (info < 0) *ierr = SVD_FAILURE;
  free(work);
  free(iwork);
  return info;
}

/* compress_off_diagonal (src/constructors.c:308-367): SVD, relative rank
 * rule (keep sigma_i >= thr*sigma_0; rank 1 when sigma_0 <= thr),
 * U = U_svd * diag(sigma), V = (V^T)^T.  `a` (m x n, lda) is overwritten.
 * sigma_out (optional) receives all min(m, n) singular values. */
ORC_API int orc_compress_block(int m, int n, double *a, long lda, double thr,
                               double **u_out, double **v_out, int *s_out,
                               double *sigma_out, int *ierr) {
  int k = m < n ? m : n;
  *u_out = NULL; *v_out = NULL; *s_out = 0;
  double *s = malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
  double *u = malloc((size_t)m * (k > 0 ? k : 1) * sizeof(double));
  double *vt = malloc((size_t)n * (k > 0 ? k : 1) * sizeof(double));
  if (!s || !u || !vt) {
    free(s); free(u); free(vt);
    *ierr = ALLOCATION_FAILURE;
    return 0;
  }
  int res = orc_svd(m, n, k, (int)lda, a, s, u, vt, ierr);
  if (*ierr != SUCCESS) {
    free(s); free(u); free(vt);
    return res;
  }
  int cut = 1;
  if (s[0] > thr) {
    for (cut = 1; cut < k; cut++)
      if (s[cut] < thr * s[0]) break;
  }
  double *uo = malloc((size_t)m * cut * sizeof(double));
  double *vo = malloc((size_t)n * cut * sizeof(double));
  if (!uo || !vo) {
    free(uo); free(vo); free(s); free(u); free(vt);
    *ierr = ALLOCATION_FAILURE;
    return res;
  }
  for (int i = 0; i < cut; i++)
    for (int j = 0; j < m; j++) uo[j + (size_t)i * m] = u[j + (size_t)i * m] * s[i];
  for (int i = 0; i < cut; i++)
    for (int j = 0; j < n; j++) vo[j + (size_t)i * n] = vt[i + (size_t)j * k];
  if (sigma_out) memcpy(sigma_out, s, (size_t)k * sizeof(double));
  *u_out = uo; *v_out = vo; *s_out = cut;
  free(s); free(u); free(vt);
  return res;
}

/* dense_to_tree_hodlr (src/constructors.c:629-704) with copy_diagonal_blocks
 * (212-243) and compress_matrix (449-546): sizes, diagonal copies, then the
 * off-diagonal blocks level by level from the bottom, top-right block at
 * A + o + lda*(o + m0), bottom-left at A + lda*o + o + m0.  The blocks of
 * one level are independent (the reference runs them as OpenMP tasks); here
 * they are distributed over OpenMP threads when built with -fopenmp. */
ORC_API int orc_dense_to_tree(struct TreeHODLR *t, int m, const int *ms,
                              double *a, double thr, int *ierr) {
  if (!t || !a) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  *ierr = SUCCESS;
  if (ms == NULL) {
    orc_block_sizes_halves(t, m);
  } else {
    orc_block_sizes_custom(t, ms);
    if (t->root->m != m) {
      *ierr = INPUT_ERROR;
      return 0;
    }
  }
  const long lda = m;
  long off = 0;
  for (long i = 0; i < (1L << t->height); i++) {
    int mi = t->innermost_leaves[i]->data.diagonal.m;
    double *d = malloc((size_t)mi * mi * sizeof(double) + 1);
    if (!d) {
      *ierr = ALLOCATION_FAILURE;
      return 0;
    }
    for (long j = 0; j < mi; j++)
      for (long r = 0; r < mi; r++) d[r + j * mi] = a[(off + r) + (off + j) * lda];
    t->innermost_leaves[i]->data.diagonal.data = d;
    off += mi;
  }
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  int result = 0, status = SUCCESS;
  for (int d = t->height - 1; d >= 0; d--) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    long cnt = 1L << d;
    long *offs = malloc(cnt * sizeof(long));
    long o = 0;
    for (long i = 0; i < cnt; i++) {
      offs[i] = o;
      o += lev[i]->m;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (long job = 0; job < 2 * cnt; job++) {
      long i = job / 2;
      int c = 1 + (int)(job % 2);
      struct NodeOffDiagonal *nd = &lev[i]->children[c].leaf->data.off_diagonal;
      int m0 = lev[i]->children[1].leaf->data.off_diagonal.m;
      double *sub = (c == 1) ? a + offs[i] + lda * (offs[i] + m0)
                             : a + lda * offs[i] + offs[i] + m0;
      int e = SUCCESS;
      int r = orc_compress_block(nd->m, nd->n, sub, lda, thr, &nd->u, &nd->v,
                                 &nd->s, NULL, &e);
      if (e != SUCCESS) {
#pragma omp critical
        {
          status = e;
          result = r;
        }
      }
    }
    free(offs);
  }
  free(all);
  *ierr = status;
  return result;
}

/* ------------------------------------------------------------------------
 * Products.
 * ---------------------------------------------------------------------- */
/* multiply_off_diagonal_vector (src/vector_algebra.c:52-85). */
static void orc_offdiag_vector(const struct HODLRInternalNode *p,
                               const double *x, double *y, double *w) {
  const double one = 1.0, zero = 0.0;
  const int inc = 1;
  const struct NodeOffDiagonal *tr = &p->children[1].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *bl = &p->children[2].leaf->data.off_diagonal;
  int m = tr->m, n = tr->n, s = tr->s;
  dgemv_("T", &n, &s, &one, tr->v, &n, x + m, &inc, &zero, w, &inc);
  dgemv_("N", &m, &s, &one, tr->u, &m, w, &inc, &one, y, &inc);
  s = bl->s;
  dgemv_("T", &m, &s, &one, bl->v, &m, x, &inc, &zero, w, &inc);
  dgemv_("N", &n, &s, &one, bl->u, &n, w, &inc, &one, y + m, &inc);
}

/* multiply_vector (src/vector_algebra.c:131-197): diagonal dgemv (beta=0),
 * then levels bottom-up, root last. */
ORC_API double *orc_multiply_vector(const struct TreeHODLR *t, const double *x,
                                    double *out, int *ierr) {
  if (!t || !x || x == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  if (!out) {
    out = malloc((size_t)t->root->m * sizeof(double));
    if (!out) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  const double one = 1.0, zero = 0.0;
  const int inc = 1;
  int smax = 1;
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  for (long k = 0; k < (1L << t->height) - 1; k++)
    for (int c = 1; c < 3; c++) {
      int s = all[k]->children[c].leaf->data.off_diagonal.s;
      if (s > smax) smax = s;
    }
  double *w = malloc((size_t)smax * sizeof(double));
  long off = 0;
  for (long i = 0; i < (1L << t->height); i++) {
    const struct NodeDiagonal *dg = &t->innermost_leaves[i]->data.diagonal;
    int mi = dg->m;
    dgemv_("N", &mi, &mi, &one, dg->data, &mi, x + off, &inc, &zero, out + off,
           &inc);
    off += mi;
  }
  for (int d = t->height - 1; d >= 0; d--) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    off = 0;
    for (long i = 0; i < (1L << d); i++) {
      orc_offdiag_vector(lev[i], x + off, out + off, w);
      off += lev[i]->m;
    }
  }
  free(w);
  free(all);
  return out;
}

/* multiply_low_rank_dense (src/dense_algebra.c:82-98) and its transpose
 * (141-157): W = V^T X (s x R), Y = U W + beta Y. */
static void orc_lowrank_dense(const struct NodeOffDiagonal *nd, int trans,
                              const double *x, int r, int ldx, double beta,
                              double *w, double *y, int ldy) {
  const double one = 1.0, zero = 0.0;
  if (!trans) {
    dgemm_("T", "N", &nd->s, &r, &nd->n, &one, nd->v, &nd->n, x, &ldx, &zero, w,
           &nd->s);
    dgemm_("N", "N", &nd->m, &r, &nd->s, &one, nd->u, &nd->m, w, &nd->s, &beta,
           y, &ldy);
  } else {
    dgemm_("T", "N", &nd->s, &r, &nd->m, &one, nd->u, &nd->m, x, &ldx, &zero, w,
           &nd->s);
    dgemm_("N", "N", &nd->n, &r, &nd->s, &one, nd->v, &nd->n, w, &nd->s, &beta,
           y, &ldy);
  }
}

/* multiply_off_diagonal_dense (src/dense_algebra.c:202-225) and the
 * transposed pair (271-295). */
static void orc_offdiag_dense(const struct HODLRInternalNode *p, int trans,
                              const double *x, int r, int ldx, double beta,
                              double *y, int ldy, double *w) {
  int m = p->children[1].leaf->data.off_diagonal.m;
  const struct NodeOffDiagonal *tr = &p->children[1].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *bl = &p->children[2].leaf->data.off_diagonal;
  if (!trans) {
    orc_lowrank_dense(tr, 0, x + m, r, ldx, beta, w, y, ldy);
    orc_lowrank_dense(bl, 0, x, r, ldx, beta, w, y + m, ldy);
  } else {
    orc_lowrank_dense(bl, 1, x + m, r, ldx, beta, w, y, ldy);
    orc_lowrank_dense(tr, 1, x, r, ldx, beta, w, y + m, ldy);
  }
}

static int orc_max_rank(const struct HODLRInternalNode *root, int height) {
  struct HODLRInternalNode **all = orc_levels(root, height);
  int smax = 0;
  for (long k = 0; k < (1L << height) - 1; k++)
    for (int c = 1; c < 3; c++) {
      int s = all[k]->children[c].leaf->data.off_diagonal.s;
      if (s > smax) smax = s;
    }
  free(all);
  return smax;
}

/* get_highest_s (src/utils.c:5-26). */
ORC_API int orc_get_highest_s(const struct TreeHODLR *t) {
  return orc_max_rank(t->root, t->height);
}

/* multiply_hodlr_dense (src/dense_algebra.c:361-429): diagonal dgemm with
 * beta=0, then levels bottom-up, root last. */
ORC_API double *orc_multiply_hodlr_dense(const struct TreeHODLR *t,
                                         const double *x, int r, int ldx,
                                         double *out, int ldo, int *ierr) {
  if (!t || !x || x == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  if (!out) {
    out = malloc((size_t)ldo * r * sizeof(double));
    if (!out) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  const double one = 1.0, zero = 0.0;
  int smax = orc_max_rank(t->root, t->height);
  double *w = malloc((size_t)(smax > 0 ? smax : 1) * r * sizeof(double));
  long off = 0;
  for (long i = 0; i < (1L << t->height); i++) {
    const struct NodeDiagonal *dg = &t->innermost_leaves[i]->data.diagonal;
    int mi = dg->m;
    dgemm_("N", "N", &mi, &r, &mi, &one, dg->data, &mi, x + off, &ldx, &zero,
           out + off, &ldo);
    off += mi;
  }
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  for (int d = t->height - 1; d >= 0; d--) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    off = 0;
    for (long i = 0; i < (1L << d); i++) {
      orc_offdiag_dense(lev[i], 0, x + off, r, ldx, one, out + off, ldo, w);
      off += lev[i]->m;
    }
  }
  free(all);
  free(w);
  return out;
}

/* multiply_internal_node_dense (src/dense_algebra.c:492-552) and
 * multiply_internal_node_transpose_dense (710-771): root pair with beta=0,
 * then the levels top-down, then the diagonal with beta=1. */
ORC_API void orc_multiply_internal_node_dense(
    const struct HODLRInternalNode *node, int height, int trans,
    const double *x, int r, int ldx, double *out, int ldo) {
  const double one = 1.0, zero = 0.0;
  int smax = orc_max_rank(node, height);
  double *w = malloc((size_t)(smax > 0 ? smax : 1) * r * sizeof(double));
  struct HODLRInternalNode **all = orc_levels(node, height);
  orc_offdiag_dense(node, trans, x, r, ldx, zero, out, ldo, w);
  for (int d = 1; d < height; d++) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    long off = 0;
    for (long i = 0; i < (1L << d); i++) {
      orc_offdiag_dense(lev[i], trans, x + off, r, ldx, one, out + off, ldo, w);
      off += lev[i]->m;
    }
  }
  struct HODLRInternalNode **bottom = ORC_LEVEL(all, height - 1);
  long off = 0;
  for (long i = 0; i < (1L << (height - 1)); i++) {
    for (int c = 0; c < 4; c += 3) {
      const struct NodeDiagonal *dg = &bottom[i]->children[c].leaf->data.diagonal;
      int mi = dg->m;
      dgemm_(trans ? "T" : "N", "N", &mi, &r, &mi, &one, dg->data, &mi, x + off,
             &ldx, &one, out + off, &ldo);
      off += mi;
    }
  }
  free(all);
  free(w);
}

/* multiply_hodlr_transpose_dense (src/dense_algebra.c:617-648). */
ORC_API double *orc_multiply_hodlr_transpose_dense(const struct TreeHODLR *t,
                                                   const double *x, int r,
                                                   int ldx, double *out,
                                                   int ldo, int *ierr) {
  if (!t || !x || x == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  if (!out) {
    out = malloc((size_t)ldo * r * sizeof(double));
    if (!out) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  orc_multiply_internal_node_dense(t->root, t->height, 1, x, r, ldx, out, ldo);
  return out;
}

/* multiply_dense_low_rank (src/dense_algebra.c:808-824) and
 * multiply_dense_off_diagonal (870-893): W = X U (rows x s), Y += W V^T. */
static void orc_dense_offdiag(const struct HODLRInternalNode *p,
                              const double *x, int rows, int ldx, double beta,
                              double *y, int ldy, double *w) {
  const double one = 1.0, zero = 0.0;
  int m = p->children[1].leaf->data.off_diagonal.m;
  for (int c = 1; c < 3; c++) {
    const struct NodeOffDiagonal *nd = &p->children[c].leaf->data.off_diagonal;
    const double *xs = (c == 1) ? x : x + (size_t)m * ldx;
    double *ys = (c == 1) ? y + (size_t)m * ldy : y;
    dgemm_("N", "N", &rows, &nd->s, &nd->m, &one, xs, &ldx, nd->u, &nd->m,
           &zero, w, &rows);
    dgemm_("N", "T", &rows, &nd->n, &nd->s, &one, w, &rows, nd->v, &nd->n,
           &beta, ys, &ldy);
  }
}

/* multiply_dense_hodlr (src/dense_algebra.c:957-1026). */
ORC_API double *orc_multiply_dense_hodlr(const struct TreeHODLR *t,
                                         const double *x, int rows, int ldx,
                                         double *out, int ldo, int *ierr) {
  if (!t || !x || x == out) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  if (!out) {
    out = malloc((size_t)ldo * t->root->m * sizeof(double));
    if (!out) {
      *ierr = ALLOCATION_FAILURE;
      return NULL;
    }
  }
  *ierr = SUCCESS;
  const double one = 1.0, zero = 0.0;
  int smax = orc_max_rank(t->root, t->height);
  double *w = malloc((size_t)(smax > 0 ? smax : 1) * rows * sizeof(double));
  long off = 0;
  for (long i = 0; i < (1L << t->height); i++) {
    const struct NodeDiagonal *dg = &t->innermost_leaves[i]->data.diagonal;
    int mi = dg->m;
    dgemm_("N", "N", &rows, &mi, &mi, &one, x + off * ldx, &ldx, dg->data, &mi,
           &zero, out + off * ldo, &ldo);
    off += mi;
  }
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  for (int d = t->height - 1; d >= 0; d--) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    off = 0;
    for (long i = 0; i < (1L << d); i++) {
      orc_dense_offdiag(lev[i], x + off * ldx, rows, ldx, one, out + off * ldo,
                        ldo, w);
      off += lev[i]->m;
    }
  }
  free(all);
  free(w);
  return out;
}

/* multiply_dense_internal_node (src/dense_algebra.c:1087-1151). */
ORC_API void orc_multiply_dense_internal_node(
    const struct HODLRInternalNode *node, int height, const double *x,
    int rows, int ldx, double *out, int ldo) {
  const double one = 1.0, zero = 0.0;
  int smax = orc_max_rank(node, height);
  double *w = malloc((size_t)(smax > 0 ? smax : 1) * rows * sizeof(double));
  struct HODLRInternalNode **all = orc_levels(node, height);
  orc_dense_offdiag(node, x, rows, ldx, zero, out, ldo, w);
  for (int d = 1; d < height; d++) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    long off = 0;
    for (long i = 0; i < (1L << d); i++) {
      orc_dense_offdiag(lev[i], x + off * ldx, rows, ldx, one, out + off * ldo,
                        ldo, w);
      off += lev[i]->m;
    }
  }
  struct HODLRInternalNode **bottom = ORC_LEVEL(all, height - 1);
  long off = 0;
  for (long i = 0; i < (1L << (height - 1)); i++) {
    for (int c = 0; c < 4; c += 3) {
      const struct NodeDiagonal *dg = &bottom[i]->children[c].leaf->data.diagonal;
      int mi = dg->m;
      dgemm_("N", "N", &rows, &mi, &mi, &one, x + off * ldx, &ldx, dg->data,
             &mi, &one, out + off * ldo, &ldo);
      off += mi;
    }
  }
  free(all);
  free(w);
}

/* Dense reconstruction of a tree (sum of its blocks); exact algebra used to
 * pin the products. */
ORC_API void orc_tree_to_dense(const struct TreeHODLR *t, double *a, long lda) {
  const double one = 1.0, zero = 0.0;
  long off = 0;
  for (long i = 0; i < (1L << t->height); i++) {
    const struct NodeDiagonal *dg = &t->innermost_leaves[i]->data.diagonal;
    for (long j = 0; j < dg->m; j++)
      for (long r = 0; r < dg->m; r++)
        a[(off + r) + (off + j) * lda] = dg->data[r + j * dg->m];
    off += dg->m;
  }
  struct HODLRInternalNode **all = orc_levels(t->root, t->height);
  for (int d = 0; d < t->height; d++) {
    struct HODLRInternalNode **lev = ORC_LEVEL(all, d);
    off = 0;
    for (long i = 0; i < (1L << d); i++) {
      const struct NodeOffDiagonal *tr = &lev[i]->children[1].leaf->data.off_diagonal;
      const struct NodeOffDiagonal *bl = &lev[i]->children[2].leaf->data.off_diagonal;
      int ld = (int)lda;
      int m0 = tr->m;
      if (tr->s > 0)
        dgemm_("N", "T", &tr->m, &tr->n, &tr->s, &one, tr->u, &tr->m, tr->v,
               &tr->n, &zero, a + off + (off + m0) * lda, &ld);
      if (bl->s > 0)
        dgemm_("N", "T", &bl->m, &bl->n, &bl->s, &one, bl->u, &bl->m, bl->v,
               &bl->n, &zero, a + (off + m0) + off * lda, &ld);
      off += lev[i]->m;
    }
  }
  free(all);
}

/* ------------------------------------------------------------------------
 * Recompression.
 * recompress (src/hodlr_algebra.c:97-211): nb = min(s, 32); dgeqrt of U and
 * V; R = triu(R_U) * R_V^T by dtrmm; SVD of R; absolute rule (stop at the
 * first sigma_i <= thr, i >= 1); U_new = Q_U [W Sigma; 0], V_new = Q_V [Z; 0]
 * by dgemqrt.
 * ---------------------------------------------------------------------- */
ORC_API int orc_recompress(struct NodeOffDiagonal *nd, int m_smaller,
                           double thr, int *ierr) {
  const int s = nd->s;
  const int nb = s < 32 ? s : 32;
  int info = 0;
  size_t tsz = (size_t)nb * (m_smaller > s ? m_smaller : s);
  double *tu = calloc(tsz, sizeof(double));
  double *tv = calloc(tsz, sizeof(double));
  double *work = calloc((size_t)nb * (s > 0 ? s : 1) + (size_t)nb * (nd->m > nd->n ? nd->m : nd->n), sizeof(double));
  double *r1 = calloc((size_t)s * s, sizeof(double));
  double *sig = calloc(s, sizeof(double));
  double *w = calloc((size_t)s * s, sizeof(double));
  double *zt = calloc((size_t)s * s, sizeof(double));
  if (!tu || !tv || !work || !r1 || !sig || !w || !zt) {
    *ierr = ALLOCATION_FAILURE;
    info = 0;
    goto done;
  }
  dgeqrt_(&nd->m, &s, &nb, nd->u, &nd->m, tu, &nb, work, &info);
  if (info < 0) { *ierr = QR_FAILURE; goto done; }
  dgeqrt_(&nd->n, &s, &nb, nd->v, &nd->n, tv, &nb, work, &info);
  if (info < 0) { *ierr = QR_FAILURE; goto done; }
  for (int j = 0; j < s; j++)
    for (int i = 0; i <= j; i++) r1[i + (size_t)j * s] = nd->u[i + (size_t)j * nd->m];
  {
    const double one = 1.0;
    dtrmm_("R", "U", "T", "N", &s, &s, &one, nd->v, &nd->n, r1, &s);
  }
  info = orc_svd(s, s, s, s, r1, sig, w, zt, ierr);
  if (info != 0 || *ierr != SUCCESS) goto done;
  {
    int cut;
    for (cut = 1; cut < s; cut++)
      if (sig[cut] <= thr) break;
    double *un = calloc((size_t)nd->m * cut, sizeof(double));
    double *vn = calloc((size_t)nd->n * cut, sizeof(double));
    if (!un || !vn) { free(un); free(vn); *ierr = ALLOCATION_FAILURE; goto done; }
    for (int j = 0; j < cut; j++)
      for (int i = 0; i < s; i++) un[i + (size_t)j * nd->m] = w[i + (size_t)j * s] * sig[j];
    int e = 0;
    dgemqrt_("L", "N", &nd->m, &cut, &s, &nb, nd->u, &nd->m, tu, &nb, un, &nd->m,
             work, &e);
    if (e < 0) { free(un); free(vn); *ierr = QR_FAILURE; info = e; goto done; }
    for (int j = 0; j < cut; j++)
      for (int i = 0; i < s; i++) vn[i + (size_t)j * nd->n] = zt[j + (size_t)i * s];
    dgemqrt_("L", "N", &nd->n, &cut, &s, &nb, nd->v, &nd->n, tv, &nb, vn, &nd->n,
             work, &e);
    if (e < 0) { free(un); free(vn); *ierr = QR_FAILURE; info = e; goto done; }
    free(nd->u); nd->u = un;
    free(nd->v); nd->v = vn;
    nd->s = cut;
  }
done:
  free(tu); free(tv); free(work); free(r1); free(sig); free(w); free(zt);
  return info;
}

/* recompress_large_s (src/hodlr_algebra.c:280-343): D = U V^T, dgesdd of D,
 * absolute rule, U_new = U_svd * Sigma, V_new = (V^T)^T. */
ORC_API int orc_recompress_large_s(struct NodeOffDiagonal *nd, int m_smaller,
                                   double thr, int *ierr) {
  const double one = 1.0, zero = 0.0;
  double *sig = malloc((size_t)m_smaller * sizeof(double));
  double *dd = malloc((size_t)nd->m * nd->n * sizeof(double));
  double *uu = malloc((size_t)nd->m * m_smaller * sizeof(double));
  double *vt = malloc((size_t)m_smaller * nd->n * sizeof(double));
  int info = 0;
  if (!sig || !dd || !uu || !vt) {
    *ierr = ALLOCATION_FAILURE;
    goto done;
  }
  dgemm_("N", "T", &nd->m, &nd->n, &nd->s, &one, nd->u, &nd->m, nd->v, &nd->n,
         &zero, dd, &nd->m);
  info = orc_svd(nd->m, nd->n, m_smaller, nd->m, dd, sig, uu, vt, ierr);
  if (info != 0 || *ierr != SUCCESS) goto done;
  {
    int cut;
    for (cut = 1; cut < m_smaller; cut++)
      if (sig[cut] <= thr) break;
    double *un = malloc((size_t)nd->m * cut * sizeof(double));
    double *vn = malloc((size_t)nd->n * cut * sizeof(double));
    if (!un || !vn) { free(un); free(vn); *ierr = ALLOCATION_FAILURE; goto done; }
    for (int i = 0; i < cut; i++)
      for (int j = 0; j < nd->m; j++) un[j + (size_t)i * nd->m] = uu[j + (size_t)i * nd->m] * sig[i];
    for (int i = 0; i < cut; i++)
      for (int j = 0; j < nd->n; j++) vn[j + (size_t)i * nd->n] = vt[i + (size_t)j * m_smaller];
    free(nd->u); nd->u = un;
    free(nd->v); nd->v = vn;
    nd->s = cut;
  }
done:
  free(sig); free(dd); free(uu); free(vt);
  return info;
}

/* ------------------------------------------------------------------------
 * HODLR x HODLR product, multiply_hodlr_hodlr (src/hodlr_algebra.c:1452-1553).
 * Every block of C = A B is formed from A and B alone:
 *   diagonal leaf c (compute_diagonal 1267-1352 with
 *   add_off_diagonal_contribution 1189-1222): C_cc = A_cc B_cc (beta 0),
 *   then for each level from the leaf's parent up to the root, with the A
 *   block holding c's rows (children[1] when the child containing c is a
 *   left child, else children[2]) and the B block holding c's columns (the
 *   other one): W1 = V1^T U2, W2 = W1 V2[c]^T, C_cc += U1[c] W2;
 *   off-diagonal pair of internal node P (compute_inner_off_diagonal
 *   780-869, compute_other_off_diagonal 1057-1127): columns
 *     for each ancestor of P from its parent up to the root (467-545):
 *       U: U1[rows of the block], V: V2[columns of the block] (V1^T U2)^T
 *     then (670-705, 930-957): U: A_qq U_B, V: V_B  and  U: U_A, V: B_pp^T V_A
 *   (q = the block's row child, p = its column child; A_qq B_pp are the
 *   diagonal leaves or subtrees, multiplied as multiply_internal_node_dense
 *   and its transpose do), then recompress when s < min(m, n), else
 *   recompress_large_s, with the absolute threshold (835-845, 1116-1126).
 * `out` must be an allocated tree of the same height and hold no data.
 * Unlike the reference, which leaves them 0, the internal-node sizes of
 * `out` are set.  Returns the LAPACK info of the first failing
 * recompression, else 0.
 * ---------------------------------------------------------------------- */
static void orc_copy_cols(int rows, int cols, const double *src, int lds,
                          double *dst, int ldd) {
  for (int j = 0; j < cols; j++)
    memcpy(dst + (size_t)j * ldd, src + (size_t)j * lds, (size_t)rows * sizeof(double));
}

/* rows x rows diagonal piece: leaf (height 0) or internal subtree */
static void orc_sub_product(const struct HODLRInternalNode *parent, int child,
                            int sub_height, int trans, const double *x, int r,
                            int ldx, double *y, int ldy) {
  const double one = 1.0, zero = 0.0;
  if (sub_height == 0) {
    const struct NodeDiagonal *dg = &parent->children[child].leaf->data.diagonal;
    int m = dg->m;
    dgemm_(trans ? "T" : "N", "N", &m, &r, &m, &one, dg->data, &m, x, &ldx,
           &zero, y, &ldy);
  } else {
    orc_multiply_internal_node_dense(parent->children[child].internal,
                                     sub_height, trans, x, r, ldx, y, ldy);
  }
}

/* Size of child 0 (left) of an internal node, from its top-right block
 * (set in every tree, also in product trees whose internal sizes are 0). */
static int orc_left_m(const struct HODLRInternalNode *p) {
  return p->children[1].leaf->data.off_diagonal.m;
}

static int orc_hh_pair(int height, int depth, long pos,
                       const struct HODLRInternalNode *pa,
                       const struct HODLRInternalNode *pb,
                       struct HODLRInternalNode *po, double thr, int *ierr) {
  const double one = 1.0, zero = 0.0;
  const int sub_h = height - depth - 1;
  const int m0 = orc_left_m(pa), m3 = pa->children[1].leaf->data.off_diagonal.n;
  struct NodeOffDiagonal *tr = &po->children[1].leaf->data.off_diagonal;
  struct NodeOffDiagonal *bl = &po->children[2].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *atr = &pa->children[1].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *abl = &pa->children[2].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *btr = &pb->children[1].leaf->data.off_diagonal;
  const struct NodeOffDiagonal *bbl = &pb->children[2].leaf->data.off_diagonal;
  /* ranks of the A blocks on the way to the root */
  int s_anc = 0;
  {
    const struct HODLRInternalNode *q = pa;
    long ps = pos;
    while (q->parent) {
      s_anc += q->parent->children[(ps % 2 == 0) ? 1 : 2].leaf->data.off_diagonal.s;
      q = q->parent;
      ps /= 2;
    }
  }
  tr->m = m0; tr->n = m3; tr->s = s_anc + atr->s + btr->s;
  bl->m = m3; bl->n = m0; bl->s = s_anc + abl->s + bbl->s;
  tr->u = malloc((size_t)tr->m * tr->s * sizeof(double) + 8);
  tr->v = malloc((size_t)tr->n * tr->s * sizeof(double) + 8);
  bl->u = malloc((size_t)bl->m * bl->s * sizeof(double) + 8);
  bl->v = malloc((size_t)bl->n * bl->s * sizeof(double) + 8);
  if (!tr->u || !tr->v || !bl->u || !bl->v) {
    *ierr = ALLOCATION_FAILURE;
    return 0;
  }
  int smax = 1;
  {
    const struct HODLRInternalNode *qa = pa, *qb = pb;
    long ps = pos;
    while (qa->parent) {
      const int w = (ps % 2 == 0) ? 1 : 2;
      const int s1 = qa->parent->children[w].leaf->data.off_diagonal.s;
      const int s2 = qb->parent->children[3 - w].leaf->data.off_diagonal.s;
      if (s1 * s2 > smax) smax = s1 * s2;
      qa = qa->parent; qb = qb->parent; ps /= 2;
    }
  }
  double *w12 = malloc((size_t)smax * sizeof(double));
  if (!w12) { *ierr = ALLOCATION_FAILURE; return 0; }
  /* ancestor terms; off = row offset of P within the ancestor's child */
  size_t ctr_u = 0, ctr_v = 0, cbl_u = 0, cbl_v = 0;
  {
    const struct HODLRInternalNode *qa = pa, *qb = pb;
    long ps = pos;
    int off = 0;
    while (qa->parent) {
      const int w = (ps % 2 == 0) ? 1 : 2;
      const struct NodeOffDiagonal *l1 = &qa->parent->children[w].leaf->data.off_diagonal;
      const struct NodeOffDiagonal *l2 = &qb->parent->children[3 - w].leaf->data.off_diagonal;
      int s1 = l1->s, s2 = l2->s, k1 = l1->n, ld2 = l1->m;
      dgemm_("T", "N", &s1, &s2, &k1, &one, l1->v, &k1, l2->u, &k1, &zero, w12, &s1);
      int nt = tr->n, nb = bl->n;
      dgemm_("N", "T", &nt, &s1, &s2, &one, l2->v + off + m0, &ld2, w12, &s1, &zero,
             tr->v + ctr_v, &nt);
      orc_copy_cols(tr->m, s1, l1->u + off, ld2, tr->u + ctr_u, tr->m);
      dgemm_("N", "T", &nb, &s1, &s2, &one, l2->v + off, &ld2, w12, &s1, &zero,
             bl->v + cbl_v, &nb);
      orc_copy_cols(bl->m, s1, l1->u + off + m0, ld2, bl->u + cbl_u, bl->m);
      ctr_u += (size_t)tr->m * s1; ctr_v += (size_t)tr->n * s1;
      cbl_u += (size_t)bl->m * s1; cbl_v += (size_t)bl->n * s1;
      if (ps % 2 == 1) off += orc_left_m(qa->parent);
      qa = qa->parent; qb = qb->parent; ps /= 2;
    }
  }
  free(w12);
  /* A_00 U_B(tr) | V_B(tr), then U_A(tr) | B_33^T V_A(tr) */
  orc_sub_product(pa, 0, sub_h, 0, btr->u, btr->s, btr->m, tr->u + ctr_u, tr->m);
  memcpy(tr->v + ctr_v, btr->v, (size_t)btr->n * btr->s * sizeof(double));
  ctr_u += (size_t)tr->m * btr->s; ctr_v += (size_t)tr->n * btr->s;
  memcpy(tr->u + ctr_u, atr->u, (size_t)atr->m * atr->s * sizeof(double));
  orc_sub_product(pb, 3, sub_h, 1, atr->v, atr->s, atr->n, tr->v + ctr_v, tr->n);
  /* A_33 U_B(bl) | V_B(bl), then U_A(bl) | B_00^T V_A(bl) */
  orc_sub_product(pa, 3, sub_h, 0, bbl->u, bbl->s, bbl->m, bl->u + cbl_u, bl->m);
  memcpy(bl->v + cbl_v, bbl->v, (size_t)bbl->n * bbl->s * sizeof(double));
  cbl_u += (size_t)bl->m * bbl->s; cbl_v += (size_t)bl->n * bbl->s;
  memcpy(bl->u + cbl_u, abl->u, (size_t)abl->m * abl->s * sizeof(double));
  orc_sub_product(pb, 0, sub_h, 1, abl->v, abl->s, abl->n, bl->v + cbl_v, bl->n);
  const int mn = m0 < m3 ? m0 : m3;
  int info = tr->s < mn ? orc_recompress(tr, mn, thr, ierr)
                        : orc_recompress_large_s(tr, mn, thr, ierr);
  if (*ierr != SUCCESS || info != 0) return info;
  return bl->s < mn ? orc_recompress(bl, mn, thr, ierr)
                    : orc_recompress_large_s(bl, mn, thr, ierr);
}

ORC_API int orc_multiply_hodlr_hodlr(const struct TreeHODLR *a,
                                     const struct TreeHODLR *b,
                                     struct TreeHODLR *out, double thr,
                                     int *ierr) {
  if (!a || !b || !out) {
    *ierr = INPUT_ERROR;
    return 0;
  }
  *ierr = SUCCESS;
  const int h = a->height;
  const double one = 1.0, zero = 0.0;
  struct HODLRInternalNode **la = orc_levels(a->root, h);
  struct HODLRInternalNode **lb = orc_levels(b->root, h);
  struct HODLRInternalNode **lo = orc_levels(out->root, h);
  const long total = (1L << h) - 1;
  for (long k = 0; k < total; k++)
    lo[k]->m = orc_left_m(la[k]) + la[k]->children[1].leaf->data.off_diagonal.n;
  /* diagonal leaves */
  int smax = orc_max_rank(a->root, h), smax2 = orc_max_rank(b->root, h), mmax = 0;
  for (long i = 0; i < (1L << h); i++)
    if (a->innermost_leaves[i]->data.diagonal.m > mmax)
      mmax = a->innermost_leaves[i]->data.diagonal.m;
  double *w1 = malloc((size_t)(smax > 0 ? smax : 1) * (smax2 > 0 ? smax2 : 1) * sizeof(double));
  double *w2 = malloc((size_t)(smax > 0 ? smax : 1) * (mmax > 0 ? mmax : 1) * sizeof(double));
  int info = 0;
  if (!w1 || !w2) {
    *ierr = ALLOCATION_FAILURE;
    goto done;
  }
  for (long c = 0; c < (1L << h); c++) {
    const struct HODLRLeafNode *da = a->innermost_leaves[c];
    const struct HODLRLeafNode *db = b->innermost_leaves[c];
    struct NodeDiagonal *dc = &out->innermost_leaves[c]->data.diagonal;
    int m = da->data.diagonal.m;
    dc->m = m;
    dc->data = malloc((size_t)m * m * sizeof(double) + 8);
    if (!dc->data) {
      *ierr = ALLOCATION_FAILURE;
      goto done;
    }
    dgemm_("N", "N", &m, &m, &m, &one, da->data.diagonal.data, &m,
           db->data.diagonal.data, &m, &zero, dc->data, &m);
    const struct HODLRInternalNode *qa = da->parent, *qb = db->parent;
    long ps = c;
    int off = 0;
    while (qa) {
      const int w = (ps % 2 == 0) ? 1 : 2;
      const struct NodeOffDiagonal *l1 = &qa->children[w].leaf->data.off_diagonal;
      const struct NodeOffDiagonal *l2 = &qb->children[3 - w].leaf->data.off_diagonal;
      int s1 = l1->s, s2 = l2->s, k1 = l1->n, ldu = l1->m, ldv = l2->n;
      dgemm_("T", "N", &s1, &s2, &k1, &one, l1->v, &k1, l2->u, &k1, &zero, w1, &s1);
      dgemm_("N", "T", &s1, &m, &s2, &one, w1, &s1, l2->v + off, &ldv, &zero, w2, &s1);
      dgemm_("N", "N", &m, &m, &s1, &one, l1->u + off, &ldu, w2, &s1, &one, dc->data, &m);
      if (ps % 2 == 1) off += orc_left_m(qa);
      qa = qa->parent; qb = qb->parent; ps /= 2;
    }
  }
  /* off-diagonal pairs, innermost internal nodes first (as the reference) */
  for (int d = h - 1; d >= 0 && *ierr == SUCCESS && info == 0; d--) {
    for (long i = 0; i < (1L << d); i++) {
      info = orc_hh_pair(h, d, i, ORC_LEVEL(la, d)[i], ORC_LEVEL(lb, d)[i],
                         ORC_LEVEL(lo, d)[i], thr, ierr);
      if (*ierr != SUCCESS || info != 0) break;
    }
  }
done:
  free(w1); free(w2);
  free(la); free(lb); free(lo);
  return info;
}

/* ------------------------------------------------------------------------
 * Synthetic kernel matrices for the workloads (same formulas as the device
 * generator): A(i,j) = k(|p_i - p_j|) + shift * delta_ij.
 *   kind 0: exp(-r/ell); kind 1: 1/(r + ell);
 *   kind 2: Matern-3/2 (1 + sqrt(3) r/ell) exp(-sqrt(3) r/ell).
 * Rows [r0, r0+nr) and columns [c0, c0+nc) only, into a (ld lda).
 * ---------------------------------------------------------------------- */
ORC_API void orc_fill_kernel_block(int kind, int dim, const double *pts,
                                   double ell, double shift, long r0, long nr,
                                   long c0, long nc, double *a, long lda) {
  for (long j = 0; j < nc; j++) {
    for (long i = 0; i < nr; i++) {
      const double *p = pts + (r0 + i) * dim, *q = pts + (c0 + j) * dim;
      double r, v;
      if (dim == 1) {
        r = fabs(p[0] - q[0]);
      } else {
        double r2 = 0.0;
        for (int k = 0; k < dim; k++) {
          double d = p[k] - q[k];
          r2 += d * d;
        }
        r = sqrt(r2);
      }
      if (kind == 0) {
        v = exp(-r / ell);
      } else if (kind == 1) {
        v = 1.0 / (r + ell);
      } else {
        double t = sqrt(3.0) * r / ell;
        v = (1.0 + t) * exp(-t);
      }
      if (r0 + i == c0 + j) v += shift;
      a[i + j * lda] = v;
    }
  }
}
