/*
 * hm_hip: the thin C-ABI HIP layer under libhmat_lib (gfx950 / MI355X).
 *
 * Every call takes plain pointers and sizes; device pointers are marked d_.
 * Streams are hipStream_t passed as void* (NULL = the library's own
 * stream).  Every int-returning call returns an enum ErrorCode value
 * (0 = SUCCESS); hm_hip_last_error() gives the message of the last
 * failure on the calling thread.  There is no host fallback: without a
 * usable GPU every compute call fails with ALLOCATION_FAILURE and a message.
 *
 * In the reference all of this arithmetic is BLAS/LAPACK
 * (include/internal/blas_wrapper.h:1-23, include/internal/lapack_wrapper.h:
 * 1-14); these entry points are what replaces that boundary.
 *
 * A device tree (hm_dtree) stores a HODLR matrix level-major and packed:
 *   - all 2^h diagonal leaves back to back, each m_i x m_i column-major;
 *   - per level l = 1..h, blocks j = 0..2^l-1 (j even: top-right child [1]
 *     of parent j/2, j odd: bottom-left child [2]); all U panels of the
 *     tree in one array and all V panels in another, each m x s / n x s
 *     column-major at a per-block offset;
 *   - int metadata (leaf offsets, block dims, ranks, offsets) on device.
 * One tree must not be used by two concurrent calls (its product scratch
 * is owned by the tree): that includes asynchronous calls on two different
 * streams from one thread, which overlap on the device and race on the
 * scratch.  Issue every product of one tree on one stream, or order the
 * streams yourself (an event between them).
 */
#ifndef HM_HIP_H
#define HM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HM_HIP_API
#define HM_HIP_API __attribute__((visibility("default")))
#endif

struct TreeHODLR;
struct HODLRInternalNode;
typedef struct hm_dtree hm_dtree;

/* ---- runtime ---------------------------------------------------------- */
HM_HIP_API int hm_hip_init(int device);
HM_HIP_API int hm_hip_device_count(int *count);
HM_HIP_API const char *hm_hip_last_error(void);
HM_HIP_API int hm_hip_synchronize(void *stream);

/* Free the device and pinned buffers the library keeps between calls so
 * that later calls need no allocation: the compression rounds' work buffer
 * (up to 8 GiB, sized by the largest compression so far) and the launch
 * programs' scratch, blob and pinned staging (up to 1 GiB + 256 MiB).  The
 * next call that needs them allocates them again.  Product scratch belongs
 * to each tree and goes with hm_hip_tree_destroy.  INPUT_ERROR (buffers kept)
 * while another thread's call is using one of them. */
HM_HIP_API int hm_hip_release_buffers(void);

/* Raw device memory helpers (for callers that bring no allocator). */
HM_HIP_API int hm_hip_malloc(void **d_ptr, size_t bytes);
HM_HIP_API int hm_hip_free(void *d_ptr);
HM_HIP_API int hm_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes, void *stream);
HM_HIP_API int hm_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes, void *stream);
HM_HIP_API int hm_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
/* 2-D strided copies of an m x n column-major block (leading dims in
 * elements of 8 bytes). */
HM_HIP_API int hm_hip_copy2d_h2d(double *d_dst, long ld_dst, const double *src, long ld_src,
                      int m, int n, void *stream);
HM_HIP_API int hm_hip_copy2d_d2h(double *dst, long ld_dst, const double *d_src, long ld_src,
                      int m, int n, void *stream);

/* ---- device trees ----------------------------------------------------- */
/* Build a device copy of a host (sub)tree.  `node` with `height` layers of
 * internal nodes below and including it; for a whole tree pass
 * tree->root and tree->height. */
HM_HIP_API int hm_hip_tree_upload(const struct HODLRInternalNode *node, int height,
                       hm_dtree **out, void *stream);
/* Copy a device tree into a host tree of the same height: sets every m, n,
 * s, and allocates (with `alloc`, or malloc when NULL) and fills every
 * diagonal `data`, `u` and `v` array. */
HM_HIP_API int hm_hip_tree_download(const hm_dtree *tree, struct TreeHODLR *host,
                         void *(*alloc)(size_t), void *stream);
HM_HIP_API int hm_hip_tree_destroy(hm_dtree *tree);

typedef struct hm_tree_info {
  int height;        /* layers of internal nodes */
  int n;             /* matrix order */
  int nleaf;         /* 2^height */
  int nblock;        /* off-diagonal blocks, 2^(height+1) - 2 */
  int max_rank;
  long long sum_rank_dims; /* sum over blocks of (m + n) * s */
  long long diag_elems;    /* sum over leaves of m_i^2 */
} hm_tree_info;
HM_HIP_API int hm_hip_tree_info(const hm_dtree *tree, hm_tree_info *info);
/* ranks[g] for block g in level-major order (level 1 first). */
HM_HIP_API int hm_hip_tree_ranks(const hm_dtree *tree, int *ranks);

/* ---- products ----------------------------------------------------------
 * y = H x (FP64).  Replaces the dgemv loops of src/vector_algebra.c:52-197.
 */
HM_HIP_API int hm_hip_matvec(hm_dtree *tree, const double *d_x, double *d_y, void *stream);

/* Y = op(H) X with op = identity (trans = 0) or transpose (trans = 1).
 * X is N x R with X(i, j) at d_X[i * x_rs + j * x_cs]; Y likewise.
 * X * H is Y^T = H^T X^T (trans = 1 with row/column strides swapped).
 * Replaces the dgemm loops of src/dense_algebra.c:82-1151. */
HM_HIP_API int hm_hip_matmat(hm_dtree *tree, int trans, int R, const double *d_X,
                  long x_rs, long x_cs, double *d_Y, long y_rs, long y_cs,
                  void *stream);

/* hm_hip_matmat without any allocation: the same product (same kernels,
 * same operands and result) with its partial products and W in a fixed
 * 64 MiB device array of the library (one per device, held until the
 * process ends) instead of the tree's scratch, which hm_hip_matmat grows
 * with hipMalloc.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Call hm_hip_init first.  It makes
 * no hipMalloc / hipFree or other allocating call: a tree whose apply
 * variant for this transpose mode and RHS width is not tuned yet runs
 * variant 1 (no timing runs); trees with sliced factors
 * (hm_hip_tree_set_precision / hm_hip_tree_set_format) give INPUT_ERROR, and so does a
 * product whose scratch (partials + W, about R * (sum over levels and
 * 64-row chunks of the input-side ranks + sum of ranks) doubles) exceeds
 * 8 Mi doubles.  Every call on one device shares the array and it keeps
 * the last call's contents (each product writes its scratch before reading
 * it): calls that overlap on one device, from any thread or stream, race. */
HM_HIP_API int hm_hip_matmat_noalloc(hm_dtree *tree, int trans, int R, const double *d_X,
                                     long x_rs, long x_cs, double *d_Y, long y_rs,
                                     long y_cs, void *stream);

/* Precision of the off-diagonal products of this tree: 0 = FP64 (default);
 * 1..4 = the factors split into that many BF16 slices (integers with a
 * shared power-of-two scale per row of 128-entry chunks), products of
 * slice pairs summed exactly in FP32 on MFMA and folded in FP64 (pairs
 * beyond the slice count dropped).  Accuracy relative to the chunk maxima
 * is ~2^-8 per slice: 2^-8 (1), 2^-16 (2), 2^-24 (3), 2^-32 (4); the
 * diagonal leaves stay FP64.  Applies to every product of the tree: H X and
 * H^T X with any operand strides (so X H as well), H x, and the tree's
 * share of a batch (hm_hip_batch_create).  Builds the sliced copies of U
 * and V on the device (those for H^T X on the first transposed product).
 * A batch created with the tree refuses to run after its precision
 * changes. */
HM_HIP_API int hm_hip_tree_set_precision(hm_dtree *tree, int slices, void *stream);

/* Formats of the off-diagonal factor copies for hm_hip_tree_set_format. */
#define HM_FMT_FP64 0 /* no copies: FP64 products (slices ignored) */
#define HM_FMT_BF16 1 /* 1..4 BF16 slices, as hm_hip_tree_set_precision */
#define HM_FMT_INT8 2 /* 1..4 INT8 slices: balanced base-254 digits per row of
                       * 256-entry chunks (|q| <= 127), products on
                       * v_mfma_i32_16x16x64_i8 summed exactly in int32 and
                       * folded in FP64; about the same bits per slice as BF16
                       * (2^-8 .. 2^-32 of the chunk maxima) with half the
                       * operand bytes and twice the MFMA rate */
#define HM_FMT_FP32 3 /* one FP32 copy (slices = 1), products on
                       * v_mfma_f32_16x16x4_f32 accumulated in FP32 over
                       * 32-entry chunks and folded in FP64: relative error
                       * near 1e-7 (FP32 rounding of operands and sums) */
/* Precision of the off-diagonal products of this tree, as
 * hm_hip_tree_set_precision but with the format chosen: the copies are
 * built on the device (those for H^T X on the first transposed product),
 * the diagonal leaves stay FP64, every product form and batches use them,
 * and a batch created with the tree refuses to run after they change.
 * INPUT_ERROR for an unknown format or a slice count outside 1..4 (1 for
 * FP32). */
HM_HIP_API int hm_hip_tree_set_format(hm_dtree *tree, int format, int slices, void *stream);

/* Apply-kernel variant the autotuner picked for this tree, transpose mode
 * and RHS width (-1 before the first product of that kind): 0-4 tile shapes
 * of the general kernel, 5 and 6 the leaf-resident kernels for uniform
 * leaves of 128/256/512 rows (H X, X column-major, R a multiple of 64).
 * HM_APPLY_VARIANT=v in the environment forces variant v (a leaf-resident
 * variant on a tree it does not accept runs variant 0); HM_TUNE=0 disables
 * tuning. */
HM_HIP_API int hm_hip_matmat_variant(const hm_dtree *tree, int trans, int R);

/* ---- compression -------------------------------------------------------
 * Dense -> HODLR on the device.  d_A is m x m column-major with leading
 * dimension lda and is not modified.  `ms` (host, 2^height ints) gives
 * custom leaf sizes or is NULL for halving (top-left gets ceil).
 * Rank rule of the reference (src/constructors.c:329-336): keep sigma_i
 * while sigma_i >= thr * sigma_0; rank 1 when sigma_0 <= thr.
 * Replaces src/constructors.c:212-546 and src/lapack_wrapper.c:8-59. */
typedef struct hm_mlp hm_mlp;
typedef struct hm_compress_opts {
  int sketch_init;      /* initial sketch width (columns), default 16 */
  int sketch_max;       /* largest sketch width before giving up, default 512 */
  int n_test;           /* posterior test vectors per block, default 6 */
  unsigned long long seed;
  const int *rank_hint; /* per block (level-major) predicted rank, or NULL */
  int oversample;       /* added to rank_hint / predicted rank, default 8 */
  double accept;        /* residual acceptance factor vs thr*sigma0, default 0.1 */
  /* Learned width: blocks that fail the first (sketch_init) round get the
   * width the model predicts from block_features (host, nblock x n_extra,
   * row-major, level-major blocks) and the first round's singular values
   * (hm_hip_rank_features): with y the model output and k = 2^y - 1 the
   * predicted count of singular values at or above accept/20 * thr * sigma_0
   * (the width the residual test needs), ceil(k (1 + margin)) + oversample
   * columns instead of doubling; the residual test still decides, and a
   * failed prediction goes on doubling. */
  const hm_mlp *rank_model;
  const float *block_features;
  int n_extra;
  double margin;        /* default 0.1 */
  double *first_sigma;  /* host out (nblock x 16): first round's singular values, or NULL */
  int *need_out;        /* host out (nblock): singular values of the final core at or above
                           accept/20 * thr * sigma_0 (the training label), or NULL */
} hm_compress_opts;

typedef struct hm_compress_stats {
  int sketch_rounds;    /* total sketch passes summed over levels */
  int blocks_retried;   /* block-rounds that failed the residual test */
  int max_sketch;       /* widest sketch used */
  int fallback_blocks;  /* blocks that hit sketch_max (then exact path) */
  double ms_total;
  int predicted_blocks; /* blocks whose second sketch width the rank model chose */
  int predicted_failed; /* ... whose residual test failed at that width (went on doubling) */
  int continued_blocks; /* block-rounds that continued the failed round's sketch (kept its
                           first columns' basis and sketched only the new columns) */
} hm_compress_stats;

HM_HIP_API void hm_hip_compress_opts_default(hm_compress_opts *opts);
HM_HIP_API int hm_hip_compress_dense(const double *d_A, long lda, int m, int height,
                          const int *ms, double thr,
                          const hm_compress_opts *opts, hm_dtree **out,
                          hm_compress_stats *stats, void *stream);

/* ---- synthetic kernel matrices (device generator) -----------------------
 * A(i, j) = k(|p_i - p_j|) + shift * delta_ij for points d_pts (n x dim,
 * row-major).  kind 0: exp(-r / ell); kind 1: 1 / (r + ell);
 * kind 2: Matern-3/2 (1 + sqrt(3) r / ell) exp(-sqrt(3) r / ell). */
HM_HIP_API int hm_hip_fill_kernel_matrix(int kind, int n, int dim, const double *d_pts,
                              double ell, double shift, double *d_A, long lda,
                              void *stream);

/* ---- recompression ------------------------------------------------------
 * Batched recompression of low-rank nodes, all on the device.  Node i has
 * U_i (m_i x s_i, ld m_i) at d_u[i] and V_i (n_i x s_i, ld n_i) at d_v[i]
 * (arrays of device pointers, host-resident).  On return the first k_i
 * columns of U_i / V_i hold the recompressed factors (U carries sigma),
 * k_out[i] = k_i (host).  Absolute rule: keep sigma_j > thr, at least 1.
 * method 0: QR method (needs s_i < min(m_i, n_i)), 1: dense U V^T SVD,
 * -1: choose per node as the reference's callers do.
 * Replaces src/hodlr_algebra.c:97-343. */
HM_HIP_API int hm_hip_recompress_batch(int nnodes, const int *m, const int *n,
                            const int *s, double *const *d_u,
                            double *const *d_v, double thr, int method,
                            int *k_out, void *stream);

/* Thin QR of device panels: Y_i (m_i x n_i, 1 <= n_i <= m_i, leading
 * dimension ldy[i]) = Q_i R_i with Q_i (m_i x n_i, ld ldq[i]) orthonormal
 * and R_i (n_i x n_i, ld ldr[i], upper triangular; R or R[i] may be NULL)
 * — the batched Householder TSQR with block Gram-Schmidt across 64-column
 * panels that compression and recompression use (LAPACK dlarfg signs: the
 * diagonal of R is -sign of the pivot).  Y_i is overwritten.  Pointer
 * arrays are host-resident, the panels device-resident.  Plays the part of
 * the dgeqrt_ / dgemqrt_ pair at src/hodlr_algebra.c:120-128, 177. */
HM_HIP_API int hm_hip_qr_batch(int njobs, const int *m, const int *n, double *const *Y,
                               const long *ldy, double *const *Q, const long *ldq,
                               double *const *R, const long *ldr, void *stream);

/* ---- binary files -------------------------------------------------------
 * The layout of hmat_amd_save_tree (include/hmat_lib_amd.h), written from /
 * read into the device pools directly, so a file saved from a host tree
 * loads as a device tree and vice versa.  Load rejects a malformed or
 * truncated file with INPUT_ERROR.  No counterpart in the reference. */
HM_HIP_API int hm_hip_tree_save(const hm_dtree *tree, const char *path, void *stream);
HM_HIP_API int hm_hip_tree_load(const char *path, hm_dtree **out, void *stream);

/* ---- HODLR x HODLR ------------------------------------------------------
 * *out = a b for two device trees of the same height and leaf sizes (ranks
 * may differ), all on the device.  Every block of the product is formed
 * from a and b (diagonal leaves: a_cc b_cc plus the low-rank terms of every
 * level; off-diagonal blocks: the ancestor terms, a_qq U_b and U_a,
 * b_pp^T V_a, in the reference's column order) and each off-diagonal block
 * is recompressed as hm_hip_recompress_batch does with method -1
 * (absolute rule: keep sigma_j > thr, at least 1).  Fails with INPUT_ERROR
 * when a block's uncompressed rank exceeds the recompression's limit
 * (1024).  Replaces src/hodlr_algebra.c:1452-1553 (multiply_hodlr_hodlr). */
HM_HIP_API int hm_hip_multiply_hodlr_hodlr(const hm_dtree *a, const hm_dtree *b, double thr,
                                           hm_dtree **out, void *stream);

/* ---- learned rank predictor --------------------------------------------
 * MLP nin -> nhidden -> nhidden -> 1 (tanh, FP32), Glorot-initialised from
 * seed; trained and run by the library's own kernels.  Inputs are row-major
 * (B x nin) FP32 device arrays; the target is log2(rank + 1). */
HM_HIP_API int hm_hip_mlp_create(int nin, int nhidden, unsigned long long seed, hm_mlp **out);
HM_HIP_API int hm_hip_mlp_destroy(hm_mlp *m);
HM_HIP_API long long hm_hip_mlp_nparams(const hm_mlp *m);
/* Flat parameters: W1 (nin x H), b1, W2 (H x H), b2, w3 (H), b3. */
HM_HIP_API int hm_hip_mlp_get_params(const hm_mlp *m, float *host);
HM_HIP_API int hm_hip_mlp_set_params(hm_mlp *m, const float *host);
HM_HIP_API int hm_hip_mlp_forward(hm_mlp *m, const float *d_X, int B, float *d_y, void *stream);
/* Full-batch mean-squared-error gradient (host out, nparams) and loss. */
HM_HIP_API int hm_hip_mlp_gradient(hm_mlp *m, const float *d_X, const float *d_t, int B,
                                   float *h_grad, float *h_loss, void *stream);
/* Mini-batch Adam; loss_out (host, epochs) = mean mini-batch loss per epoch. */
HM_HIP_API int hm_hip_mlp_train(hm_mlp *m, const float *d_X, const float *d_t, int B,
                                int epochs, int batch, float lr, unsigned long long seed,
                                float *loss_out, void *stream);
/* Model inputs of nblock blocks (host): the n_extra caller features, then
 * 17 from the first round's 16 singular values: log10(sigma_i / sigma_0) / 10
 * for i = 1..15 (floored at 1e-20), the fraction of them >= thr, and
 * log10(thr) / 10.  out is nblock x (n_extra + 17), row-major. */
HM_HIP_API int hm_hip_rank_features(int nblock, int n_extra, const float *extra,
                                    const double *first_sigma, double thr, float *out);

/* ---- batches of independent trees --------------------------------------
 * Compress nmat dense matrices (device, column-major, not modified) in one
 * batched truncated SVD over all their off-diagonal blocks; leaf sizes by
 * halving, rank rule as hm_hip_compress_dense.  out[i] receives tree i;
 * stats (optional) covers the whole batch.  The trees of one call share
 * their device pools (factors, diagonal blocks, metadata): each is freed
 * with hm_hip_tree_destroy as usual, and the pools go with the last of them.
 * Replaces nmat calls of dense_to_tree_hodlr (src/constructors.c:629-704). */
HM_HIP_API int hm_hip_compress_batch(int nmat, const double *const *d_A, const long *lda,
                                     const int *m, const int *height, double thr,
                                     const hm_compress_opts *opts, hm_dtree **out,
                                     hm_compress_stats *stats, void *stream);

/* A batched product over ntrees trees, fixed at creation: trans 0:
 * Y_i = H_i X_i, 1: Y_i = H_i^T X_i (X_i, Y_i n_i x R[i]); 2: Y_i = X_i H_i
 * (X_i, Y_i R[i] x n_i); device arrays, column-major, leading dimensions
 * ldx[i], ldy[i].  One launch program whatever the number of trees or
 * levels: low-rank trees share the launches of the single-tree product
 * kernels (partials, two reductions, one apply per RHS width group),
 * high-rank trees two batched-GEMM launches, and trees with sliced factors
 * (hm_hip_tree_set_precision) one launch per sliced stage for all of them.
 * The batch keeps pointers to the trees and operands (it owns neither).
 * Replaces per-tree loops of
 * multiply_hodlr_dense / multiply_hodlr_transpose_dense /
 * multiply_dense_hodlr (src/dense_algebra.c:361-648, 957-1026). */
typedef struct hm_batch hm_batch;
HM_HIP_API int hm_hip_batch_create(int ntrees, hm_dtree *const *trees, int trans, const int *R,
                                   const double *const *d_X, const long *ldx,
                                   double *const *d_Y, const long *ldy, hm_batch **out);
/* Time every tile variant of every launch on the live operands and keep the
 * fastest.  Leaves the Y_i undefined until the next run. */
HM_HIP_API int hm_hip_batch_tune(hm_batch *b, void *stream);
/* Run the product; use_graph = 1 captures it into a HIP graph on first use
 * and replays the graph afterwards. */
HM_HIP_API int hm_hip_batch_run(hm_batch *b, int use_graph, void *stream);
/* Tile variant chosen per launch (up to cap); returns the launch count. */
HM_HIP_API int hm_hip_batch_variants(const hm_batch *b, int *out, int cap);
HM_HIP_API int hm_hip_batch_destroy(hm_batch *b);

/* ---- diagnostics -------------------------------------------------------
 * One v_mfma_f64_16x16x4_f64 with A(i,k) = 1 + i + 16 k and
 * B(k,j) = 1 + 2 j - k, read back through the lane layout the kernels
 * assume; out (host, 256 doubles) receives D row-major. */
HM_HIP_API int hm_hip_selftest_mfma(double *out);
/* One v_mfma_f32_16x16x32_bf16 with lane L taking A[L%16][8(L/16)+e] and
 * B[8(L/16)+e][L%16] (hA 16x32, hB 32x16 row-major host floats, exact in
 * bf16); raw (host, 256) receives value v of lane L at 4 L + v. */
HM_HIP_API int hm_hip_selftest_mfma_bf16(const float *hA, const float *hB, float *raw);
/* One 16 x 16 block of the sliced GEMM of a format (HM_FMT_BF16, _INT8,
 * _FP32) from one 16-byte fragment per lane, as the kernels issue it: lane L
 * takes A[L%16][e0 + e] and B[e0 + e][L%16] for its fragment's entries e,
 * e0 = (entries per fragment) * (L/16), i.e. K = 32 (BF16: one
 * v_mfma_f32_16x16x32_bf16), 64 (INT8: one v_mfma_i32_16x16x64_i8) or 16
 * (FP32: four v_mfma_f32_16x16x4_f32, MFMA j taking entry j of every lane).
 * hA (16 x K) and hB (K x 16) are row-major host floats holding small
 * integers; raw (host, 256) receives value v of lane L at 4 L + v. */
HM_HIP_API int hm_hip_selftest_mfma_lowp(int format, const float *hA, const float *hB,
                                         float *raw);

/* Host-only check (no device needed) of the wide-panel QR launch program:
 * builds it for njobs panels of m[i] x L[i] (with the triangular factor
 * when with_r) on fake buffers and verifies that every work item's reads
 * and writes stay inside the buffers and scratch it was given.  msg
 * receives a summary or the first violation.  Returns 0 when clean. */
HM_HIP_API int hm_hip_audit_qr(int njobs, const int *m, const int *L, int with_r,
                               char *msg, int msglen);

/* Host-only check of the HODLR x HODLR launch program: built for two trees
 * of the given leaf sizes (2^height) and block ranks (level-major, block
 * (l, j) at 2^l - 2 + j) on fake buffers, every range its work items read
 * or write is checked against the operands and its scratch.  Returns 0
 * when clean; msg receives a summary or the first violation. */
HM_HIP_API int hm_hip_audit_hh(int height, const int *leaves, const int *ranks_a,
                               const int *ranks_b, char *msg, int msglen);

/* ---- kernel timing (HIP events around tracked kernels) ----------------- */
HM_HIP_API int hm_hip_kernel_timing(int enable);
/* Time only the kernel of this name (NULL: every named kernel). */
HM_HIP_API int hm_hip_kernel_timing_select(const char *name);
HM_HIP_API int hm_hip_kernel_timing_read(const char *name, double *total_ms, long *count);
HM_HIP_API int hm_hip_kernel_timing_reset(void);
/* All accumulated entries as "name\ttotal_ms\tcount\n" lines (truncated to
 * len); returns the full length. */
HM_HIP_API int hm_hip_kernel_timing_dump(char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif
