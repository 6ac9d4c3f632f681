// Dense -> HODLR compression on the device, and the synthetic kernel-matrix
// generator used for device-resident inputs.
//
// Reference: dense_to_tree_hodlr (src/constructors.c:629-704) sizes the
// blocks (44-95 halves, 133-176 custom), copies the diagonal leaves
// (212-243) and runs one dgesdd per off-diagonal block with the relative
// rank rule (308-367), level by level with OpenMP tasks (449-546).  Here
// every off-diagonal block of every level is one entry of a single batched
// truncated SVD (hm_tsvd.hip), and the diagonal leaves are gathered by one
// kernel; 64-bit offsets throughout (no N >= 32768 limit).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "hm_tsvd.h"

namespace hm {

__global__ __launch_bounds__(256) void k_fill_kernel(int kind, int n, int dim,
                                                     const double *pts,
                                                     double ell, double shift,
                                                     double *A, long long lda) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double s3 = 1.7320508075688772;
  for (int j = blockIdx.y; j < n; j += gridDim.y) {
    double r;
    if (dim == 1) {
      r = fabs(pts[i] - pts[j]);
    } else {
      double r2 = 0.0;
      for (int k = 0; k < dim; k++) {
        const double d = pts[(long long)i * dim + k] - pts[(long long)j * dim + k];
        r2 += d * d;
      }
      r = sqrt(r2);
    }
    double v;
    if (kind == 0) {
      v = exp(-r / ell);
    } else if (kind == 1) {
      v = 1.0 / (r + ell);
    } else {
      const double t = s3 * r / ell;
      v = (1.0 + t) * exp(-t);
    }
    if (i == j) v += shift;
    A[i + (long long)j * lda] = v;
  }
}

__global__ __launch_bounds__(256) void k_extract_diag(const double *A,
                                                      long long lda,
                                                      const int *leaf_off,
                                                      const long long *diag_off,
                                                      double *D) {
  const int c = blockIdx.x;
  const int lo = leaf_off[c], mc = leaf_off[c + 1] - lo;
  const long long tot = (long long)mc * mc;
  double *Dc = D + diag_off[c];
  for (long long e = threadIdx.x + (long long)blockIdx.y * 256; e < tot;
       e += 256ll * gridDim.y) {
    const long long i = e % mc, j = e / mc;
    Dc[e] = A[(lo + i) + (lo + j) * lda];
  }
}

// Leaf sizes by halving (top-left child gets the larger half).
static void halves(int m, int depth, int h, std::vector<int> &out) {
  if (depth == h) {
    out.push_back(m);
    return;
  }
  const int small = m / 2, large = m - small;
  halves(large, depth + 1, h, out);
  halves(small, depth + 1, h, out);
}

}  // namespace hm

using namespace hm;

extern "C" {

void hm_hip_compress_opts_default(hm_compress_opts *o) {
  o->sketch_init = 16;
  o->sketch_max = 512;
  o->n_test = 6;
  o->seed = 0x243f6a8885a308d3ull;
  o->rank_hint = nullptr;
  o->oversample = 8;
  o->accept = 0.1;
  o->rank_model = nullptr;
  o->block_features = nullptr;
  o->n_extra = 0;
  o->margin = 0.1;
  o->first_sigma = nullptr;
  o->need_out = nullptr;
}

int hm_hip_rank_features(int nblock, int n_extra, const float *extra, const double *sig,
                         double thr, float *out) {
  if (nblock < 0 || n_extra < 0 || (nblock > 0 && (!sig || !out || (n_extra > 0 && !extra))))
    return INPUT_ERROR;
  const int F = n_extra + 17;
  for (int g = 0; g < nblock; g++) {
    float *o = out + (size_t)g * F;
    for (int e = 0; e < n_extra; e++) o[e] = extra[(size_t)g * n_extra + e];
    const double *sv = sig + (size_t)g * FIRST_SIGMA;
    const double s0 = sv[0];
    int above = 0;
    for (int i = 1; i < FIRST_SIGMA; i++) {
      const double r = s0 > 0.0 ? sv[i] / s0 : 0.0;
      o[n_extra + i - 1] = (float)(std::log10(std::max(r, 1e-20)) / 10.0);
      above += (s0 > 0.0 && r >= thr) ? 1 : 0;
    }
    o[n_extra + 15] = (float)above / (float)FIRST_SIGMA;
    o[n_extra + 16] = (float)(std::log10(thr) / 10.0);
  }
  return SUCCESS;
}

int hm_hip_fill_kernel_matrix(int kind, int n, int dim, const double *d_pts,
                              double ell, double shift, double *d_A, long lda,
                              void *stream) {
  HM_OK(ensure_init());
  if (n <= 0) return SUCCESS;
  if (!d_pts || !d_A || dim < 1 || lda < n || kind < 0 || kind > 2 || !(ell > 0)) {
    set_error("hm_hip_fill_kernel_matrix: bad arguments");
    return INPUT_ERROR;
  }
  dim3 grid((n + 255) / 256, std::min(n, 8192));
  hipLaunchKernelGGL(k_fill_kernel, grid, dim3(256), 0, stream_of(stream), kind,
                     n, dim, d_pts, ell, shift, d_A, (long long)lda);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

}  // extern "C"

namespace hm {

// Tree skeleton (leaf offsets, block geometry) of an m x m matrix and one
// truncated-SVD entry per off-diagonal block (uid_base distinguishes the
// random sketches of different matrices).
static int plan_tree(const double *d_A, long lda, int m, int height, const int *ms,
                     double thr, const hm_compress_opts &opts, unsigned long long uid_base,
                     hm_dtree **out, std::vector<TsvdBlock> &tb) {
  *out = nullptr;
  const int nleaf = 1 << height;
  std::vector<int> leaves;
  if (ms) {
    long long sum = 0;
    for (int c = 0; c < nleaf; c++) {
      if (ms[c] < 0) {
        set_error("compress: negative leaf size");
        return INPUT_ERROR;
      }
      leaves.push_back(ms[c]);
      sum += ms[c];
    }
    if (sum != m) {
      set_error("compress: leaf sizes sum to %lld, not %d", sum, m);
      return INPUT_ERROR;
    }
  } else {
    halves(m, 0, height, leaves);
  }
  hm_dtree *t = new hm_dtree();
  t->height = height;
  t->nleaf = nleaf;
  t->nblock = (1 << (height + 1)) - 2;
  t->leaf_off.assign(nleaf + 1, 0);
  for (int c = 0; c < nleaf; c++) t->leaf_off[c + 1] = t->leaf_off[c] + leaves[c];
  t->blk.resize(t->nblock);
  tb.assign(t->nblock, TsvdBlock());
  for (int l = 1; l <= height; l++) {
    const int span = 1 << (height - l);
    for (int j = 0; j < (1 << l); j++) {
      const int g = blk_index(l, j);
      BlockMeta &b = t->blk[g];
      b.r0 = t->leaf_off[j * span];
      b.m = t->leaf_off[(j + 1) * span] - b.r0;
      b.c0 = t->leaf_off[(j ^ 1) * span];
      b.n = t->leaf_off[((j ^ 1) + 1) * span] - b.c0;
      b.level = l;
      b.j = j;
      b.s = 0;
      b.pad = 0;
      TsvdBlock &x = tb[g];
      memset(&x, 0, sizeof(x));
      x.A = d_A + b.r0 + (long long)b.c0 * lda;
      x.lda = lda;
      x.m = b.m;
      x.n = b.n;
      int ell = opts.sketch_init;
      if (opts.rank_hint) ell = std::max(1, opts.rank_hint[g] + opts.oversample);
      x.ell = ell;
      x.rule = RULE_RELATIVE;
      x.thr = thr;
      x.uid = uid_base + (unsigned long long)g;
    }
  }
  *out = t;
  return SUCCESS;
}

// Metadata upload, diagonal extraction (t->d_U, d_V already set).
static int finish_tree(hm_dtree *t, const double *d_A, long lda, hipStream_t s) {
  HM_OK(tree_finalize_meta(t, s));
  HM_OK(dalloc(&t->d_diag, (size_t)t->diag_off[t->nleaf]));
  hipLaunchKernelGGL(k_extract_diag, dim3(t->nleaf, 64), dim3(256), 0, s, d_A,
                     (long long)lda, t->d_leaf_off, t->d_diag_off, t->d_diag);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

static void fill_stats(hm_compress_stats *stats, const TsvdStats &ts, double ms) {
  if (!stats) return;
  stats->sketch_rounds = ts.rounds;
  stats->blocks_retried = ts.retried;
  stats->max_sketch = ts.max_sketch;
  stats->fallback_blocks = ts.capped;
  stats->ms_total = ms;
  stats->predicted_blocks = ts.widened;
  stats->predicted_failed = ts.widened_failed;
  stats->continued_blocks = ts.continued;
}

static TsvdOptions tsvd_options(const hm_compress_opts &opts) {
  TsvdOptions to;
  to.sketch_max = opts.sketch_max;
  to.n_test = opts.n_test;
  to.accept = opts.accept;
  to.seed = opts.seed;
  return to;
}

// Width of the second round from the learned model: the blocks in `retry`
// get ceil(2^y - 1) (1 + margin) + oversample with y the model output.
static int learned_widen(const hm_compress_opts &opts, double thr, std::vector<TsvdBlock> &blocks,
                         const std::vector<int> &retry, const std::vector<double> &sig,
                         hipStream_t s) {
  const int F = opts.n_extra + 17;
  const int nr = (int)retry.size();
  std::vector<float> ex((size_t)nr * std::max(opts.n_extra, 1), 0.0f), feats((size_t)nr * F);
  std::vector<double> sg((size_t)nr * FIRST_SIGMA);
  for (int r = 0; r < nr; r++) {
    const int g = retry[r];
    for (int e = 0; e < opts.n_extra; e++)
      ex[(size_t)r * opts.n_extra + e] = opts.block_features[(size_t)g * opts.n_extra + e];
    for (int j = 0; j < FIRST_SIGMA; j++)
      sg[(size_t)r * FIRST_SIGMA + j] = sig[(size_t)g * FIRST_SIGMA + j];
  }
  HM_OK(hm_hip_rank_features(nr, opts.n_extra, ex.data(), sg.data(), thr, feats.data()));
  float *d_x = nullptr, *d_y = nullptr;
  HM_OK(dalloc(&d_x, feats.size()));
  int rc = dalloc(&d_y, (size_t)nr);
  std::vector<float> y(nr, 0.0f);
  if (rc == SUCCESS) {
    hipError_t e = hipMemcpyAsync(d_x, feats.data(), feats.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      rc = hm_hip_mlp_forward(const_cast<hm_mlp *>(opts.rank_model), d_x, nr, d_y, s);
      if (rc == SUCCESS) {
        e = hipMemcpyAsync(y.data(), d_y, (size_t)nr * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
      }
    }
    if (rc == SUCCESS && e != hipSuccess) {
      set_error("learned width: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  dfree(d_x);
  dfree(d_y);
  if (rc != SUCCESS) return rc;
  for (int r = 0; r < nr; r++) {
    const double yy = std::min(std::max((double)y[r], 0.0), 12.0);
    const double k = std::exp2(yy) - 1.0;
    blocks[retry[r]].ell = (int)std::ceil(k * (1.0 + opts.margin)) + opts.oversample;
  }
  return SUCCESS;
}

}  // namespace hm

extern "C" {

int hm_hip_compress_dense(const double *d_A, long lda, int m, int height,
                          const int *ms, double thr,
                          const hm_compress_opts *opts_in, hm_dtree **out,
                          hm_compress_stats *stats, void *stream) {
  *out = nullptr;
  HM_OK(ensure_init());
  if (!d_A || height < 1 || height > 24 || m < 0 || lda < std::max(1, m)) {
    set_error("hm_hip_compress_dense: bad arguments (m=%d height=%d lda=%ld)",
              m, height, lda);
    return INPUT_ERROR;
  }
  hm_compress_opts opts;
  hm_hip_compress_opts_default(&opts);
  if (opts_in) opts = *opts_in;
  hipStream_t s = stream_of(stream);
  auto t0 = std::chrono::steady_clock::now();
  hm_dtree *t = nullptr;
  std::vector<TsvdBlock> tb;
  HM_OK(plan_tree(d_A, lda, m, height, ms, thr, opts, 0, &t, tb));
  if (opts.rank_model && opts.n_extra > 0 && !opts.block_features) {
    tree_free(t);
    set_error("hm_hip_compress_dense: rank_model needs block_features");
    return INPUT_ERROR;
  }
  TsvdOptions to = tsvd_options(opts);
  std::vector<double> first;
  if (opts.first_sigma) to.first_sigma = &first;
  if (opts.rank_model)
    to.widen = [&](std::vector<TsvdBlock> &bl, const std::vector<int> &retry,
                   const std::vector<double> &sig) {
      return learned_widen(opts, thr, bl, retry, sig, s);
    };
  TsvdStats ts;
  double *up = nullptr, *vp = nullptr;
  int rc = tsvd_run(tb, to, &up, &vp, &ts, nullptr, 0, s);
  if (rc == SUCCESS && opts.need_out)
    for (int g = 0; g < t->nblock; g++) opts.need_out[g] = tb[g].need;
  if (rc == SUCCESS && opts.first_sigma) {
    if (first.empty()) first.assign((size_t)t->nblock * FIRST_SIGMA, 0.0);
    memcpy(opts.first_sigma, first.data(), (size_t)t->nblock * FIRST_SIGMA * sizeof(double));
  }
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  for (int g = 0; g < t->nblock; g++) t->blk[g].s = tb[g].k;
  t->d_U = up;
  t->d_V = vp;
  rc = finish_tree(t, d_A, lda, s);
  if (rc == SUCCESS) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_compress_dense: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  fill_stats(stats, ts, std::chrono::duration<double, std::milli>(t1 - t0).count());
  *out = t;
  return SUCCESS;
}

int hm_hip_compress_batch(int nmat, const double *const *d_A, const long *lda, const int *m,
                          const int *height, double thr, const hm_compress_opts *opts_in,
                          hm_dtree **out, hm_compress_stats *stats, void *stream) {
  HM_OK(ensure_init());
  if (nmat < 0 || (nmat > 0 && (!d_A || !lda || !m || !height || !out))) {
    set_error("hm_hip_compress_batch: NULL argument");
    return INPUT_ERROR;
  }
  for (int i = 0; i < nmat; i++) {
    out[i] = nullptr;
    if (!d_A[i] || height[i] < 1 || height[i] > 24 || m[i] < 0 || lda[i] < std::max(1, m[i])) {
      set_error("hm_hip_compress_batch: matrix %d: bad arguments (m=%d height=%d)", i, m[i],
                height[i]);
      return INPUT_ERROR;
    }
  }
  hm_compress_opts opts;
  hm_hip_compress_opts_default(&opts);
  if (opts_in) opts = *opts_in;
  if (opts.rank_hint) {
    set_error("hm_hip_compress_batch: rank hints are per tree; not supported here");
    return INPUT_ERROR;
  }
  hipStream_t s = stream_of(stream);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<hm_dtree *> trees(nmat, nullptr);
  std::vector<TsvdBlock> all;
  std::vector<size_t> first(nmat + 1, 0);
  auto fail = [&](int rc) {
    for (auto *t : trees) tree_free(t);
    return rc;
  };
  for (int i = 0; i < nmat; i++) {
    std::vector<TsvdBlock> tb;
    int rc = plan_tree(d_A[i], lda[i], m[i], height[i], nullptr, thr, opts,
                       (unsigned long long)i << 32, &trees[i], tb);
    if (rc != SUCCESS) return fail(rc);
    first[i] = all.size();
    all.insert(all.end(), tb.begin(), tb.end());
  }
  first[nmat] = all.size();
  TsvdStats ts;
  double *up = nullptr, *vp = nullptr;
  int rc = tsvd_run(all, tsvd_options(opts), &up, &vp, &ts, nullptr, 0, s);
  if (rc != SUCCESS) return fail(rc);
  // The trees keep their factors where the batched SVD left them (each
  // tree's blocks are contiguous in the batch pools, in the order its
  // metadata lays them out), their diagonal blocks in one pool and their
  // metadata in one allocation: four allocations and one metadata copy for
  // the whole batch, shared by its trees and freed with the last of them.
  auto pools = std::make_shared<SharedPools>();
  pools->ptrs = {up, vp};
  for (int i = 0; i < nmat; i++) {
    hm_dtree *t = trees[i];
    for (int g = 0; g < t->nblock; g++) t->blk[g].s = all[first[i] + g].k;
    t->d_U = all[first[i]].uout;
    t->d_V = all[first[i]].vout;
    t->shared = pools;
  }
  char *meta = nullptr;
  rc = trees_finalize_meta(trees, &meta, s);
  if (meta) pools->ptrs.push_back(meta);  // also when its copy failed
  std::vector<long long> doff(nmat + 1, 0);
  for (int i = 0; i < nmat; i++) doff[i + 1] = doff[i] + trees[i]->diag_off[trees[i]->nleaf];
  double *dpool = nullptr;
  if (rc == SUCCESS) rc = dalloc(&dpool, (size_t)std::max(1LL, doff[nmat]));
  if (rc == SUCCESS) pools->ptrs.push_back(dpool);
  for (int i = 0; i < nmat && rc == SUCCESS; i++) {
    hm_dtree *t = trees[i];
    t->d_diag = dpool + doff[i];
    hipLaunchKernelGGL(k_extract_diag, dim3(t->nleaf, 64), dim3(256), 0, s, d_A[i],
                       (long long)lda[i], t->d_leaf_off, t->d_diag_off, t->d_diag);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("hm_hip_compress_batch: diagonal blocks: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  if (rc == SUCCESS) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_compress_batch: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  pools.reset();  // the trees hold it now
  if (rc != SUCCESS) return fail(rc);
  auto t1 = std::chrono::steady_clock::now();
  fill_stats(stats, ts, std::chrono::duration<double, std::milli>(t1 - t0).count());
  for (int i = 0; i < nmat; i++) out[i] = trees[i];
  return SUCCESS;
}

}  // extern "C"
