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
  const int nleaf = 1 << height;
  std::vector<int> leaves;
  if (ms) {
    long long sum = 0;
    for (int c = 0; c < nleaf; c++) {
      if (ms[c] < 0) {
        set_error("hm_hip_compress_dense: negative leaf size");
        return INPUT_ERROR;
      }
      leaves.push_back(ms[c]);
      sum += ms[c];
    }
    if (sum != m) {
      set_error("hm_hip_compress_dense: leaf sizes sum to %lld, not %d", sum, m);
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
  std::vector<TsvdBlock> tb(t->nblock);
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
      x.uid = (unsigned long long)g;
    }
  }
  TsvdOptions to;
  to.sketch_max = opts.sketch_max;
  to.n_test = opts.n_test;
  to.accept = opts.accept;
  to.seed = opts.seed;
  TsvdStats ts;
  double *up = nullptr, *vp = nullptr;
  int rc = tsvd_run(tb, to, &up, &vp, &ts, nullptr, 0, s);
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  for (int g = 0; g < t->nblock; g++) t->blk[g].s = tb[g].k;
  t->d_U = up;
  t->d_V = vp;
  rc = tree_finalize_meta(t, s);
  if (rc == SUCCESS) rc = dalloc(&t->d_diag, (size_t)t->diag_off[nleaf]);
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  hipLaunchKernelGGL(k_extract_diag, dim3(nleaf, 64), dim3(256), 0, s, d_A,
                     (long long)lda, t->d_leaf_off, t->d_diag_off, t->d_diag);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error("hm_hip_compress_dense: %s", hipGetErrorString(e));
    tree_free(t);
    return ALLOCATION_FAILURE;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (stats) {
    stats->sketch_rounds = ts.rounds;
    stats->blocks_retried = ts.retried;
    stats->max_sketch = ts.max_sketch;
    stats->fallback_blocks = ts.capped;
    stats->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  *out = t;
  return SUCCESS;
}

}  // extern "C"
