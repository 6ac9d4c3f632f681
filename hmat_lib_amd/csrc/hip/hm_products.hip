// HODLR x vector and HODLR x dense products (FP64).
//
// The reference loops over blocks calling dgemv/dgemm per block
// (src/vector_algebra.c:52-197, src/dense_algebra.c:82-1151).  Here a
// product is three launches over the whole tree:
//
//   partials  P_c = F_in(c)^T X_c for every leaf c and every level at once
//             (X_c read once; F_in = V for H X, U for H^T X)
//   reduce    W_b = sum of the partials of the leaves under the block's
//             column node (fixed order, deterministic)
//   apply     Y_c = D_c X_c (or D_c^T X_c) + sum_l F_out rows(c) W_b
//             as one K-concatenated GEMM per leaf row tile
//
// so every diagonal and U/V element is read once from HBM.  R = 1 with no
// transpose runs wavefront-reduction kernels; everything else runs FP64
// MFMA tiles (v_mfma_f64_16x16x4_f64) staged through LDS.
#include "hm_common.h"

namespace hm {

struct ProdArgs {
  int h, nleaf, R, trans;
  const int *leaf_off;
  const long long *diag_off;
  const BlockMeta *blk;
  const double *diag, *U, *V;
  const long long *pofs;
  const int *proff;
  const long long *wofs;
  double *P, *W;
  const double *X;
  long long x_rs, x_cs;
  double *Y;
  long long y_rs, y_cs;
};

// ----------------------------------------------------------- R = 1 path
__global__ __launch_bounds__(256) void k_mv_partials(ProdArgs a) {
  const int c = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  double *Pc = a.P + a.pofs[c];
  int o = 0;
  for (int l = 1; l <= a.h; l++) {
    const int q = c >> (a.h - l);
    const BlockMeta b = a.blk[blk_index(l, q ^ 1)];
    const int off = lo - a.leaf_off[q << (a.h - l)];
    const double *F = a.V + b.voff + off;
    for (int k = 0; k < b.s; k++) {
      if (((o + k) & 3) != wave) continue;
      double acc = 0.0;
      for (int i = lane; i < mc; i += 64)
        acc += F[i + (long long)k * b.n] * a.X[(long long)(lo + i) * a.x_rs];
      acc = wave_sum(acc);
      if (lane == 0) Pc[o + k] = acc;
    }
    o += b.s;
  }
}

__global__ __launch_bounds__(256) void k_mv_apply(ProdArgs a, int rtiles) {
  __shared__ double red[4][64];
  const int c = blockIdx.x / rtiles, rt = blockIdx.x % rtiles;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  const int r0 = rt * 64;
  if (r0 >= mc) return;
  const int i = r0 + lane;
  const bool valid = i < mc;
  const int ii = valid ? i : 0;
  const double *D = a.diag + a.diag_off[c];
  const int cq = (mc + 3) / 4;
  const int j0 = wave * cq, j1 = min(mc, j0 + cq);
  double acc0 = 0.0, acc1 = 0.0;
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    acc0 += D[ii + (long long)j * mc] * a.X[(long long)(lo + j) * a.x_rs];
    acc1 += D[ii + (long long)(j + 1) * mc] * a.X[(long long)(lo + j + 1) * a.x_rs];
  }
  if (j < j1) acc0 += D[ii + (long long)j * mc] * a.X[(long long)(lo + j) * a.x_rs];
  double acc = acc0 + acc1;
  for (int l = 1 + wave; l <= a.h; l += 4) {
    const int q = c >> (a.h - l);
    const int g = blk_index(l, q);
    const BlockMeta b = a.blk[g];
    const int off = lo - a.leaf_off[q << (a.h - l)];
    const double *F = a.U + b.uoff + off;
    const double *Wb = a.W + a.wofs[g];
    for (int k = 0; k < b.s; k++) acc += F[ii + (long long)k * b.m] * Wb[k];
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && valid) {
    double y = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    a.Y[(long long)(lo + i) * a.y_rs] = y;
  }
}

// ------------------------------------------------------------ reduction
__global__ __launch_bounds__(256) void k_reduce_w(ProdArgs a) {
  const int g = blockIdx.x;
  const BlockMeta b = a.blk[g];
  const int s = b.s;
  if (s == 0) return;
  const int l = b.level;
  const int src = a.trans ? b.j : (b.j ^ 1);
  const int span = 1 << (a.h - l);
  const int c_lo = src * span;
  double *Wb = a.W + a.wofs[g] * a.R;
  const long long tot = (long long)s * a.R;
  for (long long e = threadIdx.x + (long long)blockIdx.y * blockDim.x; e < tot;
       e += (long long)blockDim.x * gridDim.y) {
    const int k = (int)(e % s), r = (int)(e / s);
    double sum = 0.0;
    for (int c = c_lo; c < c_lo + span; c++) {
      const long long p0 = a.pofs[c];
      const long long sin_c = a.pofs[c + 1] - p0;
      sum += a.P[p0 * a.R + a.proff[(long long)c * a.h + (l - 1)] + k + r * sin_c];
    }
    Wb[k + (long long)r * s] = sum;
  }
}

// ------------------------------------------------------------ MFMA path
typedef MfmaTile<1, 4, 1, 1, 16> PartTile;  // 16 x 64
typedef MfmaTile<2, 2, 2, 2, 16> ApplyTile; // 64 x 64

template <bool B_NFAST>
__global__ __launch_bounds__(256) void k_mm_partials(ProdArgs a, int mtiles,
                                                     int ntiles) {
  __shared__ double lds[PartTile::LDS_DOUBLES];
  __shared__ const double *rowptr[16];
  int bid = blockIdx.x;
  const int nt = bid % ntiles;
  bid /= ntiles;
  const int mt = bid % mtiles;
  const int c = bid / mtiles;
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  const long long p0 = a.pofs[c];
  const int sin_c = (int)(a.pofs[c + 1] - p0);
  const int rho0 = mt * 16;
  if (rho0 >= sin_c) return;
  if (threadIdx.x < 16) {
    const int rho = rho0 + threadIdx.x;
    const double *ptr = nullptr;
    if (rho < sin_c) {
      for (int l = 1; l <= a.h; l++) {
        const int q = c >> (a.h - l);
        const int bi = a.trans ? q : (q ^ 1);
        const BlockMeta b = a.blk[blk_index(l, bi)];
        const int base = a.proff[(long long)c * a.h + (l - 1)];
        if (rho >= base && rho < base + b.s) {
          const int off = lo - a.leaf_off[q << (a.h - l)];
          const int k = rho - base;
          ptr = a.trans ? a.U + b.uoff + off + (long long)k * b.m
                        : a.V + b.voff + off + (long long)k * b.n;
          break;
        }
      }
    }
    rowptr[threadIdx.x] = ptr;
  }
  __syncthreads();
  const int n0 = nt * 64;
  auto la = [&](int i, int k) -> double {
    const double *p = rowptr[i];
    return (p && k < mc) ? p[k] : 0.0;
  };
  auto lb = [&](int k, int j) -> double {
    return (k < mc && n0 + j < a.R)
               ? a.X[(long long)(lo + k) * a.x_rs + (long long)(n0 + j) * a.x_cs]
               : 0.0;
  };
  PartTile tile;
  tile.zero();
  tile.template run<false, B_NFAST>(lds, mc, la, lb);
  double *Pc = a.P + p0 * a.R;
  tile.store([&](int i, int j, double v) {
    const int rho = rho0 + i;
    if (rho < sin_c && n0 + j < a.R) Pc[rho + (long long)(n0 + j) * sin_c] = v;
  });
}

template <bool B_NFAST>
__global__ __launch_bounds__(256) void k_mm_apply(ProdArgs a, int rtiles,
                                                  int ntiles) {
  __shared__ double lds[ApplyTile::LDS_DOUBLES];
  __shared__ int lp[33];
  __shared__ const double *fb[32];
  __shared__ const double *wb[32];
  __shared__ int fld[32], ws[32];
  int bid = blockIdx.x;
  const int nt = bid % ntiles;
  bid /= ntiles;
  const int rt = bid % rtiles;
  const int c = bid / rtiles;
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  const int r0 = rt * 64;
  if (r0 >= mc) return;
  const int n0 = nt * 64;
  const double *D = a.diag + a.diag_off[c];
  ApplyTile tile;
  tile.zero();
  auto lb = [&](int k, int j) -> double {
    return (k < mc && n0 + j < a.R)
               ? a.X[(long long)(lo + k) * a.x_rs + (long long)(n0 + j) * a.x_cs]
               : 0.0;
  };
  if (!a.trans) {
    auto la = [&](int i, int k) -> double {
      return (r0 + i < mc && k < mc) ? D[(r0 + i) + (long long)k * mc] : 0.0;
    };
    tile.template run<true, B_NFAST>(lds, mc, la, lb);
  } else {
    auto la = [&](int i, int k) -> double {
      return (r0 + i < mc && k < mc) ? D[k + (long long)(r0 + i) * mc] : 0.0;
    };
    tile.template run<false, B_NFAST>(lds, mc, la, lb);
  }
  // low-rank segment: K runs over (level, rank column) pairs
  if (threadIdx.x == 0) {
    int run = 0;
    for (int l = 1; l <= a.h; l++) {
      const int q = c >> (a.h - l);
      const int bo = a.trans ? (q ^ 1) : q;
      const int g = blk_index(l, bo);
      const BlockMeta b = a.blk[g];
      const int off = lo - a.leaf_off[q << (a.h - l)];
      lp[l - 1] = run;
      fb[l - 1] = a.trans ? a.V + b.voff + off : a.U + b.uoff + off;
      fld[l - 1] = a.trans ? b.n : b.m;
      wb[l - 1] = a.W + a.wofs[g] * a.R;
      ws[l - 1] = b.s;
      run += b.s;
    }
    lp[a.h] = run;
  }
  __syncthreads();
  const int K2 = lp[a.h];
  if (K2 > 0) {
    const int h = a.h;
    auto lvl = [&](int kap) -> int {
      int l = 0;
      while (l + 1 < h && kap >= lp[l + 1]) l++;
      return l;
    };
    auto la2 = [&](int i, int kap) -> double {
      if (r0 + i >= mc || kap >= K2) return 0.0;
      const int l = lvl(kap);
      return fb[l][(r0 + i) + (long long)(kap - lp[l]) * fld[l]];
    };
    auto lb2 = [&](int kap, int j) -> double {
      if (kap >= K2 || n0 + j >= a.R) return 0.0;
      const int l = lvl(kap);
      return wb[l][(kap - lp[l]) + (long long)(n0 + j) * ws[l]];
    };
    tile.template run<true, false>(lds, K2, la2, lb2);
  }
  tile.store([&](int i, int j, double v) {
    if (r0 + i < mc && n0 + j < a.R)
      a.Y[(long long)(lo + r0 + i) * a.y_rs + (long long)(n0 + j) * a.y_cs] = v;
  });
}

static ProdArgs make_args(hm_dtree *t, int trans, int R, const double *X,
                          long x_rs, long x_cs, double *Y, long y_rs,
                          long y_cs) {
  ProdArgs a;
  a.h = t->height;
  a.nleaf = t->nleaf;
  a.R = R;
  a.trans = trans;
  a.leaf_off = t->d_leaf_off;
  a.diag_off = t->d_diag_off;
  a.blk = t->d_blk;
  a.diag = t->d_diag;
  a.U = t->d_U;
  a.V = t->d_V;
  a.pofs = t->plan.d_pofs[trans];
  a.proff = t->plan.d_proff[trans];
  a.wofs = t->plan.d_wofs;
  long long pmax = std::max(t->plan.pofs[0][t->nleaf], t->plan.pofs[1][t->nleaf]);
  a.P = t->d_scratch;
  a.W = t->d_scratch + pmax * R;
  a.X = X;
  a.x_rs = x_rs;
  a.x_cs = x_cs;
  a.Y = Y;
  a.y_rs = y_rs;
  a.y_cs = y_cs;
  return a;
}

int products_run(hm_dtree *t, int trans, int R, const double *X, long x_rs,
                 long x_cs, double *Y, long y_rs, long y_cs, hipStream_t s) {
  if (!t) {
    set_error("product: NULL device tree");
    return INPUT_ERROR;
  }
  if (R <= 0) return SUCCESS;
  if (!X || !Y) {
    set_error("product: NULL operand");
    return INPUT_ERROR;
  }
  HM_OK(ensure_scratch(t, R));
  ProdArgs a = make_args(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs);
  const int nb = t->nblock;
  if (R == 1 && !trans) {
    Timed tp = timing_begin("mv_partials", s);
    hipLaunchKernelGGL(k_mv_partials, dim3(t->nleaf), dim3(256), 0, s, a);
    timing_end(tp, s);
    HM_LAUNCH_CHECK();
    Timed tr = timing_begin("reduce_w", s);
    hipLaunchKernelGGL(k_reduce_w, dim3(nb, 1), dim3(256), 0, s, a);
    timing_end(tr, s);
    HM_LAUNCH_CHECK();
    const int rtiles = (t->max_leaf + 63) / 64;
    Timed ta = timing_begin("mv_apply", s);
    hipLaunchKernelGGL(k_mv_apply, dim3(t->nleaf * rtiles), dim3(256), 0, s, a,
                       rtiles);
    timing_end(ta, s);
    HM_LAUNCH_CHECK();
    return SUCCESS;
  }
  const bool nfast = (x_rs != 1);
  const int ntiles = (R + 63) / 64;
  const int smax = t->plan.s_in_max[trans];
  if (smax > 0) {
    const int mtiles = (smax + 15) / 16;
    dim3 grid((unsigned)(t->nleaf * mtiles * ntiles));
    Timed tp = timing_begin("mm_partials", s);
    if (nfast)
      hipLaunchKernelGGL(k_mm_partials<true>, grid, dim3(256), 0, s, a, mtiles, ntiles);
    else
      hipLaunchKernelGGL(k_mm_partials<false>, grid, dim3(256), 0, s, a, mtiles, ntiles);
    timing_end(tp, s);
    HM_LAUNCH_CHECK();
    long long maxwork = (long long)t->max_rank * R;
    int ysplit = (int)std::min<long long>(64, (maxwork + 255) / 256);
    if (ysplit < 1) ysplit = 1;
    Timed tr = timing_begin("reduce_w", s);
    hipLaunchKernelGGL(k_reduce_w, dim3(nb, ysplit), dim3(256), 0, s, a);
    timing_end(tr, s);
    HM_LAUNCH_CHECK();
  }
  const int rtiles = (t->max_leaf + 63) / 64;
  dim3 grid((unsigned)(t->nleaf * rtiles * ntiles));
  Timed ta = timing_begin("mm_apply", s);
  if (nfast)
    hipLaunchKernelGGL(k_mm_apply<true>, grid, dim3(256), 0, s, a, rtiles, ntiles);
  else
    hipLaunchKernelGGL(k_mm_apply<false>, grid, dim3(256), 0, s, a, rtiles, ntiles);
  timing_end(ta, s);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_matvec(hm_dtree *t, const double *d_x, double *d_y, void *stream) {
  HM_OK(ensure_init());
  return products_run(t, 0, 1, d_x, 1, 1, d_y, 1, 1, stream_of(stream));
}

int hm_hip_matmat(hm_dtree *t, int trans, int R, const double *d_X, long x_rs,
                  long x_cs, double *d_Y, long y_rs, long y_cs, void *stream) {
  HM_OK(ensure_init());
  return products_run(t, trans ? 1 : 0, R, d_X, x_rs, x_cs, d_Y, y_rs, y_cs,
                      stream_of(stream));
}

}  // extern "C"

namespace hm {
__global__ void k_selftest_mfma(double *out) {
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const double a = 1.0 + li + 16.0 * lk;      // A[li][lk]
  const double b = 1.0 + 2.0 * li - lk;       // B[lk][li]
  d4 c = {0.0, 0.0, 0.0, 0.0};
  c = mfma_f64(a, b, c);
  for (int v = 0; v < 4; v++) {
    const int i = (lane >> 4) + 4 * v, j = lane & 15;
    out[i * 16 + j] = c[v];
  }
}
}  // namespace hm

extern "C" int hm_hip_selftest_mfma(double *out) {
  HM_OK(ensure_init());
  double *d = nullptr;
  HM_OK(dalloc(&d, 256));
  hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, 0, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
  dfree(d);
  if (e != hipSuccess) {
    set_error("selftest: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}
