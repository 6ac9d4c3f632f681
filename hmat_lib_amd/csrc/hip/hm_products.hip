// HODLR x vector and HODLR x dense products (FP64).
//
// The reference loops over blocks calling dgemv/dgemm per block
// (src/vector_algebra.c:52-197, src/dense_algebra.c:82-1151).  Here a
// product is three launches over the whole tree:
//
//   partials  P(l, k) = F_in(k)^T X_k for every row chunk k (<= 64 rows of
//             one leaf) and every level l at once (X read once; F_in = V
//             for H X, U for H^T X)
//   reduce    W_b = sum of the partials of the chunks under the block's
//             column node: one contiguous region per block, summed in a
//             fixed order (deterministic)
//   apply     Y_c = D_c X_c (or D_c^T X_c) + sum_l F_out rows(c) W_b
//             as one K-concatenated GEMM per leaf row tile
//
// so every diagonal and U/V element is read once from HBM.  R = 1 with no
// transpose runs wavefront-reduction kernels; everything else runs FP64
// MFMA tiles (v_mfma_f64_16x16x4_f64) staged through LDS.  The apply tile
// shape is chosen per tree and RHS width by a small timing autotuner.
#include <cstdlib>
#include <mutex>

#include "hm_common.h"
#include "hm_prog.h"

namespace hm {

struct ProdArgs {
  int h, nleaf, R, trans, nchunk, pad;
  const int *leaf_off;
  const long long *diag_off;
  const BlockMeta *blk;
  const double *diag, *U, *V;
  const int4 *ch;
  const PartRow *prow;
  const long long *pofs;
  const long long *wofs;
  const RedItem *fsum;  // per block: its chunk partials as one item (H x, fused)
  double *P, *W, *Wp;
  const double *X;
  long long x_rs, x_cs;
  double *Y;
  long long y_rs, y_cs;
};

// Per-level view of one leaf, preloaded into LDS by threads 0..h-1 so that
// no thread walks the levels through dependent global loads.
struct LevelInfo {
  const double *F;  // factor panel at the leaf's first row (input or output side)
  double *P;        // partial block of this chunk (input side) or W_b (output side)
  int ld, s, g, pad;
};
static constexpr int MAX_LEVELS = 24;

// side 0: input side (V for H X, U for H^T X), P = partial of chunk k;
// side 1: output side (U for H X, V for H^T X), P = W_b.
__device__ __forceinline__ void load_levels(const ProdArgs &a, int c, int k, int side,
                                            LevelInfo *li) {
  const int l = threadIdx.x + 1;
  if (l <= a.h) {
    const int q = c >> (a.h - l);
    const bool want_q = (side == 0) == (a.trans != 0);  // block (l, q) or (l, q^1)
    const int g = blk_index(l, want_q ? q : (q ^ 1));
    const BlockMeta b = a.blk[g];
    const int off = a.leaf_off[c] - a.leaf_off[q << (a.h - l)];
    const bool use_u = want_q;  // rows of block (l, q) are the leaf's rows: its U
    LevelInfo x;
    x.F = use_u ? a.U + b.uoff + off : a.V + b.voff + off;
    x.ld = use_u ? b.m : b.n;
    x.s = b.s;
    x.g = g;
    x.pad = 0;
    x.P = side == 0 ? a.P + a.pofs[(long long)(l - 1) * a.nchunk + k] * a.R
                    : a.W + a.wofs[g] * a.R;
    li[l - 1] = x;
  }
}

// ----------------------------------------------------------- R = 1 path
__global__ __launch_bounds__(256) void k_mv_partials(ProdArgs a) {
  __shared__ LevelInfo li[MAX_LEVELS];
  const int k = blockIdx.x;
  const int4 chk = a.ch[k];
  const int c = chk.x, r0 = chk.y, rows = chk.z;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  load_levels(a, c, k, 0, li);
  const int lo = a.leaf_off[c] + r0;
  const double xi = lane < rows ? a.X[(long long)(lo + lane) * a.x_rs] : 0.0;
  __syncthreads();
  int o = 0;
  for (int l = 0; l < a.h; l++) {
    const LevelInfo x = li[l];
    for (int kk = 0; kk < x.s; kk++) {
      if (((o + kk) & 3) != wave) continue;
      double acc = lane < rows ? ldg(x.F + r0 + lane + (long long)kk * x.ld) * xi : 0.0;
      acc = wave_sum(acc);
      if (lane == 0) stg(x.P + kk, acc);
    }
    o += x.s;
  }
}

// FUSED: W_b of the leaf's levels summed here from the chunk partials (one
// fixed-order sum per rank: lane-strided, then the wavefront butterfly) into
// LDS, instead of by the two reduction launches, by four more waves while
// waves 0-3 stream the diagonal block; for trees whose leaves have at most
// MV_FUSE_VALUES partial values (plan.mv_fused).
template <bool FUSED>
__global__ __launch_bounds__(FUSED ? 512 : 256) void k_mv_apply(ProdArgs a, int rtiles) {
  __shared__ double red[4][64];
  __shared__ LevelInfo li[MAX_LEVELS];
  __shared__ double wl[FUSED ? MV_FUSE_RANKS : 1];
  __shared__ int wofs_l[MAX_LEVELS + 1];
  const int c = blockIdx.x / rtiles, rt = blockIdx.x % rtiles;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  const int r0 = rt * 64;
  if (r0 >= mc) return;
  load_levels(a, c, 0, 1, li);
  if (FUSED) {
    if (threadIdx.x == 0) {
      int o = 0;
      for (int l = 1; l <= a.h; l++) {
        wofs_l[l - 1] = o;
        o += a.fsum[blk_index(l, c >> (a.h - l))].s;
      }
      wofs_l[a.h] = o;
    }
    __syncthreads();
    for (int l = wave - 4; l >= 0 && l < a.h; l += 4) {
      const RedItem r = a.fsum[blk_index(l + 1, c >> (a.h - l - 1))];
      const double *P = a.P + r.src;
      for (int kk = 0; kk < r.s; kk++) {
        double t0 = 0.0, t1 = 0.0;
        int i = lane;
        for (; i + 64 < r.n; i += 128) {
          t0 += ldg(P + (long long)i * r.s + kk);
          t1 += ldg(P + (long long)(i + 64) * r.s + kk);
        }
        if (i < r.n) t0 += ldg(P + (long long)i * r.s + kk);
        const double t = wave_sum(t0 + t1);
        if (lane == 0) wl[wofs_l[l] + kk] = t;
      }
    }
  }
  const int i = r0 + lane;
  const bool valid = i < mc;
  const int ii = valid ? i : 0;
  const double *D = a.diag + a.diag_off[c];
  const int cq = (mc + 3) / 4;
  double acc = 0.0;
  if (wave < 4) {  // the diagonal block, columns split over waves 0-3
    const int j0 = wave * cq, j1 = min(mc, j0 + cq);
    double acc0 = 0.0, acc1 = 0.0;
    int j = j0;
    for (; j + 1 < j1; j += 2) {
      acc0 += D[ii + (long long)j * mc] * a.X[(long long)(lo + j) * a.x_rs];
      acc1 += D[ii + (long long)(j + 1) * mc] * a.X[(long long)(lo + j + 1) * a.x_rs];
    }
    if (j < j1) acc0 += D[ii + (long long)j * mc] * a.X[(long long)(lo + j) * a.x_rs];
    acc = acc0 + acc1;
  }
  __syncthreads();
  if (wave < 4) {
    for (int l = wave; l < a.h; l += 4) {
      const LevelInfo x = li[l];
      if (FUSED) {
        const double *w = wl + wofs_l[l];
        for (int k = 0; k < x.s; k++) acc += ldg(x.F + ii + (long long)k * x.ld) * w[k];
      } else {
        for (int k = 0; k < x.s; k++) acc += ldg(x.F + ii + (long long)k * x.ld) * ldg(x.P + k);
      }
    }
    red[wave][lane] = acc;
  }
  __syncthreads();
  if (wave == 0 && valid) {
    double y = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    a.Y[(long long)(lo + i) * a.y_rs] = y;
  }
}

// ------------------------------------------------------------ reduction
// out[e] = sum over p < nparts of base[p * stride + e], e < E, by one
// workgroup of 256 threads; when E < 256 the parts are split over 256 / E
// thread groups and combined through LDS in a fixed order.
__device__ __forceinline__ void group_sum(const double *base, long long stride, int nparts,
                                          long long E, double *out, double *red) {
  if (E >= 256) {
    for (long long e = threadIdx.x; e < E; e += 256) {
      double s0 = 0.0, s1 = 0.0;
      int p = 0;
      for (; p + 1 < nparts; p += 2) {
        s0 += base[p * stride + e];
        s1 += base[(p + 1) * stride + e];
      }
      if (p < nparts) s0 += base[p * stride + e];
      out[e] = s0 + s1;
    }
    return;
  }
  const int G = 256 / (int)E;
  const int g = threadIdx.x / (int)E, e = threadIdx.x % (int)E;
  double s0 = 0.0, s1 = 0.0;
  if (g < G) {
    int p = g;
    for (; p + G < nparts; p += 2 * G) {
      s0 += base[p * stride + e];
      s1 += base[(p + G) * stride + e];
    }
    if (p < nparts) s0 += base[p * stride + e];
  }
  red[threadIdx.x] = s0 + s1;
  __syncthreads();
  if (threadIdx.x < E) {
    double t = 0.0;
    for (int gg = 0; gg < G; gg++) t += red[gg * E + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// One reduction item per workgroup (items precomputed on the host, so the
// only dependent load is the item itself): stage 1 sums chunk partials into
// segment sums, stage 2 segment sums (or a block's few chunk partials)
// into W_b; both in fixed order.
__device__ __forceinline__ void reduce_item(const ProdArgs &a, const RedItem &it, int to_w) {
  __shared__ double red[256];
  const long long E = (long long)it.s * a.R;
  const double *base = (it.from_seg ? a.Wp : a.P) + it.src * a.R;
  double *out = (to_w ? a.W : a.Wp) + it.dst * a.R;
  group_sum(base, E, it.n, E, out, red);
}

__global__ __launch_bounds__(256) void k_reduce_items(ProdArgs a, const RedItem *items,
                                                      int to_w) {
  reduce_item(a, items[blockIdx.x], to_w);
}

// One tree of a batched product (a device table entry): its product
// arguments, its two reduction lists (segment sums, then the sums into W)
// and its partials' rank tiles.
struct BatchProd {
  ProdArgs a;
  const RedItem *red[2];
  int mtiles, pad;
};

// Batched reduction, one wave per work item (tree, item of list to_w):
// each lane sums its elements over the item's parts in order (fixed order,
// deterministic).  A batch's items are short (at most REDUCE_SEG parts of
// s R values, mostly a few dozen), so one-wave workgroups with no LDS and no
// barrier keep the most of them in flight.
__global__ __launch_bounds__(64) void k_reduce_items_b(const BatchProd *bp, const int2 *wl,
                                                       int to_w) {
  const int2 w = wl[blockIdx.x];
  const ProdArgs a = bp[w.x].a;
  const RedItem it = bp[w.x].red[to_w][w.y];
  const long long E = (long long)it.s * a.R;
  const double *base = (it.from_seg ? a.Wp : a.P) + it.src * a.R;
  double *out = (to_w ? a.W : a.Wp) + it.dst * a.R;
  for (long long e = threadIdx.x; e < E; e += 64) {
    double s0 = 0.0, s1 = 0.0;
    int q = 0;
    for (; q + 1 < it.n; q += 2) {
      s0 += ldg(base + q * E + e);
      s1 += ldg(base + (q + 1) * E + e);
    }
    if (q < it.n) s0 += ldg(base + q * E + e);
    stg(out + e, s0 + s1);
  }
}

// ------------------------------------------------------------ MFMA path
// Partials of chunk k, rank tile mt (16 stacked factor rows), RHS tile nt
// of PartTile::TN columns (16 x 64 tile, four waves; or 16 x 16, one wave).
typedef MfmaTile<1, 4, 1, 1, 16> PartTile64;
typedef MfmaTile<1, 1, 1, 1, 16> PartTile16;
template <bool B_NFAST, class PartTile = PartTile64>
__device__ __forceinline__ void mm_partials_tile(const ProdArgs &a, int k, int mt, int nt,
                                                 int mtiles) {
  __shared__ double lds[PartTile::LDS_DOUBLES];
  __shared__ const double *rowptr[16];
  __shared__ double *outptr[16];
  __shared__ int outld[16];
  const int4 chk = a.ch[k];
  const int rows = chk.z;
  const int lo = chk.w;
  // the chunk's stacked input-factor rows of this M tile (precomputed)
  if (threadIdx.x < 16) {
    const PartRow r = a.prow[((long long)k * mtiles + mt) * 16 + threadIdx.x];
    const double *F = a.trans ? a.U : a.V;
    rowptr[threadIdx.x] = r.s ? F + r.f_off : nullptr;
    outptr[threadIdx.x] = r.s ? a.P + r.p_base * a.R + r.kk : nullptr;
    outld[threadIdx.x] = r.s ? r.s : 1;
  }
  __syncthreads();
  bool any = false;
  for (int r = 0; r < 16; r++) any |= (rowptr[r] != nullptr);
  if (!any) return;
  const int n0 = nt * PartTile::TN;
  auto la = [&](int i, int kx) -> double {
    const double *p = rowptr[i];
    return (p && kx < rows) ? ldg(p + kx) : 0.0;
  };
  auto lb = [&](int kx, int j) -> double {
    return (kx < rows && n0 + j < a.R)
               ? a.X[(long long)(lo + kx) * a.x_rs + (long long)(n0 + j) * a.x_cs]
               : 0.0;
  };
  PartTile tile;
  tile.zero();
  tile.template run<false, B_NFAST>(lds, rows, la, lb);
  tile.store([&](int i, int j, double v) {
    double *o = outptr[i];
    if (o && n0 + j < a.R) stg(o + (long long)(n0 + j) * outld[i], v);
  });
}

template <bool B_NFAST>
__global__ __launch_bounds__(256) void k_mm_partials(ProdArgs a, int mtiles,
                                                     int ntiles) {
  int bid = blockIdx.x;
  const int nt = bid % ntiles;
  bid /= ntiles;
  const int mt = bid % mtiles;
  mm_partials_tile<B_NFAST>(a, bid / mtiles, mt, nt, mtiles);
}

// Batched partials: work item (tree, chunk, rank tile, RHS tile).  The
// batch is many short chunks with few RHS each, so a workgroup's time is
// mostly its dependent round trips: the one-wave 16 x 16 tile runs four
// times as many of them per CU as the 16 x 64 one (chosen by timing).
template <bool B_NFAST, class PartTile>
__global__ __launch_bounds__(PartTile::NT) void k_mm_partials_b(const BatchProd *bp,
                                                                const int4 *wl) {
  const int4 w = wl[blockIdx.x];
  const ProdArgs a = bp[w.x].a;
  mm_partials_tile<B_NFAST, PartTile>(a, w.y, w.z, w.w, bp[w.x].mtiles);
}

// ------------------------------------------------------ block W path
// One work item = (block g, rank tile st, K segment sg, rows in segment):
// C(st rows x 64 RHS) = F_in[seg rows, st cols]^T X[src rows of seg, :].
// F_in is V (n x s) for H X and U (m x s) for H^T X, so the A operand is
// K-fast; X may be column- or row-major (B_NFAST).
typedef MfmaTile<2, 2, 2, 2, 16> WTileT;  // 64 x 64
static_assert(WTileT::TM == WTILE, "rank tile");

template <bool B_NFAST>
__global__ __launch_bounds__(256) void k_w_gemm(ProdArgs a, const int4 *items,
                                                const long long *wpofs) {
  __shared__ double lds[WTileT::LDS_DOUBLES];
  const int4 it = items[blockIdx.x];
  const BlockMeta b = a.blk[it.x];
  const int s = b.s;
  const int K = a.trans ? b.m : b.n;
  const int src = a.trans ? b.r0 : b.c0;
  const double *F = a.trans ? a.U + b.uoff : a.V + b.voff;
  const int k0 = it.z * WSEG_ROWS, kl = it.w;
  const int i0 = it.y * WTileT::TM, n0 = blockIdx.y * WTileT::TN;
  const double *Fk = F + k0;
  const double *Xk = a.X + (long long)(src + k0) * a.x_rs;
  auto la = [&](int i, int k) -> double {
    return (i0 + i < s && k < kl) ? Fk[k + (long long)(i0 + i) * K] : 0.0;
  };
  auto lb = [&](int k, int j) -> double {
    return (k < kl && n0 + j < a.R) ? Xk[(long long)k * a.x_rs + (long long)(n0 + j) * a.x_cs]
                                    : 0.0;
  };
  WTileT tile;
  tile.zero();
  tile.template run<false, B_NFAST>(lds, kl, la, lb);
  const int nseg = (K + WSEG_ROWS - 1) / WSEG_ROWS;
  double *out = nseg > 1 ? a.P + (wpofs[it.x] + (long long)it.z * s) * a.R
                         : a.W + a.wofs[it.x] * a.R;
  tile.store([&](int i, int j, double v) {
    if (i0 + i < s && n0 + j < a.R) out[(i0 + i) + (long long)(n0 + j) * s] = v;
  });
}

// W_b = sum over K segments of the partials (fixed order).
__global__ __launch_bounds__(256) void k_w_reduce(ProdArgs a, const int *multi,
                                                  const long long *wpofs) {
  const int g = multi[blockIdx.x];
  const BlockMeta b = a.blk[g];
  const int K = a.trans ? b.m : b.n;
  const int nseg = (K + WSEG_ROWS - 1) / WSEG_ROWS;
  const long long E = (long long)b.s * a.R;
  const double *base = a.P + wpofs[g] * a.R;
  double *out = a.W + a.wofs[g] * a.R;
  for (long long e = (long long)blockIdx.y * 256 + threadIdx.x; e < E;
       e += (long long)gridDim.y * 256) {
    double s0 = 0.0;
    for (int sg = 0; sg < nseg; sg++) s0 += base[(long long)sg * E + e];
    out[e] = s0;
  }
}

// Work order of the apply kernel: leaves are dealt to the 8 XCDs round
// robin (workgroup b runs on XCD b % 8), and all row/column tiles of one
// leaf stay on its XCD so that X_c and the W_b are shared through one L2.
__device__ __forceinline__ bool apply_tile_of(int bid, int nleaf, int rtiles, int ntiles,
                                              int &c, int &rt, int &nt) {
  const int xcd = bid & 7, slot = bid >> 3;
  const int per = rtiles * ntiles;
  c = (slot / per) * 8 + xcd;
  const int w = slot % per;
  rt = w / ntiles;
  nt = w % ntiles;
  return c < nleaf;
}

// FAST: every leaf is a whole number of TM x KC tiles, R is a multiple of
// TN and there is no transpose (checked on the host); only the
// interior form of the diagonal GEMM is compiled in.
// Apply tile (leaf c, row tile rt, RHS tile nt) of Tile::TM x Tile::TN.
template <bool B_NFAST, class Tile, bool FAST>
__device__ __forceinline__ void mm_apply_tile(const ProdArgs &a, int c, int rt, int nt) {
  constexpr int KC = Tile::KC_;
  __shared__ double lds[Tile::LDS_DOUBLES];
  __shared__ int lp[MAX_LEVELS + 1];
  __shared__ LevelInfo li[MAX_LEVELS];
  __shared__ const double *tabA[KC];
  __shared__ const double *tabB[KC];
  __shared__ int tabS[KC];
  const int lo = a.leaf_off[c], mc = a.leaf_off[c + 1] - lo;
  const int r0 = rt * Tile::TM;
  if (r0 >= mc) return;
  const int n0 = nt * Tile::TN;
  const double *D = a.diag + a.diag_off[c];
  Tile tile;
  tile.zero();
  auto lb = [&](int k, int j) -> double {
    return (k < mc && n0 + j < a.R)
               ? a.X[(long long)(lo + k) * a.x_rs + (long long)(n0 + j) * a.x_cs]
               : 0.0;
  };
  load_levels(a, c, 0, 1, li);
  const bool interior = FAST || ((r0 + Tile::TM <= mc) && (mc % KC == 0) &&
                                 (n0 + Tile::TN <= a.R) && !a.trans);
  if (FAST) {
    tile.template run_interior<B_NFAST>(lds, mc, D + r0, mc,
                                        a.X + (long long)lo * a.x_rs + (long long)n0 * a.x_cs,
                                        a.x_rs, a.x_cs);
  } else if (interior) {
    tile.template run_interior<B_NFAST>(lds, mc, D + r0, mc,
                                        a.X + (long long)lo * a.x_rs + (long long)n0 * a.x_cs,
                                        a.x_rs, a.x_cs);
  } else if (!a.trans) {
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
    for (int l = 0; l < a.h; l++) {
      lp[l] = run;
      run += li[l].s;
    }
    lp[a.h] = run;
  }
  __syncthreads();
  const int K2 = lp[a.h];
  // K runs over (level, rank column) pairs; each K-chunk first resolves its
  // KC pairs into LDS pointers (factor column at this tile's first row, W
  // row and stride), then streams through the MFMA tile.
  for (int k0 = 0; k0 < K2; k0 += KC) {
    if (threadIdx.x < KC) {
      const int kap = k0 + threadIdx.x;
      const double *pa = nullptr, *pb = nullptr;
      int st = 0;
      if (kap < K2) {
        int l = 0;
        while (l + 1 < a.h && kap >= lp[l + 1]) l++;
        const int kk = kap - lp[l];
        pa = li[l].F + r0 + (long long)kk * li[l].ld;
        pb = li[l].P + kk;
        st = li[l].s;
      }
      tabA[threadIdx.x] = pa;
      tabB[threadIdx.x] = pb;
      tabS[threadIdx.x] = st;
    }
    __syncthreads();
    auto la2 = [&](int i, int k) -> double {
      const double *p = tabA[k];
      return (p && (FAST || r0 + i < mc)) ? ldg(p + i) : 0.0;
    };
    auto lb2 = [&](int k, int j) -> double {
      const double *p = tabB[k];
      return (p && (FAST || n0 + j < a.R)) ? ldg(p + (long long)(n0 + j) * tabS[k]) : 0.0;
    };
    tile.template fetch<true, false>(0, la2, lb2);
    tile.template stash<true, false>(lds);
    __syncthreads();
    tile.compute(lds);
    __syncthreads();
  }
  tile.store([&](int i, int j, double v) {
    if (FAST || (r0 + i < mc && n0 + j < a.R))
      a.Y[(long long)(lo + r0 + i) * a.y_rs + (long long)(n0 + j) * a.y_cs] = v;
  });
}

template <bool B_NFAST, int WM, int FM, int FN, int KC, bool FAST>
__global__ __launch_bounds__(256) void k_mm_apply(ProdArgs a, int rtiles, int ntiles) {
  int c, rt, nt;
  if (!apply_tile_of(blockIdx.x, a.nleaf, rtiles, ntiles, c, rt, nt)) return;
  mm_apply_tile<B_NFAST, MfmaTile<WM, 4 / WM, FM, FN, KC>, FAST>(a, c, rt, nt);
}

// Batched apply: work item (tree, leaf, row tile, RHS tile); tree < 0 pads
// the per-XCD queues.
template <bool B_NFAST, class Tile>
__global__ __launch_bounds__(Tile::NT) void k_mm_apply_b(const BatchProd *bp, const int4 *wl) {
  const int4 w = wl[blockIdx.x];
  if (w.x < 0) return;
  const ProdArgs a = bp[w.x].a;
  mm_apply_tile<B_NFAST, Tile, false>(a, w.y, w.z, w.w);
}

static ProdArgs make_args(hm_dtree *t, int trans, int R, const double *X,
                          long x_rs, long x_cs, double *Y, long y_rs,
                          long y_cs, double *scratch_base = nullptr) {
  ProdArgs a;
  a.h = t->height;
  a.nleaf = t->nleaf;
  a.R = R;
  a.trans = trans;
  a.nchunk = t->plan.nchunk;
  a.pad = 0;
  a.leaf_off = t->d_leaf_off;
  a.diag_off = t->d_diag_off;
  a.blk = t->d_blk;
  a.diag = t->d_diag;
  a.U = t->d_U;
  a.V = t->d_V;
  a.ch = t->plan.d_ch;
  a.prow = t->plan.d_prow[trans ? 1 : 0];
  a.pofs = t->plan.d_pofs[trans];
  a.wofs = t->plan.d_wofs;
  a.fsum = t->plan.d_fsum;
  long long pmax = std::max(std::max(t->plan.pofs[0].back(), t->plan.pofs[1].back()),
                            std::max(t->plan.wpofs[0].back(), t->plan.wpofs[1].back()));
  a.P = scratch_base ? scratch_base : t->d_scratch;
  // no scratch (a product of the diagonal blocks only): no W either
  a.W = a.P ? a.P + pmax * R : nullptr;
  a.Wp = a.P ? a.W + t->plan.wofs[t->nblock] * R : nullptr;
  a.X = X;
  a.x_rs = x_rs;
  a.x_cs = x_cs;
  a.Y = Y;
  a.y_rs = y_rs;
  a.y_cs = y_cs;
  return a;
}

// ------------------------------------------------ leaf-resident apply
// One workgroup of NW waves per (leaf, row tile of 32 NW rows, 64 RHS);
// wave w owns 32 rows of the tile and all 64 columns (2 x 4 MFMA blocks).
// No element of D is used by two waves, so D streams from HBM straight into
// registers, one 32-column chunk ahead; X_c streams through two LDS buffers
// of 32 rows x 64 columns, one barrier per chunk.  In LDS the pair (rows
// 2p, 2p + 1) of column j sits in 16-byte slot j * 16 + (p ^ (j & 15)), so
// the 16 columns one MFMA operand reads fall in distinct banks while each
// 16 threads still load 256 contiguous bytes of a column.  NCH (32-column
// chunks per leaf) is a template argument: the diagonal pipeline is fully
// unrolled and no load inside it is conditional.  The low-rank part (K over
// the leaf's (level, rank) pairs, A = rows of U_b, B = W_b) follows as
// further chunks of at most 32, trimmed to whole MFMA k-steps; the first
// one is fetched while the last diagonal chunk computes.
// Requires (host-checked): uniform leaves of 32 NCH rows, 32 NW | leaf,
// no transpose, X column-major (x_rs = 1) with 16-byte aligned columns,
// R a multiple of 64.
static constexpr int LEAF_KC = 32, LEAF_TN = 64;
typedef double d2 __attribute__((ext_vector_type(2)));

template <int NW, int NCH>
__global__ __launch_bounds__(NW * 64) void k_leaf_apply(ProdArgs a, int rtiles, int ntiles) {
  constexpr int NT = NW * 64;
  constexpr int KC = LEAF_KC, TN = LEAF_TN, M = KC * NCH;
  constexpr int XPAIRS = KC / 2 * TN;
  constexpr int XPT = XPAIRS / NT;
  static_assert(XPAIRS % NT == 0 && M % (32 * NW) == 0, "leaf tile shape");
  __shared__ d2 xs[2][XPAIRS];
  __shared__ LevelInfo li[MAX_LEVELS];
  __shared__ int lp[MAX_LEVELS + 1];
  __shared__ const double *tabA[2][KC];  // two tables: chunk c uses slot c & 1
  __shared__ const double *tabB[2][KC];
  __shared__ int tabS[2][KC];
  int c, rt, nt;
  if (!apply_tile_of(blockIdx.x, a.nleaf, rtiles, ntiles, c, rt, nt)) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i16 = lane & 15, lk = lane >> 4;
  const int lo = a.leaf_off[c];
  const int wr = rt * (32 * NW) + 32 * w;  // this wave's first row in the leaf
  const int n0 = nt * TN;
  const double *pA = a.diag + a.diag_off[c] + wr + i16 + (long long)lk * M;
  // X pairs of this thread: p = t & 15 (row pair), columns j_e = (t >> 4) + e NT / 16
  const int xp = t & 15;
  const double *pX = a.X + lo + 2 * xp + (long long)(n0 + (t >> 4)) * a.x_cs;
  const long long xes = (long long)(NT / 16) * a.x_cs;

  d4 acc[2][4];
#pragma unroll
  for (int f = 0; f < 2; f++)
#pragma unroll
    for (int g = 0; g < 4; g++) acc[f][g] = d4{0.0, 0.0, 0.0, 0.0};
  double ar[2][16];
  d2 xr[XPT];
  double lr[16], lr2[16];  // low-rank operands of chunks 0, 2, 4 ... / 1, 3, ...
  d2 lb[XPT], lb2[XPT];
  auto loadA = [&](int ch, double *dst) {
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
#pragma unroll
      for (int f = 0; f < 2; f++) dst[ks * 2 + f] = pA[16 * f + (long long)(KC * ch + 4 * ks) * M];
  };
  auto loadX = [&](int ch) {
#pragma unroll
    for (int e = 0; e < XPT; e++)
      xr[e] = *reinterpret_cast<const d2 *>(pX + KC * ch + e * xes);
  };
  auto slot = [&](int e) {
    const int j = (t >> 4) + e * (NT / 16);
    return j * 16 + (xp ^ (j & 15));
  };
  // k-steps 0 .. nks-1 of the chunk in LDS buffer buf (nks = 8: all)
  auto compute = [&](int buf, const double *av, int nks) {
    const double *xb = reinterpret_cast<const double *>(xs[buf]);
#pragma unroll
    for (int ks = 0; ks < 8; ks++) {
      if (ks < nks) {
        const int pp = 2 * ks + (lk >> 1);
        double b[4];
#pragma unroll
        for (int g = 0; g < 4; g++)
          b[g] = xb[((16 * g + i16) * 16 + (pp ^ i16)) * 2 + (lk & 1)];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
          for (int g = 0; g < 4; g++) acc[f][g] = mfma_f64(av[ks * 2 + f], b[g], acc[f][g]);
      }
    }
  };
  // low-rank chunk operands from the table (every entry first, then every
  // load, unconditional from a valid stand-in address where empty)
  auto loadLR = [&](int ts, double *lr, d2 *lb) {
    const double *qa[8];
#pragma unroll
    for (int ks = 0; ks < 8; ks++) qa[ks] = tabA[ts][4 * ks + lk];
    const double *q0 = tabB[ts][2 * xp], *q1 = tabB[ts][2 * xp + 1];
    const long long s0 = tabS[ts][2 * xp], s1 = tabS[ts][2 * xp + 1];
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
#pragma unroll
      for (int f = 0; f < 2; f++)
        lr[ks * 2 + f] = ldg(qa[ks] ? qa[ks] + wr + 16 * f + i16 : a.diag);
#pragma unroll
    for (int e = 0; e < XPT; e++) {
      const int j = n0 + (t >> 4) + e * (NT / 16);
      lb[e].x = ldg(q0 ? q0 + j * s0 : a.diag);
      lb[e].y = ldg(q1 ? q1 + j * s1 : a.diag);
    }
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
#pragma unroll
      for (int f = 0; f < 2; f++)
        if (!qa[ks]) lr[ks * 2 + f] = 0.0;
#pragma unroll
    for (int e = 0; e < XPT; e++) {
      if (!q0) lb[e].x = 0.0;
      if (!q1) lb[e].y = 0.0;
    }
  };
  // table of low-rank chunk k0: per K index its U column at the leaf's
  // first row, its W_b row and W_b's leading dimension (threads 0..31)
  auto table = [&](int ts, int k0, int K2) {
    if (t < KC) {
      const int kap = k0 + t;
      const double *qa = nullptr, *qb = nullptr;
      int st = 0;
      if (kap < K2) {
        int l = 0;
        while (l + 1 < a.h && kap >= lp[l + 1]) l++;
        const int kk = kap - lp[l];
        qa = li[l].F + (long long)kk * li[l].ld;
        qb = li[l].P + kk;
        st = li[l].s;
      }
      tabA[ts][t] = qa;
      tabB[ts][t] = qb;
      tabS[ts][t] = st;
    }
  };

  // metadata of block (l, q) holding the leaf's rows at level l (thread
  // l - 1): its loads go out before the first diagonal chunk's, so waiting
  // for them does not wait for the chunk
  const int lvl = t + 1;
  BlockMeta bm;
  int qfirst = 0, gsel = 0;
  long long wo = 0;
  if (lvl <= a.h) {
    const int q = c >> (a.h - lvl);
    gsel = blk_index(lvl, q);
    bm = a.blk[gsel];
    qfirst = a.leaf_off[q << (a.h - lvl)];
    wo = a.wofs[gsel];
  }
  loadX(0);
  loadA(0, ar[0]);
  if (lvl <= a.h) {
    LevelInfo x;
    x.F = a.U + bm.uoff + (lo - qfirst);
    x.P = a.W + wo * a.R;
    x.ld = bm.m;
    x.s = bm.s;
    x.g = gsel;
    x.pad = 0;
    li[lvl - 1] = x;
  }
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int l = 0; l < a.h; l++) {
      lp[l] = run;
      run += li[l].s;
    }
    lp[a.h] = run;
  }
  __syncthreads();
  const int K2 = lp[a.h];
  table(0, 0, K2);

  // diagonal part: D_c (rows of this wave) X_c, chunk by chunk
#pragma unroll
  for (int e = 0; e < XPT; e++) xs[0][slot(e)] = xr[e];
  if (NCH > 1) loadX(1);
  __syncthreads();
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    if (ch + 1 < NCH) loadA(ch + 1, ar[(ch + 1) & 1]);
    if (ch + 1 == NCH) loadLR(0, lr, lb);  // unconditional: an empty table loads stand-ins
    compute(ch & 1, ar[ch & 1], 8);
    if (ch + 1 < NCH) {
#pragma unroll
      for (int e = 0; e < XPT; e++) xs[(ch + 1) & 1][slot(e)] = xr[e];
      if (ch + 2 < NCH) loadX(ch + 2);
      __syncthreads();
    }
  }

  // low-rank part, chunks of 32 alternating between the LDS buffers (chunk
  // 0 in the one the last diagonal chunk did not use): each step builds the
  // next chunk's table, stages its own operands, and after the one barrier
  // issues the next chunk's loads (always: an empty table loads stand-ins)
  // before computing, so each chunk's loads have a compute phase to land.
  // A buffer or table slot is rewritten two steps after its last use, with
  // a barrier in between.
  auto lr_step = [&](int k0, int buf, double *lrc, d2 *lbc, double *lrn, d2 *lbn) {
    const int tn = ((k0 / KC) + 1) & 1;
    table(tn, k0 + KC, K2);
#pragma unroll
    for (int e = 0; e < XPT; e++) xs[buf][slot(e)] = lbc[e];
    __syncthreads();
    loadLR(tn, lrn, lbn);
    compute(buf, lrc, (min(KC, K2 - k0) + 3) / 4);
  };
  for (int k0 = 0; k0 < K2; k0 += 2 * KC) {
    lr_step(k0, NCH & 1, lr, lb, lr2, lb2);
    if (k0 + KC >= K2) break;
    lr_step(k0 + KC, (NCH & 1) ^ 1, lr2, lb2, lr, lb);
  }
#pragma unroll
  for (int f = 0; f < 2; f++)
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = wr + 16 * f + lk + 4 * v;
        const int j = n0 + 16 * g + i16;
        a.Y[(long long)(lo + i) * a.y_rs + (long long)j * a.y_cs] = acc[f][g][v];
      }
}

// Two-pass form for leaves of 128 or 256 rows whose low-rank K is at most
// 32: all of X_c stays in LDS (32 NCH rows x 64 columns, then W in one
// more 32-row buffer), and the 8 waves sweep the leaf in passes of 128
// rows (16 rows and all 64 columns per wave), so the stores of one pass
// drain while the next pass computes.  The D stream runs one chunk ahead
// across pass boundaries; only the first pass waits on barriers (X_c
// arriving chunk by chunk).  Same LDS slot layout as k_leaf_apply.
template <int NCH>
__global__ __launch_bounds__(512) void k_leaf_apply2(ProdArgs a, int ntiles) {
  constexpr int NT = 512, KC = LEAF_KC, TN = LEAF_TN, M = KC * NCH;
  constexpr int NPASS = M / 128, XPAIRS = KC / 2 * TN, XPT = XPAIRS / NT;
  static_assert(M % 128 == 0 && XPAIRS % NT == 0, "leaf tile shape");
  __shared__ d2 xs[NCH + 1][XPAIRS];
  __shared__ LevelInfo li[MAX_LEVELS];
  __shared__ int lp[MAX_LEVELS + 1];
  __shared__ const double *tabA[KC];
  __shared__ const double *tabB[KC];
  __shared__ int tabS[KC];
  int c, rt, nt;
  if (!apply_tile_of(blockIdx.x, a.nleaf, 1, ntiles, c, rt, nt)) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i16 = lane & 15, lk = lane >> 4;
  const int lo = a.leaf_off[c];
  const int n0 = nt * TN;
  // pass p: this wave's rows are 128 p + 16 w .. + 15
  const double *pD = a.diag + a.diag_off[c] + 16 * w + i16 + (long long)lk * M;
  const int xp = t & 15;
  const double *pX = a.X + lo + 2 * xp + (long long)(n0 + (t >> 4)) * a.x_cs;
  const long long xes = (long long)(NT / 16) * a.x_cs;

  d4 acc[4];
#pragma unroll
  for (int g = 0; g < 4; g++) acc[g] = d4{0.0, 0.0, 0.0, 0.0};
  double ar[2][8];
  d2 xr[XPT];
  double lr[8];
  d2 lb[XPT];
  auto loadA = [&](int it, double *dst) {  // iteration it = pass * NCH + chunk
    const int p = it / NCH, ch = it % NCH;
#pragma unroll
    for (int ks = 0; ks < 8; ks++) dst[ks] = pD[128 * p + (long long)(KC * ch + 4 * ks) * M];
  };
  auto loadX = [&](int ch) {
#pragma unroll
    for (int e = 0; e < XPT; e++)
      xr[e] = *reinterpret_cast<const d2 *>(pX + KC * ch + e * xes);
  };
  auto slot = [&](int e) {
    const int j = (t >> 4) + e * (NT / 16);
    return j * 16 + (xp ^ (j & 15));
  };
  auto compute = [&](int buf, const double *av, int nks) {
    const double *xb = reinterpret_cast<const double *>(xs[buf]);
#pragma unroll
    for (int ks = 0; ks < 8; ks++) {
      if (ks < nks) {
        const int pp = 2 * ks + (lk >> 1);
        double b[4];
#pragma unroll
        for (int g = 0; g < 4; g++)
          b[g] = xb[((16 * g + i16) * 16 + (pp ^ i16)) * 2 + (lk & 1)];
#pragma unroll
        for (int g = 0; g < 4; g++) acc[g] = mfma_f64(av[ks], b[g], acc[g]);
      }
    }
  };
  // U rows of pass p for the low-rank chunk (unconditional loads from a
  // valid stand-in address where the table entry is empty)
  auto loadLRA = [&](int p) {
    const double *qa[8];
#pragma unroll
    for (int ks = 0; ks < 8; ks++) qa[ks] = tabA[4 * ks + lk];
#pragma unroll
    for (int ks = 0; ks < 8; ks++) lr[ks] = ldg(qa[ks] ? qa[ks] + 128 * p + 16 * w + i16 : a.diag);
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
      if (!qa[ks]) lr[ks] = 0.0;
  };
  auto loadLRB = [&]() {
    const double *q0 = tabB[2 * xp], *q1 = tabB[2 * xp + 1];
    const long long s0 = tabS[2 * xp], s1 = tabS[2 * xp + 1];
#pragma unroll
    for (int e = 0; e < XPT; e++) {
      const int j = n0 + (t >> 4) + e * (NT / 16);
      lb[e].x = ldg(q0 ? q0 + j * s0 : a.diag);
      lb[e].y = ldg(q1 ? q1 + j * s1 : a.diag);
    }
#pragma unroll
    for (int e = 0; e < XPT; e++) {
      if (!q0) lb[e].x = 0.0;
      if (!q1) lb[e].y = 0.0;
    }
  };

  const int lvl = t + 1;
  BlockMeta bm;
  int qfirst = 0, gsel = 0;
  long long wo = 0;
  if (lvl <= a.h) {
    const int q = c >> (a.h - lvl);
    gsel = blk_index(lvl, q);
    bm = a.blk[gsel];
    qfirst = a.leaf_off[q << (a.h - lvl)];
    wo = a.wofs[gsel];
  }
  loadX(0);
  loadA(0, ar[0]);
  if (lvl <= a.h) {
    LevelInfo x;
    x.F = a.U + bm.uoff + (lo - qfirst);
    x.P = a.W + wo * a.R;
    x.ld = bm.m;
    x.s = bm.s;
    x.g = gsel;
    x.pad = 0;
    li[lvl - 1] = x;
  }
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int l = 0; l < a.h; l++) {
      lp[l] = run;
      run += li[l].s;
    }
    lp[a.h] = run;
  }
  __syncthreads();
  const int K2 = lp[a.h];  // at most 32 (host-checked)
  if (t < KC) {
    const double *qa = nullptr, *qb = nullptr;
    int st = 0;
    if (t < K2) {
      int l = 0;
      while (l + 1 < a.h && t >= lp[l + 1]) l++;
      const int kk = t - lp[l];
      qa = li[l].F + (long long)kk * li[l].ld;
      qb = li[l].P + kk;
      st = li[l].s;
    }
    tabA[t] = qa;
    tabB[t] = qb;
    tabS[t] = st;
  }
#pragma unroll
  for (int e = 0; e < XPT; e++) xs[0][slot(e)] = xr[e];
  if (NCH > 1) loadX(1);
  __syncthreads();

  const int nks_lr = (K2 + 3) / 4;
#pragma unroll
  for (int p = 0; p < NPASS; p++) {
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
      const int it = p * NCH + ch;
      if (it + 1 < NPASS * NCH) loadA(it + 1, ar[(it + 1) & 1]);
      if (ch + 1 == NCH) {
        loadLRA(p);
        if (p == 0) loadLRB();
      }
      compute(ch, ar[it & 1], 8);
      if (p == 0 && ch + 1 < NCH) {
#pragma unroll
        for (int e = 0; e < XPT; e++) xs[ch + 1][slot(e)] = xr[e];
        if (ch + 2 < NCH) loadX(ch + 2);
        __syncthreads();
      }
    }
    if (p == 0) {
#pragma unroll
      for (int e = 0; e < XPT; e++) xs[NCH][slot(e)] = lb[e];
      __syncthreads();
    }
    compute(NCH, lr, nks_lr);
#pragma unroll
    for (int g = 0; g < 4; g++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = 128 * p + 16 * w + lk + 4 * v;
        const int j = n0 + 16 * g + i16;
        a.Y[(long long)(lo + i) * a.y_rs + (long long)j * a.y_cs] = acc[g][v];
      }
      acc[g] = d4{0.0, 0.0, 0.0, 0.0};
    }
  }
}

// Apply-kernel variants: (WM, FM, FN, KC) -> tile TM = WM*FM*16 rows x
// TN = (4/WM)*FN*16 columns, K-chunk KC.
struct ApplyVariant {
  int tm, tn, kc;
  void (*launch)(const ProdArgs &, bool nfast, bool fast, int rtiles, int ntiles,
                 int grid, hipStream_t s);
};

template <int WM, int FM, int FN, int KC>
static void launch_apply(const ProdArgs &a, bool nfast, bool fast, int rtiles, int ntiles,
                         int grid, hipStream_t s) {
  if (fast) {
    if (nfast)
      hipLaunchKernelGGL((k_mm_apply<true, WM, FM, FN, KC, true>), dim3(grid), dim3(256),
                         0, s, a, rtiles, ntiles);
    else
      hipLaunchKernelGGL((k_mm_apply<false, WM, FM, FN, KC, true>), dim3(grid), dim3(256),
                         0, s, a, rtiles, ntiles);
  } else {
    if (nfast)
      hipLaunchKernelGGL((k_mm_apply<true, WM, FM, FN, KC, false>), dim3(grid), dim3(256),
                         0, s, a, rtiles, ntiles);
    else
      hipLaunchKernelGGL((k_mm_apply<false, WM, FM, FN, KC, false>), dim3(grid),
                         dim3(256), 0, s, a, rtiles, ntiles);
  }
}

// Measured at N = 65536, leaf 256, R = 64: 128 x 64 with four 32 x 64 wave
// tiles (60.6 us) and 64 x 64 (62.9 us) lead; smaller tiles (more waves per
// SIMD), one 64 x 64 block per wave (one wave per SIMD), K-chunks of 32 at
// 128 x 64 and non-temporal loads of D were all slower.
static const ApplyVariant kVariants[] = {
    {64, 64, 16, launch_apply<2, 2, 2, 16>},   // 64 x 64, KC 16
    {64, 64, 32, launch_apply<2, 2, 2, 32>},   // 64 x 64, KC 32
    {128, 64, 32, launch_apply<2, 4, 2, 32>},  // 128 x 64, KC 32
    {128, 32, 32, launch_apply<4, 2, 2, 32>},  // 128 x 32, KC 32
    {128, 64, 16, launch_apply<4, 2, 4, 16>},  // 128 x 64, KC 16, 32 x 64 per wave
};
static constexpr int kTileVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// Variants kTileVariants and kTileVariants + 1 are the leaf-resident
// kernels (one chunk ring / X_c resident in two passes); on trees they do
// not accept they run variant 0.
static constexpr int kNumVariants = kTileVariants + 2;

// Leaf size the leaf-resident kernel takes this product at (0: not eligible).
static int leaf_apply_size(const hm_dtree *t, const ProdArgs &a, bool nfast) {
  if (a.trans || nfast || a.R % LEAF_TN != 0 || t->nleaf <= 0) return 0;
  if ((reinterpret_cast<uintptr_t>(a.X) & 15) != 0 || (a.x_cs & 1) != 0) return 0;
  const int m = t->leaf_off[1] - t->leaf_off[0];
  if (m != 128 && m != 256 && m != 512) return 0;
  for (int c = 1; c < t->nleaf; c++)
    if (t->leaf_off[c + 1] - t->leaf_off[c] != m) return 0;
  return m;
}

template <int NW, int NCH>
static void launch_leaf(const hm_dtree *t, const ProdArgs &a, hipStream_t s) {
  const int rtiles = (LEAF_KC * NCH) / (32 * NW);
  const int ntiles = a.R / LEAF_TN;
  const int grid = ((t->nleaf + 7) / 8) * 8 * rtiles * ntiles;
  hipLaunchKernelGGL((k_leaf_apply<NW, NCH>), dim3(grid), dim3(NW * 64), 0, s, a, rtiles,
                     ntiles);
}

// Largest low-rank K of any leaf: the sum of its ancestors' block ranks.
static int max_leaf_rank_sum(const hm_dtree *t) {
  int best = 0;
  for (int c = 0; c < t->nleaf; c++) {
    int k = 0;
    for (int l = 1; l <= t->height; l++) k += t->blk[blk_index(l, c >> (t->height - l))].s;
    best = std::max(best, k);
  }
  return best;
}

static void run_apply(hm_dtree *t, const ProdArgs &a, int v, bool nfast, hipStream_t s) {
  if (v == kTileVariants + 1) {
    const int m = leaf_apply_size(t, a, nfast);
    if ((m == 128 || m == 256) && max_leaf_rank_sum(t) <= LEAF_KC) {
      const int ntiles = a.R / LEAF_TN;
      const int grid = ((t->nleaf + 7) / 8) * 8 * ntiles;
      if (m == 128)
        hipLaunchKernelGGL(k_leaf_apply2<4>, dim3(grid), dim3(512), 0, s, a, ntiles);
      else
        hipLaunchKernelGGL(k_leaf_apply2<8>, dim3(grid), dim3(512), 0, s, a, ntiles);
      return;
    }
    v = 0;
  }
  if (v == kTileVariants) {
    switch (leaf_apply_size(t, a, nfast)) {
      case 128: launch_leaf<4, 4>(t, a, s); return;
      case 256: launch_leaf<8, 8>(t, a, s); return;
      case 512: launch_leaf<8, 16>(t, a, s); return;
      default: v = 0;
    }
  }
  const ApplyVariant &av = kVariants[v];
  const int rtiles = (t->max_leaf + av.tm - 1) / av.tm;
  const int ntiles = (a.R + av.tn - 1) / av.tn;
  const int grid = ((t->nleaf + 7) / 8) * 8 * rtiles * ntiles;
  bool fast = !a.trans && (a.R % av.tn == 0);
  for (int c = 0; fast && c < t->nleaf; c++) {
    const int m = t->leaf_off[c + 1] - t->leaf_off[c];
    fast = (m % av.tm == 0) && (m % av.kc == 0);
  }
  av.launch(a, nfast, fast, rtiles, ntiles, grid, s);
}

static int r_bucket(int R) { return R <= 16 ? 0 : (R <= 128 ? 1 : 2); }

// Pick the fastest apply variant for this tree, transpose mode and RHS
// bucket by timing each once (HIP events) on the live operands.  HM_TUNE=0
// disables tuning (variant 1), HM_APPLY_VARIANT=v forces one.  With
// no_tune (the allocation-free product) an untuned tree runs variant 1
// and nothing is recorded.
static int choose_variant(hm_dtree *t, const ProdArgs &a, bool nfast, hipStream_t s,
                          bool no_tune = false) {
  int &slot = t->plan.apply_variant[a.trans][r_bucket(a.R)];
  if (slot >= 0) return slot;
  const char *force = getenv("HM_APPLY_VARIANT");
  if (force) {
    slot = std::max(0, std::min(kNumVariants - 1, atoi(force)));
    return slot;
  }
  if (no_tune) return 1;
  const char *tune = getenv("HM_TUNE");
  if (tune && tune[0] == '0') {
    slot = 1;
    return slot;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    slot = 1;
    return slot;
  }
  float best = 1e30f;
  int bestv = 1;
  for (int v = 0; v < kNumVariants; v++) {
    run_apply(t, a, v, nfast, s);  // warm
    (void)hipEventRecord(e0, s);
    for (int rep = 0; rep < 3; rep++) run_apply(t, a, v, nfast, s);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) {
      best = ms;
      bestv = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  slot = bestv;
  return slot;
}

// W path per transpose mode: the chunk path (partials of every level per
// 64-row chunk, X read once) suits small rank sums; the block path (split-K
// GEMM per block) keeps the partial buffer small when ranks are large.
// HM_WPATH=0/1 forces one.
static int choose_wpath(const hm_dtree *t, int trans) {
  const char *force = getenv("HM_WPATH");
  if (force && (force[0] == '0' || force[0] == '1')) return force[0] - '0';
  return t->plan.s_in_max[trans] > 64 ? 1 : 0;
}

static int run_block_w(hm_dtree *t, const ProdArgs &a, bool nfast, hipStream_t s) {
  const ProductPlan &p = t->plan;
  const int tr = a.trans;
  if (p.witems[tr].empty()) return SUCCESS;
  dim3 grid((unsigned)p.witems[tr].size(), (unsigned)((a.R + WTileT::TN - 1) / WTileT::TN));
  Timed tp = timing_begin("w_gemm", s);
  if (nfast)
    hipLaunchKernelGGL(k_w_gemm<true>, grid, dim3(256), 0, s, a, p.d_witems[tr],
                       p.d_wpofs[tr]);
  else
    hipLaunchKernelGGL(k_w_gemm<false>, grid, dim3(256), 0, s, a, p.d_witems[tr],
                       p.d_wpofs[tr]);
  timing_end(tp, s);
  HM_LAUNCH_CHECK();
  if (!p.wmulti[tr].empty()) {
    const long long E = (long long)p.wmulti_smax[tr] * a.R;
    const int gy = (int)std::min<long long>(64, (E + 255) / 256);
    Timed tr_ = timing_begin("w_reduce", s);
    hipLaunchKernelGGL(k_w_reduce, dim3((unsigned)p.wmulti[tr].size(), gy), dim3(256), 0, s,
                       a, p.d_wmulti[tr], p.d_wpofs[tr]);
    timing_end(tr_, s);
    HM_LAUNCH_CHECK();
  }
  return SUCCESS;
}

// HM_MV_FUSE=0 forces H x through the reduction launches (for tests that
// compare both paths).
static bool mv_fuse_disabled() {
  const char *e = getenv("HM_MV_FUSE");
  return e && e[0] == '0';
}

static int run_reduce(hm_dtree *t, const ProdArgs &a, hipStream_t s) {
  const ProductPlan &p = t->plan;
  const int tr_ = a.trans ? 1 : 0;
  Timed tr = timing_begin("reduce_w", s);
  if (!p.red_seg[tr_].empty())
    hipLaunchKernelGGL(k_reduce_items, dim3((unsigned)p.red_seg[tr_].size()), dim3(256), 0, s,
                       a, p.d_red_seg[tr_], 0);
  if (!p.red_w[tr_].empty())
    hipLaunchKernelGGL(k_reduce_items, dim3((unsigned)p.red_w[tr_].size()), dim3(256), 0, s,
                       a, p.d_red_w[tr_], 1);
  timing_end(tr, s);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

int products_diag(hm_dtree *t, int trans, int R, const double *X, long x_rs, long x_cs,
                  double *Y, long y_rs, long y_cs, hipStream_t s) {
  if (R <= 0) return SUCCESS;
  HM_OK(ensure_scratch(t, R));
  ProdArgs a = make_args(t, trans ? 1 : 0, R, X, x_rs, x_cs, Y, y_rs, y_cs);
  a.h = 0;  // no off-diagonal levels: the tile computes op(D_c) X_c only
  Timed ta = timing_begin("mm_diag", s);
  run_apply(t, a, 0, x_rs != 1, s);
  timing_end(ta, s);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

// ext_scratch: product scratch supplied by the caller (the allocation-free
// entry point), else the tree's own, grown on demand.
static int products_run_ex(hm_dtree *t, int trans, int R, const double *X, long x_rs,
                           long x_cs, double *Y, long y_rs, long y_cs, hipStream_t s,
                           double *ext_scratch);

int products_run(hm_dtree *t, int trans, int R, const double *X, long x_rs,
                 long x_cs, double *Y, long y_rs, long y_cs, hipStream_t s) {
  return products_run_ex(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs, s, nullptr);
}

static int products_run_ex(hm_dtree *t, int trans, int R, const double *X, long x_rs,
                           long x_cs, double *Y, long y_rs, long y_cs, hipStream_t s,
                           double *ext_scratch) {
  if (!t) {
    set_error("product: NULL device tree");
    return INPUT_ERROR;
  }
  if (R <= 0) return SUCCESS;
  if (!X || !Y) {
    set_error("product: NULL operand");
    return INPUT_ERROR;
  }
  if (t->slices > 0) {
    if (ext_scratch) {
      set_error("allocation-free product: trees with sliced (mixed-precision) factors "
                "are not supported (hm_hip_tree_set_precision(tree, 0, ...) first)");
      return INPUT_ERROR;
    }
    return mixed_matmat(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs, s);
  }
  if (!ext_scratch) HM_OK(ensure_scratch(t, R));
  ProdArgs a = make_args(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs, ext_scratch);
  const int nchunk = t->plan.nchunk;
  if (R == 1 && !trans) {
    if (nchunk > 0) {
      Timed tp = timing_begin("mv_partials", s);
      hipLaunchKernelGGL(k_mv_partials, dim3(nchunk), dim3(256), 0, s, a);
      timing_end(tp, s);
      HM_LAUNCH_CHECK();
    }
    const bool fused = t->plan.mv_fused && !mv_fuse_disabled();
    if (!fused) HM_OK(run_reduce(t, a, s));
    const int rtiles = (t->max_leaf + 63) / 64;
    Timed ta = timing_begin("mv_apply", s);
    if (fused)
      hipLaunchKernelGGL(k_mv_apply<true>, dim3(t->nleaf * rtiles), dim3(512), 0, s, a, rtiles);
    else
      hipLaunchKernelGGL(k_mv_apply<false>, dim3(t->nleaf * rtiles), dim3(256), 0, s, a, rtiles);
    timing_end(ta, s);
    HM_LAUNCH_CHECK();
    return SUCCESS;
  }
  const bool nfast = (x_rs != 1);
  const int smax = t->plan.s_in_max[trans];
  if (smax > 0 && nchunk > 0 && choose_wpath(t, trans) == 1) {
    HM_OK(run_block_w(t, a, nfast, s));
  } else if (smax > 0 && nchunk > 0) {
    const int ntiles = (R + 63) / 64;
    const int mtiles = (smax + 15) / 16;
    dim3 grid((unsigned)(nchunk * mtiles * ntiles));
    Timed tp = timing_begin("mm_partials", s);
    if (nfast)
      hipLaunchKernelGGL(k_mm_partials<true>, grid, dim3(256), 0, s, a, mtiles, ntiles);
    else
      hipLaunchKernelGGL(k_mm_partials<false>, grid, dim3(256), 0, s, a, mtiles, ntiles);
    timing_end(tp, s);
    HM_LAUNCH_CHECK();
    HM_OK(run_reduce(t, a, s));
  }
  const int v = choose_variant(t, a, nfast, s, ext_scratch != nullptr);
  Timed ta = timing_begin("mm_apply", s);
  run_apply(t, a, v, nfast, s);
  timing_end(ta, s);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

// ------------------------------------------------ batched tree products
// Apply tile shapes of a batch; the narrow ones serve small RHS counts,
// whose 64-wide tiles would be mostly padding.
struct BatchApplyVariant {
  int tm, tn;
  void (*launch)(bool nfast, unsigned grid, const BatchProd *bp, const int4 *wl,
                 hipStream_t s);
};

template <int WM, int WN, int FM, int FN, int KC>
static void launch_apply_b(bool nfast, unsigned grid, const BatchProd *bp, const int4 *wl,
                           hipStream_t s) {
  typedef MfmaTile<WM, WN, FM, FN, KC> Tile;
  if (nfast)
    hipLaunchKernelGGL((k_mm_apply_b<true, Tile>), dim3(grid), dim3(Tile::NT), 0, s, bp, wl);
  else
    hipLaunchKernelGGL((k_mm_apply_b<false, Tile>), dim3(grid), dim3(Tile::NT), 0, s, bp, wl);
}

// (WM, WN, FM, FN, KC): WM x WN waves, tile 16 WM FM x 16 WN FN.
static const BatchApplyVariant kBatchVariants[] = {
    {64, 64, launch_apply_b<2, 2, 2, 2, 16>},   // 64 x 64
    {128, 64, launch_apply_b<4, 1, 2, 4, 16>},  // 128 x 64, 32 x 64 per wave
    {128, 32, launch_apply_b<4, 1, 2, 2, 32>},  // 128 x 32, KC 32
    {64, 32, launch_apply_b<2, 2, 2, 1, 16>},   // 64 x 32
    {64, 16, launch_apply_b<4, 1, 1, 1, 16>},   // 64 x 16
    {32, 32, launch_apply_b<1, 1, 2, 2, 16>},   // 32 x 32, one wave
};
static constexpr int kBatchVariantCount =
    sizeof(kBatchVariants) / sizeof(kBatchVariants[0]);

bool tree_prod_batchable(const hm_dtree *t, int trans, bool diag_only) {
  if (!t || t->nleaf <= 0) return false;
  if (diag_only) return true;
  return t->slices == 0 && t->height >= 1 && t->plan.nchunk > 0 &&
         choose_wpath(t, trans ? 1 : 0) == 0;
}

// Bytes spanned by a rows x cols operand with element strides rs, cs.
static size_t span_bytes(long long rows, long long cols, long long rs, long long cs) {
  if (rows <= 0 || cols <= 0) return 0;
  return (size_t)((rows - 1) * rs + (cols - 1) * cs + 1) * sizeof(double);
}

int tree_prod_add(Program &p, Arena &ar, const std::vector<TreeProdJob> &jobs, bool nfast) {
  if (jobs.empty()) return SUCCESS;
  const int nj = (int)jobs.size();
  std::vector<BatchProd> bp(nj);
  std::vector<int4> pw[2];  // partials work: 64-column tiles, 16-column tiles
  std::vector<int2> rw[2];
  for (int i = 0; i < nj; i++) {
    const TreeProdJob &j = jobs[i];
    hm_dtree *t = j.t;
    const int tr = j.trans ? 1 : 0;
    const long long need = j.diag_only ? 0 : scratch_need(t, j.R);
    double *scr = j.diag_only ? nullptr : ar.take<double>((size_t)need);
    BatchProd &b = bp[i];
    memset(&b, 0, sizeof(b));
    b.a = make_args(t, tr, j.R, j.X, j.x_rs, j.x_cs, j.Y, j.y_rs, j.y_cs, scr);
    if (j.diag_only) b.a.h = 0;  // no off-diagonal levels: op(D_c) X_c only
    b.red[0] = t->plan.d_red_seg[tr];
    b.red[1] = t->plan.d_red_w[tr];
    b.mtiles = t->plan.part_mtiles[tr];
    long long u = 0, v = 0;
    for (const auto &blk : t->blk) {
      u += (long long)blk.m * blk.s;
      v += (long long)blk.n * blk.s;
    }
    p.touch(j.X, span_bytes(t->n, j.R, j.x_rs, j.x_cs), "batch tree products");
    p.touch(j.Y, span_bytes(t->n, j.R, j.y_rs, j.y_cs), "batch tree products");
    p.touch(t->d_diag, (size_t)t->diag_off[t->nleaf] * sizeof(double), "batch tree products");
    if (j.diag_only) continue;
    p.touch(scr, (size_t)need * sizeof(double), "batch tree products");
    p.touch(t->d_U, (size_t)u * sizeof(double), "batch tree products");
    p.touch(t->d_V, (size_t)v * sizeof(double), "batch tree products");
    if (t->plan.s_in_max[tr] <= 0) continue;
    // same order as the single-tree partials grid: chunk, rank tile, RHS tile
    for (int v = 0; v < 2; v++) {
      const int nt = v ? (j.R + 15) / 16 : (j.R + 63) / 64;
      for (int k = 0; k < t->plan.nchunk; k++)
        for (int mt = 0; mt < b.mtiles; mt++)
          for (int n = 0; n < nt; n++) pw[v].push_back(make_int4(i, k, mt, n));
    }
    for (size_t q = 0; q < t->plan.red_seg[tr].size(); q++) rw[0].push_back(make_int2(i, (int)q));
    for (size_t q = 0; q < t->plan.red_w[tr].size(); q++) rw[1].push_back(make_int2(i, (int)q));
  }
  const size_t bp_off = p.push(bp);
  const size_t pw0_off = p.push(pw[0]), pw1_off = p.push(pw[1]);
  const size_t r0_off = p.push(rw[0]), r1_off = p.push(rw[1]);
  const unsigned npw0 = (unsigned)pw[0].size(), npw1 = (unsigned)pw[1].size();
  const unsigned nr0 = (unsigned)rw[0].size(), nr1 = (unsigned)rw[1].size();
  if (npw0) {
    auto pchoice = std::make_shared<int>(1);
    p.tunables.push_back({p.ops.size(), 2, pchoice});
    p.add("batch partials", [=](hipStream_t s, const unsigned char *blob) {
      const BatchProd *B = reinterpret_cast<const BatchProd *>(blob + bp_off);
      const int4 *W0 = reinterpret_cast<const int4 *>(blob + pw0_off);
      const int4 *W1 = reinterpret_cast<const int4 *>(blob + pw1_off);
      if (*pchoice == 0) {
        if (nfast)
          hipLaunchKernelGGL((k_mm_partials_b<true, PartTile64>), dim3(npw0), dim3(256), 0, s,
                             B, W0);
        else
          hipLaunchKernelGGL((k_mm_partials_b<false, PartTile64>), dim3(npw0), dim3(256), 0, s,
                             B, W0);
      } else {
        if (nfast)
          hipLaunchKernelGGL((k_mm_partials_b<true, PartTile16>), dim3(npw1), dim3(64), 0, s,
                             B, W1);
        else
          hipLaunchKernelGGL((k_mm_partials_b<false, PartTile16>), dim3(npw1), dim3(64), 0, s,
                             B, W1);
      }
    });
  }
  if (nr0 || nr1)
    p.add("batch reduce", [=](hipStream_t s, const unsigned char *blob) {
      const BatchProd *B = reinterpret_cast<const BatchProd *>(blob + bp_off);
      if (nr0)
        hipLaunchKernelGGL(k_reduce_items_b, dim3(nr0), dim3(64), 0, s, B,
                           reinterpret_cast<const int2 *>(blob + r0_off), 0);
      if (nr1)
        hipLaunchKernelGGL(k_reduce_items_b, dim3(nr1), dim3(64), 0, s, B,
                           reinterpret_cast<const int2 *>(blob + r1_off), 1);
    });
  // apply: one launch per RHS width group, tile shape tuned per group; the
  // leaves of a group are dealt to the eight XCD queues in turn (workgroup
  // b runs on XCD b % 8) with all tiles of one leaf in one queue, so X_c
  // and its W_b are shared through one L2
  static const int kDefault[3] = {4, 3, 0};
  static const char *kName[3] = {"batch apply R<=16", "batch apply R<=32", "batch apply R>32"};
  for (int g = 0; g < 3; g++) {
    std::vector<int> mem;
    for (int i = 0; i < nj; i++) {
      const int R = jobs[i].R;
      if ((R <= 16 ? 0 : (R <= 32 ? 1 : 2)) == g) mem.push_back(i);
    }
    if (mem.empty()) continue;
    size_t toff[kBatchVariantCount];
    unsigned ntile[kBatchVariantCount];
    for (int v = 0; v < kBatchVariantCount; v++) {
      const int tm = kBatchVariants[v].tm, tn = kBatchVariants[v].tn;
      std::vector<std::vector<int4>> q(8);
      long long leaves = 0;
      for (int i : mem) {
        const hm_dtree *t = jobs[i].t;
        const int ntl = (jobs[i].R + tn - 1) / tn;
        for (int c = 0; c < t->nleaf; c++) {
          const int mc = t->leaf_off[c + 1] - t->leaf_off[c];
          if (mc <= 0) continue;
          std::vector<int4> &dst = q[leaves++ % 8];
          for (int rt = 0; rt < (mc + tm - 1) / tm; rt++)
            for (int n = 0; n < ntl; n++) dst.push_back(make_int4(i, c, rt, n));
        }
      }
      size_t len = 0;
      for (const auto &x : q) len = std::max(len, x.size());
      std::vector<int4> wl(len * 8, make_int4(-1, 0, 0, 0));
      for (int x = 0; x < 8; x++)
        for (size_t u = 0; u < q[x].size(); u++) wl[u * 8 + x] = q[x][u];
      toff[v] = p.push(wl);
      ntile[v] = (unsigned)wl.size();
    }
    // HM_BATCH_APPLY_VARIANT=v fixes every group's tile shape (not tuned)
    const char *force = getenv("HM_BATCH_APPLY_VARIANT");
    const int fv = force ? std::max(0, std::min(kBatchVariantCount - 1, atoi(force))) : -1;
    auto choice = std::make_shared<int>(fv >= 0 ? fv : kDefault[g]);
    if (fv < 0) p.tunables.push_back({p.ops.size(), kBatchVariantCount, choice});
    std::vector<size_t> to(toff, toff + kBatchVariantCount);
    std::vector<unsigned> nt(ntile, ntile + kBatchVariantCount);
    p.add(kName[g], [=](hipStream_t s, const unsigned char *blob) {
      const int v = *choice;
      if (nt[v])
        kBatchVariants[v].launch(nfast, nt[v], reinterpret_cast<const BatchProd *>(blob + bp_off),
                                 reinterpret_cast<const int4 *>(blob + to[v]), s);
    });
  }
  return SUCCESS;
}

// --------------------------------------------- allocation-free products
// Product scratch of hm_hip_matmat_noalloc: one fixed array per device,
// 64 MiB, held from the load of this code object until the process ends.
// Host code never uses its name as an address: the device address comes
// from hipGetSymbolAddress, looked up once per device with that device
// current (on the entry point's first call there) and kept per device.
static constexpr long long kNoAllocScratch = 8LL << 20;  // doubles
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
double *g_scratch_addr[kMaxDevices] = {};
}  // namespace

static int noalloc_scratch(double **out) {
  int dev = -1;
  HM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("allocation-free product: device %d beyond the supported %d", dev, kMaxDevices);
    return INPUT_ERROR;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("allocation-free product: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                hipGetErrorString(e));
      return ALLOCATION_FAILURE;
    }
    g_scratch_addr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_addr[dev];
  return SUCCESS;
}

int products_run_noalloc(hm_dtree *t, int trans, int R, const double *X, long x_rs,
                         long x_cs, double *Y, long y_rs, long y_cs, hipStream_t s) {
  if (!t) {
    set_error("product: NULL device tree");
    return INPUT_ERROR;
  }
  if (R <= 0) return SUCCESS;
  const long long need = scratch_need(t, R);
  if (need > kNoAllocScratch) {
    set_error("allocation-free product: needs %lld doubles of scratch for R = %d, more than "
              "the fixed %lld (64 MiB); use hm_hip_matmat or fewer right-hand sides",
              need, R, kNoAllocScratch);
    return INPUT_ERROR;
  }
  double *base = nullptr;
  HM_OK(noalloc_scratch(&base));
  return products_run_ex(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs, s, base);
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_matmat_noalloc(hm_dtree *t, int trans, int R, const double *d_X, long x_rs,
                          long x_cs, double *d_Y, long y_rs, long y_cs, void *stream) {
  HM_OK(ensure_init());
  return products_run_noalloc(t, trans ? 1 : 0, R, d_X, x_rs, x_cs, d_Y, y_rs, y_cs,
                              stream_of(stream));
}

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

int hm_hip_matmat_variant(const hm_dtree *t, int trans, int R) {
  if (!t) return -1;
  return t->plan.apply_variant[trans ? 1 : 0][r_bucket(R)];
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

namespace hm {
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
// One v_mfma_f32_16x16x32_bf16: lane L takes A[L%16][8(L/16)+e] and
// B[8(L/16)+e][L%16], e = 0..7 (A 16x32, B 32x16 row-major, values exact in
// bf16); raw[4 L + v] receives accumulator value v of lane L.
__global__ void k_selftest_bf16(const float *A, const float *B, float *raw) {
  const int L = threadIdx.x;
  bf8v a, b;
  for (int e = 0; e < 8; e++) {
    a[e] = (__bf16)A[(L & 15) * 32 + 8 * (L >> 4) + e];
    b[e] = (__bf16)B[(8 * (L >> 4) + e) * 16 + (L & 15)];
  }
  f4v c = {0.f, 0.f, 0.f, 0.f};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  for (int v = 0; v < 4; v++) raw[4 * L + v] = c[v];
}
}  // namespace hm

extern "C" int hm_hip_selftest_mfma_bf16(const float *hA, const float *hB, float *raw) {
  HM_OK(ensure_init());
  float *d = nullptr;
  HM_OK(dalloc(&d, 512 + 512 + 256));
  hipError_t e = hipMemcpy(d, hA, 512 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 512, hB, 512 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_selftest_bf16, dim3(1), dim3(64), 0, 0, d, d + 512, d + 1024);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(raw, d + 1024, 256 * sizeof(float), hipMemcpyDeviceToHost);
  dfree(d);
  if (e != hipSuccess) {
    set_error("selftest bf16: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

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
