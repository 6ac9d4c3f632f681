// Internal header of libhmat_hip (gfx950).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "hm_hip.h"
#include "hmat_lib/error.h"

namespace hm {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
const char *hip_name(hipError_t e);

#define HM_HIP(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      hm::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,      \
                    hm::hip_name(e_));                                      \
      return ALLOCATION_FAILURE;                                            \
    }                                                                       \
  } while (0)

#define HM_OK(expr)                                                         \
  do {                                                                      \
    int r_ = (expr);                                                        \
    if (r_ != SUCCESS) return r_;                                           \
  } while (0)

#define HM_LAUNCH_CHECK() HM_HIP(hipGetLastError())

int ensure_init();
hipStream_t stream_of(void *s);

// -------------------------------------------------------- kernel timing
// Optional HIP-event bracketing of named kernels (bench instrumentation).
struct Timed {
  int slot = -1;
};
Timed timing_begin(const char *name, hipStream_t s);
void timing_end(const Timed &t, hipStream_t s);
// Adds host-side milliseconds under a name (when timing is on).
void timing_host(const char *name, double ms);

// ------------------------------------------------------------ memory
int dmalloc(void **p, size_t bytes);
template <class T>
int dalloc(T **p, size_t count) {
  return dmalloc(reinterpret_cast<void **>(p), count * sizeof(T) + 16);
}
void dfree(void *p);

// -------------------------------------------------------- device tree
// One off-diagonal block: rows [r0, r0+m) x cols [c0, c0+n) of the matrix,
// factors U (m x s, ld m) at U + uoff, V (n x s, ld n) at V + voff.
struct BlockMeta {
  int r0, m, c0, n;
  int s, level, j, pad;
  long long uoff, voff;
};

// Product scratch layout.  Leaves are cut into row chunks of at most
// CHUNK_ROWS rows.  For chunk k of leaf c and level l (1..h) let
// q = c >> (h - l).  In Y = H X the chunk's rows are multiplied into block
// (l, q^1) [input side] and receive block (l, q) [output side]; in
// Y = H^T X it is (l, q) and (l, q^1).  P(l, k) = F_in rows(k)^T X_k is an
// s x R column-major block; blocks are stored level by level and chunk by
// chunk, so the partials of one off-diagonal block are contiguous.
static constexpr int CHUNK_ROWS = 64;
static constexpr int REDUCE_SEG = 16;
// H x: the apply kernel sums a leaf's chunk partials itself (no reduction
// launches) when they are at most this many values, of at most this many
// ranks over the leaf's levels
static constexpr int MV_FUSE_VALUES = 4096;
static constexpr int MV_FUSE_RANKS = 256;

// Block path (trees whose per-chunk rank sum is large): W_b = F_in^T X_src
// is computed per block as a split-K GEMM over K segments of WSEG_ROWS
// rows; blocks with more than one segment write s x R partials that a
// second kernel sums in segment order.
static constexpr int WSEG_ROWS = 4096;
static constexpr int WTILE = 64;  // rank rows per work item

// One stacked input-factor row of a chunk's partial product (k_mm_partials):
// factor column kk of level l's input block at the chunk's first row
// (f_off, into V for H X / U for H^T X, leading dimension ld), its output
// row in the chunk partial (p_base x R + kk, column stride s).  s == 0
// marks a padding row.
struct PartRow {
  long long f_off, p_base;
  int kk, ld, s, pad;
};

// One reduction workgroup: out = sum of n consecutive partial blocks of
// s x R doubles starting at src (units of R doubles, in the chunk-partial
// buffer P, or in the segment-sum buffer when from_seg), written at dst
// (units of R doubles, segment-sum buffer or W).
struct RedItem {
  long long src, dst;
  int n, s;
  int from_seg, pad;
};

struct ProductPlan {
  int nchunk = 0;
  std::vector<int> ch_leaf, ch_r0, ch_rows;  // per chunk
  std::vector<int> leaf_chunk0;              // nleaf + 1
  std::vector<long long> pofs[2];            // [(l-1)*nchunk + k], x R
  std::vector<long long> wofs;               // per block prefix of s, x R
  int s_in_max[2] = {0, 0};
  int s_out_max[2] = {0, 0};
  // two-stage reduction: block g's chunks are cut into segments of
  // REDUCE_SEG chunks; seg_list holds (block, segment) pairs of blocks
  // with more than one segment, part_ofs[g] the offset (x s_g x R) of block
  // g's segment sums in the partial buffer.  The reduction kernels read
  // flat per-workgroup items (red_seg: chunk partials -> segment sums,
  // red_w: segment sums or chunk partials -> W_b), one table per
  // transpose mode, so each workgroup costs one metadata load.
  std::vector<int2> seg_list;
  std::vector<long long> part_ofs;
  std::vector<RedItem> red_seg[2], red_w[2];
  RedItem *d_red_seg[2] = {nullptr, nullptr};
  RedItem *d_red_w[2] = {nullptr, nullptr};
  // H x (R = 1) with the reduction in the apply kernel: per block, its chunk
  // partials as one item (src, n chunks, s); usable when every leaf's sum of
  // n * s over its levels is at most MV_FUSE_VALUES and of s at most
  // MV_FUSE_RANKS (mv_fused)
  std::vector<RedItem> fsum;
  RedItem *d_fsum = nullptr;
  bool mv_fused = false;
  int4 *d_ch = nullptr;                      // (leaf, r0, rows, first global row)
  // stacked input-factor rows of each chunk, 16 per M tile, per transpose
  // mode: [(k * part_mtiles[tr] + mt) * 16 + r]
  int part_mtiles[2] = {0, 0};
  std::vector<PartRow> prow[2];
  PartRow *d_prow[2] = {nullptr, nullptr};
  long long *d_pofs[2] = {nullptr, nullptr};
  long long *d_wofs = nullptr;
  int apply_variant[2][3] = {{-1, -1, -1}, {-1, -1, -1}};  // autotuned, by R bucket
  // block path, per transpose mode: items (block, rank tile, K segment, rows
  // in segment), partial offsets per block (x R), blocks with > 1 segment
  std::vector<int4> witems[2];
  std::vector<long long> wpofs[2];
  std::vector<int> wmulti[2];
  int wmulti_smax[2] = {0, 0};
  int4 *d_witems[2] = {nullptr, nullptr};
  long long *d_wpofs[2] = {nullptr, nullptr};
  int *d_wmulti[2] = {nullptr, nullptr};
};

}  // namespace hm

struct MixedFactors;
namespace hm {
// Device pools of the trees of one batched compression (U, V, diagonal
// blocks, metadata of all of them in one allocation each), freed with the
// last of those trees.
struct SharedPools {
  std::vector<void *> ptrs;
  ~SharedPools();
};
}  // namespace hm

struct hm_dtree {
  int height = 0;
  int n = 0;
  int nleaf = 0;
  int nblock = 0;
  int max_rank = 0;
  int max_leaf = 0;
  std::vector<int> leaf_off;          // nleaf + 1
  std::vector<long long> diag_off;    // nleaf + 1
  std::vector<hm::BlockMeta> blk;     // nblock, index (1 << l) - 2 + j
  double *d_diag = nullptr;
  double *d_U = nullptr;
  double *d_V = nullptr;
  int *d_leaf_off = nullptr;
  long long *d_diag_off = nullptr;
  hm::BlockMeta *d_blk = nullptr;
  hm::ProductPlan plan;
  // one device allocation holding every metadata and plan array above (the
  // d_* pointers of the tree and of the plan point into it)
  char *d_meta = nullptr;
  double *d_scratch = nullptr;        // P then W, grown on demand
  size_t scratch_elems = 0;
  // low-precision copies of U and V for mixed-precision products (format
  // slice_fmt: 0 BF16, 1 INT8, 2 FP32 slices); slices = 0 means FP64 products
  int slices = 0;
  int slice_fmt = 0;
  struct MixedFactors *mixed = nullptr;
  // set when d_diag, d_U, d_V and the metadata point into pools shared with
  // other trees (hm_hip_compress_batch); null: the tree owns them
  std::shared_ptr<hm::SharedPools> shared;
};

namespace hm {

__host__ __device__ inline int blk_index(int level, int j) { return (1 << level) - 2 + j; }

// Fill sizes/offsets of a tree from its leaf sizes and block ranks;
// allocates and uploads metadata and the product plan (not the data pools).
int tree_finalize_meta(hm_dtree *t, hipStream_t s);
// The same for several trees, their metadata in one device allocation and
// one copy, returned in *blob (the caller owns it; the trees' d_meta stay
// null).
int trees_finalize_meta(const std::vector<hm_dtree *> &ts, char **blob, hipStream_t s);
int tree_alloc_pools(hm_dtree *t);
int ensure_scratch(hm_dtree *t, int R);
// Doubles of product scratch (partials, W, segment sums) for R columns.
long long scratch_need(const hm_dtree *t, int R);
void tree_free(hm_dtree *t);
void mixed_free(MixedFactors *m);
// Y = op(H) X with sliced factors (t->slices > 0; op = transpose
// when trans); X, Y element (row, col) at [row * rs + col * cs].
int mixed_matmat(hm_dtree *t, int trans, int R, const double *X, long x_rs, long x_cs,
                 double *Y, long y_rs, long y_cs, hipStream_t s);
// the sliced factors of product mode trans exist (sliced on first use)
int mixed_ensure_mode(hm_dtree *t, int trans, hipStream_t s);
// the FP64 diagonal part only: Y = op(D) X
int products_diag(hm_dtree *t, int trans, int R, const double *X, long x_rs, long x_cs,
                  double *Y, long y_rs, long y_cs, hipStream_t s);

// ------------------------------------------------- global memory access
// Kernels reach global memory through pointers read from work-item tables
// or LDS, which the compiler types as generic: a generic (flat) access
// counts against the same completion counter as LDS, so every LDS wait
// after it (each MFMA step reading its staged operands) would also wait
// for the global load in flight.  Accesses through these helpers are
// global ones and leave LDS waits alone.
typedef __attribute__((address_space(1))) double gdbl;
// Orders one wavefront's LDS writes before its lanes' later LDS reads of
// them: the wave barrier keeps the scheduler from moving instructions across
// it, the wavefront-scope release / acquire fences keep the compiler from
// moving the memory operations themselves.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ gdbl *gmem(double *p) { return (gdbl *)p; }
__device__ __forceinline__ const gdbl *gmem(const double *p) { return (const gdbl *)p; }
__device__ __forceinline__ double ldg(const double *p) { return *gmem(p); }
__device__ __forceinline__ void stg(double *p, double v) { *gmem(p) = v; }

// ----------------------------------------------------------- MFMA tiles
typedef double d4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4_f64: lane L holds A[L%16][L/16] and B[L/16][L%16];
// the 16x16 result lives as 4 values per lane, value v of lane L being
// D[L/16 + 4*v][L%16] (rows interleaved by 4; checked on the device by
// hm_hip_selftest_mfma).
__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// A workgroup of WM*WN waves computes a TM x TN tile C += A B over
// K-chunks of KC, staging each chunk of A (TM x KC) and B (KC x TN) in LDS.
// Loaders: la(i, k) and lb(k, j) return the (bounds-checked) element.
// A_MFAST: consecutive threads fetch consecutive rows i (A column-major);
// otherwise consecutive k.  B_NFAST: consecutive threads fetch consecutive
// columns j; otherwise consecutive k (B column-major).
template <int WM, int WN, int FM, int FN, int KC>
struct MfmaTile {
  static constexpr int NW = WM * WN;
  static constexpr int NT = NW * 64;
  static constexpr int TM = WM * FM * 16;
  static constexpr int TN = WN * FN * 16;
  static constexpr int LDA = TM + 1;
  static constexpr int LDB = TN + 1;
  static constexpr int LDS_DOUBLES = KC * LDA + KC * LDB;
  static constexpr int KC_ = KC;
  static constexpr int AREG = (TM * KC + NT - 1) / NT;
  static constexpr int BREG = (TN * KC + NT - 1) / NT;
  // every thread owns whole elements (threadIdx.x < NT by the launch)
  static constexpr bool AEXACT = (TM * KC) % NT == 0;
  static constexpr bool BEXACT = (TN * KC) % NT == 0;

  d4 acc[FM][FN];
  double ra[AREG], rb[BREG];

  __device__ void zero() {
#pragma unroll
    for (int f = 0; f < FM; f++)
#pragma unroll
      for (int g = 0; g < FN; g++) acc[f][g] = d4{0.0, 0.0, 0.0, 0.0};
  }

  template <bool A_MFAST, bool B_NFAST, class LA, class LB>
  __device__ void fetch(int k0, const LA &la, const LB &lb) {
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < AREG; e++) {
      int idx = t + e * NT;
      int i, k;
      if (A_MFAST) { i = idx % TM; k = idx / TM; }
      else         { k = idx % KC; i = idx / KC; }
      ra[e] = (AEXACT || idx < TM * KC) ? la(i, k0 + k) : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BREG; e++) {
      int idx = t + e * NT;
      int j, k;
      if (B_NFAST) { j = idx % TN; k = idx / TN; }
      else         { k = idx % KC; j = idx / KC; }
      rb[e] = (BEXACT || idx < TN * KC) ? lb(k0 + k, j) : 0.0;
    }
  }

  template <bool A_MFAST, bool B_NFAST>
  __device__ void stash_regs(double *lds, const double *xa, const double *xb) const {
    double *As = lds, *Bs = lds + KC * LDA;
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < AREG; e++) {
      int idx = t + e * NT;
      int i, k;
      if (A_MFAST) { i = idx % TM; k = idx / TM; }
      else         { k = idx % KC; i = idx / KC; }
      if (AEXACT || idx < TM * KC) As[k * LDA + i] = xa[e];
    }
#pragma unroll
    for (int e = 0; e < BREG; e++) {
      int idx = t + e * NT;
      int j, k;
      if (B_NFAST) { j = idx % TN; k = idx / TN; }
      else         { k = idx % KC; j = idx / KC; }
      if (BEXACT || idx < TN * KC) Bs[k * LDB + j] = xb[e];
    }
  }

  template <bool A_MFAST, bool B_NFAST>
  __device__ void stash(double *lds) {
    stash_regs<A_MFAST, B_NFAST>(lds, ra, rb);
  }

  __device__ void compute(const double *lds) {
    const double *As = lds, *Bs = lds + KC * LDA;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w % WM, wn = w / WM;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) {
      double a[FM], b[FN];
#pragma unroll
      for (int f = 0; f < FM; f++)
        a[f] = As[(ks * 4 + lk) * LDA + (wm * FM + f) * 16 + li];
#pragma unroll
      for (int g = 0; g < FN; g++)
        b[g] = Bs[(ks * 4 + lk) * LDB + (wn * FN + g) * 16 + li];
#pragma unroll
      for (int f = 0; f < FM; f++)
#pragma unroll
        for (int g = 0; g < FN; g++) acc[f][g] = mfma_f64(a[f], b[g], acc[f][g]);
    }
  }

  // C += A[:, 0:K) B[0:K, :), software-pipelined over K-chunks.
  template <bool A_MFAST, bool B_NFAST, class LA, class LB>
  __device__ void run(double *lds, int K, const LA &la, const LB &lb) {
    if (K <= 0) return;
    fetch<A_MFAST, B_NFAST>(0, la, lb);
    for (int k0 = 0; k0 < K; k0 += KC) {
      __syncthreads();
      stash<A_MFAST, B_NFAST>(lds);
      __syncthreads();
      if (k0 + KC < K) fetch<A_MFAST, B_NFAST>(k0 + KC, la, lb);
      compute(lds);
    }
    __syncthreads();
  }

  // Interior form of run(): A is column-major with leading dimension lda
  // (rows contiguous), B(k, j) sits at B[k * brs + j * bcs]; the tile lies
  // fully inside both operands and K is a multiple of KC, so each thread
  // walks fixed pointers with no index arithmetic or bounds checks.  Two
  // register sets alternate: the loads of chunk c + 2 are issued as soon as
  // chunk c is in LDS, so each chunk has two compute phases to arrive.
  template <bool B_NFAST>
  __device__ void run_interior(double *lds, int K, const double *A, long long lda,
                               const double *B, long long brs, long long bcs) {
    static_assert(NT % TM == 0 && NT % KC == 0 && NT % TN == 0, "tile shape");
    if (K <= 0) return;
    const int t = threadIdx.x;
    const gdbl *pa = gmem(A) + (t % TM) + (long long)(t / TM) * lda;
    const long long aes = (long long)(NT / TM) * lda, astep = (long long)KC * lda;
    const gdbl *pb;
    long long bes;
    if (B_NFAST) {
      pb = gmem(B) + (long long)(t % TN) * bcs + (long long)(t / TN) * brs;
      bes = (long long)(NT / TN) * brs;
    } else {
      pb = gmem(B) + (long long)(t % KC) * brs + (long long)(t / KC) * bcs;
      bes = (long long)(NT / KC) * bcs;
    }
    const long long bstep = (long long)KC * brs;
    double ra2[AREG], rb2[BREG];
    auto load = [&](double *xa, double *xb) {
#pragma unroll
      for (int e = 0; e < AREG; e++) xa[e] = pa[e * aes];
#pragma unroll
      for (int e = 0; e < BREG; e++) xb[e] = pb[e * bes];
      pa += astep;
      pb += bstep;
    };
    const int nch = K / KC;
    load(ra, rb);
    // second set unconditionally (chunk 0 again when there is one chunk):
    // with no conditional loads before or inside the loop the compiler's
    // wait counts stay exact and each stash waits for its own set only
    if (nch == 1) {
      pa -= astep;
      pb -= bstep;
    }
    load(ra2, rb2);
    int c = 0;
    for (; c + 3 < nch; c += 2) {
      __syncthreads();
      stash_regs<true, B_NFAST>(lds, ra, rb);
      __syncthreads();
      load(ra, rb);
      compute(lds);
      __syncthreads();
      stash_regs<true, B_NFAST>(lds, ra2, rb2);
      __syncthreads();
      load(ra2, rb2);
      compute(lds);
    }
    // last one to three chunks (c and c + 1 already in flight)
    __syncthreads();
    stash_regs<true, B_NFAST>(lds, ra, rb);
    __syncthreads();
    if (c + 2 < nch) load(ra, rb);
    compute(lds);
    if (c + 1 < nch) {
      __syncthreads();
      stash_regs<true, B_NFAST>(lds, ra2, rb2);
      __syncthreads();
      compute(lds);
    }
    if (c + 2 < nch) {
      __syncthreads();
      stash_regs<true, B_NFAST>(lds, ra, rb);
      __syncthreads();
      compute(lds);
    }
    __syncthreads();
  }

  // Strided interior form: A(i, k) at A[i * ars + k * acs], B(k, j) at
  // B[k * brs + j * bcs] (B_NFAST false: consecutive threads take
  // consecutive k).  The tile's rows lie inside A and K is a multiple of
  // KC; B's tile columns j >= bcols are read as zero (their loads are
  // clamped to column bcols - 1, bcols >= 1).  Two register sets alternate
  // as in run_interior.
  template <bool A_MFAST>
  __device__ void run_strided(double *lds, int K, const double *A, long long ars,
                              long long acs, const double *B, long long brs, long long bcs,
                              int bcols) {
    static_assert(NT % TM == 0 && NT % KC == 0, "tile shape");
    if (K <= 0) return;
    const int t = threadIdx.x;
    const gdbl *pa;
    long long aes;
    if (A_MFAST) {
      pa = gmem(A) + (long long)(t % TM) * ars + (long long)(t / TM) * acs;
      aes = (long long)(NT / TM) * acs;
    } else {
      pa = gmem(A) + (long long)(t % KC) * acs + (long long)(t / KC) * ars;
      aes = (long long)(NT / KC) * ars;
    }
    const long long astep = (long long)KC * acs, bstep = (long long)KC * brs;
    const gdbl *pb[BREG];
    bool bok[BREG];
#pragma unroll
    for (int e = 0; e < BREG; e++) {
      const int idx = t + e * NT;
      const int k = idx % KC, j = idx / KC;
      bok[e] = (BEXACT || idx < TN * KC) && j < bcols;
      pb[e] = gmem(B) + (long long)k * brs + (long long)min(j, bcols - 1) * bcs;
    }
    double ra2[AREG], rb2[BREG];
    int c0 = 0;
    auto load = [&](double *xa, double *xb) {
#pragma unroll
      for (int e = 0; e < AREG; e++) xa[e] = pa[e * aes];
#pragma unroll
      for (int e = 0; e < BREG; e++) {
        const double v = pb[e][(long long)c0 * bstep];
        xb[e] = bok[e] ? v : 0.0;
      }
      pa += astep;
      c0++;
    };
    const int nch = K / KC;
    load(ra, rb);
    if (nch == 1) {
      pa -= astep;
      c0--;
    }
    load(ra2, rb2);
    int c = 0;
    for (; c + 3 < nch; c += 2) {
      __syncthreads();
      stash_regs<A_MFAST, false>(lds, ra, rb);
      __syncthreads();
      load(ra, rb);
      compute(lds);
      __syncthreads();
      stash_regs<A_MFAST, false>(lds, ra2, rb2);
      __syncthreads();
      load(ra2, rb2);
      compute(lds);
    }
    __syncthreads();
    stash_regs<A_MFAST, false>(lds, ra, rb);
    __syncthreads();
    if (c + 2 < nch) load(ra, rb);
    compute(lds);
    if (c + 1 < nch) {
      __syncthreads();
      stash_regs<A_MFAST, false>(lds, ra2, rb2);
      __syncthreads();
      compute(lds);
    }
    if (c + 2 < nch) {
      __syncthreads();
      stash_regs<A_MFAST, false>(lds, ra, rb);
      __syncthreads();
      compute(lds);
    }
    __syncthreads();
  }

  // st(i, j, value) for every element of the tile (caller bounds-checks).
  template <class ST>
  __device__ void store(const ST &st) const {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w % WM, wn = w / WM;
#pragma unroll
    for (int f = 0; f < FM; f++)
#pragma unroll
      for (int g = 0; g < FN; g++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          int i = (wm * FM + f) * 16 + (lane >> 4) + 4 * v;
          int j = (wn * FN + g) * 16 + (lane & 15);
          st(i, j, acc[f][g][v]);
        }
  }
};

// ----------------------------------------------------- small reductions
// Sum over the 64 lanes of a wavefront, returned to every lane; every lane
// must be active.  Within each row of 16 lanes: DPP moves (quad_perm
// [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror) leave the row sum in
// all its lanes; the four row sums are then read with v_readlane.  No LDS
// round trips (a shuffle-based butterfly costs 12 dependent ds_bpermute).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

}  // namespace hm
