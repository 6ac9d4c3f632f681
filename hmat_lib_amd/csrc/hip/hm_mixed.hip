// Mixed-precision HODLR x dense: the off-diagonal factors and products on
// low-precision MFMA with exact (or FP32) accumulation folded into FP64,
// diagonal leaves in FP64.
//
// Formats of the factor copies (MxFmt):
//  * BF16 slices.  An FP64 operand is split, per row of a K-chunk (at most
//    128 entries along the summed dimension), into S integer slices q_a with
//    |q_a| <= 128:
//      x = 2^e (q_0 2^-7 + q_1 2^-15 + ... + q_{S-1} 2^-(8S-1)) + r,
//      |r| <= 2^(e - 8S), with 2^e > max |x| over the chunk row.
//    Integers of magnitude <= 256 are exact in BF16, products of two slices
//    are integers <= 2^14, and v_mfma_f32_16x16x32_bf16 sums them in FP32
//    exactly as long as the running sum stays below 2^24: the slice pairs
//    (a, b) with a + b = d share one FP32 accumulator, at most S of them,
//    over a chunk of 128 entries (S * 2^14 * 128 <= 2^23).
//  * INT8 slices.  The same split with balanced base-254 digits, per row of
//    a 256-entry chunk: q_0 = rint(127 x / 2^e), then each remainder
//    (in [-1/2, 1/2]) times 254 rounded, so |q_a| <= 127 fits int8 and
//      x = 2^e / 127 (q_0 + q_1 / 254 + ... + q_{S-1} / 254^(S-1)) + r,
//      |r| <= 2^e / (254 * 254^(S-1)),
//    about 8 bits per slice as for BF16.  v_mfma_i32_16x16x64_i8 sums the
//    products in int32 exactly (S * 127^2 * 256 < 2^31), with twice the
//    BF16 rate per instruction and one byte per slice entry.
//  * FP32 (one "slice"): x / 2^e rounded to FP32 per row of a 32-entry
//    chunk (the scale is exact and only keeps the layout shared), products on
//    v_mfma_f32_16x16x4_f32 accumulated in FP32 over the chunk.  FP32
//    rounding of the operands and of the accumulation bounds the relative
//    error near 1e-7: FP32 MFMA runs at twice the FP64 rate, and no
//    correction term helps, since the FP32 accumulation rounds at 2^-24
//    whatever the operands carry.
// At the end of every chunk the S accumulators d = S-1 .. 0 are combined in
// FP64 (Horner in the digit base) and added with the row and column scales.
// Pairs with a + b >= S are dropped (their size is that of the
// representation error), so S slices cost S(S+1)/2 MFMAs per k-step and
// carry about 8S bits of each chunk row's maximum: S = 1 is BF16-like,
// S = 3 FP32-like (2^-24) and S = 4 reaches 2^-32.
//
// Operand storage ("chunk-padded", bytes): along K every chunk starts at a
// multiple of one k-step (64 bytes: 32 BF16, 64 INT8 or 16 FP32 entries)
// and is zero-padded to one, so MFMA k-steps never straddle chunks and loads
// are 16-byte aligned; slices of one operand follow each other at a fixed
// stride.  Scales are stored per (chunk, row).
//
// The product Y = H X (no transpose, X and Y column-major) runs:
//   slice X   (per leaf chunk and column)
//   W_b = V_b^T X_src   sliced GEMM, split-K over segments of chunks
//   slice W_b (per rank chunk and column; the segment partials summed in
//             order first, so the result is deterministic)
//   Y = D X   FP64 MFMA (the diagonal part of the FP64 apply kernel)
//   Y += sum over levels U_b W_b   sliced GEMM with the levels concatenated
//                                  along K, one pass over Y
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "hm_prog.h"

namespace hm {

enum MxFmt { MX_BF16 = 0, MX_I8 = 1, MX_F32 = 2 };
// Entry size (bytes), entries per 64-byte k-step, chunk length (entries)
// of each format, on the host.
static inline int fmt_esz(int f) { return f == MX_I8 ? 1 : (f == MX_F32 ? 4 : 2); }
static inline int fmt_ks(int f) { return 64 / fmt_esz(f); }
static inline int fmt_kch(int f) { return f == MX_I8 ? 256 : (f == MX_F32 ? 32 : 128); }
static constexpr int MX_SEG_K = 4096;  // W split-K segment (entries along K)

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
// global (not generic) loads, so that LDS waits do not wait for them
__device__ __forceinline__ i4 ldg16(const unsigned char *p) {
  return *(const __attribute__((address_space(1))) i4 *)(p);
}
__device__ __forceinline__ int ldgi(const int *p) {
  return *(const __attribute__((address_space(1))) int *)(p);
}

// Per-format device side: accumulator type, the MFMAs of one 16 x 16 block
// for one 16-byte fragment per lane of A (row L % 16) and of B (column
// L % 16), and the fold constants (digit base, scale of a product).
template <int F> struct MxDev;
template <> struct MxDev<MX_BF16> {
  static constexpr int KS = 32;
  using acc_t = f4;
  static constexpr double BASE = 0x1p-8, PROD = 0x1p-14;
  __device__ static __forceinline__ acc_t mma(i4 a, i4 b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                   __builtin_bit_cast(bf8, b), c, 0, 0, 0);
  }
};
template <> struct MxDev<MX_I8> {
  static constexpr int KS = 64;
  using acc_t = i4;
  static constexpr double BASE = 1.0 / 254.0, PROD = 1.0 / (127.0 * 127.0);
  __device__ static __forceinline__ acc_t mma(i4 a, i4 b, acc_t c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
  }
};
template <> struct MxDev<MX_F32> {
  static constexpr int KS = 16;
  using acc_t = f4;
  static constexpr double BASE = 1.0, PROD = 1.0;
  // lane L holds entries 4(L/16) .. 4(L/16)+3 of its row (A) / column (B);
  // MFMA j takes entry j of every lane, so each of the 16 k is summed once
  __device__ static __forceinline__ acc_t mma(i4 a, i4 b, acc_t c) {
    const f4 fa = __builtin_bit_cast(f4, a), fb = __builtin_bit_cast(f4, b);
#pragma unroll
    for (int j = 0; j < 4; j++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[j], fb[j], c, 0, 0, 0);
    return c;
  }
};

// ------------------------------------------------------------ slicing
// One workgroup per item: columns p .. p + ncol - 1 (ncol <= 16, four per
// wave, their loads issued before their reductions), each the chunk
// [k0, k0 + klen) (klen <= 256) of one row of the source, element (p, k) =
// sum over t < nsum (in order) of src[t * sum_stride + p * sp + k * sk]
// (nsum > 1: the split-K partials of a W block, summed here instead of in a
// separate pass) -> S slices at byte
// dst + (a * dst_slice + p * dst_ld + dk0) * entry size (klen padded to a
// whole k-step with zeros), its scale to scale[p - first p].
struct SliceItem {
  const double *src;
  long long sp, sk, sum_stride;
  unsigned char *dst;
  long long dst_slice, dst_ld, dk0;
  double *scale;
  int p, k0, klen, nsum;
  int ncol, pad;
};
static constexpr int SLICE_COLS = 16, SLICE_CPW = 4;

template <int F, int S>
__global__ __launch_bounds__(256) void k_slice(const SliceItem *items) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const SliceItem it = items[blockIdx.x];
  constexpr int ks = F == MX_I8 ? 64 : (F == MX_F32 ? 16 : 32);
  const int kp = (it.klen + ks - 1) / ks * ks;
  // the columns' entries, four per lane (k = lane + 64 i)
  double v[SLICE_CPW][4], mx[SLICE_CPW];
#pragma unroll
  for (int c = 0; c < SLICE_CPW; c++) {
    const int jj = wv * SLICE_CPW + c;
    mx[c] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = lane + 64 * i;
      double x = 0.0;
      if (jj < it.ncol && k < it.klen) {
        const double *q =
            it.src + (long long)(it.p + jj) * it.sp + (long long)(it.k0 + k) * it.sk;
        x = q[0];
        for (int t = 1; t < it.nsum; t++) x += q[(long long)t * it.sum_stride];
      }
      v[c][i] = x;
      mx[c] = fmax(mx[c], fabs(x));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int c = 0; c < SLICE_CPW; c++) mx[c] = fmax(mx[c], __shfl_xor(mx[c], o, 64));
#pragma unroll
  for (int c = 0; c < SLICE_CPW; c++) {
    const int jj = wv * SLICE_CPW + c;
    if (jj >= it.ncol) break;
    int e = 0;
    if (mx[c] > 0.0) {
      (void)frexp(mx[c], &e);  // mx < 2^e
    }
    // BF16 and FP32: the power of two 2^e; INT8: the maximum itself, so q_0
    // spans the whole [-127, 127] (steps of max / 127 instead of up to twice
    // that: the INT8 products' error is about half the BF16 one)
    const bool exact = F == MX_I8 && mx[c] > 0.0;
    const double sc = exact ? mx[c] : ldexp(1.0, e);
    if (lane == 0) it.scale[jj] = sc;
    const double inv = exact ? 1.0 / mx[c] : ldexp(1.0, -e);
    const long long base = (long long)(it.p + jj) * it.dst_ld + it.dk0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = lane + 64 * i;
      if (k >= kp) break;
      double x = v[c][i] * inv;  // zero beyond klen
      // x in (-1, 1) ([-1, 1] for INT8)
      if constexpr (F == MX_BF16) {
        // q_0 = rint(x 2^7), then the remainder scaled by 2^8
        __bf16 *d = reinterpret_cast<__bf16 *>(it.dst) + base;
        x *= 128.0;
#pragma unroll
        for (int a = 0; a < S; a++) {
          const double q = rint(x);
          d[(long long)a * it.dst_slice + k] = (__bf16)(float)q;
          x = (x - q) * 256.0;
        }
      } else if constexpr (F == MX_I8) {
        // q_0 = rint(127 x), then the remainder (in [-1/2, 1/2]) times 254
        signed char *d = reinterpret_cast<signed char *>(it.dst) + base;
        x *= 127.0;
#pragma unroll
        for (int a = 0; a < S; a++) {
          const double q = rint(x);
          d[(long long)a * it.dst_slice + k] = (signed char)(int)q;
          x = (x - q) * 254.0;
        }
      } else {
        float *d = reinterpret_cast<float *>(it.dst) + base;
        d[k] = (float)x;
      }
    }
  }
}

template <int F>
static void launch_slice_f(int S, unsigned n, const SliceItem *items, hipStream_t st) {
  if (S == 1) hipLaunchKernelGGL((k_slice<F, 1>), dim3(n), dim3(256), 0, st, items);
  if constexpr (F != MX_F32) {
    if (S == 2) hipLaunchKernelGGL((k_slice<F, 2>), dim3(n), dim3(256), 0, st, items);
    if (S == 3) hipLaunchKernelGGL((k_slice<F, 3>), dim3(n), dim3(256), 0, st, items);
    if (S == 4) hipLaunchKernelGGL((k_slice<F, 4>), dim3(n), dim3(256), 0, st, items);
  }
}
static void launch_slice(int fmt, int S, unsigned n, const SliceItem *items, hipStream_t st) {
  if (n == 0) return;
  if (fmt == MX_BF16) launch_slice_f<MX_BF16>(S, n, items, st);
  if (fmt == MX_I8) launch_slice_f<MX_I8>(S, n, items, st);
  if (fmt == MX_F32) launch_slice_f<MX_F32>(S, n, items, st);
}

// ---------------------------------------------------------- sliced GEMM
// One segment: C += op(A) op(B) over K entries given as chunks
// [cs[c], cs[c+1]) (padded, whole k-steps) of the chunk table.  Offsets and
// strides count entries of the format.
struct MxSeg {
  const unsigned char *A;  // entry (row p, k) of slice a: A[a * a_slice + p * lda + k]
  long long lda, a_slice;
  const double *sA;  // scale of (chunk c, row p): sA[(a_c0 + c) * ldsa + p]
  long long ldsa;
  const unsigned char *B;  // entry (k, col j) of slice b: B[b * b_slice + j * ldb + k]
  long long ldb, b_slice;
  const double *sB;
  long long ldsb;
  const int *cs;     // chunk boundaries (padded K), nch + 1 entries
  int nch, a_c0, b_c0, pad;
};

// One work item: a 64 x 64 tile of C (rows i0.., cols j0.., within M x N)
// accumulated over segments [s0, s1).  mode 0: C = acc; 1: C += acc.
struct MxItem {
  double *C;          // element (row, col) at C[row * crs + col * ldc]
  long long ldc, crs;
  int M, N, i0, j0, s0, s1, mode, pad;
};

// MX_TM x TN tile per workgroup (2 x TN/32 waves of 32 x 32; TN = MX_TN, 4
// waves, measured faster than 128 wide with 8 waves), k-steps of 64 bytes per
// row: each step's S slices of the A rows and B columns go global
// -> registers (fetched one step ahead) -> one of two LDS stages, and every
// wave reads its MFMA fragments from LDS; the stage written during a step is
// read in the next, so one barrier per step orders both.  The row and
// column scales of a chunk are loaded into registers when the chunk starts
// and used by its fold.  LDS rows are padded to MX_RSB bytes.
static constexpr int MX_RSB = 80;
static constexpr int MX_TM = 64, MX_TN = 64;

template <int F, int S, int TN>
__global__ __launch_bounds__(2 * TN * 2) void k_sliced_gemm(const MxItem *items,
                                                             const MxSeg *segs) {
  using D = MxDev<F>;
  using acc_t = typename D::acc_t;
  constexpr int KS = D::KS, ESZ = 64 / KS;
  constexpr int FM = 2, FN = 2;  // 16 x 16 blocks per wave
  __shared__ __attribute__((aligned(16))) unsigned char As[2][S][MX_TM * MX_RSB];
  __shared__ __attribute__((aligned(16))) unsigned char Bs[2][S][TN * MX_RSB];
  const MxItem it = items[blockIdx.x];
  const int tid = threadIdx.x;
  const int L = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int li = L & 15, lk = L >> 4;
  // global loads: B row lr (every thread), A row lr (the first 256 threads);
  // 16 bytes at byte kq
  const int lr = tid >> 2, kq = (tid & 3) * 16;
  const bool ldA = TN == MX_TM || lr < MX_TM;  // uniform per wave
  double acc64[FM][FN][4];
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc64[f][g][v] = 0.0;
  for (int sg = it.s0; sg < it.s1; sg++) {
    const MxSeg sd = segs[sg];
    if (sd.nch <= 0) continue;
    // Operands through raw buffer loads: one descriptor per slice whose
    // range ends after row M - 1 (A) / column N - 1 (B), so rows beyond the
    // tile's edge read zeros without a branch, and a 32-bit lane offset
    // (the host keeps M * lda and N * ldb bytes below 2^31).
    const unsigned nA = (unsigned)((long long)it.M * sd.lda * ESZ);
    const unsigned nB = (unsigned)((long long)it.N * sd.ldb * ESZ);
    __amdgpu_buffer_rsrc_t rA[S], rB[S];
#pragma unroll
    for (int a = 0; a < S; a++) {
      rA[a] = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(sd.A + (long long)a * sd.a_slice * ESZ), 0, (int)nA, 0x00020000);
      rB[a] = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(sd.B + (long long)a * sd.b_slice * ESZ), 0, (int)nB, 0x00020000);
    }
    const unsigned offA = (unsigned)((long long)(it.i0 + lr) * sd.lda * ESZ) + kq;
    const unsigned offB = (unsigned)((long long)(it.j0 + lr) * sd.ldb * ESZ) + kq;
    i4 ra0[S], rb[S];
    auto fetch = [&](int k0) {
      const unsigned kb = (unsigned)k0 * ESZ;
#pragma unroll
      for (int a = 0; a < S; a++) {
        if (ldA)
          ra0[a] = __builtin_bit_cast(
              i4, __builtin_amdgcn_raw_buffer_load_b128(rA[a], (int)(offA + kb), 0, 0));
        rb[a] = __builtin_bit_cast(
            i4, __builtin_amdgcn_raw_buffer_load_b128(rB[a], (int)(offB + kb), 0, 0));
      }
    };
    auto stage = [&](int buf) {
#pragma unroll
      for (int a = 0; a < S; a++) {
        if (ldA) *reinterpret_cast<i4 *>(&As[buf][a][lr * MX_RSB + kq]) = ra0[a];
        *reinterpret_cast<i4 *>(&Bs[buf][a][lr * MX_RSB + kq]) = rb[a];
      }
    };
    double sa[FM][4], sb[FN];
    auto fetch_scales = [&](int c) {
#pragma unroll
      for (int f = 0; f < FM; f++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = it.i0 + wm * 32 + f * 16 + 4 * lk + v;
          sa[f][v] = ldg(sd.sA + (long long)(sd.a_c0 + c) * sd.ldsa + (row < it.M ? row : it.i0));
        }
#pragma unroll
      for (int g = 0; g < FN; g++) {
        const int col = it.j0 + wn * 32 + g * 16 + li;
        sb[g] = ldg(sd.sB + (long long)(sd.b_c0 + c) * sd.ldsb + (col < it.N ? col : it.j0));
      }
    };
    const int kbeg = ldgi(sd.cs), kend = ldgi(sd.cs + sd.nch);
    int c = 0, cend = ldgi(sd.cs + 1);
    acc_t acc[S][FM][FN];
#pragma unroll
    for (int d = 0; d < S; d++)
#pragma unroll
      for (int f = 0; f < FM; f++)
#pragma unroll
        for (int g = 0; g < FN; g++) acc[d][f][g] = acc_t{};
    fetch(kbeg);
    fetch_scales(0);
    __syncthreads();  // the previous segment's last stage has been read
    stage(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += KS) {
      const bool more = k0 + KS < kend;
      if (more) fetch(k0 + KS);
      i4 fb[S][FN];
#pragma unroll
      for (int a = 0; a < S; a++)
#pragma unroll
        for (int g = 0; g < FN; g++)
          fb[a][g] =
              *reinterpret_cast<const i4 *>(&Bs[buf][a][(wn * 32 + g * 16 + li) * MX_RSB + 16 * lk]);
#pragma unroll
      for (int f = 0; f < FM; f++) {
        i4 fa[S];
#pragma unroll
        for (int a = 0; a < S; a++)
          fa[a] =
              *reinterpret_cast<const i4 *>(&As[buf][a][(wm * 32 + f * 16 + li) * MX_RSB + 16 * lk]);
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
          for (int b = 0; a + b < S; b++)
#pragma unroll
            for (int g = 0; g < FN; g++)
              acc[a + b][f][g] = D::mma(fa[a], fb[b][g], acc[a + b][f][g]);
      }
      if (k0 + KS == cend) {
        // fold the chunk into FP64 with its row and column scales
#pragma unroll
        for (int f = 0; f < FM; f++)
#pragma unroll
          for (int g = 0; g < FN; g++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
              double t = 0.0;
#pragma unroll
              for (int d = S - 1; d >= 0; d--) t = t * D::BASE + (double)acc[d][f][g][v];
              acc64[f][g][v] += sa[f][v] * sb[g] * D::PROD * t;
            }
#pragma unroll
        for (int d = 0; d < S; d++)
#pragma unroll
          for (int f = 0; f < FM; f++)
#pragma unroll
            for (int g = 0; g < FN; g++) acc[d][f][g] = acc_t{};
        c++;
        if (c < sd.nch) {
          cend = ldgi(sd.cs + c + 1);
          fetch_scales(c);
        }
      }
      if (more) stage(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int f = 0; f < FM; f++)
#pragma unroll
    for (int g = 0; g < FN; g++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = it.i0 + wm * 32 + f * 16 + 4 * lk + v;
        const int col = it.j0 + wn * 32 + g * 16 + li;
        if (row < it.M && col < it.N) {
          double *cp = it.C + (long long)row * it.crs + (long long)col * it.ldc;
          stg(cp, it.mode ? ldg(cp) + acc64[f][g][v] : acc64[f][g][v]);
        }
      }
}

}  // namespace hm

// ---------------------------------------------------------- tree factors
// Per tree: the format and slice count, the K-chunk grid over the rows
// (chunks of at most the format's chunk length inside leaves, shared by both
// modes) and, per product mode (0: Y = H X, 1:
// Y = H^T X), the sliced factors: the W side F_in^T (V^T for H X, U^T for
// H^T X) sliced along the rows of its source node by that grid, the apply
// side F_out (U for H X, V for H^T X) sliced along the rank by chunks of
// the chunk length.  Mode 1 is sliced on its first use.
struct MixedFactors {
  int S = 0, fmt = 0;
  int esz = 2, ks = 32, kch = 128;  // entry bytes, k-step and chunk (entries)
  // chunk q covers rows [row[q], row[q] + len[q]) and padded coordinates
  // [pad[q], pad[q + 1])
  std::vector<int> row, len, leaf_chunk0;
  std::vector<long long> pad;
  long long npad = 0;
  struct Blk {
    long long w_off = 0, w_kp = 0;  // F_in^T slices [S][s][w_kp] at w_off
    long long w_soff = 0;           // their scales [node chunk][s]
    int q0 = 0, q1 = 0;             // chunk range of the source node
    long long a_off = 0;            // F_out slices [S][rows][s_pad]
    long long a_soff = 0;           // their scales [rank chunk][rows]
    int rows = 0;                   // rows of F_out
  };
  struct Mode {
    bool ready = false;
    std::vector<Blk> blk;
    unsigned char *d_ws = nullptr, *d_as = nullptr;
    double *d_sw = nullptr, *d_sa = nullptr;
  } mode[2];
  std::vector<int> s_pad, nrk;  // per block: rank padded to a k-step, rank chunks
  std::vector<long long> rk_cs;  // per block: offset of its rank chunk table
  int *d_rk = nullptr;           // rank chunk tables (nrk + 1 ints per block)
  // cached single-tree product program (operands fixed)
  hm::Program *prog = nullptr;
  char *arena = nullptr;
  int trans = -1, R = 0;
  const double *X = nullptr;
  double *Y = nullptr;
  long x_rs = 0, x_cs = 0, y_rs = 0, y_cs = 0;
};

namespace hm {

void mixed_free(MixedFactors *m) {
  if (!m) return;
  (void)hipDeviceSynchronize();
  for (auto &md : m->mode) {
    dfree(md.d_ws);
    dfree(md.d_as);
    dfree(md.d_sw);
    dfree(md.d_sa);
  }
  dfree(m->d_rk);
  delete m->prog;
  dfree(m->arena);
  delete m;
}

static long long pad_to(long long x, int ks) { return (x + ks - 1) / ks * ks; }

static int run_slices(const std::vector<SliceItem> &items, int S, int fmt, hipStream_t s) {
  if (items.empty()) return SUCCESS;
  SliceItem *d_items = nullptr;
  HM_OK(dalloc(&d_items, items.size()));
  hipError_t e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(SliceItem),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    const int n = (int)items.size();
    launch_slice(fmt, S, (unsigned)n, d_items, s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(d_items);
  if (e != hipSuccess) {
    set_error("mixed: slicing factors: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

// Slice the factors of product mode tr (0: H X, 1: H^T X).
static int prepare_mode(hm_dtree *t, MixedFactors *m, int tr, hipStream_t s) {
  MixedFactors::Mode &md = m->mode[tr];
  const int S = m->S;
  md.blk.assign(t->nblock, MixedFactors::Blk());
  long long woff = 0, wsoff = 0, aoff = 0, asoff = 0;
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    MixedFactors::Blk &x = md.blk[g];
    const int span = 1 << (t->height - b.level);
    const int src = tr ? b.j : (b.j ^ 1);
    x.q0 = m->leaf_chunk0[src * span];
    x.q1 = m->leaf_chunk0[(src + 1) * span];
    x.w_kp = m->pad[x.q1] - m->pad[x.q0];
    x.w_off = woff;
    woff += (long long)S * b.s * x.w_kp;
    x.w_soff = wsoff;
    wsoff += (long long)(x.q1 - x.q0) * b.s;
    x.rows = tr ? b.n : b.m;
    x.a_off = aoff;
    aoff += (long long)S * x.rows * m->s_pad[g];
    x.a_soff = asoff;
    asoff += (long long)m->nrk[g] * x.rows;
  }
  HM_OK(dalloc(&md.d_ws, (size_t)woff * m->esz + 64));
  HM_OK(dalloc(&md.d_as, (size_t)aoff * m->esz + 64));
  HM_OK(dalloc(&md.d_sw, (size_t)wsoff + 1));
  HM_OK(dalloc(&md.d_sa, (size_t)asoff + 1));
  std::vector<SliceItem> items;
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    const MixedFactors::Blk &x = md.blk[g];
    if (b.s <= 0) continue;
    // W side: row r of F_in^T = column r of F_in (V: n x s, U: m x s),
    // K = rows of the source node
    const double *fin = tr ? t->d_U + b.uoff : t->d_V + b.voff;
    const int fin_ld = tr ? b.m : b.n;
    const int node_lo = tr ? b.r0 : b.c0;
    for (int q = x.q0; q < x.q1; q++)
      for (int r = 0; r < b.s; r += SLICE_COLS) {
        SliceItem it;
        it.src = fin;
        it.sp = fin_ld;
        it.sk = 1;
        it.dst = md.d_ws + x.w_off * m->esz;
        it.dst_slice = (long long)b.s * x.w_kp;
        it.dst_ld = x.w_kp;
        it.dk0 = m->pad[q] - m->pad[x.q0];
        it.scale = md.d_sw + x.w_soff + (long long)(q - x.q0) * b.s + r;
        it.p = r;
        it.k0 = m->row[q] - node_lo;
        it.klen = m->len[q];
        it.nsum = 1;
        it.sum_stride = 0;
        it.ncol = std::min(SLICE_COLS, b.s - r);
        it.pad = 0;
        items.push_back(it);
      }
    // apply side: row i of F_out, K = rank
    const double *fout = tr ? t->d_V + b.voff : t->d_U + b.uoff;
    for (int c = 0; c < m->nrk[g]; c++)
      for (int i = 0; i < x.rows; i += SLICE_COLS) {
        SliceItem it;
        it.src = fout;
        it.sp = 1;
        it.sk = x.rows;
        it.dst = md.d_as + x.a_off * m->esz;
        it.dst_slice = (long long)x.rows * m->s_pad[g];
        it.dst_ld = m->s_pad[g];
        it.dk0 = (long long)c * m->kch;
        it.scale = md.d_sa + x.a_soff + (long long)c * x.rows + i;
        it.p = i;
        it.k0 = c * m->kch;
        it.klen = std::max(0, std::min(m->kch, b.s - c * m->kch));
        it.nsum = 1;
        it.sum_stride = 0;
        it.ncol = std::min(SLICE_COLS, x.rows - i);
        it.pad = 0;
        items.push_back(it);
      }
  }
  HM_OK(run_slices(items, S, m->fmt, s));
  md.ready = true;
  return SUCCESS;
}

static int mixed_prepare(hm_dtree *t, int fmt, int S, hipStream_t s) {
  MixedFactors *m = new MixedFactors();
  m->S = S;
  m->fmt = fmt;
  m->esz = fmt_esz(fmt);
  m->ks = fmt_ks(fmt);
  m->kch = fmt_kch(fmt);
  const int nl = t->nleaf;
  m->leaf_chunk0.assign(nl + 1, 0);
  m->pad.push_back(0);
  for (int c = 0; c < nl; c++) {
    m->leaf_chunk0[c] = (int)m->row.size();
    const int lo = t->leaf_off[c], mc = t->leaf_off[c + 1] - lo;
    for (int r = 0; r < mc; r += m->kch) {
      m->row.push_back(lo + r);
      m->len.push_back(std::min(m->kch, mc - r));
      m->pad.push_back(m->pad.back() + pad_to(m->len.back(), m->ks));
    }
  }
  m->leaf_chunk0[nl] = (int)m->row.size();
  m->npad = m->pad.back();
  std::vector<int> rktab;
  m->s_pad.assign(t->nblock, 0);
  m->nrk.assign(t->nblock, 0);
  m->rk_cs.assign(t->nblock, 0);
  for (int g = 0; g < t->nblock; g++) {
    m->s_pad[g] = (int)pad_to(t->blk[g].s, m->ks);
    m->nrk[g] = (m->s_pad[g] + m->kch - 1) / m->kch;
    m->rk_cs[g] = (long long)rktab.size();
    for (int c = 0; c <= m->nrk[g]; c++) rktab.push_back(std::min(c * m->kch, m->s_pad[g]));
  }
  int rc = dalloc(&m->d_rk, rktab.size() + 1);
  if (rc == SUCCESS && !rktab.empty()) {
    hipError_t e = hipMemcpyAsync(m->d_rk, rktab.data(), rktab.size() * sizeof(int),
                                  hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      set_error("mixed: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  if (rc == SUCCESS) rc = prepare_mode(t, m, 0, s);
  if (rc != SUCCESS) {
    mixed_free(m);
    return rc;
  }
  mixed_free(t->mixed);
  t->mixed = m;
  t->slices = S;
  t->slice_fmt = fmt;
  return SUCCESS;
}

// Chunk tables of program-built segments live in the program's own blob,
// whose device address is known only at launch: fix[i] >= 0 is the byte
// offset of segment i's table in the blob.
__global__ __launch_bounds__(256) void k_mx_fix_segs(MxSeg *segs, const long long *fix,
                                                     const unsigned char *blob, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && fix[i] >= 0) segs[i].cs = reinterpret_cast<const int *>(blob + fix[i]);
}

// Work items of the sliced stages of one or more products (trees of one
// slice count), gathered tree by tree and launched together.
struct MixedCollect {
  std::vector<SliceItem> sl, wsl;
  std::vector<MxSeg> segs;
  std::vector<MxItem> wit, ait;
  std::vector<int> ctab;           // chunk tables of the W segments
  std::vector<long long> seg_ctab;  // per segment: its table in ctab, or -1
};

// Bytes spanned by an n x R operand with the given strides.
static size_t operand_bytes(long long n, int R, long rs, long cs) {
  if (n <= 0 || R <= 0) return 0;
  return (size_t)((n - 1) * rs + (long long)(R - 1) * cs + 1) * sizeof(double);
}

// The sliced part of op(H) X for one tree (op = transpose when tr): slice
// X, W_b = F_in^T X_src per block (split-K segments, then their sum), slice
// W, Y += F_out W over all levels per leaf.  X, Y element (row, col) at
// [row * rs + col * cs].  The diagonal FP64 part is not included.
static int mixed_collect(MixedCollect &mc, Program &p, Arena &ar, hm_dtree *t, int tr, int R,
                         const double *X, long x_rs, long x_cs, double *Y, long y_rs,
                         long y_cs) {
  MixedFactors *m = t->mixed;
  const MixedFactors::Mode &md = m->mode[tr];
  const int S = m->S;
  const int nq = (int)m->row.size();
  const int esz = m->esz, kch = m->kch;
  // the sliced GEMM addresses an operand's rows (or columns) with 32-bit byte
  // offsets below 2^31
  auto fits = [&](long long rows, long long ld) { return rows * ld * esz < (1LL << 31) - (1LL << 24); };
  if (!fits(R, m->npad)) {
    set_error("mixed product: %d right-hand sides x %lld padded rows exceed the sliced "
              "kernels' 2 GiB operand range", R, m->npad);
    return INPUT_ERROR;
  }
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    if (b.s > 0 && (!fits(b.s, md.blk[g].w_kp) || !fits(md.blk[g].rows, m->s_pad[g]) ||
                    !fits(R, m->s_pad[g]))) {
      set_error("mixed product: block %d exceeds the sliced kernels' 2 GiB operand range", g);
      return INPUT_ERROR;
    }
  }
  unsigned char *xs = ar.take<unsigned char>((size_t)S * R * m->npad * esz + 64);
  double *sx = ar.take<double>((size_t)nq * R + 1);
  p.touch(xs, (size_t)S * R * m->npad * esz, "mixed X slices");
  for (int q = 0; q < nq; q++)
    for (int j = 0; j < R; j += SLICE_COLS) {
      SliceItem it;
      it.src = X;
      it.sp = x_cs;
      it.sk = x_rs;
      it.dst = xs;
      it.dst_slice = (long long)R * m->npad;
      it.dst_ld = m->npad;
      it.dk0 = m->pad[q];
      it.scale = sx + (long long)q * R + j;
      it.p = j;
      it.k0 = m->row[q];
      it.klen = m->len[q];
      it.nsum = 1;
      it.sum_stride = 0;
      it.ncol = std::min(SLICE_COLS, R - j);
      it.pad = 0;
      mc.sl.push_back(it);
    }
  // W = F_in^T X per block and segment
  long long wtot = 0;
  std::vector<long long> woff(t->nblock, 0);
  for (int g = 0; g < t->nblock; g++) {
    woff[g] = wtot;
    wtot += (long long)t->blk[g].s * R;
  }
  double *W = ar.take<double>((size_t)wtot + 1);
  // per block: its W (the GEMM's output for one segment) or the segments'
  // partials, summed by the slicing of W
  std::vector<const double *> wsrc(t->nblock, nullptr);
  std::vector<int> wnsum(t->nblock, 1);
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    const MixedFactors::Blk &x = md.blk[g];
    if (b.s <= 0) continue;
    double *Wg = W + woff[g];
    const int nch = x.q1 - x.q0;
    const int seg_chunks = std::max(1, MX_SEG_K / kch);
    const int nseg = std::max(1, (nch + seg_chunks - 1) / seg_chunks);
    double *part = nseg > 1 ? ar.take<double>((size_t)nseg * b.s * R) : nullptr;
    wsrc[g] = nseg > 1 ? part : Wg;
    wnsum[g] = nseg;
    if (nch == 0) {  // empty source: W_b = 0
      mc.segs.push_back(
          {nullptr, 0, 0, nullptr, 0, nullptr, 0, 0, nullptr, 0, nullptr, 0, 0, 0, 0});
      mc.seg_ctab.push_back((long long)mc.ctab.size());
      mc.ctab.push_back(0);
      const int sgi = (int)mc.segs.size() - 1;
      for (int i0 = 0; i0 < b.s; i0 += MX_TM)
        for (int j0 = 0; j0 < R; j0 += MX_TN)
          mc.wit.push_back({Wg, b.s, 1, b.s, R, i0, j0, sgi, sgi, 0, 0});
      p.touch(Wg, (size_t)b.s * R * sizeof(double), "mixed W");
      continue;
    }
    for (int sg = 0; sg < nseg; sg++) {
      const int c0 = sg * seg_chunks, c1 = std::min(nch, c0 + seg_chunks);
      const long long base = m->pad[x.q0 + c0] - m->pad[x.q0];
      mc.seg_ctab.push_back((long long)mc.ctab.size());
      for (int c = c0; c <= c1; c++)
        mc.ctab.push_back((int)(m->pad[x.q0 + c] - m->pad[x.q0 + c0]));
      MxSeg sd;
      sd.A = md.d_ws + (x.w_off + base) * esz;
      sd.lda = x.w_kp;
      sd.a_slice = (long long)b.s * x.w_kp;
      sd.sA = md.d_sw + x.w_soff;
      sd.ldsa = b.s;
      sd.B = xs + m->pad[x.q0 + c0] * esz;
      sd.ldb = m->npad;
      sd.b_slice = (long long)R * m->npad;
      sd.sB = sx;
      sd.ldsb = R;
      sd.cs = nullptr;
      sd.nch = c1 - c0;
      sd.a_c0 = c0;
      sd.b_c0 = x.q0 + c0;
      sd.pad = 0;
      mc.segs.push_back(sd);
      const int sgi = (int)mc.segs.size() - 1;
      double *C = nseg > 1 ? part + (long long)sg * b.s * R : Wg;
      for (int i0 = 0; i0 < b.s; i0 += MX_TM)
        for (int j0 = 0; j0 < R; j0 += MX_TN)
          mc.wit.push_back({C, b.s, 1, b.s, R, i0, j0, sgi, sgi + 1, 0, 0});
      p.touch(C, (size_t)b.s * R * sizeof(double), "mixed W");
    }
  }
  // slices of W, then the apply segments (one per level with a rank)
  std::vector<unsigned char *> Ws(t->nblock, nullptr);
  std::vector<double *> sW(t->nblock, nullptr);
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    if (b.s <= 0) continue;
    Ws[g] = ar.take<unsigned char>((size_t)S * R * m->s_pad[g] * esz + 64);
    sW[g] = ar.take<double>((size_t)m->nrk[g] * R + 1);
    p.touch(Ws[g], (size_t)S * R * m->s_pad[g] * esz, "mixed W slices");
    for (int c = 0; c < m->nrk[g]; c++)
      for (int j = 0; j < R; j += SLICE_COLS) {
        SliceItem it;
        it.src = wsrc[g];
        it.sp = b.s;
        it.sk = 1;
        it.dst = Ws[g];
        it.dst_slice = (long long)R * m->s_pad[g];
        it.dst_ld = m->s_pad[g];
        it.dk0 = (long long)c * kch;
        it.scale = sW[g] + (long long)c * R + j;
        it.p = j;
        it.k0 = c * kch;
        it.klen = std::max(0, std::min(kch, b.s - c * kch));
        it.nsum = wnsum[g];
        it.sum_stride = (long long)b.s * R;
        it.ncol = std::min(SLICE_COLS, R - j);
        it.pad = 0;
        mc.wsl.push_back(it);
      }
  }
  for (int c = 0; c < t->nleaf; c++) {
    const int lo = t->leaf_off[c], rows = t->leaf_off[c + 1] - lo;
    if (rows <= 0) continue;
    const int s0 = (int)mc.segs.size();
    for (int l = 1; l <= t->height; l++) {
      const int q = c >> (t->height - l);
      // H X: block (l, q) has the leaf's rows (U); H^T X: block (l, q^1)
      // has them as columns (V)
      const int g = blk_index(l, tr ? (q ^ 1) : q);
      const BlockMeta &b = t->blk[g];
      const MixedFactors::Blk &x = md.blk[g];
      if (b.s <= 0) continue;
      const int roff = lo - (tr ? b.c0 : b.r0);
      MxSeg sd;
      sd.A = md.d_as + (x.a_off + (long long)roff * m->s_pad[g]) * esz;
      sd.lda = m->s_pad[g];
      sd.a_slice = (long long)x.rows * m->s_pad[g];
      sd.sA = md.d_sa + x.a_soff + roff;
      sd.ldsa = x.rows;
      sd.B = Ws[g];
      sd.ldb = m->s_pad[g];
      sd.b_slice = (long long)R * m->s_pad[g];
      sd.sB = sW[g];
      sd.ldsb = R;
      sd.cs = m->d_rk + m->rk_cs[g];
      sd.nch = m->nrk[g];
      sd.a_c0 = 0;
      sd.b_c0 = 0;
      sd.pad = 0;
      mc.seg_ctab.push_back(-1);  // device table already resident
      mc.segs.push_back(sd);
    }
    const int s1 = (int)mc.segs.size();
    if (s1 == s0) continue;
    for (int i0 = 0; i0 < rows; i0 += MX_TM)
      for (int j0 = 0; j0 < R; j0 += MX_TN)
        mc.ait.push_back({Y + (long long)lo * y_rs, y_cs, y_rs, rows, R, i0, j0, s0, s1, 1, 0});
  }
  p.touch(X, operand_bytes(t->n, R, x_rs, x_cs), "mixed X");
  p.touch(Y, operand_bytes(t->n, R, y_rs, y_cs), "mixed Y");
  return SUCCESS;
}

// The sliced GEMM over nitems work items.
template <int F, int TN>
static void launch_sliced_f(int S, unsigned nitems, const MxItem *I, const MxSeg *G,
                            hipStream_t st) {
  const dim3 b(4 * TN);
  if (S == 1) hipLaunchKernelGGL((k_sliced_gemm<F, 1, TN>), dim3(nitems), b, 0, st, I, G);
  if constexpr (F != MX_F32) {  // FP32 has one slice
    if (S == 2) hipLaunchKernelGGL((k_sliced_gemm<F, 2, TN>), dim3(nitems), b, 0, st, I, G);
    if (S == 3) hipLaunchKernelGGL((k_sliced_gemm<F, 3, TN>), dim3(nitems), b, 0, st, I, G);
    if (S == 4) hipLaunchKernelGGL((k_sliced_gemm<F, 4, TN>), dim3(nitems), b, 0, st, I, G);
  }
}
static void launch_sliced(int fmt, int S, unsigned nitems, const MxItem *I, const MxSeg *G,
                          hipStream_t st) {
  if (fmt == MX_BF16) launch_sliced_f<MX_BF16, MX_TN>(S, nitems, I, G, st);
  if (fmt == MX_I8) launch_sliced_f<MX_I8, MX_TN>(S, nitems, I, G, st);
  if (fmt == MX_F32) launch_sliced_f<MX_F32, MX_TN>(S, nitems, I, G, st);
}

// Work order of the sliced GEMM launches: workgroup b runs on XCD b % 8, and
// each XCD has its own L2.  Items that share operands have the same first
// segment (the rank and column tiles of one W segment read the same slices
// of F_in^T and X; the tiles of one leaf's apply the same rows of F_out and
// the same W slices), so each such group goes whole to one XCD (the least
// loaded so far, load = items x chunks along K) and the eight XCD queues are
// interleaved; queues shorter than the longest are padded with empty items
// (no segments, no rows).
static std::vector<MxItem> xcd_order(const std::vector<MxItem> &items,
                                     const std::vector<MxSeg> &segs) {
  constexpr int NX = 8;
  std::vector<std::vector<MxItem>> q(NX);
  // groups (runs of items with the same segments) and their loads
  std::vector<std::pair<size_t, size_t>> grp;
  std::vector<long long> gl;
  for (size_t i = 0; i < items.size();) {
    size_t j = i;
    while (j < items.size() && items[j].s0 == items[i].s0 && items[j].s1 == items[i].s1) j++;
    long long k = 0;
    for (int sg = items[i].s0; sg < items[i].s1; sg++) k += std::max(1, segs[sg].nch);
    grp.push_back({i, j});
    gl.push_back((long long)(j - i) * std::max(1LL, k));
    i = j;
  }
  long long load[NX] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t g = 0; g < grp.size(); g++) {
    int x = 0;
    for (int y = 1; y < NX; y++)
      if (load[y] < load[x]) x = y;
    for (size_t u = grp[g].first; u < grp[g].second; u++) q[x].push_back(items[u]);
    load[x] += gl[g];
  }
  size_t len = 0;
  for (const auto &v : q) len = std::max(len, v.size());
  const MxItem none{nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<MxItem> out(len * NX, none);
  for (int x = 0; x < NX; x++)
    for (size_t u = 0; u < q[x].size(); u++) out[u * NX + x] = q[x][u];
  return out;
}

// The launches of the gathered stages: segment fix-up, slice X, W GEMMs,
// slice W (summing the split-K partials), the apply GEMMs (Y += F_out W).
static void mixed_emit(Program &p, const MixedCollect &mc, int F_, int S_) {
  const size_t ctab_off = p.push(mc.ctab);
  const size_t sl_off = p.push(mc.sl), wsl_off = p.push(mc.wsl);
  const std::vector<MxItem> wit = xcd_order(mc.wit, mc.segs);
  const std::vector<MxItem> ait = xcd_order(mc.ait, mc.segs);
  const size_t wit_off = p.push(wit), ait_off = p.push(ait);
  const size_t seg_off = p.push(mc.segs);
  const int nsl = (int)mc.sl.size(), nwsl = (int)mc.wsl.size();
  const unsigned nwit = (unsigned)wit.size(), nait = (unsigned)ait.size();
  const int nsegs = (int)mc.segs.size();
  std::vector<long long> fix;
  for (int i = 0; i < nsegs; i++)
    fix.push_back(mc.seg_ctab[i] < 0 ? -1
                                     : (long long)(ctab_off + mc.seg_ctab[i] * sizeof(int)));
  const size_t fix_off = p.push(fix);
  p.add("mixed fixup", [=](hipStream_t st, const unsigned char *blob) {
    if (nsegs)
      hipLaunchKernelGGL(k_mx_fix_segs, dim3((nsegs + 255) / 256), dim3(256), 0, st,
                         reinterpret_cast<MxSeg *>(const_cast<unsigned char *>(blob) + seg_off),
                         reinterpret_cast<const long long *>(blob + fix_off), blob, nsegs);
  });
  p.add("slice X", [=](hipStream_t st, const unsigned char *blob) {
    launch_slice(F_, S_, (unsigned)nsl, reinterpret_cast<const SliceItem *>(blob + sl_off), st);
  });
  p.add("mixed W", [=](hipStream_t st, const unsigned char *blob) {
    if (!nwit) return;
    launch_sliced(F_, S_, nwit, reinterpret_cast<const MxItem *>(blob + wit_off),
                  reinterpret_cast<const MxSeg *>(blob + seg_off), st);
  });
  p.add("slice W", [=](hipStream_t st, const unsigned char *blob) {
    launch_slice(F_, S_, (unsigned)nwsl, reinterpret_cast<const SliceItem *>(blob + wsl_off),
                 st);
  });
  p.add("mixed apply", [=](hipStream_t st, const unsigned char *blob) {
    if (!nait) return;
    launch_sliced(F_, S_, nait, reinterpret_cast<const MxItem *>(blob + ait_off),
                  reinterpret_cast<const MxSeg *>(blob + seg_off), st);
  });
}

int mixed_ensure_mode(hm_dtree *t, int tr, hipStream_t s) {
  MixedFactors *m = t->mixed;
  if (!m) {
    set_error("mixed product: no sliced factors");
    return INPUT_ERROR;
  }
  if (!m->mode[tr].ready) HM_OK(prepare_mode(t, m, tr, s));
  return SUCCESS;
}

int mixed_add(Program &p, Arena &ar, const std::vector<MixedJob> &jobs) {
  for (int F = MX_BF16; F <= MX_F32; F++)
    for (int S = 1; S <= 4; S++) {
      MixedCollect mc;
      bool any = false;
      for (const MixedJob &j : jobs) {
        if (!j.t->mixed || j.t->mixed->fmt != F || j.t->mixed->S != S || j.R <= 0) continue;
        HM_OK(mixed_collect(mc, p, ar, j.t, j.trans, j.R, j.X, j.x_rs, j.x_cs, j.Y, j.y_rs,
                            j.y_cs));
        any = true;
      }
      if (any) mixed_emit(p, mc, F, S);
    }
  return SUCCESS;
}

int mixed_matmat(hm_dtree *t, int trans, int R, const double *X, long x_rs, long x_cs,
                 double *Y, long y_rs, long y_cs, hipStream_t s) {
  MixedFactors *m = t->mixed;
  HM_OK(mixed_ensure_mode(t, trans, s));
  HM_OK(products_diag(t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs, s));
  if (!m->prog || m->trans != trans || m->R != R || m->X != X || m->Y != Y ||
      m->x_rs != x_rs || m->x_cs != x_cs || m->y_rs != y_rs || m->y_cs != y_cs) {
    delete m->prog;
    m->prog = nullptr;
    dfree(m->arena);
    m->arena = nullptr;
    const std::vector<MixedJob> jobs{{t, trans, R, X, x_rs, x_cs, Y, y_rs, y_cs}};
    Program dry;
    Arena count;
    dry.record = true;
    HM_OK(mixed_add(dry, count, jobs));
    dry.allow(Arena::fake_base(), count.off + 256);
    dry.allow(X, operand_bytes(t->n, R, x_rs, x_cs));
    dry.allow(Y, operand_bytes(t->n, R, y_rs, y_cs));
    std::string msg;
    if (dry.audit(&msg) != SUCCESS) {
      set_error("mixed product: audit: %s", msg.c_str());
      return INPUT_ERROR;
    }
    Arena ar;
    ar.counting = false;
    ar.cap = count.off + 256;
    HM_OK(dalloc(&ar.base, ar.cap));
    m->arena = ar.base;
    m->prog = new Program();
    int rc = mixed_add(*m->prog, ar, jobs);
    if (rc == SUCCESS) rc = m->prog->upload(s);
    if (rc != SUCCESS) {
      delete m->prog;
      m->prog = nullptr;
      return rc;
    }
    m->trans = trans;
    m->R = R;
    m->X = X;
    m->Y = Y;
    m->x_rs = x_rs;
    m->x_cs = x_cs;
    m->y_rs = y_rs;
    m->y_cs = y_cs;
  }
  return m->prog->run(s);
}

}  // namespace hm

using namespace hm;

namespace hm {
// One block of format F: lane L loads its 16-byte fragment of A row L % 16
// and B column L % 16 (entries KS/4 * (L/16) ..), one MxDev<F>::mma.
template <int F>
__global__ void k_selftest_lowp(const float *A, const float *B, float *raw) {
  constexpr int KS = MxDev<F>::KS, E = KS / 4;
  const int L = threadIdx.x;
  unsigned char fa[16], fb[16];
  for (int e = 0; e < E; e++) {
    const float a = A[(L & 15) * KS + E * (L >> 4) + e];
    const float b = B[(E * (L >> 4) + e) * 16 + (L & 15)];
    if constexpr (F == MX_BF16) {
      reinterpret_cast<__bf16 *>(fa)[e] = (__bf16)a;
      reinterpret_cast<__bf16 *>(fb)[e] = (__bf16)b;
    } else if constexpr (F == MX_I8) {
      reinterpret_cast<signed char *>(fa)[e] = (signed char)(int)a;
      reinterpret_cast<signed char *>(fb)[e] = (signed char)(int)b;
    } else {
      reinterpret_cast<float *>(fa)[e] = a;
      reinterpret_cast<float *>(fb)[e] = b;
    }
  }
  i4 ra, rb;
  memcpy(&ra, fa, 16);
  memcpy(&rb, fb, 16);
  typename MxDev<F>::acc_t c{};
  c = MxDev<F>::mma(ra, rb, c);
  for (int v = 0; v < 4; v++) raw[4 * L + v] = (float)c[v];
}
}  // namespace hm

extern "C" int hm_hip_selftest_mfma_lowp(int format, const float *hA, const float *hB,
                                         float *raw) {
  HM_OK(ensure_init());
  if (format < HM_FMT_BF16 || format > HM_FMT_FP32) {
    set_error("selftest lowp: unknown format %d", format);
    return INPUT_ERROR;
  }
  const int ks = format == HM_FMT_BF16 ? 32 : (format == HM_FMT_INT8 ? 64 : 16);
  float *d = nullptr;
  HM_OK(dalloc(&d, 2 * 16 * ks + 256));
  hipError_t e = hipMemcpy(d, hA, 16 * ks * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d + 16 * ks, hB, 16 * ks * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    float *a = d, *b = d + 16 * ks, *r = d + 32 * ks;
    if (format == HM_FMT_BF16)
      hipLaunchKernelGGL(k_selftest_lowp<MX_BF16>, dim3(1), dim3(64), 0, 0, a, b, r);
    if (format == HM_FMT_INT8)
      hipLaunchKernelGGL(k_selftest_lowp<MX_I8>, dim3(1), dim3(64), 0, 0, a, b, r);
    if (format == HM_FMT_FP32)
      hipLaunchKernelGGL(k_selftest_lowp<MX_F32>, dim3(1), dim3(64), 0, 0, a, b, r);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(raw, d + 32 * ks, 256 * sizeof(float), hipMemcpyDeviceToHost);
  dfree(d);
  if (e != hipSuccess) {
    set_error("selftest lowp: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

extern "C" int hm_hip_tree_set_format(hm_dtree *t, int format, int slices, void *stream) {
  HM_OK(ensure_init());
  if (!t || format < HM_FMT_FP64 || format > HM_FMT_FP32) {
    set_error("hm_hip_tree_set_format: unknown format %d", format);
    return INPUT_ERROR;
  }
  if (format == HM_FMT_FP64) {
    mixed_free(t->mixed);
    t->mixed = nullptr;
    t->slices = 0;
    t->slice_fmt = 0;
    return SUCCESS;
  }
  const int maxs = format == HM_FMT_FP32 ? 1 : 4;
  if (slices < 1 || slices > maxs) {
    set_error("hm_hip_tree_set_format: %d slices (1..%d for this format)", slices, maxs);
    return INPUT_ERROR;
  }
  const int fmt = format == HM_FMT_BF16 ? MX_BF16 : (format == HM_FMT_INT8 ? MX_I8 : MX_F32);
  return mixed_prepare(t, fmt, slices, stream_of(stream));
}

extern "C" int hm_hip_tree_set_precision(hm_dtree *t, int slices, void *stream) {
  if (slices < 0 || slices > 4) {
    set_error("hm_hip_tree_set_precision: slices must be 0 (FP64) or 1..4");
    return INPUT_ERROR;
  }
  return hm_hip_tree_set_format(t, slices ? HM_FMT_BF16 : HM_FMT_FP64, slices, stream);
}
