// Batched truncated SVD of dense blocks on the GPU (FP64).
//
// The reference factors every off-diagonal block with a full LAPACK dgesdd
// (src/constructors.c:308-367, src/lapack_wrapper.c:8-59) and keeps the
// leading singular triplets by its rank rule.  Here each block A (m x n)
// goes through a randomized range finder with a posterior error test
// (Halko, Martinsson & Tropp 2011, Alg. 4.2 and Lemma 4.1):
//
//   1. Y = A [Omega | Omega_t]        Gaussian sketch, ell + ntest columns
//   2. Y = Q R                        TSQR (hm_tsqr.hip); Q keeps the first
//                                     ell columns, and the trailing block of
//                                     R gives |(I - Q Q^T) A w_t| exactly
//   3. B^T = A^T Q = Q2 R2            second pass over A, TSQR
//   4. R2^T = W S Z^T                 one-sided Jacobi SVD (ell x ell)
//   5. err = 10 sqrt(2/pi) max_t |(I - Q Q^T) A w_t|; accept when
//      err <= accept * thr * sigma_0 (relative rule; sigma_0 from step 4)
//      or <= accept * thr (absolute rule); otherwise ell doubles
//   6. U = Q (R2^T Z)[:, :k] = Q W S,  V = Q2 Z[:, :k]
//
// When ell reaches min(m, n) the sketch spans the whole range of A and the
// result is an exact SVD (no test needed), so small blocks are factored
// exactly.  sigma_j(B)^2 >= sigma_j(A)^2 - err^2, so with err <= 0.1 thr
// sigma0 a rank decision can differ from a full SVD only when a singular
// value lies within 0.5% above the threshold.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "hm_tsqr.h"
#include "hm_tsvd.h"

namespace hm {

static constexpr int MAX_ELL = 1024;

// ------------------------------------------------------------ RNG
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double gauss(unsigned long long key, long long i) {
  unsigned long long a = mix64(key + 2ull * (unsigned long long)i);
  unsigned long long b = mix64(key + 2ull * (unsigned long long)i + 1ull);
  double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
  double u2 = (double)(b >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ------------------------------------------------- block reductions
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *red,
                                          double *out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) red[i * NW + w] = v[i];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int ww = 0; ww < NW; ww++) s += red[threadIdx.x * NW + ww];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// first row index >= r that thread `tid` owns (rows i with i % NT == tid)
__device__ __forceinline__ int own_from(int tid, int r, int NT) {
  if (tid >= r) return tid;
  return tid + ((r - tid + NT - 1) / NT) * NT;
}

// --------------------------------------------------- kernels
struct TsvdBufs {
  double *Y, *Om, *Q, *Bt, *Q2, *M, *Vj, *G, *Z, *Sig;
};

__global__ __launch_bounds__(256) void k_gen_omega(TsvdBlock *blocks,
                                                   const int2 *work,
                                                   TsvdBufs bf,
                                                   unsigned long long seed) {
  const int2 w = work[blockIdx.x];
  const TsvdBlock &B = blocks[w.x];
  const int j = w.y;
  const int n = B.n;
  const unsigned long long key = mix64(seed ^ mix64(B.uid * 0x100000001b3ull + (unsigned long long)j));
  double *col = bf.Om + B.ooff + (long long)j * n;
  for (int i = threadIdx.x; i < n; i += 256) col[i] = gauss(key, i);
}

// Y = A Omega, tiles of 128 rows x 32 sketch columns.
typedef MfmaTile<4, 1, 2, 2, 16> SketchTile;

__global__ __launch_bounds__(256) void k_sketch(TsvdBlock *blocks,
                                                const int4 *work, TsvdBufs bf) {
  __shared__ double lds[SketchTile::LDS_DOUBLES];
  const int4 w = work[blockIdx.x];
  if (w.x < 0) return;  // padding of a shorter XCD sequence
  const TsvdBlock B = blocks[w.x];
  const int m = B.m, n = B.n, L = B.ell + B.ntest;
  const int i0 = w.y * SketchTile::TM, j0 = w.z * SketchTile::TN;
  const double *A = B.A;
  const double *Om = bf.Om + B.ooff;
  auto la = [&](int i, int k) -> double {
    return (i0 + i < m && k < n) ? ldg(A + (long long)(i0 + i) + (long long)k * B.lda) : 0.0;
  };
  auto lb = [&](int k, int j) -> double {
    return (k < n && j0 + j < L) ? Om[k + (long long)(j0 + j) * n] : 0.0;
  };
  SketchTile tile;
  tile.zero();
  if (i0 + SketchTile::TM <= m && n % SketchTile::KC_ == 0)
    tile.run_strided<true>(lds, n, A + i0, 1, B.lda, Om + (long long)j0 * n, 1, n,
                           min(SketchTile::TN, L - j0));
  else
    tile.run<true, false>(lds, n, la, lb);
  double *Y = bf.Y + B.yoff;
  tile.store([&](int i, int j, double v) {
    if (i0 + i < m && j0 + j < L) Y[(i0 + i) + (long long)(j0 + j) * m] = v;
  });
}

// B^T = A^T Q, tiles of 128 (columns of A) x 32 (ell); ProjTile16 (128 x
// 16) for sketches of at most 16 columns, which would leave half of the
// 32-wide tile's MFMA work on zero columns (the pass is close to MFMA-bound:
// a 128 x 32 x 16 FP64 step per 16 KB of A).
typedef MfmaTile<4, 1, 2, 1, 16> ProjTile16;

template <class Tile>
__global__ __launch_bounds__(256) void k_project(TsvdBlock *blocks,
                                                 const int4 *work, TsvdBufs bf) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  const int4 w = work[blockIdx.x];
  if (w.x < 0) return;  // padding of a shorter XCD sequence
  const TsvdBlock B = blocks[w.x];
  const int m = B.m, n = B.n, ell = B.ell;
  const int i0 = w.y * Tile::TM, j0 = w.z * Tile::TN;
  const double *A = B.A;
  const double *Q = bf.Q + B.qoff;
  auto la = [&](int i, int k) -> double {
    return (i0 + i < n && k < m) ? ldg(A + (long long)k + (long long)(i0 + i) * B.lda) : 0.0;
  };
  auto lb = [&](int k, int j) -> double {
    return (k < m && j0 + j < ell) ? Q[k + (long long)(j0 + j) * m] : 0.0;
  };
  Tile tile;
  tile.zero();
  if (i0 + Tile::TM <= n && m % Tile::KC_ == 0)
    tile.template run_strided<false>(lds, m, A + (long long)i0 * B.lda, B.lda, 1,
                                     Q + (long long)j0 * m, 1, m, min(Tile::TN, ell - j0));
  else
    tile.template run<false, false>(lds, m, la, lb);
  double *Bt = bf.Bt + B.boff;
  tile.store([&](int i, int j, double v) {
    if (i0 + i < n && j0 + j < ell) Bt[(i0 + i) + (long long)(j0 + j) * n] = v;
  });
}

// One-sided Jacobi SVD of M (nc x nc); V accumulates the rotations.
// Outputs G = (M V)[:, perm[:k]] (U carries sigma), Z = V[:, perm[:k]],
// sorted singular values and the rank k by the block's rule.
// finish_only: M and V were already rotated by the block sweeps (columns of
// M V-orthogonal); only the ordering, rank rule and gathers remain.
__global__ __launch_bounds__(256) void k_jacobi(TsvdBlock *blocks,
                                                const int *list, TsvdBufs bf,
                                                double accept, int finish_only) {
  __shared__ double sig[MAX_ELL];
  __shared__ int perm[MAX_ELL];
  __shared__ int rotated;
  TsvdBlock &B = blocks[list[blockIdx.x]];
  const int nc = B.ell;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = 4;
  double *M = bf.M + B.moff;
  double *V = bf.Vj + B.moff;
  if (!finish_only)
    for (int e = threadIdx.x; e < nc * nc; e += 256)
      V[e] = ((e % nc) == (e / nc)) ? 1.0 : 0.0;
  __syncthreads();
  const int np = nc + (nc & 1);
  const double tol = fmax(1.0, sqrt((double)nc)) * 2.220446049250313e-16;
  for (int sweep = 0; sweep < (finish_only ? 0 : 60) && nc > 1; sweep++) {
    // every thread has read the previous sweep's flag before it is cleared,
    // so all waves take the same exit decision
    __syncthreads();
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < np - 1; r++) {
      for (int pi = wave; pi < np / 2; pi += NW) {
        const int pa = pi, pb = np - 1 - pi;
        const int a = (pa == 0) ? 0 : ((pa - 1 + r) % (np - 1)) + 1;
        const int b = (pb == 0) ? 0 : ((pb - 1 + r) % (np - 1)) + 1;
        if (a >= nc || b >= nc) continue;
        const int p = min(a, b), q = max(a, b);
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < nc; i += 64) {
          const double mp = M[i + (long long)p * nc], mq = M[i + (long long)q * nc];
          al += mp * mp;
          be += mq * mq;
          ga += mp * mq;
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (ga == 0.0 || fabs(ga) <= tol * sqrt(al * be)) continue;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = lane; i < nc; i += 64) {
          const double mp = M[i + (long long)p * nc], mq = M[i + (long long)q * nc];
          M[i + (long long)p * nc] = c * mp - s * mq;
          M[i + (long long)q * nc] = s * mp + c * mq;
          const double vp = V[i + (long long)p * nc], vq = V[i + (long long)q * nc];
          V[i + (long long)p * nc] = c * vp - s * vq;
          V[i + (long long)q * nc] = s * vp + c * vq;
        }
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    if (!rotated) break;
  }
  __syncthreads();
  for (int j = wave; j < nc; j += NW) {
    double ss = 0.0;
    for (int i = lane; i < nc; i += 64) {
      const double x = M[i + (long long)j * nc];
      ss += x * x;
    }
    ss = wave_sum(ss);
    if (lane == 0) sig[j] = sqrt(ss);
  }
  __syncthreads();
  // descending order, ties by index; a NaN sorts last, so perm is always a
  // permutation of 0..nc-1
  for (int j = threadIdx.x; j < nc; j += 256) {
    int rank = 0;
    const double sj = sig[j] == sig[j] ? sig[j] : -1.0;
    for (int i = 0; i < nc; i++) {
      const double si = sig[i] == sig[i] ? sig[i] : -1.0;
      rank += (si > sj) || (si == sj && i < j);
    }
    perm[rank] = j;
  }
  __syncthreads();
  __shared__ int kk;
  if (threadIdx.x == 0) {
    const double s0 = sig[perm[0]];
    int cut = 1;
    if (B.rule == RULE_RELATIVE) {
      if (s0 > B.thr)
        for (cut = 1; cut < nc; cut++)
          if (sig[perm[cut]] < B.thr * s0) break;
    } else {
      for (cut = 1; cut < nc; cut++)
        if (sig[perm[cut]] <= B.thr) break;
    }
    B.k = cut;
    B.sigma0 = s0;
    kk = cut;
    // posterior test: B.terr = max_t |(I - Q Q^T) A w_t| bounds |(I - Q Q^T) A|_2
    // by 10 sqrt(2/pi) B.terr (Halko et al. Lemma 4.1; loose by up to the
    // square root of the tail length when the tail is flat); B.perr = the
    // same residual after one power step, |E E^T z| / |E^T z| ~ sigma_1(E),
    // taken with a safety factor 2 (+inf until the power step has run: it
    // runs only for blocks the first test rejects).  Either passing accepts.
    const double scale = accept * B.thr * (B.rule == RULE_RELATIVE ? s0 : 1.0);
    bool ok = true;
    if (B.ntest > 0) {
      const double est = 10.0 * 0.7978845608028654 * B.terr;
      const double pest = 2.0 * B.perr;
      const bool tiny = (B.rule == RULE_RELATIVE) && !(s0 > B.thr);
      ok = tiny || est <= scale || pest <= scale;
      if (ok && !tiny && cut == nc) ok = false;  // sketch saturated: widen
      B.err = fmin(est, pest);
    }
    int need = 0;
    while (need < nc && sig[perm[need]] >= 0.05 * scale) need++;
    B.need = need;
    B.status = ok ? 1 : 0;
  }
  __syncthreads();
  double *S = bf.Sig + B.soff;
  for (int j = threadIdx.x; j < nc; j += 256) S[j] = sig[perm[j]];
  double *G = bf.G + B.moff, *Z = bf.Z + B.moff;
  for (int e = threadIdx.x; e < nc * kk; e += 256) {
    const int i = e % nc, j = e / nc;
    G[i + (long long)j * nc] = M[i + (long long)perm[j] * nc];
    Z[i + (long long)j * nc] = V[i + (long long)perm[j] * nc];
  }
}

// Several host arrays in one device allocation and one copy (each at a
// 256-byte boundary); the host staging buffer is kept until the copy ends.
struct HostPack {
  struct Part {
    void **dst;
    const void *src;
    size_t off, bytes;
  };
  std::vector<Part> parts;
  size_t total = 0;
  template <class T>
  void add(T **dst, const std::vector<T> &h) {
    total = (total + 255) & ~(size_t)255;
    parts.push_back({reinterpret_cast<void **>(dst), h.data(), total, h.size() * sizeof(T)});
    total += h.size() * sizeof(T);
  }
  int upload(char **blob, hipStream_t s) {
    HM_OK(dalloc(blob, total + 256));
    std::vector<char> host(total + 256, 0);
    for (const Part &p : parts)
      if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
    HM_HIP(hipMemcpyAsync(*blob, host.data(), host.size(), hipMemcpyHostToDevice, s));
    for (const Part &p : parts) *p.dst = *blob + p.off;
    HM_HIP(hipStreamSynchronize(s));
    return SUCCESS;
  }
};

template <class T>
static int to_device(T **d, const std::vector<T> &h, hipStream_t s) {
  HM_OK(dalloc(d, h.size()));
  if (!h.empty())
    HM_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T),
                          hipMemcpyHostToDevice, s));
  return SUCCESS;
}

// ------------------------------------------------ block one-sided Jacobi
// For cores wider than BJ_SCALAR_MAX columns: the columns of M (nc x nc)
// are cut into blocks of BJ_W; each round of a sweep pairs the blocks
// round robin and one workgroup per pair (a, b) (1) forms the Gram matrix
// G = M_S^T M_S of the <= 2 BJ_W columns S = a u b on MFMA, (2) diagonalises G
// in LDS by cyclic two-sided Jacobi (J accumulates the rotations); then
// k_bjac_apply, one workgroup per 64-row tile, (3) applies J to the columns
// S of M and of V (a tile loop inside the pair's workgroup was its longest
// phase).  A pair whose first inner sweep rotates anything marks its matrix
// as not converged.  Blocks of 16 columns measured 3-10 % faster than
// blocks of 32 on whole compressions and H x H products: twice the pair
// visits per sweep, but each inner sweep has 31 rounds over a 32 x 32 Gram
// instead of 63 over 64 x 64, so the LDS rotation work per sweep halves
// while the Gram and apply flops stay the same.
static constexpr int BJ_W = 16;
static constexpr int BJ_SCALAR_MAX = 64;
// inner cyclic sweeps per pair visit: the outer sweeps revisit every pair
// until none rotates, so a partly diagonalised pair only costs another
// visit, while each inner sweep costs 2 (2 BJ_W - 1) barriers (one sweep
// measured faster than two or four)
static constexpr int BJ_INNER = 1;

__global__ __launch_bounds__(256) void k_eye(TsvdBlock *blocks, const int *list, TsvdBufs bf) {
  const TsvdBlock B = blocks[list[blockIdx.x]];
  const int nc = B.ell;
  double *V = bf.Vj + B.moff;
  const long long E = (long long)nc * nc;
  for (long long e = (long long)blockIdx.y * 256 + threadIdx.x; e < E;
       e += (long long)gridDim.y * 256)
    V[e] = ((e % nc) == (e / nc)) ? 1.0 : 0.0;
}

// Tiles of the block Jacobi for blocks of BW columns (pairs of PW = 2 BW):
// the pair's Gram (PW x PW) and the application of J to 64-row tiles.
template <int BW> struct BjTiles;
template <> struct BjTiles<32> {
  typedef MfmaTile<2, 2, 2, 2, 32> Gram;   // 64 x 64, K-chunks of 32
  typedef MfmaTile<2, 2, 2, 2, 64> Apply;  // 64 x 64, K = 64 in one chunk
};
template <> struct BjTiles<16> {
  typedef MfmaTile<2, 2, 1, 1, 32> Gram;   // 32 x 32, K-chunks of 32
  typedef MfmaTile<2, 2, 2, 1, 32> Apply;  // 64 x 32, K = 32 in one chunk
};

template <int BW>
__global__ __launch_bounds__(256) void k_bjac_pair(TsvdBlock *blocks, const int4 *items,
                                                   TsvdBufs bf, int *flags, double *jout,
                                                   int *rotated) {
  constexpr int PW = 2 * BW, NP = PW / 2, LG = PW + 1;
  typedef typename BjTiles<BW>::Gram GramTile;
  // the MFMA staging area lives inside G: the Gram tile is stored to G only
  // after its last K-chunk, and G is dead while J is applied
  constexpr int GSZ = GramTile::LDS_DOUBLES > PW * LG ? GramTile::LDS_DOUBLES : PW * LG;
  __shared__ double Gs[GSZ];
  __shared__ double J[PW][LG];
  double *lds = Gs;
  auto G = [&](int r, int c) -> double & { return Gs[r * LG + c]; };
  __shared__ double cs[NP][2];
  __shared__ int pq[NP][2];
  __shared__ int rot;
  const int4 it = items[blockIdx.x];
  const TsvdBlock B = blocks[it.x];
  const int nc = B.ell;
  const int ca = it.y * BW, wa = min(BW, nc - ca);
  const int cb = it.z * BW, wb = min(BW, nc - cb);
  const int W = wa + wb;
  auto col = [&](int i) -> int { return i < wa ? ca + i : cb + (i - wa); };
  const double *M = bf.M + B.moff;
  const int tid = threadIdx.x;
  // (1) Gram matrix of the pair's columns
  {
    GramTile tile;
    tile.zero();
    auto la = [&](int i, int k) -> double {
      return (i < W && k < nc) ? M[k + (long long)col(i) * nc] : 0.0;
    };
    auto lb = [&](int k, int j) -> double {
      return (j < W && k < nc) ? M[k + (long long)col(j) * nc] : 0.0;
    };
    tile.template run<false, false>(lds, nc, la, lb);
    __syncthreads();  // the staging area is read before G overwrites it
    tile.store([&](int i, int j, double v) { G(i, j) = v; });
  }
  const double tol = fmax(1.0, sqrt((double)nc)) * 2.220446049250313e-16;
  __syncthreads();
  // A pair whose Gram passes every rotation test is left alone: with no
  // rotation G never changes, so the sweep below would test exactly these
  // entries and rotate nothing (later outer sweeps revisit mostly such pairs).
  {
    int mine = 0;
    for (int e = tid; e < PW * PW; e += 256) {
      const int p = e / PW, q = e % PW;
      if (p < q && q < W) {
        const double gpq = G(p, q);
        mine |= (gpq != 0.0 && fabs(gpq) > tol * sqrt(G(p, p) * G(q, q)));
      }
    }
    if (!__syncthreads_or(mine)) {
      if (tid == 0) rotated[blockIdx.x] = 0;
      return;
    }
  }
  for (int e = tid; e < PW * PW; e += 256) J[e / PW][e % PW] = ((e / PW) == (e % PW)) ? 1.0 : 0.0;
  if (tid == 0) rot = 0;
  __syncthreads();
  int rot_any = 0;
  // (2) cyclic Jacobi on G (round robin over PW indices, NP disjoint pairs)
  for (int sweep = 0; sweep < BJ_INNER; sweep++) {
    __syncthreads();
    int any = 0;
    for (int r = 0; r < PW - 1; r++) {
      if (tid < NP) {
        const int pa = tid, pb = PW - 1 - tid;
        const int a = (pa == 0) ? 0 : ((pa - 1 + r) % (PW - 1)) + 1;
        const int b = (pb == 0) ? 0 : ((pb - 1 + r) % (PW - 1)) + 1;
        const int p = min(a, b), q = max(a, b);
        double c = 1.0, sn = 0.0;
        if (q < W) {
          const double gpp = G(p, p), gqq = G(q, q), gpq = G(p, q);
          if (gpq != 0.0 && fabs(gpq) > tol * sqrt(gpp * gqq)) {
            const double zeta = (gqq - gpp) / (2.0 * gpq);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
            any = 1;
          }
        }
        cs[tid][0] = c;
        cs[tid][1] = sn;
        pq[tid][0] = p;
        pq[tid][1] = q;
      }
      __syncthreads();
      // G <- R^T G R in one pass: thread owns whole 2 x 2 blocks (rows
      // p_i, q_i of pair i x columns p_j, q_j of pair j), so the update is
      // in place with no barrier between the row and the column rotation
      for (int e = tid; e < NP * NP; e += 256) {
        const int ti = e / NP, tj = e % NP;
        const double si = cs[ti][1], sj = cs[tj][1];
        if (si != 0.0 || sj != 0.0) {
          const double ci = cs[ti][0], cj = cs[tj][0];
          const int pi = pq[ti][0], qi = pq[ti][1], pj = pq[tj][0], qj = pq[tj][1];
          const double b00 = G(pi, pj), b01 = G(pi, qj), b10 = G(qi, pj), b11 = G(qi, qj);
          const double r00 = ci * b00 - si * b10, r01 = ci * b01 - si * b11;
          const double r10 = si * b00 + ci * b10, r11 = si * b01 + ci * b11;
          G(pi, pj) = cj * r00 - sj * r01;
          G(pi, qj) = sj * r00 + cj * r01;
          G(qi, pj) = cj * r10 - sj * r11;
          G(qi, qj) = sj * r10 + cj * r11;
        }
      }
      for (int e = tid; e < NP * PW; e += 256) {  // columns p, q of J
        const int t = e / PW, k = e % PW;
        const double sn = cs[t][1];
        if (sn != 0.0) {
          const double c = cs[t][0];
          const int p = pq[t][0], q = pq[t][1];
          const double jp = J[k][p], jq = J[k][q];
          J[k][p] = c * jp - sn * jq;
          J[k][q] = sn * jp + c * jq;
        }
      }
      __syncthreads();
    }
    // any rotation in this sweep (only threads < NP decide rotations)
    if (any) rot = 1;
    __syncthreads();
    const int go = rot;
    __syncthreads();
    if (sweep == 0 && go && tid == 0) flags[it.x] = 1;
    rot_any |= go;
    if (!go) break;
    if (tid == 0) rot = 0;
  }
  __syncthreads();
  // (3) J goes out for k_bjac_apply (one workgroup per 64-row tile of M_S
  // and of V_S), with whether it differs from the identity
  double *Jg = jout + (size_t)blockIdx.x * (PW * PW);
  for (int e = tid; e < PW * PW; e += 256) Jg[e] = J[e / PW][e % PW];
  if (tid == 0) rotated[blockIdx.x] = rot_any;
}

// X_S <- X_S J for one 64-row tile of M (side 0) or V (side 1) of one pair:
// in place (the tile's rows are read whole before they are written; pairs
// of one round touch disjoint columns).
template <int BW>
__global__ __launch_bounds__(256) void k_bjac_apply(TsvdBlock *blocks, const int4 *items,
                                                    TsvdBufs bf, const double *jin,
                                                    const int *rotated) {
  constexpr int PW = 2 * BW;
  typedef typename BjTiles<BW>::Apply ApplyTile;
  __shared__ double lds[ApplyTile::LDS_DOUBLES];
  if (!rotated[blockIdx.x]) return;
  const int4 it = items[blockIdx.x];
  const TsvdBlock B = blocks[it.x];
  const int nc = B.ell;
  const int r0 = blockIdx.y * 64;
  if (r0 >= nc) return;
  const int ca = it.y * BW, wa = min(BW, nc - ca);
  const int cb = it.z * BW, wb = min(BW, nc - cb);
  const int W = wa + wb;
  auto col = [&](int i) -> int { return i < wa ? ca + i : cb + (i - wa); };
  double *X = (blockIdx.z ? bf.Vj : bf.M) + B.moff;
  const double *Jg = jin + (size_t)blockIdx.x * (PW * PW);
  ApplyTile tile;
  tile.zero();
  auto la = [&](int i, int k) -> double {
    return (r0 + i < nc && k < W) ? X[(r0 + i) + (long long)col(k) * nc] : 0.0;
  };
  auto lb = [&](int k, int j) -> double { return Jg[k * PW + j]; };
  tile.template run<true, false>(lds, PW, la, lb);
  tile.store([&](int i, int j, double v) {
    if (r0 + i < nc && j < W) X[(r0 + i) + (long long)col(j) * nc] = v;
  });
}

// SVD cores of the listed blocks (M = the core, nc = ell): scalar Jacobi for
// narrow cores, block sweeps + finish for wide ones.  Synchronizes.
static int svd_cores(TsvdBlock *d_blocks, const std::vector<TsvdBlock> &blocks,
                     const std::vector<int> &list, const TsvdBufs &bf, double accept,
                     hipStream_t s) {
  std::vector<int> small, big;
  for (int i : list) (blocks[i].ell > BJ_SCALAR_MAX ? big : small).push_back(i);
  std::vector<void *> mem;
  auto cleanup = [&]() {
    for (void *p : mem) dfree(p);
  };
  int rc = SUCCESS;
  if (!big.empty()) {
    // round-robin schedule per matrix; items of round r of every matrix
    int maxr = 0;
    std::vector<std::vector<int4>> rounds;
    for (int i : big) {
      const int nb = (blocks[i].ell + BJ_W - 1) / BJ_W;
      const int np = nb + (nb & 1);
      maxr = std::max(maxr, np - 1);
    }
    rounds.resize(maxr);
    for (int i : big) {
      const int nb = (blocks[i].ell + BJ_W - 1) / BJ_W;
      const int np = nb + (nb & 1);
      for (int r = 0; r < np - 1; r++)
        for (int t = 0; t < np / 2; t++) {
          const int pa = t, pb = np - 1 - t;
          const int a = (pa == 0) ? 0 : ((pa - 1 + r) % (np - 1)) + 1;
          const int b = (pb == 0) ? 0 : ((pb - 1 + r) % (np - 1)) + 1;
          if (a >= nb || b >= nb) continue;
          rounds[r].push_back(make_int4(i, std::min(a, b), std::max(a, b), 0));
        }
    }
    std::vector<int4> all;
    std::vector<size_t> roff(maxr + 1, 0);
    for (int r = 0; r < maxr; r++) {
      roff[r] = all.size();
      all.insert(all.end(), rounds[r].begin(), rounds[r].end());
    }
    roff[maxr] = all.size();
    int4 *d_items = nullptr;
    int *d_big = nullptr, *d_flags = nullptr, *d_rot = nullptr;
    double *d_j = nullptr;
    size_t maxn = 1;
    int maxnc = 1;
    for (int r = 0; r < maxr; r++) maxn = std::max(maxn, roff[r + 1] - roff[r]);
    for (int i : big) maxnc = std::max(maxnc, blocks[i].ell);
    rc = to_device(&d_items, all, s);
    if (rc == SUCCESS) { mem.push_back(d_items); rc = to_device(&d_big, big, s); }
    if (rc == SUCCESS) { mem.push_back(d_big); rc = dalloc(&d_flags, blocks.size()); }
    if (rc == SUCCESS) { mem.push_back(d_flags); rc = dalloc(&d_rot, maxn); }
    if (rc == SUCCESS) { mem.push_back(d_rot); rc = dalloc(&d_j, maxn * 4 * BJ_W * BJ_W); }
    if (rc == SUCCESS) mem.push_back(d_j);
    if (rc != SUCCESS) {
      cleanup();
      return rc;
    }
    const unsigned row_tiles = (unsigned)((maxnc + 63) / 64);
    Timed tb = timing_begin("block jacobi", s);
    hipLaunchKernelGGL(k_eye, dim3((unsigned)big.size(), 64), dim3(256), 0, s, d_blocks, d_big, bf);
    std::vector<int> hflags(blocks.size());
    for (int sweep = 0; sweep < 30; sweep++) {
      if (hipMemsetAsync(d_flags, 0, blocks.size() * sizeof(int), s) != hipSuccess) break;
      for (int r = 0; r < maxr; r++) {
        const unsigned n = (unsigned)(roff[r + 1] - roff[r]);
        if (n) {
          hipLaunchKernelGGL(k_bjac_pair<BJ_W>, dim3(n), dim3(256), 0, s, d_blocks,
                             d_items + roff[r], bf, d_flags, d_j, d_rot);
          hipLaunchKernelGGL(k_bjac_apply<BJ_W>, dim3(n, row_tiles, 2), dim3(256), 0, s,
                             d_blocks, d_items + roff[r], bf, d_j, d_rot);
        }
      }
      hipError_t e = hipMemcpyAsync(hflags.data(), d_flags, blocks.size() * sizeof(int),
                                    hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        set_error("block Jacobi: %s", hipGetErrorString(e));
        cleanup();
        return ALLOCATION_FAILURE;
      }
      bool any = false;
      for (int i : big) any |= hflags[i] != 0;
      if (!any) break;
    }
    hipLaunchKernelGGL(k_jacobi, dim3((unsigned)big.size()), dim3(256), 0, s, d_blocks, d_big,
                       bf, accept, 1);
    timing_end(tb, s);
  }
  if (!small.empty()) {
    int *d_small = nullptr;
    rc = to_device(&d_small, small, s);
    if (rc != SUCCESS) {
      cleanup();
      return rc;
    }
    mem.push_back(d_small);
    hipLaunchKernelGGL(k_jacobi, dim3((unsigned)small.size()), dim3(256), 0, s, d_blocks,
                       d_small, bf, accept, 0);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) {
    set_error("svd cores: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

struct CopyItem {
  const double *src;
  double *dst;
  long long count;
};

__global__ __launch_bounds__(256) void k_copy_list(const CopyItem *items) {
  const CopyItem it = items[blockIdx.x];
  for (long long e = threadIdx.x + (long long)blockIdx.y * 256; e < it.count;
       e += 256ll * gridDim.y)
    it.dst[e] = it.src[e];
}

// D = U V^T (m x n, ld m) for the dense recompression method.
typedef MfmaTile<2, 2, 2, 2, 16> DenseTile;
struct OuterJob {
  const double *U, *V;
  double *D;
  int m, n, s, pad;
};
__global__ __launch_bounds__(256) void k_outer(const OuterJob *jobs, const int4 *work) {
  __shared__ double lds[DenseTile::LDS_DOUBLES];
  const int4 w = work[blockIdx.x];
  const OuterJob J = jobs[w.x];
  const int i0 = w.y * DenseTile::TM, j0 = w.z * DenseTile::TN;
  auto la = [&](int i, int k) -> double {
    return (i0 + i < J.m && k < J.s) ? ldg(J.U + (i0 + i) + (long long)k * J.m) : 0.0;
  };
  auto lb = [&](int k, int j) -> double {
    return (k < J.s && j0 + j < J.n) ? ldg(J.V + (j0 + j) + (long long)k * J.n) : 0.0;
  };
  DenseTile tile;
  tile.zero();
  tile.run<true, true>(lds, J.s, la, lb);
  tile.store([&](int i, int j, double v) {
    if (i0 + i < J.m && j0 + j < J.n) stg(J.D + (i0 + i) + (long long)(j0 + j) * J.m, v);
  });
}

// ------------------------------------------------------------ driver

namespace {
// Work order for the sketch and projection passes: the items of one (block,
// row tile) — its column tiles, which all read the same rows of A — go to
// one XCD (workgroup b runs on XCD b % 8), so that the A tile is fetched into
// that XCD's L2 once instead of once per XCD.  Groups are dealt greedily to
// the XCD with the least work so far; the shorter XCD sequences are padded
// with empty items (x = -1).
std::vector<int4> xcd_grouped(const std::vector<int4> &items) {
  constexpr int NX = 8;
  std::vector<std::vector<int4>> seq(NX);
  for (size_t a = 0; a < items.size();) {
    size_t b = a + 1;
    while (b < items.size() && items[b].x == items[a].x && items[b].y == items[a].y) b++;
    int x = 0;
    for (int k = 1; k < NX; k++)
      if (seq[k].size() < seq[x].size()) x = k;
    seq[x].insert(seq[x].end(), items.begin() + a, items.begin() + b);
    a = b;
  }
  size_t len = 0;
  for (auto &q : seq) len = std::max(len, q.size());
  std::vector<int4> out(len * NX, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < NX; x++)
    for (size_t t = 0; t < seq[x].size(); t++) out[t * NX + x] = seq[x][t];
  return out;
}

// The round's work buffers (sketch, Omega, Q, B^T, Q2, cores, sigma) carved
// from one device buffer that is kept between rounds and calls: each round
// ends with its stream synchronized, so the buffer is idle in between.  A
// round that finds it taken (another host thread), runs on another device
// or needs more than kRoundKeep bytes allocates its own.
struct RoundCache {
  std::mutex mu;
  bool busy = false;
  int device = -1;
  char *buf = nullptr;
  size_t cap = 0;
};
RoundCache g_round;
constexpr size_t kRoundKeep = (size_t)8 << 30;

// Holds the cached buffer for one round (or owns a buffer of its own).  The
// destructor waits for the round's stream before the buffer is freed or
// handed back, so a round left early on an error path (kernels still in
// flight) never lets another thread reuse memory that is being written.
struct RoundBuffer {
  char *base = nullptr;
  bool cached = false, own = false;
  hipStream_t stream = nullptr;
  int get(size_t need, hipStream_t s) {
    stream = s;
    int dev = -1;
    if (need <= kRoundKeep && hipGetDevice(&dev) == hipSuccess) {
      std::lock_guard<std::mutex> lock(g_round.mu);
      if (!g_round.busy && (g_round.device < 0 || g_round.device == dev)) {
        g_round.busy = cached = true;
        g_round.device = dev;
      }
    }
    if (cached) {
      if (g_round.cap < need) {
        const size_t want = std::max(need, g_round.cap + g_round.cap / 2);
        dfree(g_round.buf);
        g_round.buf = nullptr;
        g_round.cap = 0;
        HM_OK(dalloc(&g_round.buf, want));
        g_round.cap = want;
      }
      base = g_round.buf;
      return SUCCESS;
    }
    HM_OK(dalloc(&base, need));
    own = true;
    return SUCCESS;
  }
  ~RoundBuffer() {
    if (own || cached) (void)hipStreamSynchronize(stream);
    if (own) dfree(base);
    if (cached) {
      std::lock_guard<std::mutex> lock(g_round.mu);
      g_round.busy = false;
    }
  }
};

}  // namespace

// Frees the kept round buffer unless a round holds it (then BUSY: returns
// INPUT_ERROR and keeps it).
int release_round_buffer() {
  std::lock_guard<std::mutex> lock(g_round.mu);
  if (g_round.busy) return INPUT_ERROR;
  dfree(g_round.buf);
  g_round.buf = nullptr;
  g_round.cap = 0;
  g_round.device = -1;
  return SUCCESS;
}

namespace {
struct Scoped {
  std::vector<void *> ptrs;
  ~Scoped() {
    for (void *p : ptrs) dfree(p);
  }
  template <class T>
  T *keep(T *p) {
    ptrs.push_back((void *)p);
    return p;
  }
};
}  // namespace

int tsvd_run(std::vector<TsvdBlock> &blocks, const TsvdOptions &opt,
             double **u_pool, double **v_pool, TsvdStats *stats,
             std::vector<double> *sigma, int sigma_cap, hipStream_t s) {
  *u_pool = nullptr;
  *v_pool = nullptr;
  TsvdStats st;
  const int nb = (int)blocks.size();
  const int cap = std::min(opt.sketch_max, MAX_ELL);
  std::vector<int> active;
  for (int i = 0; i < nb; i++) {
    TsvdBlock &b = blocks[i];
    b.k = 0;
    b.status = 0;
    b.uout = b.vout = nullptr;
    const int mn = std::min(b.m, b.n);
    if (mn <= 0) {
      b.k = 1;  // empty block: rank 1, zero factors (handled on the host side)
      b.ell = 0;
      continue;
    }
    b.ell = std::max(1, std::min(std::min(b.ell, mn), cap));
    b.ntest = (b.ell >= mn) ? 0 : std::min(opt.n_test, b.m - b.ell);
    active.push_back(i);
  }
  if (sigma) sigma->assign((size_t)nb * sigma_cap, 0.0);
  struct RoundOut {
    double *u, *v;
  };
  std::vector<RoundOut> round_pools;
  Scoped round_mem;
  std::vector<long long> uoff_final(nb, 0), voff_final(nb, 0);
  // Rounds after a failed one continue its range finder: column j of the
  // sketch depends only on (block, j), so a retried block whose failed
  // width w is a multiple of the 64-column QR panel keeps the first w
  // columns of Q (the sketch's basis), of Q2 (B^T's) and R2 (carried out
  // of the round buffer before the next round reuses it), and only the new
  // columns are sketched, orthonormalised and projected; the continued
  // QRs recompute exactly what a fresh round computes for those columns.
  struct Carry {
    int cols = 0;
    const double *Q = nullptr, *Q2 = nullptr, *Mt = nullptr;
  };
  std::vector<Carry> carry(nb);
  struct CarryBuf {
    double *p = nullptr;
    size_t bytes = 0;
    ~CarryBuf() { dfree(p); }
  } carry_cur;
  auto can_carry = [&](const TsvdBlock &b) {
    return b.ell >= 64 && b.ell % 64 == 0 && b.ell < std::min(std::min(b.m, b.n), cap);
  };
  while (!active.empty()) {
    st.rounds++;
    const auto r0 = std::chrono::steady_clock::now();
    Scoped mem;
    // offsets for this round
    long long ytot = 0, otot = 0, qtot = 0, btot = 0, mtot = 0, stot = 0;
    for (int i : active) {
      TsvdBlock &b = blocks[i];
      const int L = b.ell + b.ntest;
      b.yoff = ytot; ytot += (long long)b.m * L;
      b.ooff = otot; otot += (long long)b.n * L;
      b.qoff = qtot; qtot += (long long)b.m * b.ell;
      b.boff = btot; btot += (long long)b.n * b.ell;
      b.moff = mtot; mtot += (long long)b.ell * b.ell;
      b.soff = stot; stot += b.ell;
      b.status = 0;
      b.err = 0.0;
      b.terr = 0.0;
      b.perr = INFINITY;
      st.max_sketch = std::max(st.max_sketch, b.ell);
    }
    TsvdBufs bf;
    RoundBuffer rbuf;
    // R2^T of every block (M before the SVD overwrites it), kept when a
    // block of this round could be continued
    bool any_carry = false;
    for (int i : active) any_carry |= can_carry(blocks[i]);
    double *Ms = nullptr;
    {
      const long long sizes[11] = {ytot, otot, qtot, btot, btot, mtot, mtot, mtot, mtot, stot,
                                   any_carry ? mtot : 0};
      double **dst[11] = {&bf.Y, &bf.Om, &bf.Q, &bf.Bt, &bf.Q2, &bf.M,
                          &bf.Vj, &bf.G, &bf.Z, &bf.Sig, &Ms};
      size_t offs[11], need = 0;
      for (int q = 0; q < 11; q++) {
        offs[q] = need;
        need += ((size_t)sizes[q] * sizeof(double) + 16 + 255) & ~(size_t)255;
      }
      HM_OK(rbuf.get(need, s));
      for (int q = 0; q < 11; q++) *dst[q] = reinterpret_cast<double *>(rbuf.base + offs[q]);
      if (!any_carry) Ms = nullptr;  // no region was carved for it
    }
    // work lists, biggest K first for balance
    std::vector<int> order = active;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
      return (long long)blocks[x].n > (long long)blocks[y].n;
    });
    std::vector<int2> wo;
    std::vector<int4> ws, wp, wp16;
    for (int i : order) {
      const TsvdBlock &b = blocks[i];
      const int L = b.ell + b.ntest;
      // a continued block: only its new columns (and test columns)
      const int c0 = carry[i].cols;
      for (int j = c0; j < L; j++) wo.push_back(make_int2(i, j));
      const int mt = (b.m + SketchTile::TM - 1) / SketchTile::TM;
      const int ntl = (L + SketchTile::TN - 1) / SketchTile::TN;
      for (int a = 0; a < mt; a++)
        for (int c = c0 / SketchTile::TN; c < ntl; c++) ws.push_back(make_int4(i, a, c, 0));
    }
    std::vector<int> order_m = order;
    std::sort(order_m.begin(), order_m.end(), [&](int x, int y) {
      return (long long)blocks[x].m > (long long)blocks[y].m;
    });
    for (int i : order_m) {
      const TsvdBlock &b = blocks[i];
      const bool narrow = b.ell <= ProjTile16::TN;
      const int mt = (b.n + SketchTile::TM - 1) / SketchTile::TM;
      const int ntl = narrow ? 1 : (b.ell + SketchTile::TN - 1) / SketchTile::TN;
      const int c0 = narrow ? 0 : carry[i].cols / SketchTile::TN;
      for (int a = 0; a < mt; a++)
        for (int c = c0; c < ntl; c++) (narrow ? wp16 : wp).push_back(make_int4(i, a, c, 0));
    }
    ws = xcd_grouped(ws);
    wp = xcd_grouped(wp);
    wp16 = xcd_grouped(wp16);
    // the block table and the work lists: one allocation, one copy
    TsvdBlock *d_blocks = nullptr;
    int2 *d_wo;
    int4 *d_ws, *d_wp;
    int *d_list;
    int4 *d_wp16;
    {
      HostPack hp;
      hp.add(&d_blocks, blocks);
      hp.add(&d_wo, wo);
      hp.add(&d_ws, ws);
      hp.add(&d_wp, wp);
      hp.add(&d_wp16, wp16);
      hp.add(&d_list, order);
      char *pack = nullptr;
      HM_OK(hp.upload(&pack, s));
      mem.keep(pack);
    }
    timing_host("round setup (host)", std::chrono::duration<double, std::milli>(
                                          std::chrono::steady_clock::now() - r0).count());
    // the round as one launch program: sketch, QR of the sketch, posterior
    // test, projection, QR of B^T with R, Jacobi SVD of R^T
    std::vector<MemRange> inputs = {
        {bf.Y, (size_t)ytot * 8 + 16}, {bf.Om, (size_t)otot * 8 + 16},
        {bf.Q, (size_t)qtot * 8 + 16}, {bf.Bt, (size_t)btot * 8 + 16},
        {bf.Q2, (size_t)btot * 8 + 16}, {bf.M, (size_t)mtot * 8 + 16},
        {d_blocks, (size_t)nb * sizeof(TsvdBlock)}};
    for (int i : order) {
      const TsvdBlock &b = blocks[i];
      inputs.push_back({b.A, ((size_t)(b.n - 1) * b.lda + b.m) * sizeof(double)});
    }
    if (carry_cur.p) inputs.push_back({carry_cur.p, carry_cur.bytes});
    const unsigned long long seed = opt.seed;
    const unsigned nwo = (unsigned)wo.size(), nws = (unsigned)ws.size(),
                   nwp = (unsigned)wp.size(), nwp16 = (unsigned)wp16.size();
    const double accept = opt.accept;
    auto build = [&](Program &pg, Arena &ar) -> int {
      std::vector<QrJob> qy, qb;
      std::vector<GemmJob> t1, t2;
      std::vector<ColNorm> cn;
      std::vector<Copy2D> tr;
      for (int i : order) {
        const TsvdBlock &b = blocks[i];
        double *Y = bf.Y + b.yoff;
        QrJob jy{Y, b.m, b.m, b.ell, bf.Q + b.qoff, b.m, nullptr, 0};
        double *R2 = ar.take<double>((size_t)b.ell * b.ell);
        QrJob jb{bf.Bt + b.boff, b.n, b.n, b.ell, bf.Q2 + b.boff, b.n, R2, b.ell};
        const Carry &cy = carry[i];
        if (cy.cols > 0) {
          jy.c_start = cy.cols;
          jy.Qprev = cy.Q;
          jy.ldqp = b.m;
          jb.c_start = cy.cols;
          jb.Qprev = cy.Q2;
          jb.ldqp = b.n;
          jb.Rprev_t = cy.Mt;
          jb.ldrp = cy.cols;
        }
        qy.push_back(jy);
        qb.push_back(jb);
        tr.push_back({R2, bf.M + b.moff, b.ell, b.ell, b.ell, b.ell, 1, 0});
        if (b.ntest > 0) {
          double *Yt = Y + (long long)b.ell * b.m;
          double *T = ar.take<double>((size_t)b.ell * b.ntest);
          t1.push_back({bf.Q + b.qoff, Yt, T, nullptr, b.m, b.m, b.ell, b.ell, b.ntest, b.m,
                        1, 0, 0, 1.0, 0.0});
          t2.push_back({bf.Q + b.qoff, T, Yt, nullptr, b.m, b.ell, b.m, b.m, b.ntest, b.ell,
                        1, 0, 0, -1.0, 1.0});
          double *terr = reinterpret_cast<double *>(reinterpret_cast<char *>(d_blocks + i) +
                                                    offsetof(TsvdBlock, terr));
          cn.push_back({Yt, b.m, b.m, b.ntest, terr, nullptr, 0, 0, 0});
        }
      }
      HM_OK(qr_add(pg, ar, qy));
      for (int rep = 0; rep < 2; rep++) {  // residual of the test columns, twice
        HM_OK(gemm_add(pg, ar, t1, true, false, "test proj", GEMM_BY_SHAPE));
        HM_OK(gemm_add(pg, ar, t2, false, false, "test upd", GEMM_BY_SHAPE));
      }
      HM_OK(colnorm_add(pg, cn));
      pg.add("project", [=](hipStream_t st, const unsigned char *) {
        Timed t3 = timing_begin("tsvd_project", st);
        if (nwp)
          hipLaunchKernelGGL(k_project<SketchTile>, dim3(nwp), dim3(256), 0, st, d_blocks, d_wp,
                             bf);
        if (nwp16)
          hipLaunchKernelGGL(k_project<ProjTile16>, dim3(nwp16), dim3(256), 0, st, d_blocks,
                             d_wp16, bf);
        timing_end(t3, st);
      });
      HM_OK(qr_add(pg, ar, qb));
      HM_OK(copy2d_add(pg, tr));
      return SUCCESS;
    };
    // the sketch needs none of the program's scratch: it is launched first
    // and runs while the host builds the rest of the round's program (all
    // of the round's host-to-device table copies are queued above, so
    // nothing below waits for it before the program's own upload)
    {
      Timed t0 = timing_begin("tsvd_omega", s);
      hipLaunchKernelGGL(k_gen_omega, dim3(nwo), dim3(256), 0, s, d_blocks, d_wo, bf, seed);
      timing_end(t0, s);
      Timed t1 = timing_begin("tsvd_sketch", s);
      hipLaunchKernelGGL(k_sketch, dim3(nws), dim3(256), 0, s, d_blocks, d_ws, bf);
      timing_end(t1, s);
      HM_HIP(hipGetLastError());
    }
    HM_OK(program_exec(build, inputs, s));
    if (Ms && mtot > 0)
      HM_HIP(hipMemcpyAsync(Ms, bf.M, (size_t)mtot * sizeof(double), hipMemcpyDeviceToDevice, s));
    {
      Timed t5 = timing_begin("tsvd_jacobi", s);
      HM_OK(svd_cores(d_blocks, blocks, order, bf, accept, s));
      timing_end(t5, s);
    }
    std::vector<TsvdBlock> hb(nb);
    HM_HIP(hipMemcpyAsync(hb.data(), d_blocks, nb * sizeof(TsvdBlock),
                          hipMemcpyDeviceToHost, s));
    HM_HIP(hipStreamSynchronize(s));
    // one power step on the residual E = (I - Q Q^T) A, only for the blocks
    // the plain test rejected below a saturated sketch (two more passes over
    // their A): W = A^T Z, U = (I - Q Q^T) A W, perr = max_t |U_t| / |W_t|;
    // then their decision is taken again with both estimates
    std::vector<int> power;
    for (int i : order)
      if (!hb[i].status && blocks[i].ntest > 0 && hb[i].k < blocks[i].ell) power.push_back(i);
    if (!power.empty()) {
      std::vector<MemRange> pin = {{bf.Y, (size_t)ytot * 8 + 16},
                                   {bf.Q, (size_t)qtot * 8 + 16},
                                   {d_blocks, (size_t)nb * sizeof(TsvdBlock)}};
      for (int i : power) {
        const TsvdBlock &b = blocks[i];
        pin.push_back({b.A, ((size_t)(b.n - 1) * b.lda + b.m) * sizeof(double)});
      }
      HM_OK(program_exec(
          [&](Program &pg, Arena &ar) -> int {
            std::vector<GemmJob> p1, p2, p3, p4;
            std::vector<ColNorm> pn;
            for (int i : power) {
              const TsvdBlock &b = blocks[i];
              double *Yt = bf.Y + b.yoff + (long long)b.ell * b.m;
              double *Wt = ar.take<double>((size_t)b.n * b.ntest);
              double *Ut = ar.take<double>((size_t)b.m * b.ntest);
              double *T2 = ar.take<double>((size_t)b.ell * b.ntest);
              p1.push_back({b.A, Yt, Wt, nullptr, b.lda, b.m, b.n, b.n, b.ntest, b.m, 1, 0, 0,
                            1.0, 0.0});
              p2.push_back({b.A, Wt, Ut, nullptr, b.lda, b.n, b.m, b.m, b.ntest, b.n, 1, 0, 0,
                            1.0, 0.0});
              p3.push_back({bf.Q + b.qoff, Ut, T2, nullptr, b.m, b.m, b.ell, b.ell, b.ntest,
                            b.m, 1, 0, 0, 1.0, 0.0});
              p4.push_back({bf.Q + b.qoff, T2, Ut, nullptr, b.m, b.ell, b.m, b.m, b.ntest,
                            b.ell, 1, 0, 0, -1.0, 1.0});
              double *perr = reinterpret_cast<double *>(
                  reinterpret_cast<char *>(d_blocks + i) + offsetof(TsvdBlock, perr));
              pn.push_back({Ut, b.m, b.m, b.ntest, perr, Wt, b.n, b.n, 0});
            }
            HM_OK(gemm_add(pg, ar, p1, true, false, "power At", GEMM_BY_SHAPE));
            HM_OK(gemm_add(pg, ar, p2, false, false, "power A", GEMM_BY_SHAPE));
            for (int rep = 0; rep < 2; rep++) {
              HM_OK(gemm_add(pg, ar, p3, true, false, "power proj", GEMM_BY_SHAPE));
              HM_OK(gemm_add(pg, ar, p4, false, false, "power upd", GEMM_BY_SHAPE));
            }
            return colnorm_add(pg, pn);
          },
          pin, s));
      int *d_power = nullptr;
      HM_OK(to_device(&d_power, power, s));
      mem.keep(d_power);
      hipLaunchKernelGGL(k_jacobi, dim3((unsigned)power.size()), dim3(256), 0, s, d_blocks,
                         d_power, bf, accept, 1);
      HM_LAUNCH_CHECK();
      HM_HIP(hipMemcpyAsync(hb.data(), d_blocks, nb * sizeof(TsvdBlock),
                            hipMemcpyDeviceToHost, s));
      HM_HIP(hipStreamSynchronize(s));
    }
    const bool first_round = (st.rounds == 1);
    std::vector<double> sig16;
    if (first_round && (opt.widen || opt.first_sigma)) {
      std::vector<double> hs(stot);
      HM_HIP(hipMemcpyAsync(hs.data(), bf.Sig, stot * sizeof(double), hipMemcpyDeviceToHost, s));
      HM_HIP(hipStreamSynchronize(s));
      sig16.assign((size_t)nb * FIRST_SIGMA, 0.0);
      for (int i : order) {
        const TsvdBlock &b = blocks[i];
        for (int j = 0; j < std::min(b.ell, FIRST_SIGMA); j++)
          sig16[(size_t)i * FIRST_SIGMA + j] = hs[b.soff + j];
      }
      if (opt.first_sigma) *opt.first_sigma = sig16;
    }
    std::vector<int> done, retry;
    for (int i : order) {
      TsvdBlock &b = blocks[i];
      b.k = hb[i].k;
      b.need = hb[i].need;
      b.sigma0 = hb[i].sigma0;
      b.err = hb[i].err;
      b.perr = hb[i].perr;
      b.status = hb[i].status;
      if (b.status || b.ell >= cap) {
        if (!b.status) st.capped++;
        if (b.ntest == 0) st.exact_blocks++;
        done.push_back(i);
      } else {
        retry.push_back(i);
      }
    }
    st.retried += (int)retry.size();
    if (st.rounds == 2 && st.widened > 0) st.widened_failed = (int)retry.size();
    if (!done.empty()) {
      long long ut = 0, vt = 0;
      for (int i : done) {
        ut += (long long)blocks[i].m * blocks[i].k;
        vt += (long long)blocks[i].n * blocks[i].k;
      }
      RoundOut ro;
      HM_OK(dalloc(&ro.u, ut)); round_mem.keep(ro.u);
      HM_OK(dalloc(&ro.v, vt)); round_mem.keep(ro.v);
      round_pools.push_back(ro);
      long long uo = 0, vo = 0;
      for (int i : done) {
        blocks[i].uout = ro.u + uo;
        blocks[i].vout = ro.v + vo;
        uo += (long long)blocks[i].m * blocks[i].k;
        vo += (long long)blocks[i].n * blocks[i].k;
      }
      HM_HIP(hipMemcpyAsync(d_blocks, blocks.data(), nb * sizeof(TsvdBlock),
                            hipMemcpyHostToDevice, s));
      {
        std::vector<GemmJob> gu;
        for (int i : done) {
          const TsvdBlock &b = blocks[i];
          gu.push_back({bf.Q + b.qoff, bf.G + b.moff, b.uout, nullptr, b.m, b.ell, b.m, b.m, b.k,
                        b.ell, 1, 0, 0, 1.0, 0.0});
          gu.push_back({bf.Q2 + b.boff, bf.Z + b.moff, b.vout, nullptr, b.n, b.ell, b.n, b.n,
                        b.k, b.ell, 1, 0, 0, 1.0, 0.0});
        }
        std::vector<MemRange> fin = {{bf.Q, (size_t)qtot * 8 + 16},
                                     {bf.Q2, (size_t)btot * 8 + 16},
                                     {bf.G, (size_t)mtot * 8 + 16},
                                     {bf.Z, (size_t)mtot * 8 + 16},
                                     {ro.u, (size_t)ut * 8 + 16},
                                     {ro.v, (size_t)vt * 8 + 16}};
        HM_OK(program_exec(
            [&](Program &pg, Arena &ar) {
              pg.add("form", [](hipStream_t st, const unsigned char *) {});
              return gemm_add(pg, ar, gu, false, false, "tsvd form", GEMM_BY_SHAPE);
            },
            fin, s));
      }
      if (sigma) {
        std::vector<double> hs(stot);
        HM_HIP(hipMemcpyAsync(hs.data(), bf.Sig, stot * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        HM_HIP(hipStreamSynchronize(s));
        for (int i : done) {
          const TsvdBlock &b = blocks[i];
          for (int j = 0; j < std::min(b.ell, sigma_cap); j++)
            (*sigma)[(size_t)i * sigma_cap + j] = hs[b.soff + j];
        }
      }
      HM_HIP(hipStreamSynchronize(s));
    }
    std::vector<int> prev_ell;
    for (int i : retry) prev_ell.push_back(blocks[i].ell);
    if (first_round && opt.widen && !retry.empty()) {
      HM_OK(opt.widen(blocks, retry, sig16));
      st.widened = (int)retry.size();
    } else {
      for (int i : retry) blocks[i].ell = 2 * blocks[i].ell;
    }
    for (size_t r = 0; r < retry.size(); r++) {
      TsvdBlock &b = blocks[retry[r]];
      const int mn = std::min(b.m, b.n);
      // never narrower than the sketch that just failed
      b.ell = std::min(std::min(std::max(b.ell, prev_ell[r] + 1), mn), cap);
      b.ntest = (b.ell >= mn) ? 0 : std::min(opt.n_test, b.m - b.ell);
    }
    // carry the done columns of the continued blocks out of the round
    // buffer (blocks[] already holds the next widths; the offsets are still
    // this round's)
    {
      std::vector<Carry> next(nb);
      size_t ctot = 0;
      std::vector<int> cont;
      for (size_t r = 0; r < retry.size(); r++) {
        const int i = retry[r], w = prev_ell[r];
        const TsvdBlock &b = blocks[i];
        if (!Ms || w < 64 || w % 64 != 0 || b.ell <= w || w >= std::min(b.m, b.n)) continue;
        cont.push_back((int)r);
        ctot += ((size_t)b.m + b.n + w) * w;
      }
      double *cbuf = nullptr;
      if (ctot > 0) {
        HM_OK(dalloc(&cbuf, ctot));
        size_t off = 0;
        for (int r : cont) {
          const int i = retry[r], w = prev_ell[r];
          const TsvdBlock &b = blocks[i];
          Carry c;
          c.cols = w;
          double *q = cbuf + off, *q2 = q + (size_t)b.m * w, *mt = q2 + (size_t)b.n * w;
          off += ((size_t)b.m + b.n + w) * w;
          hipError_t e = hipMemcpyAsync(q, bf.Q + b.qoff, (size_t)b.m * w * sizeof(double),
                                        hipMemcpyDeviceToDevice, s);
          if (e == hipSuccess)
            e = hipMemcpyAsync(q2, bf.Q2 + b.boff, (size_t)b.n * w * sizeof(double),
                               hipMemcpyDeviceToDevice, s);
          if (e == hipSuccess)
            e = hipMemcpyAsync(mt, Ms + b.moff, (size_t)w * w * sizeof(double),
                               hipMemcpyDeviceToDevice, s);
          if (e != hipSuccess) {
            dfree(cbuf);
            set_error("tsvd: carrying a round: %s", hipGetErrorString(e));
            return ALLOCATION_FAILURE;
          }
          c.Q = q;
          c.Q2 = q2;
          c.Mt = mt;
          next[i] = c;
        }
      }
      HM_HIP(hipStreamSynchronize(s));
      dfree(carry_cur.p);  // the round that used it has finished
      carry_cur.p = cbuf;
      carry_cur.bytes = ctot * sizeof(double);
      carry.swap(next);
      st.continued += (int)cont.size();
    }
    active = retry;
  }
  // gather into two packed pools
  long long ut = 0, vt = 0;
  for (int i = 0; i < nb; i++) {
    ut += (long long)blocks[i].m * blocks[i].k;
    vt += (long long)blocks[i].n * blocks[i].k;
  }
  double *up = nullptr, *vp = nullptr;
  HM_OK(dalloc(&up, ut));
  int rc = dalloc(&vp, vt);
  if (rc != SUCCESS) {
    dfree(up);
    return rc;
  }
  HM_HIP(hipMemsetAsync(up, 0, (ut + 1) * sizeof(double), s));
  HM_HIP(hipMemsetAsync(vp, 0, (vt + 1) * sizeof(double), s));
  std::vector<CopyItem> items;
  long long uo = 0, vo = 0;
  for (int i = 0; i < nb; i++) {
    TsvdBlock &b = blocks[i];
    if (b.uout) {
      items.push_back({b.uout, up + uo, (long long)b.m * b.k});
      items.push_back({b.vout, vp + vo, (long long)b.n * b.k});
    } else if (b.n > 0 && b.k > 0) {
      // empty block: V = e_0
      const double one = 1.0;
      HM_HIP(hipMemcpyAsync(vp + vo, &one, sizeof(double), hipMemcpyHostToDevice, s));
      HM_HIP(hipStreamSynchronize(s));
    }
    b.uout = up + uo;
    b.vout = vp + vo;
    uo += (long long)b.m * b.k;
    vo += (long long)b.n * b.k;
  }
  if (!items.empty()) {
    CopyItem *d_items = nullptr;
    rc = to_device(&d_items, items, s);
    if (rc == SUCCESS) {
      hipLaunchKernelGGL(k_copy_list, dim3((unsigned)items.size(), 8), dim3(256),
                         0, s, d_items);
      (void)hipGetLastError();
      (void)hipStreamSynchronize(s);
    }
    dfree(d_items);
  }
  HM_HIP(hipStreamSynchronize(s));
  *u_pool = up;
  *v_pool = vp;
  if (stats) *stats = st;
  return SUCCESS;
}

// QR-method recompression of factored nodes (s < min(m, n)).
static int recompress_qr(int nn, const int *m, const int *n, const int *sr,
                         double *const *d_u, double *const *d_v, double thr,
                         int *k_out, hipStream_t s) {
  std::vector<TsvdBlock> blocks(nn);
  long long ytot = 0, btot = 0, mtot = 0, stot = 0;
  for (int i = 0; i < nn; i++) {
    TsvdBlock &b = blocks[i];
    memset(&b, 0, sizeof(b));
    b.m = m[i];
    b.n = n[i];
    b.ell = sr[i];
    b.ntest = 0;
    b.rule = RULE_ABSOLUTE;
    b.thr = thr;
    b.yoff = b.qoff = ytot;
    ytot += (long long)b.m * b.ell;
    b.boff = btot;
    btot += (long long)b.n * b.ell;
    b.moff = mtot;
    mtot += (long long)b.ell * b.ell;
    b.soff = stot;
    stot += b.ell;
  }
  Scoped mem;
  TsvdBufs bf{};
  HM_OK(dalloc(&bf.Y, ytot)); mem.keep(bf.Y);
  HM_OK(dalloc(&bf.Q, ytot)); mem.keep(bf.Q);
  HM_OK(dalloc(&bf.Bt, btot)); mem.keep(bf.Bt);
  HM_OK(dalloc(&bf.Q2, btot)); mem.keep(bf.Q2);
  HM_OK(dalloc(&bf.M, mtot)); mem.keep(bf.M);
  HM_OK(dalloc(&bf.Vj, mtot)); mem.keep(bf.Vj);
  HM_OK(dalloc(&bf.G, mtot)); mem.keep(bf.G);
  HM_OK(dalloc(&bf.Z, mtot)); mem.keep(bf.Z);
  HM_OK(dalloc(&bf.Sig, stot)); mem.keep(bf.Sig);
  double *RU, *RV;
  HM_OK(dalloc(&RU, mtot)); mem.keep(RU);
  HM_OK(dalloc(&RV, mtot)); mem.keep(RV);
  bf.Om = nullptr;
  std::vector<CopyItem> in;
  for (int i = 0; i < nn; i++) {
    in.push_back({d_u[i], bf.Y + blocks[i].yoff, (long long)m[i] * sr[i]});
    in.push_back({d_v[i], bf.Bt + blocks[i].boff, (long long)n[i] * sr[i]});
  }
  CopyItem *d_in;
  HM_OK(to_device(&d_in, in, s)); mem.keep(d_in);
  hipLaunchKernelGGL(k_copy_list, dim3((unsigned)in.size(), 8), dim3(256), 0, s, d_in);
  HM_LAUNCH_CHECK();
  std::vector<int> list(nn);
  for (int i = 0; i < nn; i++) list[i] = i;
  TsvdBlock *d_blocks;
  HM_OK(to_device(&d_blocks, blocks, s)); mem.keep(d_blocks);
  Timed t0 = timing_begin("recompress_core", s);
  {
    std::vector<QrJob> qj;
    for (int i = 0; i < nn; i++) {
      const TsvdBlock &b = blocks[i];
      qj.push_back({bf.Y + b.yoff, b.m, b.m, b.ell, bf.Q + b.qoff, b.m, RU + b.moff, b.ell});
      qj.push_back({bf.Bt + b.boff, b.n, b.n, b.ell, bf.Q2 + b.boff, b.n, RV + b.moff, b.ell});
    }
    // M = R_U R_V^T (both upper triangular; qr_add zeroes R first)
    std::vector<GemmJob> mj;
    for (int i = 0; i < nn; i++) {
      const TsvdBlock &b = blocks[i];
      mj.push_back({RU + b.moff, RV + b.moff, bf.M + b.moff, nullptr, b.ell, b.ell, b.ell, b.ell,
                    b.ell, b.ell, 1, 0, 0, 1.0, 0.0});
    }
    std::vector<MemRange> inputs = {{bf.Y, (size_t)ytot * 8 + 16}, {bf.Q, (size_t)ytot * 8 + 16},
                                    {bf.Bt, (size_t)btot * 8 + 16}, {bf.Q2, (size_t)btot * 8 + 16},
                                    {RU, (size_t)mtot * 8 + 16}, {RV, (size_t)mtot * 8 + 16},
                                    {bf.M, (size_t)mtot * 8 + 16}};
    HM_OK(program_exec([&](Program &pg, Arena &ar) {
      HM_OK(qr_add(pg, ar, qj));
      return gemm_add(pg, ar, mj, false, true, "recompress core");
    }, inputs, s));
  }
  HM_OK(svd_cores(d_blocks, blocks, list, bf, 0.0, s));
  timing_end(t0, s);
  HM_LAUNCH_CHECK();
  HM_HIP(hipMemcpyAsync(blocks.data(), d_blocks, nn * sizeof(TsvdBlock),
                        hipMemcpyDeviceToHost, s));
  HM_HIP(hipStreamSynchronize(s));
  long long ut = 0, vt = 0;
  for (auto &b : blocks) {
    ut += (long long)b.m * b.k;
    vt += (long long)b.n * b.k;
  }
  double *uo, *vo;
  HM_OK(dalloc(&uo, ut)); mem.keep(uo);
  HM_OK(dalloc(&vo, vt)); mem.keep(vo);
  ut = vt = 0;
  std::vector<CopyItem> outc;
  for (int i = 0; i < nn; i++) {
    TsvdBlock &b = blocks[i];
    b.uout = uo + ut;
    b.vout = vo + vt;
    outc.push_back({b.uout, d_u[i], (long long)b.m * b.k});
    outc.push_back({b.vout, d_v[i], (long long)b.n * b.k});
    ut += (long long)b.m * b.k;
    vt += (long long)b.n * b.k;
    k_out[i] = b.k;
  }
  HM_HIP(hipMemcpyAsync(d_blocks, blocks.data(), nn * sizeof(TsvdBlock),
                        hipMemcpyHostToDevice, s));
  CopyItem *d_out;
  HM_OK(to_device(&d_out, outc, s)); mem.keep(d_out);
  {
    std::vector<GemmJob> gu;
    for (int i = 0; i < nn; i++) {
      const TsvdBlock &b = blocks[i];
      gu.push_back({bf.Q + b.qoff, bf.G + b.moff, b.uout, nullptr, b.m, b.ell, b.m, b.m, b.k,
                    b.ell, 1, 0, 0, 1.0, 0.0});
      gu.push_back({bf.Q2 + b.boff, bf.Z + b.moff, b.vout, nullptr, b.n, b.ell, b.n, b.n, b.k,
                    b.ell, 1, 0, 0, 1.0, 0.0});
    }
    std::vector<MemRange> fin = {{bf.Q, (size_t)ytot * 8 + 16}, {bf.Q2, (size_t)btot * 8 + 16},
                                 {bf.G, (size_t)mtot * 8 + 16}, {bf.Z, (size_t)mtot * 8 + 16},
                                 {uo, (size_t)ut * 8 + 16}, {vo, (size_t)vt * 8 + 16}};
    HM_OK(program_exec([&](Program &pg, Arena &ar) {
      return gemm_add(pg, ar, gu, false, false, "recompress form", GEMM_BY_SHAPE);
    }, fin, s));
  }
  hipLaunchKernelGGL(k_copy_list, dim3((unsigned)outc.size(), 8), dim3(256), 0, s, d_out);
  HM_LAUNCH_CHECK();
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

// Dense-method recompression: D = U V^T, exact SVD, absolute rule.
static int recompress_dense(int nn, const int *m, const int *n, const int *sr,
                            double *const *d_u, double *const *d_v, double thr,
                            int *k_out, hipStream_t s) {
  Scoped mem;
  std::vector<OuterJob> jobs(nn);
  long long dtot = 0;
  std::vector<long long> doff(nn);
  for (int i = 0; i < nn; i++) {
    doff[i] = dtot;
    dtot += (long long)m[i] * n[i];
  }
  double *D;
  HM_OK(dalloc(&D, dtot)); mem.keep(D);
  std::vector<int4> work;
  for (int i = 0; i < nn; i++) {
    jobs[i] = {d_u[i], d_v[i], D + doff[i], m[i], n[i], sr[i], 0};
    for (int a = 0; a < (m[i] + DenseTile::TM - 1) / DenseTile::TM; a++)
      for (int c = 0; c < (n[i] + DenseTile::TN - 1) / DenseTile::TN; c++)
        work.push_back(make_int4(i, a, c, 0));
  }
  OuterJob *d_jobs;
  int4 *d_work;
  HM_OK(to_device(&d_jobs, jobs, s)); mem.keep(d_jobs);
  HM_OK(to_device(&d_work, work, s)); mem.keep(d_work);
  if (!work.empty()) {
    hipLaunchKernelGGL(k_outer, dim3((unsigned)work.size()), dim3(256), 0, s, d_jobs, d_work);
    HM_LAUNCH_CHECK();
  }
  std::vector<TsvdBlock> blocks(nn);
  for (int i = 0; i < nn; i++) {
    TsvdBlock &b = blocks[i];
    memset(&b, 0, sizeof(b));
    b.A = D + doff[i];
    b.lda = m[i];
    b.m = m[i];
    b.n = n[i];
    b.ell = std::min(m[i], n[i]);
    b.rule = RULE_ABSOLUTE;
    b.thr = thr;
    b.uid = 0x5eed0000ull + (unsigned long long)i;
  }
  TsvdOptions opt;
  double *up = nullptr, *vp = nullptr;
  HM_OK(tsvd_run(blocks, opt, &up, &vp, nullptr, nullptr, 0, s));
  mem.keep(up);
  mem.keep(vp);
  std::vector<CopyItem> outc;
  for (int i = 0; i < nn; i++) {
    outc.push_back({blocks[i].uout, d_u[i], (long long)m[i] * blocks[i].k});
    outc.push_back({blocks[i].vout, d_v[i], (long long)n[i] * blocks[i].k});
    k_out[i] = blocks[i].k;
  }
  CopyItem *d_out;
  HM_OK(to_device(&d_out, outc, s)); mem.keep(d_out);
  hipLaunchKernelGGL(k_copy_list, dim3((unsigned)outc.size(), 8), dim3(256), 0, s, d_out);
  HM_LAUNCH_CHECK();
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

}  // namespace hm

using namespace hm;

extern "C" int hm_hip_recompress_batch(int nnodes, const int *m, const int *n,
                                       const int *s, double *const *d_u,
                                       double *const *d_v, double thr,
                                       int method, int *k_out, void *stream) {
  HM_OK(ensure_init());
  if (nnodes <= 0) return SUCCESS;
  if (!m || !n || !s || !d_u || !d_v || !k_out) {
    set_error("hm_hip_recompress_batch: NULL argument");
    return INPUT_ERROR;
  }
  std::vector<int> qr_idx, dn_idx;
  for (int i = 0; i < nnodes; i++) {
    if (s[i] <= 0 || m[i] <= 0 || n[i] <= 0 || s[i] > MAX_ELL) {
      set_error("hm_hip_recompress_batch: node %d has m=%d n=%d s=%d", i, m[i], n[i], s[i]);
      return INPUT_ERROR;
    }
    const int mn = std::min(m[i], n[i]);
    int meth = method;
    if (meth < 0) meth = (s[i] < mn) ? 0 : 1;
    if (meth == 0 && s[i] >= mn) {
      set_error("hm_hip_recompress_batch: QR method needs s < min(m, n) (node %d)", i);
      return INPUT_ERROR;
    }
    (meth == 0 ? qr_idx : dn_idx).push_back(i);
  }
  hipStream_t st = stream_of(stream);
  for (int pass = 0; pass < 2; pass++) {
    const std::vector<int> &idx = pass ? dn_idx : qr_idx;
    if (idx.empty()) continue;
    std::vector<int> mm, nn_, ss, kk(idx.size());
    std::vector<double *> uu, vv;
    for (int i : idx) {
      mm.push_back(m[i]);
      nn_.push_back(n[i]);
      ss.push_back(s[i]);
      uu.push_back(d_u[i]);
      vv.push_back(d_v[i]);
    }
    int rc = pass ? recompress_dense((int)idx.size(), mm.data(), nn_.data(), ss.data(),
                                     uu.data(), vv.data(), thr, kk.data(), st)
                  : recompress_qr((int)idx.size(), mm.data(), nn_.data(), ss.data(),
                                  uu.data(), vv.data(), thr, kk.data(), st);
    if (rc != SUCCESS) return rc;
    for (size_t t = 0; t < idx.size(); t++) k_out[idx[t]] = kk[t];
  }
  return SUCCESS;
}
