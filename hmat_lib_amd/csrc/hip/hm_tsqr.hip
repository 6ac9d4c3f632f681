// Batched tall-skinny Householder QR (TSQR) with explicit Q, FP64.
//
// A panel (m x ncol, ncol <= 64, column-major) is cut into chunks of at most
// 128 rows (rows spread evenly, so every chunk of a multi-chunk level has
// >= ncol rows); each chunk is factored by one workgroup (Householder,
// LAPACK dlarfg sign convention: beta = -sign(alpha) ||x||, tau = 0 when the
// tail is zero), its R is stacked into the next level's panel, and so on
// until one chunk remains.  The explicit Q (first nq columns) is then built
// top-down: each chunk applies its reflectors to [C; 0] where C is its
// ncol x nq slice of the parent level's Q.  All chunks of all panels of one
// level run in one launch.  Extras of the top chunk: R (ncol x ncol), R^T,
// and the norms of the trailing columns of R past column `ell` (for the
// randomized range finder's posterior test: they equal
// |(I - Q_1 Q_1^T) y_t|).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include <mutex>
#include <thread>

#include "hm_tsqr.h"

namespace hm {

static constexpr int TQ_NT = 256;

// One reduction level of one panel.  Its chunks are work items of 8 bytes
// (level index, chunk index); chunk i covers rows [i rows / nch,
// (i + 1) rows / nch) of the level.
struct QrLevel {
  double *panel;        // the level's rows (the panel, or the stacked R's)
  long long ld;
  long long rows;
  int nch, ncol, nq, is_top;
  double *tau;          // nch x ncol
  double *rout;         // up: R_i to rows i ncol.. of the next level (top: R)
  long long ldr;
  const double *cin;    // down: the next level's Q (NULL at the top)
  long long ldc;
  double *qout;         // down: this level's Q (ld ldq)
  long long ldq;
  double *est;          // top: trailing-column norms of R past ell
  double *mt;           // top: R^T (ncol x ncol, ld ncol)
  int ell, pad;
};

// the chunk's share of its level
struct QrChunk {
  const QrLevel *L;
  int i, r0, rows;
};
__device__ __forceinline__ QrChunk qr_chunk(const QrLevel *levels, const int2 *chunks,
                                            int idx) {
  const int2 c = chunks[idx];
  QrChunk q;
  q.L = levels + c.x;
  q.i = c.y;
  const long long rows = q.L->rows, nch = q.L->nch;
  q.r0 = (int)((long long)c.y * rows / nch);
  q.rows = (int)(((long long)c.y + 1) * rows / nch) - q.r0;
  return q;
}

// ------------------------------------------------ LDS-resident chunks
// Panels of at most TQ_FAST_COLS columns use chunks of at most TQ_CR rows
// (rows spread evenly, so every chunk of a multi-chunk level has >= ncol
// rows).  Column per lane: lane j of wave w keeps rows 32w .. 32w + 31 of
// column j in registers.  Applying a reflector to every column is then a
// per-lane dot product over the wave's 32 rows (the reflector read from LDS
// as a broadcast) and a sum of the four waves' partials through LDS: no
// cross-lane reductions per column, one or two barriers per reflector.
// LDS stages the panel so that global reads and writes stay coalesced.
static constexpr int TQ_CR = 128;
static constexpr int TQ_FAST_COLS = 64;
static constexpr int TQ_NW = TQ_NT / 64;     // waves per chunk
static constexpr int TQ_RW = TQ_CR / TQ_NW;  // rows per wave
static constexpr int TQ_LS = TQ_CR + 2;      // LDS column stride (16-byte aligned columns)
static_assert(TQ_NT == 256 && TQ_RW == 32, "chunk layout");

// LDS panel (column stride TQ_LS) <- rows x ncol of a global panel: thread
// t holds row t % 128 of columns t / 128, + 2, ...; eight loads in flight.
__device__ __forceinline__ void lds_load_panel(double *S, const double *A, long long ld, int rows,
                                               int ncol) {
  const int i = threadIdx.x & 127, j0 = threadIdx.x >> 7;
  if (i >= rows) return;
  const gdbl *g = gmem(A) + i;
  for (int j = j0; j < ncol; j += 16) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = g[(long long)min(j + 2 * u, ncol - 1) * ld];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j + 2 * u < ncol) S[i + (j + 2 * u) * TQ_LS] = v[u];
  }
}

// global panel <- LDS panel (the inverse of lds_load_panel)
__device__ __forceinline__ void lds_store_panel(const double *S, double *A, long long ld, int rows,
                                                int ncol) {
  const int i = threadIdx.x & 127, j0 = threadIdx.x >> 7;
  if (i >= rows) return;
  gdbl *g = gmem(A) + i;
  for (int j = j0; j < ncol; j += 2) g[(long long)j * ld] = S[i + j * TQ_LS];
}

__global__ __launch_bounds__(TQ_NT) void k_tsqr_up_lds(const QrLevel *levels, const int2 *chunks) {
  __shared__ __attribute__((aligned(16))) double S[TQ_LS * TQ_FAST_COLS];
  __shared__ __attribute__((aligned(16))) double vb[TQ_CR];  // column k below the diagonal
  __shared__ double part[TQ_NW][64];
  __shared__ double tauS[TQ_FAST_COLS], scalS[TQ_FAST_COLS], betaS[TQ_FAST_COLS];
  __shared__ double alphaS;
  const QrChunk ck = qr_chunk(levels, chunks, blockIdx.x);
  const QrLevel &it = *ck.L;
  const int rows = ck.rows, nc = it.ncol, tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, rw0 = w * TQ_RW;
  double *G = it.panel + ck.r0;
  const long long ld = it.ld;
  lds_load_panel(S, G, ld, rows, nc);
  __syncthreads();
  double y[TQ_RW];
  {
    const double *sj = S + min(lane, nc - 1) * TQ_LS + rw0;
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      const double x = sj[r];
      y[r] = (rw0 + r < rows && lane < nc) ? x : 0.0;
    }
  }
  // Reflector k (LAPACK dlarfg convention): column k's rows below the
  // diagonal x, alpha on it; beta = -sign(alpha) |(alpha, x)|, tau =
  // (beta - alpha) / beta, v = [1; x / (alpha - beta)]; tau = 0 when x = 0.
  // Column k itself keeps x and alpha until the end (its scale and beta are
  // applied when it is written back).
  const int nref = min(rows, nc);
  for (int k = 0; k < nref; k++) {
    const int kw = k / TQ_RW, kr = k % TQ_RW;
    if (lane == k && w >= kw) {
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) vb[rw0 + r] = y[r];
      if (w == kw) {
        alphaS = vb[k];
        for (int r = 0; r <= kr; r++) vb[rw0 + r] = 0.0;
      }
    }
    __syncthreads();
    const double a0 = vb[lane], a1 = vb[lane + 64];
    const double x0 = lane > k ? a0 : 0.0, x1 = lane + 64 > k ? a1 : 0.0;
    const double x2 = wave_sum(x0 * x0 + x1 * x1);
    const double alpha = alphaS;
    double tau = 0.0, beta = alpha, scal = 1.0;
    if (x2 > 0.0) {
      const double hh = hypot(alpha, sqrt(x2));
      beta = (alpha >= 0.0) ? -hh : hh;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) {
      tauS[k] = tau;
      scalS[k] = scal;
      betaS[k] = beta;
    }
    // v = scal * vb with the unit entry: vb[k] = alpha - beta = 1 / scal,
    // written by the wave that reads it (its LDS accesses stay in order;
    // the other waves' norms never look at row k)
    if (w == kw && lane == 0) vb[k] = alpha - beta;
    double v[TQ_RW];
    double pd = 0.0;
    if (w >= kw) {
      double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) {
        v[r] = vb[rw0 + r];
        p4[r & 3] += v[r] * y[r];
      }
      pd = (p4[0] + p4[1]) + (p4[2] + p4[3]);
    }
    part[w][lane] = pd;
    __syncthreads();
    if (tau != 0.0 && lane > k && w >= kw) {
      double dot = part[0][lane];
#pragma unroll
      for (int ww = 1; ww < TQ_NW; ww++) dot += part[ww][lane];
      const double fs = tau * scal * scal * dot;
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) y[r] -= fs * v[r];
    }
  }
  // back to LDS in factored form: R on and above the diagonal, the scaled
  // reflectors below it
#pragma unroll
  for (int r = 0; r < TQ_RW; r++) {
    const int i = rw0 + r;
    double x = y[r];
    if (lane < nref) {
      if (i > lane) x *= scalS[lane];
      else if (i == lane) x = betaS[lane];
    }
    S[i + lane * TQ_LS] = x;
  }
  __syncthreads();
  double *tau_out = it.tau + (long long)ck.i * nc;
  for (int k = tid; k < nc; k += TQ_NT) gmem(tau_out)[k] = k < nref ? tauS[k] : 0.0;
  lds_store_panel(S, G, ld, rows, nc);
  {
    // R (upper nc x nc, zero below the diagonal and past `rows`)
    const int i = tid & 63;
    gdbl *ro = gmem(it.rout) + (it.is_top ? 0 : ck.i * nc) + i;
    gdbl *mt = (it.is_top && it.mt) ? gmem(it.mt) + (long long)i * nc : nullptr;
    if (i < nc)
      for (int j = tid >> 6; j < nc; j += 4) {
        const double r = (i <= j && i < rows) ? S[i + j * TQ_LS] : 0.0;
        ro[(long long)j * it.ldr] = r;
        if (mt) mt[j] = r;
      }
  }
  if (it.is_top && it.est && tid == 0) {
    double mx = 0.0;
    for (int t = it.ell; t < nc; t++) {
      double ss = 0.0;
      for (int i = it.ell; i <= t && i < rows; i++) ss += S[i + t * TQ_LS] * S[i + t * TQ_LS];
      mx = fmax(mx, sqrt(ss));
    }
    *gmem(it.est) = mx;
  }
}

// Panels of at most 16 columns: one wave per chunk, four chunks per
// workgroup and no workgroup barrier.  Lane 16 q + j keeps rows
// 32 q .. 32 q + 31 of column j in registers; the four row quarters of a
// column are summed with lane shuffles.  Same reflectors, same output
// layout as k_tsqr_up_lds (whose comments give the conventions).
static constexpr int TQ_NARROW = 16;
static constexpr int TQ_NLS = TQ_CR + 1;  // LDS column stride of the narrow staging
__global__ __launch_bounds__(TQ_NT) void k_tsqr_up_narrow(const QrLevel *levels,
                                                          const int2 *chunks, int nchunks) {
  __shared__ double Sw[TQ_NW][TQ_NLS * TQ_NARROW];
  __shared__ __attribute__((aligned(16))) double vbw[TQ_NW][TQ_CR];
  __shared__ double tauW[TQ_NW][TQ_NARROW], scalW[TQ_NW][TQ_NARROW], betaW[TQ_NW][TQ_NARROW];
  __shared__ double alphaW[TQ_NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int idx = blockIdx.x * TQ_NW + w;
  if (idx >= nchunks) return;  // no workgroup barriers below
  const QrChunk ck = qr_chunk(levels, chunks, idx);
  const QrLevel &it = *ck.L;
  const int rows = ck.rows, nc = it.ncol;
  const int q = lane >> 4, j = lane & 15, rq0 = q * TQ_RW;
  double *S = Sw[w], *vb = vbw[w];
  double *G = it.panel + ck.r0;
  const long long ld = it.ld;
  {
    const gdbl *g = gmem(G);
    for (int jj = 0; jj < nc; jj++) {
      const double a = g[min(lane, rows - 1) + jj * ld], b = g[min(lane + 64, rows - 1) + jj * ld];
      if (lane < rows) S[lane + jj * TQ_NLS] = a;
      if (lane + 64 < rows) S[lane + 64 + jj * TQ_NLS] = b;
    }
  }
  wave_lds_sync();
  double y[TQ_RW];
  {
    const double *sj = S + min(j, nc - 1) * TQ_NLS + rq0;
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      const double x = sj[r];
      y[r] = (rq0 + r < rows && j < nc) ? x : 0.0;
    }
  }
  const int nref = min(rows, nc);
  for (int k = 0; k < nref; k++) {
    const int kq = k / TQ_RW, kr = k % TQ_RW;
    wave_lds_sync();
    if (j == k && q >= kq) {
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) vb[rq0 + r] = y[r];
    }
    wave_lds_sync();
    if (j == k && q == kq) {
      alphaW[w] = vb[k];
      for (int r = 0; r <= kr; r++) vb[rq0 + r] = 0.0;
    }
    wave_lds_sync();
    const double a0 = vb[lane], a1 = vb[lane + 64];
    const double x0 = lane > k ? a0 : 0.0, x1 = lane + 64 > k ? a1 : 0.0;
    const double x2 = wave_sum(x0 * x0 + x1 * x1);
    const double alpha = alphaW[w];
    double tau = 0.0, beta = alpha, scal = 1.0;
    if (x2 > 0.0) {
      const double hh = hypot(alpha, sqrt(x2));
      beta = (alpha >= 0.0) ? -hh : hh;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    wave_lds_sync();
    if (lane == 0) {
      tauW[w][k] = tau;
      scalW[w][k] = scal;
      betaW[w][k] = beta;
      vb[k] = alpha - beta;  // the unit entry of v = scal * vb
    }
    wave_lds_sync();
    double v[TQ_RW];
    double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      v[r] = vb[rq0 + r];
      p4[r & 3] += v[r] * y[r];
    }
    double pd = q >= kq ? (p4[0] + p4[1]) + (p4[2] + p4[3]) : 0.0;
    pd += __shfl_xor(pd, 16, 64);
    pd += __shfl_xor(pd, 32, 64);
    if (tau != 0.0 && j > k && q >= kq) {
      const double fs = tau * scal * scal * pd;
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) y[r] -= fs * v[r];
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < TQ_RW; r++) {
    const int i = rq0 + r;
    double x = y[r];
    if (j < nref) {
      if (i > j) x *= scalW[w][j];
      else if (i == j) x = betaW[w][j];
    }
    if (i < TQ_CR) S[i + j * TQ_NLS] = x;
  }
  wave_lds_sync();
  {
    double *tau_out = it.tau + (long long)ck.i * nc;
    if (lane < nc) gmem(tau_out)[lane] = lane < nref ? tauW[w][lane] : 0.0;
    gdbl *g = gmem(G);
    for (int jj = 0; jj < nc; jj++) {
      if (lane < rows) g[lane + jj * ld] = S[lane + jj * TQ_NLS];
      if (lane + 64 < rows) g[lane + 64 + jj * ld] = S[lane + 64 + jj * TQ_NLS];
    }
    // R (upper nc x nc, zero below the diagonal and past `rows`)
    gdbl *ro = gmem(it.rout) + (it.is_top ? 0 : ck.i * nc);
    gdbl *mt = (it.is_top && it.mt) ? gmem(it.mt) : nullptr;
    for (int e = lane; e < nc * nc; e += 64) {
      const int i = e % nc, jj = e / nc;
      const double r = (i <= jj && i < rows) ? S[i + jj * TQ_NLS] : 0.0;
      ro[i + (long long)jj * it.ldr] = r;
      if (mt) mt[jj + (long long)i * nc] = r;
    }
    if (it.is_top && it.est && lane == 0) {
      double mx = 0.0;
      for (int t = it.ell; t < nc; t++) {
        double ss = 0.0;
        for (int i = it.ell; i <= t && i < rows; i++) ss += S[i + t * TQ_NLS] * S[i + t * TQ_NLS];
        mx = fmax(mx, sqrt(ss));
      }
      *gmem(it.est) = mx;
    }
  }
}

// Q of one chunk: [C; 0] (C the parent's ncol x nq slice, the identity at
// the top) times the chunk's reflectors in reverse order, column per lane.
__global__ __launch_bounds__(TQ_NT) void k_tsqr_down_lds(const QrLevel *levels,
                                                         const int2 *chunks) {
  __shared__ __attribute__((aligned(16))) double S[TQ_LS * TQ_FAST_COLS];
  __shared__ double part[2][TQ_NW][64];
  __shared__ double tauS[TQ_FAST_COLS];
  const QrChunk ck = qr_chunk(levels, chunks, blockIdx.x);
  const QrLevel &it = *ck.L;
  const int rows = ck.rows, nc = it.ncol, nq = it.nq, tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, rw0 = w * TQ_RW;
  const int nref = min(rows, nc);
  {
    // reflectors as LDS columns: unit diagonal, zero above it and past rows
    const int i = tid & 127, j0 = tid >> 7;
    const gdbl *g = gmem(it.panel + ck.r0) + min(i, rows - 1);
    const long long ld = it.ld;
    for (int j = j0; j < nref; j += 16) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = g[(long long)min(j + 2 * u, nref - 1) * ld];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int jj = j + 2 * u;
        if (jj < nref) S[i + jj * TQ_LS] = (i > jj && i < rows) ? x[u] : (i == jj ? 1.0 : 0.0);
      }
    }
  }
  for (int k = tid; k < nref; k += TQ_NT) tauS[k] = gmem(it.tau + (long long)ck.i * nc)[k];
  double q[TQ_RW];
  if (it.cin) {
    const gdbl *c = gmem(it.cin) + (long long)ck.i * nc + (long long)min(lane, nq - 1) * it.ldc;
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      const int i = rw0 + r;
      const double x = c[min(i, nc - 1)];
      q[r] = (i < nc && lane < nq) ? x : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) q[r] = (rw0 + r == lane && lane < nq) ? 1.0 : 0.0;
  }
  __syncthreads();
  int par = 0;
  for (int k = nref - 1; k >= 0; k--) {
    const double tau = tauS[k];
    if (tau == 0.0) continue;
    const int kw = k / TQ_RW;
    double v[TQ_RW];
    double pd = 0.0;
    if (w >= kw) {
      const double *vk = S + k * TQ_LS + rw0;
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) {
        v[r] = vk[r];
        pd += v[r] * q[r];
      }
    }
    part[par][w][lane] = pd;
    __syncthreads();
    if (w >= kw) {
      double dot = part[par][0][lane];
#pragma unroll
      for (int ww = 1; ww < TQ_NW; ww++) dot += part[par][ww][lane];
      const double f = tau * dot;
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) q[r] -= f * v[r];
    }
    par ^= 1;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < TQ_RW; r++) S[rw0 + r + lane * TQ_LS] = q[r];
  __syncthreads();
  lds_store_panel(S, it.qout + ck.r0, it.ldq, rows, nq);
}

// Q of a chunk of a panel of at most 16 columns: one wave per chunk, four
// chunks per workgroup, lane 16 q + j holding rows 32 q .. 32 q + 31 of
// column j of Q (as in k_tsqr_up_narrow); no workgroup barriers.
__global__ __launch_bounds__(TQ_NT) void k_tsqr_down_narrow(const QrLevel *levels,
                                                            const int2 *chunks, int nchunks) {
  __shared__ __attribute__((aligned(16))) double Sw[TQ_NW][TQ_CR * TQ_NARROW];
  __shared__ double tauW[TQ_NW][TQ_NARROW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int idx = blockIdx.x * TQ_NW + w;
  if (idx >= nchunks) return;  // no workgroup barriers below
  const QrChunk ck = qr_chunk(levels, chunks, idx);
  const QrLevel &it = *ck.L;
  const int rows = ck.rows, nc = it.ncol, nq = it.nq;
  const int q = lane >> 4, j = lane & 15, rq0 = q * TQ_RW;
  const int nref = min(rows, nc);
  double *S = Sw[w];
  {
    // reflectors as LDS columns: unit diagonal, zero above it and past rows
    const gdbl *g = gmem(it.panel + ck.r0);
    const long long ld = it.ld;
    for (int jj = 0; jj < nref; jj++) {
      const double a = g[min(lane, rows - 1) + jj * ld], b = g[min(lane + 64, rows - 1) + jj * ld];
      S[lane + jj * TQ_CR] = (lane > jj && lane < rows) ? a : (lane == jj ? 1.0 : 0.0);
      S[lane + 64 + jj * TQ_CR] = (lane + 64 < rows) ? b : 0.0;
    }
  }
  if (lane < nref) tauW[w][lane] = gmem(it.tau + (long long)ck.i * nc)[lane];
  double qv[TQ_RW];
  if (it.cin) {
    const gdbl *c = gmem(it.cin) + (long long)ck.i * nc + (long long)min(j, nq - 1) * it.ldc;
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      const int i = rq0 + r;
      const double x = c[min(i, nc - 1)];
      qv[r] = (i < nc && j < nq) ? x : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) qv[r] = (rq0 + r == j && j < nq) ? 1.0 : 0.0;
  }
  wave_lds_sync();
  for (int k = nref - 1; k >= 0; k--) {
    const double tau = tauW[w][k];
    if (tau == 0.0) continue;
    const int kq = k / TQ_RW;
    const double *vk = S + k * TQ_CR + rq0;
    double v[TQ_RW];
    double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < TQ_RW; r++) {
      v[r] = vk[r];
      p4[r & 3] += v[r] * qv[r];
    }
    double pd = q >= kq ? (p4[0] + p4[1]) + (p4[2] + p4[3]) : 0.0;
    pd += __shfl_xor(pd, 16, 64);
    pd += __shfl_xor(pd, 32, 64);
    if (q >= kq) {
      const double f = tau * pd;
#pragma unroll
      for (int r = 0; r < TQ_RW; r++) qv[r] -= f * v[r];
    }
  }
  wave_lds_sync();
  {
    // Q rows of this chunk, one column per 16 lanes' group: straight from
    // registers (lane (q, j) writes its 32 rows of column j)
    gdbl *Q = gmem(it.qout + ck.r0) + (long long)j * it.ldq;
    if (j < nq) {
#pragma unroll
      for (int r = 0; r < TQ_RW; r++)
        if (rq0 + r < rows) Q[rq0 + r] = qv[r];
    }
  }
}

// chunks of one level: rows spread evenly over ceil(rows / TQ_CR) chunks
static long long tq_nchunks(long long rows) {
  return std::max<long long>(1, (rows + TQ_CR - 1) / TQ_CR);
}

int tsqr_add(Program &prog, Arena &ar, const std::vector<TsqrPanel> &panels) {
  if (panels.empty()) return SUCCESS;
  // per panel: rows of each level (level 0 = the panel itself)
  std::vector<std::vector<long long>> rows_per_level(panels.size());
  for (size_t p = 0; p < panels.size(); p++) {
    const TsqrPanel &P = panels[p];
    if (P.m < P.ncol || P.ncol < 1 || P.ncol > TQ_FAST_COLS || P.nq > P.ncol) {
      set_error("tsqr: panel %zu has m=%d ncol=%d nq=%d (needs ncol <= %d)", p, P.m, P.ncol,
                P.nq, TQ_FAST_COLS);
      return INPUT_ERROR;
    }
    long long rows = P.m;
    for (;;) {
      rows_per_level[p].push_back(rows);
      const long long nch = tq_nchunks(rows);
      if (nch == 1) break;
      rows = nch * P.ncol;
    }
  }
  size_t maxL = 0;
  for (auto &r : rows_per_level) maxL = std::max(maxL, r.size());
  std::vector<QrLevel> levels;
  std::vector<std::vector<int2>> up(maxL), up_narrow(maxL), down(maxL), down_narrow(maxL);
  for (size_t p = 0; p < panels.size(); p++) {
    const TsqrPanel &P = panels[p];
    const size_t nl = rows_per_level[p].size();
    // scratch of this panel: tau per level, stacked R and level Q for
    // levels >= 1, the top R when the caller wants none
    std::vector<double *> tau(nl), st(nl, nullptr), qs(nl, nullptr);
    for (size_t L = 0; L < nl; L++) {
      const long long rows = rows_per_level[p][L];
      const long long nch = tq_nchunks(rows);
      tau[L] = ar.take<double>((size_t)nch * P.ncol);
      prog.touch(tau[L], (size_t)nch * P.ncol * sizeof(double), "tsqr tau");
      if (L > 0) {
        st[L] = ar.take<double>((size_t)rows * P.ncol);
        qs[L] = ar.take<double>((size_t)rows * std::max(P.nq, 1));
        prog.touch(st[L], (size_t)rows * P.ncol * sizeof(double), "tsqr stack");
        prog.touch(qs[L], (size_t)rows * std::max(P.nq, 1) * sizeof(double), "tsqr q");
      }
    }
    double *rtop = P.R ? P.R : ar.take<double>((size_t)P.ncol * P.ncol);
    const long long ldr_top = P.R ? P.ldr : P.ncol;
    prog.touch(rtop, ((size_t)(P.ncol - 1) * ldr_top + P.ncol) * sizeof(double), "tsqr R");
    prog.touch(P.A, ((size_t)(P.ncol - 1) * P.lda + P.m) * sizeof(double), "tsqr panel");
    if (P.Q)
      prog.touch(P.Q, ((size_t)(std::max(P.nq, 1) - 1) * P.ldq + P.m) * sizeof(double),
                 "tsqr Q out");
    if (P.Mt) prog.touch(P.Mt, (size_t)P.ncol * P.ncol * sizeof(double), "tsqr R^T");
    if (P.est) prog.touch(P.est, sizeof(double), "tsqr est");
    for (size_t L = 0; L < nl; L++) {
      const long long rows = rows_per_level[p][L];
      const long long nch = tq_nchunks(rows);
      const bool top = (L + 1 == nl);
      QrLevel lv;
      memset(&lv, 0, sizeof(lv));
      lv.panel = L == 0 ? P.A : st[L];
      lv.ld = L == 0 ? P.lda : rows;
      lv.rows = rows;
      lv.nch = (int)nch;
      lv.ncol = P.ncol;
      lv.nq = P.nq;
      lv.is_top = top ? 1 : 0;
      lv.tau = tau[L];
      if (top) {
        lv.rout = rtop;
        lv.ldr = ldr_top;
        lv.mt = P.Mt;
        lv.est = P.est;
        lv.ell = P.ell;
      } else {
        const long long rows_next = rows_per_level[p][L + 1];
        lv.rout = st[L + 1];
        lv.ldr = rows_next;
        lv.cin = qs[L + 1];
        lv.ldc = rows_next;
      }
      lv.qout = L == 0 ? P.Q : qs[L];
      lv.ldq = L == 0 ? P.ldq : rows;
      const int li = (int)levels.size();
      levels.push_back(lv);
      for (long long i = 0; i < nch; i++) {
        (P.ncol <= TQ_NARROW ? up_narrow[L] : up[L]).push_back(make_int2(li, (int)i));
        if (P.Q) (P.ncol <= TQ_NARROW ? down_narrow[L] : down[L]).push_back(make_int2(li, (int)i));
      }
    }
  }
  const size_t lev_off = prog.push(levels);
  std::vector<size_t> up_off(maxL), upn_off(maxL), down_off(maxL), downn_off(maxL);
  for (size_t L = 0; L < maxL; L++) {
    up_off[L] = prog.push(up[L]);
    upn_off[L] = prog.push(up_narrow[L]);
    down_off[L] = prog.push(down[L]);
    downn_off[L] = prog.push(down_narrow[L]);
  }
  for (size_t L = 0; L < maxL; L++) {
    const unsigned n = (unsigned)up[L].size(), nn = (unsigned)up_narrow[L].size();
    const size_t off = up_off[L], offn = upn_off[L];
    if (n || nn)
      prog.add("tsqr_up", [=](hipStream_t s, const unsigned char *b) {
        if (n)
          hipLaunchKernelGGL(k_tsqr_up_lds, dim3(n), dim3(TQ_NT), 0, s,
                             reinterpret_cast<const QrLevel *>(b + lev_off),
                             reinterpret_cast<const int2 *>(b + off));
        if (nn)
          hipLaunchKernelGGL(k_tsqr_up_narrow, dim3((nn + TQ_NW - 1) / TQ_NW), dim3(TQ_NT), 0, s,
                             reinterpret_cast<const QrLevel *>(b + lev_off),
                             reinterpret_cast<const int2 *>(b + offn), (int)nn);
      });
  }
  for (size_t L = maxL; L-- > 0;) {
    const unsigned n = (unsigned)down[L].size(), nn = (unsigned)down_narrow[L].size();
    const size_t off = down_off[L], offn = downn_off[L];
    if (n || nn)
      prog.add("tsqr_down", [=](hipStream_t s, const unsigned char *b) {
        if (n)
          hipLaunchKernelGGL(k_tsqr_down_lds, dim3(n), dim3(TQ_NT), 0, s,
                             reinterpret_cast<const QrLevel *>(b + lev_off),
                             reinterpret_cast<const int2 *>(b + off));
        if (nn)
          hipLaunchKernelGGL(k_tsqr_down_narrow, dim3((nn + TQ_NW - 1) / TQ_NW), dim3(TQ_NT), 0,
                             s, reinterpret_cast<const QrLevel *>(b + lev_off),
                             reinterpret_cast<const int2 *>(b + offn), (int)nn);
      });
  }
  return SUCCESS;
}

int qr_add(Program &prog, Arena &ar, const std::vector<QrJob> &jobs, int bw) {
  int maxp = 0;
  for (const QrJob &J : jobs) {
    if (J.L > J.m || J.L < 0) {
      set_error("qr: job with m=%d L=%d", J.m, J.L);
      return INPUT_ERROR;
    }
    if (J.c_start < 0 || J.c_start % bw != 0 || (J.c_start > 0 && (J.c_start >= J.L || !J.Qprev ||
                                                                 (J.R && !J.Rprev_t)))) {
      set_error("qr: continued job with c_start=%d (L=%d, panel width %d)", J.c_start, J.L, bw);
      return INPUT_ERROR;
    }
    maxp = std::max(maxp, (J.L + bw - 1) / bw);
  }
  // per-job temporaries: C1 (when R is not wanted) and C2 (L x bw), R1, R2
  struct Tmp {
    double *c1, *c2, *c3, *r1, *r2;
  };
  std::vector<Tmp> tmp(jobs.size());
  std::vector<Copy2D> zero;
  for (size_t q = 0; q < jobs.size(); q++) {
    const QrJob &J = jobs[q];
    if (J.L <= 0) continue;
    tmp[q].c1 = J.R ? nullptr : ar.take<double>((size_t)J.L * bw);
    tmp[q].c2 = ar.take<double>((size_t)J.L * bw);
    tmp[q].c3 = ar.take<double>((size_t)J.L * bw);
    tmp[q].r1 = ar.take<double>((size_t)bw * bw);
    tmp[q].r2 = ar.take<double>((size_t)bw * bw);
    if (J.R) zero.push_back({nullptr, J.R, 0, J.ldr, J.L, J.L, 0, 0});
  }
  // the done columns of continued jobs (after R's zero fill)
  std::vector<Copy2D> carried;
  for (const QrJob &J : jobs) {
    if (J.c_start <= 0 || J.L <= 0) continue;
    carried.push_back({J.Qprev, J.Q, J.ldqp, J.ldq, J.m, J.c_start, 0, 0});
    if (J.R) carried.push_back({J.Rprev_t, J.R, J.ldrp, J.ldr, J.c_start, J.c_start, 1, 0});
  }
  HM_OK(copy2d_add(prog, zero));
  HM_OK(copy2d_add(prog, carried));
  for (int p = 0; p < maxp; p++) {
    const int c0 = p * bw;
    std::vector<GemmJob> g1, g2, g3, g4, g3b, g4b, g5, g5b, g6;
    std::vector<TsqrPanel> t1, t2;
    std::vector<Copy2D> cp;
    for (size_t q = 0; q < jobs.size(); q++) {
      const QrJob &J = jobs[q];
      if (J.L <= c0 || c0 < J.c_start) continue;
      const int w = std::min(bw, J.L - c0);
      double *Yp = J.Y + (long long)c0 * J.ldy;
      double *Qp = J.Q + (long long)c0 * J.ldq;
      double *C1 = J.R ? J.R + (long long)c0 * J.ldr : tmp[q].c1;
      const long long ldc1 = J.R ? J.ldr : J.L;
      if (p > 0) {
        // C1 = Q_prev^T Y_p ; Y_p -= Q_prev C1
        g1.push_back({J.Q, Yp, C1, nullptr, J.ldq, J.ldy, ldc1, c0, w, J.m, 1, 0, 0, 1.0, 0.0});
        g2.push_back({J.Q, C1, Yp, nullptr, J.ldq, ldc1, J.ldy, J.m, w, c0, 1, 0, 0, -1.0, 1.0});
      }
      TsqrPanel P;
      memset(&P, 0, sizeof(P));
      P.A = Yp;
      P.lda = J.ldy;
      P.m = J.m;
      P.ncol = w;
      P.nq = w;
      P.Q = Qp;
      P.ldq = J.ldq;
      if (p == 0 && J.R) {
        P.R = J.R;
        P.ldr = J.ldr;
      } else {
        P.R = tmp[q].r1;
        P.ldr = bw;
      }
      t1.push_back(P);
      if (p > 0) {
        // reorthogonalise: C2 = Q_prev^T Q_p ; Q_p -= Q_prev C2, twice
        // (C3 the second time), then TSQR again.  When the panel lies in the
        // span of the previous ones to rounding (a numerically rank-deficient
        // Y), the first projection leaves |Q_p| tiny and a single second
        // projection is not enough: the TSQR would magnify what is left in
        // that span and Q would lose orthogonality.
        g3.push_back({J.Q, Qp, tmp[q].c2, nullptr, J.ldq, J.ldq, J.L, c0, w, J.m, 1, 0, 0, 1.0, 0.0});
        g4.push_back({J.Q, tmp[q].c2, Qp, nullptr, J.ldq, J.L, J.ldq, J.m, w, c0, 1, 0, 0, -1.0, 1.0});
        g3b.push_back({J.Q, Qp, tmp[q].c3, nullptr, J.ldq, J.ldq, J.L, c0, w, J.m, 1, 0, 0, 1.0, 0.0});
        g4b.push_back({J.Q, tmp[q].c3, Qp, nullptr, J.ldq, J.L, J.ldq, J.m, w, c0, 1, 0, 0, -1.0, 1.0});
        TsqrPanel P2;
        memset(&P2, 0, sizeof(P2));
        P2.A = Qp;
        P2.lda = J.ldq;
        P2.m = J.m;
        P2.ncol = w;
        P2.nq = w;
        P2.Q = Yp;
        P2.ldq = J.ldy;
        P2.R = tmp[q].r2;
        P2.ldr = bw;
        t2.push_back(P2);
        cp.push_back({Yp, Qp, J.ldy, J.ldq, J.m, w, 0, 0});
        if (J.R) {
          // R[:c0, p] = C1 + (C2 + C3) R1 ; R[p, p] = R2 R1
          g5.push_back({tmp[q].c2, tmp[q].r1, J.R + (long long)c0 * J.ldr, nullptr, J.L, bw,
                        J.ldr, c0, w, w, 1, 0, 0, 1.0, 1.0});
          g5b.push_back({tmp[q].c3, tmp[q].r1, J.R + (long long)c0 * J.ldr, nullptr, J.L, bw,
                        J.ldr, c0, w, w, 1, 0, 0, 1.0, 1.0});
          g6.push_back({tmp[q].r2, tmp[q].r1, J.R + c0 + (long long)c0 * J.ldr, nullptr, bw, bw,
                        J.ldr, w, w, w, 1, 0, 0, 1.0, 0.0});
        }
      }
    }
    HM_OK(gemm_add(prog, ar, g1, true, false, "qr proj1", GEMM_BY_SHAPE));
    HM_OK(gemm_add(prog, ar, g2, false, false, "qr upd1", GEMM_BY_SHAPE));
    HM_OK(tsqr_add(prog, ar, t1));
    HM_OK(gemm_add(prog, ar, g3, true, false, "qr proj2", GEMM_BY_SHAPE));
    HM_OK(gemm_add(prog, ar, g4, false, false, "qr upd2", GEMM_BY_SHAPE));
    HM_OK(gemm_add(prog, ar, g3b, true, false, "qr proj3", GEMM_BY_SHAPE));
    HM_OK(gemm_add(prog, ar, g4b, false, false, "qr upd3", GEMM_BY_SHAPE));
    HM_OK(tsqr_add(prog, ar, t2));
    HM_OK(copy2d_add(prog, cp));
    HM_OK(gemm_add(prog, ar, g5, false, false, "qr r12", GEMM_BY_SHAPE));
    HM_OK(gemm_add(prog, ar, g5b, false, false, "qr r12", GEMM_BY_SHAPE));  // same output: own launch
    HM_OK(gemm_add(prog, ar, g6, false, false, "qr r22", GEMM_BY_SHAPE));
  }
  return SUCCESS;
}

// Buffers reused by consecutive program_exec calls (it synchronizes its
// stream before returning, so they are idle in between): device scratch,
// device work-item blob and a pinned host staging copy of the blob, all on
// the first device that used them.  A call that finds them taken (another
// host thread) or runs on another device allocates its own.
namespace {
struct ExecBuffers {
  std::mutex mu;
  bool busy = false;
  int device = -1;  // device the buffers live on
  char *arena = nullptr;
  size_t arena_cap = 0;
  unsigned char *blob = nullptr;
  size_t blob_cap = 0;
  unsigned char *host = nullptr;
  size_t host_cap = 0;
};
ExecBuffers g_exec;

// Programs whose scratch is larger than this allocate and free their own
// (the kept buffers live until the process ends).
constexpr size_t kKeepBytes = (size_t)1 << 30;

template <class T>
int grow_device(T **p, size_t *cap, size_t need) {
  if (*cap >= need) return SUCCESS;
  const size_t want = std::max(need, *cap + *cap / 2);
  dfree(*p);
  *p = nullptr;
  *cap = 0;
  HM_OK(dalloc(p, want));
  *cap = want;
  return SUCCESS;
}
}  // namespace

int release_exec_buffers() {
  std::lock_guard<std::mutex> lock(g_exec.mu);
  if (g_exec.busy) return INPUT_ERROR;
  dfree(g_exec.arena);
  dfree(g_exec.blob);
  if (g_exec.host) (void)hipHostFree(g_exec.host);
  g_exec.arena = nullptr;
  g_exec.blob = nullptr;
  g_exec.host = nullptr;
  g_exec.arena_cap = g_exec.blob_cap = g_exec.host_cap = 0;
  g_exec.device = -1;
  return SUCCESS;
}

int program_exec(const std::function<int(Program &, Arena &)> &build,
                 const std::vector<MemRange> &inputs, hipStream_t s) {
  const auto h0 = std::chrono::steady_clock::now();
  Program dry;
  Arena count;
  dry.record = true;
  HM_OK(build(dry, count));
  for (const auto &r : inputs) dry.allow(r.lo, r.bytes);
  dry.allow(Arena::fake_base(), count.off + 256);
  // The audit only reads `dry`: for large programs (a batched compression
  // records hundreds of thousands of ranges) it runs on a helper thread while
  // the real program is built, and is joined before anything is uploaded or
  // launched; small ones are audited here (a thread costs more than they do).
  std::string msg;
  int audit_rc = SUCCESS;
  std::thread auditor;
  struct Join {
    std::thread &t;
    ~Join() {
      if (t.joinable()) t.join();
    }
  } join_auditor{auditor};
  if (dry.touched.size() >= 100000) {
    auditor = std::thread([&]() { audit_rc = dry.audit(&msg); });
  } else if (dry.audit(&msg) != SUCCESS) {
    set_error("launch program audit: %s", msg.c_str());
    return INPUT_ERROR;
  }
  bool cached = false;
  int dev = -1;
  if (count.off + 256 <= kKeepBytes && dry.blob.size() + 64 <= kKeepBytes / 4 &&
      hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_exec.mu);
    if (!g_exec.busy && (g_exec.device < 0 || g_exec.device == dev)) {
      g_exec.busy = cached = true;
      g_exec.device = dev;
    }
  }
  auto release = [&]() {
    if (!cached) return;
    std::lock_guard<std::mutex> lock(g_exec.mu);
    g_exec.busy = false;
  };
  Arena ar;
  ar.counting = false;
  ar.cap = count.off + 256;
  int rc = SUCCESS;
  if (cached) {
    rc = grow_device(&g_exec.arena, &g_exec.arena_cap, ar.cap);
    ar.base = g_exec.arena;
  } else {
    rc = dalloc(&ar.base, ar.cap);
  }
  if (rc != SUCCESS) {
    release();
    return rc;
  }
  Program prog;
  prog.blob.reserve(dry.blob.size() + 64);  // the same items: no regrowth copies
  rc = build(prog, ar);
  if (auditor.joinable()) {
    auditor.join();
    if (audit_rc != SUCCESS) {
      set_error("launch program audit: %s", msg.c_str());
      rc = INPUT_ERROR;
    }
  }
  if (rc == SUCCESS && ar.off > ar.cap) {
    set_error("launch program: scratch grew between passes");
    rc = INPUT_ERROR;
  }
  if (rc == SUCCESS) {
    if (cached) {
      const size_t nb = prog.blob.size() + 64;
      rc = grow_device(&g_exec.blob, &g_exec.blob_cap, nb);
      if (rc == SUCCESS && g_exec.host_cap < nb) {
        if (g_exec.host) (void)hipHostFree(g_exec.host);
        g_exec.host = nullptr;
        g_exec.host_cap = 0;
        const size_t want = std::max(nb, (size_t)1 << 20);
        if (hipHostMalloc(reinterpret_cast<void **>(&g_exec.host), want) == hipSuccess)
          g_exec.host_cap = want;
        else
          g_exec.host = nullptr;
      }
      if (rc == SUCCESS) {
        prog.d_blob = g_exec.blob;
        prog.own_blob = false;
        if (!prog.blob.empty()) {
          const unsigned char *src = prog.blob.data();
          if (g_exec.host) {
            memcpy(g_exec.host, prog.blob.data(), prog.blob.size());
            src = g_exec.host;
          }
          hipError_t e = hipMemcpyAsync(prog.d_blob, src, prog.blob.size(),
                                        hipMemcpyHostToDevice, s);
          if (e != hipSuccess) {
            set_error("launch program upload: %s", hipGetErrorString(e));
            rc = ALLOCATION_FAILURE;
          }
        }
      }
    } else {
      rc = prog.upload(s);
    }
  }
  timing_host("program build (host)",
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0)
                  .count());
  if (rc == SUCCESS) rc = prog.run(s);
  if (rc == SUCCESS) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("launch program: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  (void)hipStreamSynchronize(s);
  if (!cached) dfree(ar.base);
  release();
  return rc;
}

}  // namespace hm

using namespace hm;

// Host-only check of the QR program builder: builds the program for jobs of
// the given shapes on fake buffers (nothing is allocated or launched) and
// verifies that every GEMM, TSQR, copy and norm item stays inside them.
extern "C" int hm_hip_audit_qr(int njobs, const int *m, const int *L, int with_r,
                               char *msg, int msglen) {
  std::vector<QrJob> jobs;
  std::vector<MemRange> bufs;
  char *fake = reinterpret_cast<char *>((uintptr_t)1 << 40);
  size_t off = 0;
  auto take = [&](size_t n) {
    char *p = fake + off;
    off += (n * sizeof(double) + 4095) & ~(size_t)4095;
    bufs.push_back({p, n * sizeof(double)});
    return reinterpret_cast<double *>(p);
  };
  for (int i = 0; i < njobs; i++) {
    QrJob J;
    J.m = m[i];
    J.L = L[i];
    J.ldy = m[i] + (i % 3);
    J.Y = take((size_t)J.ldy * std::max(1, J.L));
    J.ldq = m[i] + (i % 2);
    J.Q = take((size_t)J.ldq * std::max(1, J.L));
    J.ldr = J.L + 1;
    J.R = with_r ? take((size_t)J.ldr * std::max(1, J.L)) : nullptr;
    jobs.push_back(J);
  }
  Program p;
  Arena ar;
  p.record = true;
  int rc = qr_add(p, ar, jobs);
  std::string text = rc == SUCCESS ? "" : "builder failed";
  if (rc == SUCCESS) {
    for (const auto &b : bufs) p.allow(b.lo, b.bytes);
    p.allow(Arena::fake_base(), ar.off + 256);
    rc = p.audit(&text);
    if (rc == SUCCESS) {
      char buf[128];
      snprintf(buf, sizeof(buf), "%zu ops, %zu items bytes, %zu scratch bytes", p.ops.size(),
               p.blob.size(), ar.off);
      text = buf;
    }
  }
  if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", text.c_str());
  return rc;
}

// Thin QR of device panels: the batched Householder/TSQR + BCGS2 program the
// compression uses.  Each Y_i (m_i x n_i, n_i <= m_i) is overwritten.
extern "C" int hm_hip_qr_batch(int njobs, const int *m, const int *n, double *const *Y,
                               const long *ldy, double *const *Q, const long *ldq,
                               double *const *R, const long *ldr, void *stream) {
  if (njobs <= 0) return SUCCESS;
  if (!m || !n || !Y || !ldy || !Q || !ldq || (R && !ldr)) {
    set_error("hm_hip_qr_batch: NULL argument");
    return INPUT_ERROR;
  }
  std::vector<QrJob> jobs;
  std::vector<MemRange> inputs;
  for (int i = 0; i < njobs; i++) {
    if (n[i] < 1 || m[i] < n[i] || !Y[i] || !Q[i] || ldy[i] < m[i] || ldq[i] < m[i] ||
        (R && R[i] && ldr[i] < n[i])) {
      set_error("hm_hip_qr_batch: panel %d has m=%d n=%d (needs 1 <= n <= m, ld >= rows)", i,
                m[i], n[i]);
      return INPUT_ERROR;
    }
    QrJob J;
    J.Y = Y[i];
    J.m = m[i];
    J.ldy = ldy[i];
    J.L = n[i];
    J.Q = Q[i];
    J.ldq = ldq[i];
    J.R = (R && R[i]) ? R[i] : nullptr;
    J.ldr = J.R ? ldr[i] : 0;
    jobs.push_back(J);
    inputs.push_back({J.Y, ((size_t)(J.L - 1) * J.ldy + J.m) * sizeof(double)});
    inputs.push_back({J.Q, ((size_t)(J.L - 1) * J.ldq + J.m) * sizeof(double)});
    if (J.R) inputs.push_back({J.R, ((size_t)(J.L - 1) * J.ldr + J.L) * sizeof(double)});
  }
  return program_exec([&](Program &p, Arena &ar) { return qr_add(p, ar, jobs); }, inputs,
                      stream_of(stream));
}
