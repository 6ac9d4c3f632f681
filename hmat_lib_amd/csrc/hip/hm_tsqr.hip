// Batched tall-skinny Householder QR (TSQR) with explicit Q, FP64.
//
// A panel (m x ncol, column-major) is cut into chunks of CR rows (the last
// chunk takes the remainder, so every chunk has >= ncol rows); each chunk is
// factored by one workgroup (Householder, LAPACK dlarfg sign convention:
// beta = -sign(alpha) ||x||, tau = 0 when the tail is zero), its R is
// stacked into the next level's panel, and so on until one chunk remains.
// The explicit Q (first nq columns) is then built top-down: each chunk
// applies its reflectors to [C; 0] where C is its ncol x nq slice of the
// parent level's Q.  All chunks of all panels of one level run in one
// launch.  Extras of the top chunk: R (ncol x ncol), R^T, and the norms of
// the trailing columns of R past column `ell` (for the randomized range
// finder's posterior test: they equal |(I - Q_1 Q_1^T) y_t|).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hm_tsqr.h"

namespace hm {

static constexpr int TQ_NT = 256;

struct QrItem {
  double *panel;
  long long ld;
  int r0, rows, ncol, nq;
  double *tau;
  double *rout;       // up: R_i destination (rows rrow0.., ld ldr)
  long long ldr;
  int rrow0;
  int is_top;
  const double *cin;  // down: parent Q slice (NULL: identity)
  long long ldc;
  int crow0;
  int ell;            // top: trailing-column norms past ell -> *est
  double *qout;       // down: Q rows [r0, r0+rows) with ld ldq
  long long ldq;
  double *est;
  double *mt;         // top: R^T (ncol x ncol, ld ncol)
};

template <int NV>
__device__ __forceinline__ void bsum(double (&v)[NV], double *red, double *out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int NW = TQ_NT / 64;
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

__device__ __forceinline__ int own_first(int tid, int r) {
  return tid >= r ? tid : tid + ((r - tid + TQ_NT - 1) / TQ_NT) * TQ_NT;
}

__global__ __launch_bounds__(TQ_NT) void k_tsqr_up(const QrItem *items) {
  __shared__ double red[8 * (TQ_NT / 64)], out[8];
  const QrItem it = items[blockIdx.x];
  double *Y = it.panel + it.r0;
  const long long ld = it.ld;
  const int rows = it.rows, nc = it.ncol, tid = threadIdx.x;
  const int nref = min(rows, nc);
  for (int k = 0; k < nref; k++) {
    double p[1] = {0.0};
    for (int i = own_first(tid, k + 1); i < rows; i += TQ_NT) {
      const double x = Y[i + k * ld];
      p[0] += x * x;
    }
    bsum<1>(p, red, out);
    const double xn2 = out[0];
    const double alpha = Y[k + k * ld];
    double tau = 0.0, beta = alpha, scal = 1.0;
    if (xn2 > 0.0) {
      const double hh = hypot(alpha, sqrt(xn2));
      beta = (alpha >= 0.0) ? -hh : hh;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int i = own_first(tid, k); i < rows; i += TQ_NT) {
      if (i == k) Y[i + k * ld] = beta;
      else if (tau != 0.0) Y[i + k * ld] *= scal;
    }
    if (tid == 0) it.tau[k] = tau;
    if (tau == 0.0) continue;
    for (int jb = k + 1; jb < nc; jb += 8) {
      const int nv = min(8, nc - jb);
      double q[8];
#pragma unroll
      for (int v = 0; v < 8; v++) q[v] = 0.0;
      for (int i = own_first(tid, k); i < rows; i += TQ_NT) {
        const double vi = (i == k) ? 1.0 : Y[i + k * ld];
#pragma unroll
        for (int v = 0; v < 8; v++)
          if (v < nv) q[v] += vi * Y[i + (jb + v) * ld];
      }
      bsum<8>(q, red, out);
      for (int i = own_first(tid, k); i < rows; i += TQ_NT) {
        const double vi = (i == k) ? 1.0 : Y[i + k * ld];
#pragma unroll
        for (int v = 0; v < 8; v++)
          if (v < nv) Y[i + (jb + v) * ld] -= tau * out[v] * vi;
      }
    }
  }
  if (tid < nc && nref < nc) {
    for (int k = nref; k < nc; k++)
      if (tid == 0) it.tau[k] = 0.0;
  }
  __syncthreads();
  // R (upper ncol x ncol; rows past `rows` are zero) to the next level
  for (int e = tid; e < nc * nc; e += TQ_NT) {
    const int i = e % nc, j = e / nc;
    const double r = (i <= j && i < rows) ? Y[i + j * ld] : 0.0;
    it.rout[it.rrow0 + i + j * it.ldr] = r;
    if (it.is_top && it.mt) it.mt[j + i * nc] = r;
  }
  if (it.is_top && it.est) {
    __syncthreads();
    if (tid == 0) {
      double mx = 0.0;
      for (int t = it.ell; t < nc; t++) {
        double ss = 0.0;
        for (int i = it.ell; i <= t && i < rows; i++) {
          const double r = Y[i + t * ld];
          ss += r * r;
        }
        mx = fmax(mx, sqrt(ss));
      }
      *it.est = mx;
    }
  }
}

__global__ __launch_bounds__(TQ_NT) void k_tsqr_down(const QrItem *items) {
  __shared__ double red[8 * (TQ_NT / 64)], out[8];
  const QrItem it = items[blockIdx.x];
  const double *Y = it.panel + it.r0;
  const long long ld = it.ld;
  double *Q = it.qout + it.r0;
  const long long ldq = it.ldq;
  const int rows = it.rows, nc = it.ncol, nq = it.nq, tid = threadIdx.x;
  for (int i = tid; i < rows; i += TQ_NT)
    for (int j = 0; j < nq; j++) {
      double v = 0.0;
      if (i < nc) v = it.cin ? it.cin[it.crow0 + i + j * it.ldc] : (i == j ? 1.0 : 0.0);
      Q[i + j * ldq] = v;
    }
  __syncthreads();
  const int nref = min(rows, nc);
  for (int k = nref - 1; k >= 0; k--) {
    const double tau = it.tau[k];
    if (tau == 0.0) continue;
    for (int jb = 0; jb < nq; jb += 8) {
      const int nv = min(8, nq - jb);
      double q[8];
#pragma unroll
      for (int v = 0; v < 8; v++) q[v] = 0.0;
      for (int i = own_first(tid, k); i < rows; i += TQ_NT) {
        const double vi = (i == k) ? 1.0 : Y[i + k * ld];
#pragma unroll
        for (int v = 0; v < 8; v++)
          if (v < nv) q[v] += vi * Q[i + (jb + v) * ldq];
      }
      bsum<8>(q, red, out);
      for (int i = own_first(tid, k); i < rows; i += TQ_NT) {
        const double vi = (i == k) ? 1.0 : Y[i + k * ld];
#pragma unroll
        for (int v = 0; v < 8; v++)
          if (v < nv) Q[i + (jb + v) * ldq] -= tau * out[v] * vi;
      }
    }
  }
}

int tsqr_run(const std::vector<TsqrPanel> &panels, hipStream_t s) {
  if (panels.empty()) return SUCCESS;
  // plan: per panel, levels of chunks
  struct Level {
    std::vector<QrItem> up, down;
  };
  std::vector<Level> levels;
  long long scratch = 0;
  struct Pending {
    int panel, level, kind;  // kind 0 tau, 1 storage S, 2 storage QS, 3 top R
    long long off;
  };
  // first pass: sizes
  std::vector<std::vector<long long>> rows_per_level(panels.size());
  for (size_t p = 0; p < panels.size(); p++) {
    const TsqrPanel &P = panels[p];
    const int CR = std::max(128, 2 * P.ncol);
    long long rows = P.m;
    for (;;) {
      rows_per_level[p].push_back(rows);
      const long long nch = std::max<long long>(1, rows / CR);
      if (nch == 1) break;
      rows = nch * P.ncol;
    }
  }
  // scratch layout
  std::vector<std::vector<long long>> s_off(panels.size()), q_off(panels.size()),
      t_off(panels.size());
  std::vector<long long> rtop_off(panels.size());
  for (size_t p = 0; p < panels.size(); p++) {
    const TsqrPanel &P = panels[p];
    const int CR = std::max(128, 2 * P.ncol);
    for (size_t L = 0; L < rows_per_level[p].size(); L++) {
      const long long rows = rows_per_level[p][L];
      const long long nch = std::max<long long>(1, rows / CR);
      t_off[p].push_back(scratch);
      scratch += nch * P.ncol;
      if (L > 0) {
        s_off[p].push_back(scratch);
        scratch += rows * P.ncol;
        q_off[p].push_back(scratch);
        scratch += rows * P.nq;
      } else {
        s_off[p].push_back(-1);
        q_off[p].push_back(-1);
      }
    }
    rtop_off[p] = scratch;
    scratch += (long long)P.ncol * P.ncol;
  }
  double *sc = nullptr;
  HM_OK(dalloc(&sc, (size_t)scratch + 1));
  size_t maxL = 0;
  for (auto &r : rows_per_level) maxL = std::max(maxL, r.size());
  levels.resize(maxL);
  for (size_t p = 0; p < panels.size(); p++) {
    const TsqrPanel &P = panels[p];
    const int CR = std::max(128, 2 * P.ncol);
    const size_t nl = rows_per_level[p].size();
    for (size_t L = 0; L < nl; L++) {
      const long long rows = rows_per_level[p][L];
      const long long nch = std::max<long long>(1, rows / CR);
      const bool top = (L + 1 == nl);
      double *store = L == 0 ? P.A : sc + s_off[p][L];
      const long long ld = L == 0 ? P.lda : rows;
      double *qstore = L == 0 ? P.Q : sc + q_off[p][L];
      const long long ldq = L == 0 ? P.ldq : rows;
      const long long rows_next = top ? 0 : rows_per_level[p][L + 1];
      for (long long i = 0; i < nch; i++) {
        QrItem it;
        memset(&it, 0, sizeof(it));
        it.panel = store;
        it.ld = ld;
        it.r0 = (int)(i * CR);
        it.rows = (int)((i == nch - 1) ? rows - i * CR : CR);
        it.ncol = P.ncol;
        it.nq = P.nq;
        it.tau = sc + t_off[p][L] + i * P.ncol;
        it.is_top = top ? 1 : 0;
        if (top) {
          it.rout = P.R ? P.R : sc + rtop_off[p];
          it.ldr = P.R ? P.ldr : P.ncol;
          it.rrow0 = 0;
          it.mt = P.Mt;
          it.est = P.est;
          it.ell = P.ell;
          it.cin = nullptr;
        } else {
          it.rout = sc + s_off[p][L + 1];
          it.ldr = rows_next;
          it.rrow0 = (int)(i * P.ncol);
          it.cin = sc + q_off[p][L + 1];
          it.ldc = rows_next;
          it.crow0 = (int)(i * P.ncol);
        }
        it.qout = qstore;
        it.ldq = ldq;
        levels[L].up.push_back(it);
        if (P.Q) levels[L].down.push_back(it);
      }
    }
  }
  int rc = SUCCESS;
  std::vector<QrItem *> dev(levels.size(), nullptr);
  for (size_t L = 0; L < levels.size() && rc == SUCCESS; L++) {
    auto &v = levels[L].up;
    rc = dalloc(&dev[L], v.size() * 2 + 1);
    if (rc == SUCCESS) {
      std::vector<QrItem> both(v);
      both.insert(both.end(), levels[L].down.begin(), levels[L].down.end());
      hipError_t e = hipMemcpyAsync(dev[L], both.data(), both.size() * sizeof(QrItem),
                                    hipMemcpyHostToDevice, s);
      if (e != hipSuccess) {
        set_error("tsqr: copy items: %s", hipGetErrorString(e));
        rc = ALLOCATION_FAILURE;
      }
    }
  }
  if (rc == SUCCESS) {
    Timed tu = timing_begin("tsqr_up", s);
    for (size_t L = 0; L < levels.size(); L++)
      hipLaunchKernelGGL(k_tsqr_up, dim3((unsigned)levels[L].up.size()), dim3(TQ_NT), 0, s,
                         dev[L]);
    timing_end(tu, s);
    Timed td = timing_begin("tsqr_down", s);
    for (size_t L = levels.size(); L-- > 0;)
      if (!levels[L].down.empty())
        hipLaunchKernelGGL(k_tsqr_down, dim3((unsigned)levels[L].down.size()), dim3(TQ_NT),
                           0, s, dev[L] + levels[L].up.size());
    timing_end(td, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("tsqr: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  for (auto *d : dev) dfree(d);
  dfree(sc);
  return rc;
}

}  // namespace hm
