// Launch programs and the batched kernels they are built from (FP64):
// MFMA GEMM tiles with deterministic split-K, strided copies, column norms.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hm_prog.h"

namespace hm {

// ------------------------------------------------------------- program
int Program::audit(std::string *msg) const {
  // allowed ranges sorted by start, with the running maximum of their ends:
  // a touched range is covered iff some allowed range starting at or before
  // it ends at or after it; ranges covered only by a union of several
  // allowed ranges are rejected (every item stays in one buffer)
  std::vector<std::pair<const char *, const char *>> al;
  for (const auto &a : allowed) {
    const char *alo = static_cast<const char *>(a.lo);
    al.push_back({alo, alo + a.bytes});
  }
  std::sort(al.begin(), al.end());
  std::vector<const char *> maxend(al.size());
  for (size_t i = 0; i < al.size(); i++)
    maxend[i] = i ? std::max(maxend[i - 1], al[i].second) : al[i].second;
  for (const auto &t : touched) {
    const char *lo = static_cast<const char *>(t.first.lo);
    const char *hi = lo + t.first.bytes;
    // the allowed range with the furthest end among those starting at or
    // before lo covers [lo, hi) iff that end reaches hi
    auto it = std::upper_bound(al.begin(), al.end(),
                               std::make_pair(lo, (const char *)UINTPTR_MAX));
    const bool ok = it != al.begin() && hi <= maxend[(it - al.begin()) - 1];
    if (!ok) {
      if (msg) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s touches [%p, %p) outside every buffer",
                 t.second, (const void *)lo, (const void *)hi);
        *msg = buf;
      }
      return INPUT_ERROR;
    }
  }
  return SUCCESS;
}

int Program::upload(hipStream_t s) {
  dfree(d_blob);
  d_blob = nullptr;
  HM_OK(dalloc(&d_blob, blob.size() + 64));
  if (!blob.empty())
    HM_HIP(hipMemcpyAsync(d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  return SUCCESS;
}

int Program::run(hipStream_t s, bool timed) {
  for (size_t i = 0; i < ops.size(); i++) {
    Timed t;
    if (timed) t = timing_begin(names[i].c_str(), s);
    ops[i](s, d_blob);
    if (timed) timing_end(t, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("program op %zu (%s): launch failed: %s", i, names[i].c_str(),
                hipGetErrorString(e));
      return ALLOCATION_FAILURE;
    }
  }
  return SUCCESS;
}

// ------------------------------------------------------------------ GEMM
typedef MfmaTile<2, 2, 2, 2, 16> GTile0;  // 64 x 64
typedef MfmaTile<4, 1, 1, 1, 16> GTile1;  // 64 x 16
typedef MfmaTile<1, 4, 1, 1, 16> GTile2;  // 16 x 64
static_assert(GTile0::NT == 256 && GTile1::NT == 256 && GTile2::NT == 256, "256 threads");

template <class Tile, bool TA, bool TB>
__device__ __forceinline__ void gemm_tile(const GemmJob &J, int4 t, double *lds) {
  const int i0 = t.y * Tile::TM, j0 = t.z * Tile::TN;
  const int kb = t.w * J.kseg;
  const int kl = min(J.K, kb + J.kseg) - kb;
  const double *A = J.A + (TA ? (long long)kb : (long long)kb * J.lda);
  const double *B = J.B + (TB ? (long long)kb * J.ldb : (long long)kb);
  const long long lda = J.lda, ldb = J.ldb;
  const int M = J.M, N = J.N;
  auto la = [&](int i, int k) -> double {
    if (i0 + i >= M || k >= kl) return 0.0;
    return ldg(A + (TA ? k + (long long)(i0 + i) * lda : (i0 + i) + (long long)k * lda));
  };
  auto lb = [&](int k, int j) -> double {
    if (k >= kl || j0 + j >= N) return 0.0;
    return ldg(B + (TB ? (j0 + j) + (long long)k * ldb : k + (long long)(j0 + j) * ldb));
  };
  Tile tile;
  tile.zero();
  tile.template run<!TA, TB>(lds, kl, la, lb);
  if (J.ksplit == 1) {
    const double alpha = J.alpha, beta = J.beta;
    double *C = J.C;
    const long long ldc = J.ldc;
    tile.store([&](int i, int j, double v) {
      if (i0 + i < M && j0 + j < N) {
        double *c = C + (i0 + i) + (long long)(j0 + j) * ldc;
        stg(c, beta != 0.0 ? alpha * v + beta * ldg(c) : alpha * v);
      }
    });
  } else {
    double *P = J.part + (long long)t.w * M * N;
    tile.store([&](int i, int j, double v) {
      if (i0 + i < M && j0 + j < N) stg(P + (i0 + i) + (long long)(j0 + j) * M, v);
    });
  }
  __syncthreads();
}

template <class Tile, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(const GemmJob *jobs, const int4 *tiles,
                                              int ntiles) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  for (int b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const int4 t = tiles[b];
    gemm_tile<Tile, TA, TB>(jobs[t.x], t, lds);
  }
}

__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmJob *jobs, const int *list) {
  const GemmJob J = jobs[list[blockIdx.x]];
  const long long E = (long long)J.M * J.N;
  for (long long e = (long long)blockIdx.y * 256 + threadIdx.x; e < E;
       e += (long long)gridDim.y * 256) {
    double s = 0.0;
    for (int k = 0; k < J.ksplit; k++) s += J.part[(long long)k * E + e];
    const long long i = e % J.M, j = e / J.M;
    double *c = J.C + i + j * J.ldc;
    *c = J.beta != 0.0 ? J.alpha * s + J.beta * *c : J.alpha * s;
  }
}

template <class Tile>
static void launch_gemm(bool ta, bool tb, unsigned grid, hipStream_t s, const GemmJob *J,
                        const int4 *T, int n) {
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<Tile, false, false>), dim3(grid), dim3(256), 0, s, J, T, n);
  if (!ta && tb) hipLaunchKernelGGL((k_gemm<Tile, false, true>), dim3(grid), dim3(256), 0, s, J, T, n);
  if (ta && !tb) hipLaunchKernelGGL((k_gemm<Tile, true, false>), dim3(grid), dim3(256), 0, s, J, T, n);
  if (ta && tb) hipLaunchKernelGGL((k_gemm<Tile, true, true>), dim3(grid), dim3(256), 0, s, J, T, n);
}

static const int kTileTM[GEMM_VARIANTS] = {GTile0::TM, GTile1::TM, GTile2::TM};
static const int kTileTN[GEMM_VARIANTS] = {GTile0::TN, GTile1::TN, GTile2::TN};

int gemm_add(Program &p, Arena &ar, std::vector<GemmJob> jobs, bool ta, bool tb,
             const char *name, int variant) {
  if (variant == GEMM_BY_SHAPE) {
    int maxm = 0, maxn = 0;
    for (const GemmJob &J : jobs) {
      maxm = std::max(maxm, J.M);
      maxn = std::max(maxn, J.N);
    }
    variant = maxn <= 16 ? (maxm <= 16 ? 2 : 1) : (maxm <= 16 ? 2 : 0);
  }
  std::vector<int> split, live;
  long long max_e = 0;
  for (size_t q = 0; q < jobs.size(); q++) {
    GemmJob &J = jobs[q];
    if (J.M <= 0 || J.N <= 0) continue;
    live.push_back((int)q);
    const int mt = (J.M + 63) / 64, nt = (J.N + 63) / 64;
    int ks = 1;
    if (J.K >= 4096 && mt * nt < 128) ks = std::min(64, J.K / 2048);
    J.kseg = ks > 1 ? ((J.K + ks - 1) / ks + 15) / 16 * 16 : std::max(J.K, 1);
    ks = std::max(1, (J.K + J.kseg - 1) / J.kseg);
    J.ksplit = ks;
    J.part = nullptr;
    if (ks > 1) {
      J.part = ar.take<double>((size_t)ks * J.M * J.N);
      split.push_back((int)q);
      max_e = std::max(max_e, (long long)J.M * J.N);
      p.touch(J.part, (size_t)ks * J.M * J.N * sizeof(double), name);
    }
    if (J.K > 0) {
      const size_t abytes = ta ? ((size_t)(J.M - 1) * J.lda + J.K) : ((size_t)(J.K - 1) * J.lda + J.M);
      const size_t bbytes = tb ? ((size_t)(J.K - 1) * J.ldb + J.N) : ((size_t)(J.N - 1) * J.ldb + J.K);
      p.touch(J.A, abytes * sizeof(double), name);
      p.touch(J.B, bbytes * sizeof(double), name);
    }
    p.touch(J.C, ((size_t)(J.N - 1) * J.ldc + J.M) * sizeof(double), name);
  }
  if (live.empty()) return SUCCESS;
  const size_t joff = p.push(jobs);
  // tile lists only for the variants that can run: all of them when the op
  // is tunable, else the one asked for
  const int fixed = variant < 0 ? -1 : std::min(variant, GEMM_VARIANTS - 1);
  size_t toff[GEMM_VARIANTS] = {0, 0, 0};
  int ntile[GEMM_VARIANTS] = {0, 0, 0};
  for (int v = 0; v < GEMM_VARIANTS; v++) {
    if (fixed >= 0 && v != fixed) continue;
    std::vector<int4> tiles;
    for (int q : live) {
      const GemmJob &J = jobs[q];
      const int mt = (J.M + kTileTM[v] - 1) / kTileTM[v];
      const int nt = (J.N + kTileTN[v] - 1) / kTileTN[v];
      for (int w = 0; w < J.ksplit; w++)
        for (int a = 0; a < mt; a++)
          for (int b = 0; b < nt; b++) tiles.push_back(make_int4(q, a, b, w));
    }
    toff[v] = p.push(tiles);
    ntile[v] = (int)tiles.size();
  }
  auto choice = std::make_shared<int>(fixed < 0 ? 0 : fixed);
  if (variant < 0) p.tunables.push_back({p.ops.size(), GEMM_VARIANTS, choice});
  const size_t t0 = toff[0], t1 = toff[1], t2 = toff[2];
  const int n0 = ntile[0], n1 = ntile[1], n2 = ntile[2];
  p.add(name, [=](hipStream_t s, const unsigned char *b) {
    const GemmJob *J = reinterpret_cast<const GemmJob *>(b + joff);
    const int v = *choice;
    const int n = v == 0 ? n0 : (v == 1 ? n1 : n2);
    const int4 *T = reinterpret_cast<const int4 *>(b + (v == 0 ? t0 : (v == 1 ? t1 : t2)));
    const unsigned grid = (unsigned)std::min(n, GEMM_MAX_GRID);
    if (v == 0) launch_gemm<GTile0>(ta, tb, grid, s, J, T, n);
    if (v == 1) launch_gemm<GTile1>(ta, tb, grid, s, J, T, n);
    if (v == 2) launch_gemm<GTile2>(ta, tb, grid, s, J, T, n);
  });
  if (!split.empty()) {
    const size_t loff = p.push(split);
    const unsigned ns = (unsigned)split.size();
    const unsigned gy = (unsigned)std::min<long long>(64, (max_e + 255) / 256);
    p.add(name, [=](hipStream_t s, const unsigned char *b) {
      hipLaunchKernelGGL(k_gemm_reduce, dim3(ns, gy), dim3(256), 0, s,
                         reinterpret_cast<const GemmJob *>(b + joff),
                         reinterpret_cast<const int *>(b + loff));
    });
  }
  return SUCCESS;
}

template <class Tile>
__global__ __launch_bounds__(256) void k_gemm_chain(const GemmChain *chains,
                                                    const GemmLink *links, const int4 *tiles,
                                                    int ntiles) {
  __shared__ double lds[Tile::LDS_DOUBLES];
  for (int b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const int4 t = tiles[b];
    const GemmChain ch = chains[t.x];
    const int i0 = t.y * Tile::TM, j0 = t.z * Tile::TN;
    Tile tile;
    tile.zero();
    const bool full = i0 + Tile::TM <= ch.M && j0 + Tile::TN <= ch.N;
    for (int q = ch.link0; q < ch.link0 + ch.nlink; q++) {
      const GemmLink lk = links[q];
      // a link whose tile lies inside its operands with K a whole number of
      // chunks (the diagonal leaves of uniform trees): the branch-free,
      // double-buffered interior loop
      if (full && !lk.ta && lk.K > 0 && lk.K % Tile::KC_ == 0) {
        if (lk.tb)
          tile.template run_interior<true>(lds, lk.K, lk.A + i0, lk.lda, lk.B + j0, lk.ldb, 1);
        else
          tile.template run_interior<false>(lds, lk.K, lk.A + i0, lk.lda,
                                            lk.B + (long long)j0 * lk.ldb, 1, lk.ldb);
        continue;
      }
      auto la_n = [&](int i, int k) -> double {
        return (i0 + i < ch.M && k < lk.K) ? ldg(lk.A + (i0 + i) + (long long)k * lk.lda) : 0.0;
      };
      auto la_t = [&](int i, int k) -> double {
        return (i0 + i < ch.M && k < lk.K) ? ldg(lk.A + k + (long long)(i0 + i) * lk.lda) : 0.0;
      };
      if (lk.tb) {
        auto lb = [&](int k, int j) -> double {
          return (k < lk.K && j0 + j < ch.N) ? ldg(lk.B + (j0 + j) + (long long)k * lk.ldb) : 0.0;
        };
        if (lk.ta)
          tile.template run<false, true>(lds, lk.K, la_t, lb);
        else
          tile.template run<true, true>(lds, lk.K, la_n, lb);
      } else {
        auto lb = [&](int k, int j) -> double {
          return (k < lk.K && j0 + j < ch.N) ? ldg(lk.B + k + (long long)(j0 + j) * lk.ldb) : 0.0;
        };
        if (lk.ta)
          tile.template run<false, false>(lds, lk.K, la_t, lb);
        else
          tile.template run<true, false>(lds, lk.K, la_n, lb);
      }
    }
    if (ch.tc)
      tile.store([&](int i, int j, double v) {
        if (i0 + i < ch.M && j0 + j < ch.N) stg(ch.C + (j0 + j) + (long long)(i0 + i) * ch.ldc, v);
      });
    else
      tile.store([&](int i, int j, double v) {
        if (i0 + i < ch.M && j0 + j < ch.N) stg(ch.C + (i0 + i) + (long long)(j0 + j) * ch.ldc, v);
      });
    __syncthreads();
  }
}

int gemm_chain_add(Program &p, const std::vector<GemmChain> &chains,
                   const std::vector<GemmLink> &links, const char *name, int variant) {
  std::vector<int> live;
  for (size_t q = 0; q < chains.size(); q++) {
    const GemmChain &c = chains[q];
    if (c.M <= 0 || c.N <= 0) continue;
    live.push_back((int)q);
    const size_t cb = c.tc ? ((size_t)(c.M - 1) * c.ldc + c.N) : ((size_t)(c.N - 1) * c.ldc + c.M);
    p.touch(c.C, cb * sizeof(double), name);
    for (int i = c.link0; i < c.link0 + c.nlink; i++) {
      const GemmLink &l = links[i];
      if (l.K <= 0) continue;
      const size_t ab = l.ta ? ((size_t)(c.M - 1) * l.lda + l.K) : ((size_t)(l.K - 1) * l.lda + c.M);
      const size_t bb = l.tb ? ((size_t)(l.K - 1) * l.ldb + c.N) : ((size_t)(c.N - 1) * l.ldb + l.K);
      p.touch(l.A, ab * sizeof(double), name);
      p.touch(l.B, bb * sizeof(double), name);
    }
  }
  if (live.empty()) return SUCCESS;
  const size_t coff = p.push(chains), loff = p.push(links);
  // tile lists only for the variants that can run: all of them when the op
  // is tunable, else the one asked for
  const int fixed = variant < 0 ? -1 : std::min(variant, GEMM_VARIANTS - 1);
  size_t toff[GEMM_VARIANTS] = {0, 0, 0};
  int ntile[GEMM_VARIANTS] = {0, 0, 0};
  for (int v = 0; v < GEMM_VARIANTS; v++) {
    if (fixed >= 0 && v != fixed) continue;
    std::vector<int4> tiles;
    for (int q : live) {
      const GemmChain &c = chains[q];
      for (int a = 0; a < (c.M + kTileTM[v] - 1) / kTileTM[v]; a++)
        for (int b = 0; b < (c.N + kTileTN[v] - 1) / kTileTN[v]; b++)
          tiles.push_back(make_int4(q, a, b, 0));
    }
    toff[v] = p.push(tiles);
    ntile[v] = (int)tiles.size();
  }
  auto choice = std::make_shared<int>(fixed < 0 ? 0 : fixed);
  if (variant < 0) p.tunables.push_back({p.ops.size(), GEMM_VARIANTS, choice});
  const size_t t0 = toff[0], t1 = toff[1], t2 = toff[2];
  const int n0 = ntile[0], n1 = ntile[1], n2 = ntile[2];
  p.add(name, [=](hipStream_t s, const unsigned char *b) {
    const GemmChain *C = reinterpret_cast<const GemmChain *>(b + coff);
    const GemmLink *L = reinterpret_cast<const GemmLink *>(b + loff);
    const int v = *choice;
    const int n = v == 0 ? n0 : (v == 1 ? n1 : n2);
    const int4 *T = reinterpret_cast<const int4 *>(b + (v == 0 ? t0 : (v == 1 ? t1 : t2)));
    const unsigned grid = (unsigned)std::min(n, GEMM_MAX_GRID);
    if (v == 0) hipLaunchKernelGGL(k_gemm_chain<GTile0>, dim3(grid), dim3(256), 0, s, C, L, T, n);
    if (v == 1) hipLaunchKernelGGL(k_gemm_chain<GTile1>, dim3(grid), dim3(256), 0, s, C, L, T, n);
    if (v == 2) hipLaunchKernelGGL(k_gemm_chain<GTile2>, dim3(grid), dim3(256), 0, s, C, L, T, n);
  });
  return SUCCESS;
}

int program_tune(Program &p, hipStream_t s) {
  if (p.tunables.empty()) return SUCCESS;
  // HM_TUNE_VERBOSE=1 prints every variant's time to stderr
  const char *ev = getenv("HM_TUNE_VERBOSE");
  const bool verbose = ev && ev[0] == '1';
  hipEvent_t e0, e1;
  HM_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("program_tune: event creation failed");
    return ALLOCATION_FAILURE;
  }
  int rc = SUCCESS;
  for (const Tunable &t : p.tunables) {
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < t.nvariants && rc == SUCCESS; v++) {
      *t.choice = v;
      p.ops[t.op](s, p.d_blob);
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < 3; r++) p.ops[t.op](s, p.d_blob);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
        set_error("program_tune: variant %d of op %s failed", v, p.names[t.op].c_str());
        rc = ALLOCATION_FAILURE;
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (verbose)
        fprintf(stderr, "program_tune: %s variant %d: %.1f us\n", p.names[t.op].c_str(), v,
                ms * 1e3f / 3);
      if (ms < best) {
        best = ms;
        bestv = v;
      }
    }
    *t.choice = bestv;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// -------------------------------------------------------- copy, norms
__global__ __launch_bounds__(256) void k_copy2d(const Copy2D *items) {
  const Copy2D it = items[blockIdx.x];
  const long long E = (long long)it.rows * it.cols;
  for (long long e = (long long)blockIdx.y * 256 + threadIdx.x; e < E;
       e += (long long)gridDim.y * 256) {
    const long long i = e % it.rows, j = e / it.rows;
    const double v = it.src ? it.src[i + j * it.lds] : 0.0;
    if (it.trans)
      it.dst[j + i * it.ldd] = v;
    else
      it.dst[i + j * it.ldd] = v;
  }
}

int copy2d_add(Program &p, const std::vector<Copy2D> &items) {
  std::vector<Copy2D> v;
  for (const auto &c : items) {
    if (c.rows <= 0 || c.cols <= 0) continue;
    v.push_back(c);
    if (c.src)
      p.touch(c.src, ((size_t)(c.cols - 1) * c.lds + c.rows) * sizeof(double), "copy2d src");
    if (c.trans)
      p.touch(c.dst, ((size_t)(c.rows - 1) * c.ldd + c.cols) * sizeof(double), "copy2d dst");
    else
      p.touch(c.dst, ((size_t)(c.cols - 1) * c.ldd + c.rows) * sizeof(double), "copy2d dst");
  }
  if (v.empty()) return SUCCESS;
  const size_t off = p.push(v);
  const unsigned n = (unsigned)v.size();
  p.add("copy2d", [=](hipStream_t s, const unsigned char *b) {
    hipLaunchKernelGGL(k_copy2d, dim3(n, 32), dim3(256), 0, s,
                       reinterpret_cast<const Copy2D *>(b + off));
  });
  return SUCCESS;
}

__device__ __forceinline__ double block_norm2(const double *a, int rows, double *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double ss = 0.0;
  for (int i = threadIdx.x; i < rows; i += 256) ss += a[i] * a[i];
  ss = wave_sum(ss);
  __syncthreads();
  if (lane == 0) red[w] = ss;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_colnorm(const ColNorm *items) {
  __shared__ double red[4];
  const ColNorm it = items[blockIdx.x];
  double mx = 0.0;
  for (int c = 0; c < it.cols; c++) {
    const double na = sqrt(block_norm2(it.a + (long long)c * it.ld, it.rows, red));
    if (it.b) {
      const double nb = sqrt(block_norm2(it.b + (long long)c * it.ldb, it.b_rows, red));
      if (nb > 0.0) mx = fmax(mx, na / nb);
    } else {
      mx = fmax(mx, na);
    }
  }
  if (threadIdx.x == 0) *it.out = mx;
}

int colnorm_add(Program &p, const std::vector<ColNorm> &items) {
  std::vector<ColNorm> v;
  for (const auto &c : items) {
    v.push_back(c);
    if (c.rows > 0 && c.cols > 0)
      p.touch(c.a, ((size_t)(c.cols - 1) * c.ld + c.rows) * sizeof(double), "colnorm");
    if (c.b && c.b_rows > 0 && c.cols > 0)
      p.touch(c.b, ((size_t)(c.cols - 1) * c.ldb + c.b_rows) * sizeof(double), "colnorm den");
    p.touch(c.out, sizeof(double), "colnorm out");
  }
  if (v.empty()) return SUCCESS;
  const size_t off = p.push(v);
  const unsigned n = (unsigned)v.size();
  p.add("colnorm", [=](hipStream_t s, const unsigned char *b) {
    hipLaunchKernelGGL(k_colnorm, dim3(n), dim3(256), 0, s,
                       reinterpret_cast<const ColNorm *>(b + off));
  });
  return SUCCESS;
}

}  // namespace hm
