// Launch programs and the batched kernels they are built from (FP64):
// MFMA GEMM tiles with deterministic split-K, strided copies, column norms.
#include <algorithm>
#include <cstring>

#include "hm_prog.h"

namespace hm {

// ------------------------------------------------------------- program
int Program::audit(std::string *msg) const {
  for (const auto &t : touched) {
    const char *lo = static_cast<const char *>(t.first.lo);
    const char *hi = lo + t.first.bytes;
    bool ok = false;
    for (const auto &a : allowed) {
      const char *alo = static_cast<const char *>(a.lo);
      if (lo >= alo && hi <= alo + a.bytes) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      if (msg) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s touches [%p, %p) outside every buffer",
                 t.second.c_str(), (const void *)lo, (const void *)hi);
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

int Program::run(hipStream_t s) {
  for (size_t i = 0; i < ops.size(); i++) {
    ops[i](s, d_blob);
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
typedef MfmaTile<2, 2, 2, 2, 16> GTile;  // 64 x 64, K-chunk 16

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(const GemmJob *jobs, const int4 *tiles) {
  __shared__ double lds[GTile::LDS_DOUBLES];
  const int4 t = tiles[blockIdx.x];
  const GemmJob J = jobs[t.x];
  const int i0 = t.y * GTile::TM, j0 = t.z * GTile::TN;
  const int kb = t.w * J.kseg;
  const int kl = min(J.K, kb + J.kseg) - kb;
  const double *A = J.A + (TA ? (long long)kb : (long long)kb * J.lda);
  const double *B = J.B + (TB ? (long long)kb * J.ldb : (long long)kb);
  const long long lda = J.lda, ldb = J.ldb;
  const int M = J.M, N = J.N;
  auto la = [&](int i, int k) -> double {
    if (i0 + i >= M || k >= kl) return 0.0;
    return TA ? A[k + (long long)(i0 + i) * lda] : A[(i0 + i) + (long long)k * lda];
  };
  auto lb = [&](int k, int j) -> double {
    if (k >= kl || j0 + j >= N) return 0.0;
    return TB ? B[(j0 + j) + (long long)k * ldb] : B[k + (long long)(j0 + j) * ldb];
  };
  GTile tile;
  tile.zero();
  tile.template run<!TA, TB>(lds, kl, la, lb);
  if (J.ksplit == 1) {
    const double alpha = J.alpha, beta = J.beta;
    double *C = J.C;
    const long long ldc = J.ldc;
    tile.store([&](int i, int j, double v) {
      if (i0 + i < M && j0 + j < N) {
        double *c = C + (i0 + i) + (long long)(j0 + j) * ldc;
        *c = beta != 0.0 ? alpha * v + beta * *c : alpha * v;
      }
    });
  } else {
    double *P = J.part + (long long)t.w * M * N;
    tile.store([&](int i, int j, double v) {
      if (i0 + i < M && j0 + j < N) P[(i0 + i) + (long long)(j0 + j) * M] = v;
    });
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

int gemm_add(Program &p, Arena &ar, std::vector<GemmJob> jobs, bool ta, bool tb,
             const char *name) {
  std::vector<int4> tiles;
  std::vector<int> split;
  long long max_e = 0;
  for (size_t q = 0; q < jobs.size(); q++) {
    GemmJob &J = jobs[q];
    if (J.M <= 0 || J.N <= 0) continue;
    const int mt = (J.M + GTile::TM - 1) / GTile::TM;
    const int nt = (J.N + GTile::TN - 1) / GTile::TN;
    int ks = 1;
    if (J.K >= 4096 && mt * nt < 128) ks = std::min(64, J.K / 2048);
    J.ksplit = ks;
    J.kseg = ks > 1 ? ((J.K + ks - 1) / ks + 15) / 16 * 16 : std::max(J.K, 1);
    ks = (J.K + J.kseg - 1) / J.kseg;
    if (ks < 1) ks = 1;
    J.ksplit = ks;
    J.part = nullptr;
    if (ks > 1) {
      J.part = ar.take<double>((size_t)ks * J.M * J.N);
      split.push_back((int)q);
      max_e = std::max(max_e, (long long)J.M * J.N);
      p.touch(J.part, (size_t)ks * J.M * J.N * sizeof(double), name);
    }
    for (int w = 0; w < ks; w++)
      for (int a = 0; a < mt; a++)
        for (int b = 0; b < nt; b++) tiles.push_back(make_int4((int)q, a, b, w));
    // ranges read and written
    if (J.K > 0) {
      const size_t abytes = ta ? ((size_t)(J.M - 1) * J.lda + J.K) : ((size_t)(J.K - 1) * J.lda + J.M);
      const size_t bbytes = tb ? ((size_t)(J.K - 1) * J.ldb + J.N) : ((size_t)(J.N - 1) * J.ldb + J.K);
      p.touch(J.A, abytes * sizeof(double), name);
      p.touch(J.B, bbytes * sizeof(double), name);
    }
    p.touch(J.C, ((size_t)(J.N - 1) * J.ldc + J.M) * sizeof(double), name);
  }
  if (tiles.empty()) return SUCCESS;
  const size_t joff = p.push(jobs), toff = p.push(tiles);
  const unsigned nt = (unsigned)tiles.size();
  p.add(name, [=](hipStream_t s, const unsigned char *b) {
    const GemmJob *J = reinterpret_cast<const GemmJob *>(b + joff);
    const int4 *T = reinterpret_cast<const int4 *>(b + toff);
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), dim3(nt), dim3(256), 0, s, J, T);
    if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), dim3(nt), dim3(256), 0, s, J, T);
    if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), dim3(nt), dim3(256), 0, s, J, T);
    if (ta && tb) hipLaunchKernelGGL((k_gemm<true, true>), dim3(nt), dim3(256), 0, s, J, T);
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

__global__ __launch_bounds__(256) void k_colnorm(const ColNorm *items) {
  __shared__ double red[4];
  const ColNorm it = items[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double mx = 0.0;
  for (int c = 0; c < it.cols; c++) {
    double ss = 0.0;
    for (int i = threadIdx.x; i < it.rows; i += 256) {
      const double x = it.a[i + (long long)c * it.ld];
      ss += x * x;
    }
    ss = wave_sum(ss);
    __syncthreads();
    if (lane == 0) red[w] = ss;
    __syncthreads();
    mx = fmax(mx, sqrt(red[0] + red[1] + red[2] + red[3]));
  }
  if (threadIdx.x == 0) *it.out = mx;
}

int colnorm_add(Program &p, const std::vector<ColNorm> &items) {
  std::vector<ColNorm> v;
  for (const auto &c : items) {
    v.push_back(c);
    if (c.rows > 0 && c.cols > 0)
      p.touch(c.a, ((size_t)(c.cols - 1) * c.ld + c.rows) * sizeof(double), "colnorm");
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
