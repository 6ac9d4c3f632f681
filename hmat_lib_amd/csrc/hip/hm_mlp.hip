// Rank/sketch predictor: a small MLP (nin -> H -> H -> 1, tanh, FP32)
// trained and evaluated by hand-written kernels.
//
// Parameters are one flat FP32 vector, in order
//   W1 (nin x H, row i = input i), b1 (H), W2 (H x H), b2 (H), w3 (H), b3 (1)
// and the loss is the mean squared error over the samples.  Forward and
// backward run one thread per sample with the parameters staged in LDS;
// weight gradients are summed over samples by one thread per parameter in
// a fixed order, so training is deterministic.  Adam updates in place.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "hm_common.h"

struct hm_mlp {
  int nin = 0, nh = 0;
  long long np = 0;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  long long step = 0;
  // scratch for up to cap samples
  int cap = 0;
  float *d_a1 = nullptr, *d_a2 = nullptr, *d_y = nullptr, *d_dz1 = nullptr, *d_dz2 = nullptr,
        *d_dy = nullptr;
  int *d_idx = nullptr;
  float *d_loss = nullptr;
};

namespace hm {

static constexpr int MLP_MAX_H = 64;
static constexpr int MLP_MAX_IN = 64;

struct MlpView {
  int nin, nh;
  const float *W1, *b1, *W2, *b2, *w3, *b3;
};

__device__ __forceinline__ MlpView mlp_view(const float *P, int nin, int nh) {
  MlpView v;
  v.nin = nin;
  v.nh = nh;
  v.W1 = P;
  v.b1 = v.W1 + nin * nh;
  v.W2 = v.b1 + nh;
  v.b2 = v.W2 + nh * nh;
  v.w3 = v.b2 + nh;
  v.b3 = v.w3 + nh;
  return v;
}

__host__ __device__ static inline long long mlp_nparams(int nin, int nh) {
  return (long long)nin * nh + nh + (long long)nh * nh + nh + nh + 1;
}

// Forward for samples idx[0..B) (idx NULL: 0..B), rows of X (ld = nin).
// Stores the activations when A1 != NULL (row b of A1/A2 = sample idx[b]).
__global__ __launch_bounds__(256) void k_mlp_fwd(const float *P, int nin, int nh,
                                                 const float *X, const int *idx, int B,
                                                 float *A1, float *A2, float *y) {
  extern __shared__ float sp[];
  const long long np = mlp_nparams(nin, nh);
  for (long long e = threadIdx.x; e < np; e += blockDim.x) sp[e] = P[e];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const MlpView w = mlp_view(sp, nin, nh);
  const int r = idx ? idx[b] : b;
  const float *x = X + (long long)r * nin;
  float a1[MLP_MAX_H], a2[MLP_MAX_H];
  for (int j = 0; j < nh; j++) a1[j] = w.b1[j];
  for (int i = 0; i < nin; i++) {
    const float xi = x[i];
    for (int j = 0; j < nh; j++) a1[j] += xi * w.W1[i * nh + j];
  }
  for (int j = 0; j < nh; j++) a1[j] = tanhf(a1[j]);
  for (int j = 0; j < nh; j++) a2[j] = w.b2[j];
  for (int i = 0; i < nh; i++) {
    const float ai = a1[i];
    for (int j = 0; j < nh; j++) a2[j] += ai * w.W2[i * nh + j];
  }
  float out = w.b3[0];
  for (int j = 0; j < nh; j++) {
    a2[j] = tanhf(a2[j]);
    out += a2[j] * w.w3[j];
  }
  y[b] = out;
  if (A1) {
    for (int j = 0; j < nh; j++) {
      A1[(long long)b * nh + j] = a1[j];
      A2[(long long)b * nh + j] = a2[j];
    }
  }
}

// Per-sample backward: dy = 2 (y - t) / B, dZ2 = dy w3 o (1 - a2^2),
// dZ1 = (W2 dZ2) o (1 - a1^2); squared error per sample into sq.
__global__ __launch_bounds__(256) void k_mlp_bwd(const float *P, int nin, int nh,
                                                 const float *T, const int *idx, int B,
                                                 const float *A1, const float *A2,
                                                 const float *y, float *dZ1, float *dZ2,
                                                 float *dy, float *sq) {
  extern __shared__ float sp[];
  const long long np = mlp_nparams(nin, nh);
  for (long long e = threadIdx.x; e < np; e += blockDim.x) sp[e] = P[e];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const MlpView w = mlp_view(sp, nin, nh);
  const int r = idx ? idx[b] : b;
  const float err = y[b] - T[r];
  const float g = 2.0f * err / (float)B;
  dy[b] = g;
  sq[b] = err * err;
  float d2[MLP_MAX_H];
  for (int j = 0; j < nh; j++) {
    const float a = A2[(long long)b * nh + j];
    d2[j] = g * w.w3[j] * (1.0f - a * a);
    dZ2[(long long)b * nh + j] = d2[j];
  }
  for (int i = 0; i < nh; i++) {
    float s = 0.0f;
    for (int j = 0; j < nh; j++) s += w.W2[i * nh + j] * d2[j];
    const float a = A1[(long long)b * nh + i];
    dZ1[(long long)b * nh + i] = s * (1.0f - a * a);
  }
}

// Gradient of every parameter: one thread per parameter, samples summed in
// order.  Element sq[B] of the loss buffer receives the mean squared error.
__global__ __launch_bounds__(256) void k_mlp_grad(int nin, int nh, const float *X,
                                                  const int *idx, int B, const float *A1,
                                                  const float *A2, const float *dZ1,
                                                  const float *dZ2, const float *dy,
                                                  float *G, float *sq) {
  const long long np = mlp_nparams(nin, nh);
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p == np) {  // the loss
    float s = 0.0f;
    for (int b = 0; b < B; b++) s += sq[b];
    sq[B] = s / (float)B;
    return;
  }
  if (p > np) return;
  const long long oW1 = 0, ob1 = oW1 + (long long)nin * nh, oW2 = ob1 + nh,
                  ob2 = oW2 + (long long)nh * nh, ow3 = ob2 + nh, ob3 = ow3 + nh;
  float s = 0.0f;
  if (p < ob1) {
    const int i = (int)(p / nh), j = (int)(p % nh);
    for (int b = 0; b < B; b++) {
      const int r = idx ? idx[b] : b;
      s += X[(long long)r * nin + i] * dZ1[(long long)b * nh + j];
    }
  } else if (p < oW2) {
    const int j = (int)(p - ob1);
    for (int b = 0; b < B; b++) s += dZ1[(long long)b * nh + j];
  } else if (p < ob2) {
    const long long q = p - oW2;
    const int i = (int)(q / nh), j = (int)(q % nh);
    for (int b = 0; b < B; b++) s += A1[(long long)b * nh + i] * dZ2[(long long)b * nh + j];
  } else if (p < ow3) {
    const int j = (int)(p - ob2);
    for (int b = 0; b < B; b++) s += dZ2[(long long)b * nh + j];
  } else if (p < ob3) {
    const int j = (int)(p - ow3);
    for (int b = 0; b < B; b++) s += A2[(long long)b * nh + j] * dy[b];
  } else {
    for (int b = 0; b < B; b++) s += dy[b];
  }
  G[p] = s;
}

__global__ __launch_bounds__(256) void k_adam(float *P, const float *G, float *Mo, float *Vo,
                                              long long np, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float g = G[p];
  const float m = b1 * Mo[p] + (1.0f - b1) * g;
  const float v = b2 * Vo[p] + (1.0f - b2) * g * g;
  Mo[p] = m;
  Vo[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

static int mlp_reserve(hm_mlp *m, int B) {
  if (B <= m->cap) return SUCCESS;
  for (float *p : {m->d_a1, m->d_a2, m->d_y, m->d_dz1, m->d_dz2, m->d_dy, m->d_loss}) dfree(p);
  dfree(m->d_idx);
  m->d_a1 = m->d_a2 = m->d_y = m->d_dz1 = m->d_dz2 = m->d_dy = m->d_loss = nullptr;
  m->d_idx = nullptr;
  m->cap = 0;
  const size_t bh = (size_t)B * m->nh;
  HM_OK(dalloc(&m->d_a1, bh));
  HM_OK(dalloc(&m->d_a2, bh));
  HM_OK(dalloc(&m->d_dz1, bh));
  HM_OK(dalloc(&m->d_dz2, bh));
  HM_OK(dalloc(&m->d_y, (size_t)B));
  HM_OK(dalloc(&m->d_dy, (size_t)B));
  HM_OK(dalloc(&m->d_loss, (size_t)B + 1));
  HM_OK(dalloc(&m->d_idx, (size_t)B));
  m->cap = B;
  return SUCCESS;
}

static size_t mlp_lds(const hm_mlp *m) { return (size_t)m->np * sizeof(float); }

// forward + backward + gradient for samples idx (or 0..B), loss into d_loss[B]
static void mlp_grad_launch(hm_mlp *m, const float *X, const float *T, const int *idx, int B,
                            hipStream_t s) {
  const unsigned g = (unsigned)((B + 255) / 256);
  hipLaunchKernelGGL(k_mlp_fwd, dim3(g), dim3(256), mlp_lds(m), s, m->d_p, m->nin, m->nh, X,
                     idx, B, m->d_a1, m->d_a2, m->d_y);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(g), dim3(256), mlp_lds(m), s, m->d_p, m->nin, m->nh, T,
                     idx, B, m->d_a1, m->d_a2, m->d_y, m->d_dz1, m->d_dz2, m->d_dy, m->d_loss);
  const unsigned gp = (unsigned)((m->np + 1 + 255) / 256);
  hipLaunchKernelGGL(k_mlp_grad, dim3(gp), dim3(256), 0, s, m->nin, m->nh, X, idx, B,
                     m->d_a1, m->d_a2, m->d_dz1, m->d_dz2, m->d_dy, m->d_g, m->d_loss);
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_mlp_create(int nin, int nhidden, unsigned long long seed, hm_mlp **out) {
  HM_OK(ensure_init());
  if (!out || nin < 1 || nin > MLP_MAX_IN || nhidden < 1 || nhidden > MLP_MAX_H) {
    set_error("hm_hip_mlp_create: sizes must be 1..%d inputs, 1..%d hidden", MLP_MAX_IN,
              MLP_MAX_H);
    return INPUT_ERROR;
  }
  *out = nullptr;
  hm_mlp *m = new hm_mlp();
  m->nin = nin;
  m->nh = nhidden;
  m->np = mlp_nparams(nin, nhidden);
  // Glorot-uniform weights, zero biases (host generator, seeded)
  std::vector<float> p((size_t)m->np, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, int fan_in, int fan_out, size_t count) {
    const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (size_t e = 0; e < count; e++) p[off + e] = u(rng);
  };
  const size_t oW1 = 0, oW2 = (size_t)nin * nhidden + nhidden,
               ow3 = oW2 + (size_t)nhidden * nhidden + nhidden;
  fill(oW1, nin, nhidden, (size_t)nin * nhidden);
  fill(oW2, nhidden, nhidden, (size_t)nhidden * nhidden);
  fill(ow3, nhidden, 1, (size_t)nhidden);
  int rc = dalloc(&m->d_p, (size_t)m->np);
  if (rc == SUCCESS) rc = dalloc(&m->d_g, (size_t)m->np);
  if (rc == SUCCESS) rc = dalloc(&m->d_m, (size_t)m->np);
  if (rc == SUCCESS) rc = dalloc(&m->d_v, (size_t)m->np);
  if (rc == SUCCESS) {
    hipStream_t s = stream_of(nullptr);
    std::vector<float> z((size_t)m->np, 0.0f);
    hipError_t e = hipMemcpyAsync(m->d_p, p.data(), p.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(m->d_m, z.data(), z.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(m->d_v, z.data(), z.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_mlp_create: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  if (rc != SUCCESS) {
    dfree(m->d_p);
    dfree(m->d_g);
    dfree(m->d_m);
    dfree(m->d_v);
    delete m;
    return rc;
  }
  *out = m;
  return SUCCESS;
}

int hm_hip_mlp_destroy(hm_mlp *m) {
  if (!m) return SUCCESS;
  (void)hipStreamSynchronize(stream_of(nullptr));
  for (float *p : {m->d_p, m->d_g, m->d_m, m->d_v, m->d_a1, m->d_a2, m->d_y, m->d_dz1,
                   m->d_dz2, m->d_dy, m->d_loss})
    dfree(p);
  dfree(m->d_idx);
  delete m;
  return SUCCESS;
}

long long hm_hip_mlp_nparams(const hm_mlp *m) { return m ? m->np : -1; }

int hm_hip_mlp_get_params(const hm_mlp *m, float *host) {
  if (!m || !host) return INPUT_ERROR;
  HM_HIP(hipMemcpy(host, m->d_p, (size_t)m->np * 4, hipMemcpyDeviceToHost));
  return SUCCESS;
}

int hm_hip_mlp_set_params(hm_mlp *m, const float *host) {
  if (!m || !host) return INPUT_ERROR;
  HM_HIP(hipMemcpy(m->d_p, host, (size_t)m->np * 4, hipMemcpyHostToDevice));
  return SUCCESS;
}

int hm_hip_mlp_forward(hm_mlp *m, const float *d_X, int B, float *d_y, void *stream) {
  if (!m || (B > 0 && (!d_X || !d_y))) return INPUT_ERROR;
  if (B <= 0) return SUCCESS;
  hipStream_t s = stream_of(stream);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((B + 255) / 256), dim3(256), mlp_lds(m), s, m->d_p,
                     m->nin, m->nh, d_X, (const int *)nullptr, B, (float *)nullptr,
                     (float *)nullptr, d_y);
  HM_LAUNCH_CHECK();
  return SUCCESS;
}

int hm_hip_mlp_gradient(hm_mlp *m, const float *d_X, const float *d_t, int B, float *h_grad,
                        float *h_loss, void *stream) {
  if (!m || B <= 0 || !d_X || !d_t) return INPUT_ERROR;
  HM_OK(mlp_reserve(m, B));
  hipStream_t s = stream_of(stream);
  mlp_grad_launch(m, d_X, d_t, nullptr, B, s);
  HM_LAUNCH_CHECK();
  if (h_grad)
    HM_HIP(hipMemcpyAsync(h_grad, m->d_g, (size_t)m->np * 4, hipMemcpyDeviceToHost, s));
  if (h_loss) HM_HIP(hipMemcpyAsync(h_loss, m->d_loss + B, 4, hipMemcpyDeviceToHost, s));
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

// Mini-batch Adam over `epochs` passes of the B samples (order reshuffled
// per epoch by a seeded host generator).  loss_out[e] = mean loss of the
// mini-batches of epoch e.
int hm_hip_mlp_train(hm_mlp *m, const float *d_X, const float *d_t, int B, int epochs,
                     int batch, float lr, unsigned long long seed, float *loss_out,
                     void *stream) {
  if (!m || B <= 0 || !d_X || !d_t || epochs < 0 || batch < 1) return INPUT_ERROR;
  batch = std::min(batch, B);
  HM_OK(mlp_reserve(m, batch));
  int *d_perm = nullptr;
  HM_OK(dalloc(&d_perm, (size_t)B));
  hipStream_t s = stream_of(stream);
  std::vector<int> perm(B);
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int rc = SUCCESS;
  std::vector<float> lb;
  for (int ep = 0; ep < epochs && rc == SUCCESS; ep++) {
    for (int i = 0; i < B; i++) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    hipError_t e = hipMemcpyAsync(d_perm, perm.data(), (size_t)B * 4, hipMemcpyHostToDevice, s);
    const int nb = B / batch;  // drop the ragged tail
    lb.assign(std::max(nb, 1), 0.0f);
    for (int k = 0; k < nb && e == hipSuccess; k++) {
      mlp_grad_launch(m, d_X, d_t, d_perm + (long long)k * batch, batch, s);
      m->step++;
      const float c1 = 1.0f - std::pow(b1, (float)m->step);
      const float c2 = 1.0f - std::pow(b2, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, s,
                         m->d_p, m->d_g, m->d_m, m->d_v, m->np, lr, b1, b2, eps, c1, c2);
      if (loss_out)
        e = hipMemcpyAsync(&lb[k], m->d_loss + batch, 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_mlp_train: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
      break;
    }
    if (loss_out) {
      double acc = 0.0;
      for (int k = 0; k < nb; k++) acc += lb[k];
      loss_out[ep] = (float)(acc / std::max(nb, 1));
    }
  }
  (void)hipStreamSynchronize(s);
  dfree(d_perm);
  return rc;
}

}  // extern "C"
