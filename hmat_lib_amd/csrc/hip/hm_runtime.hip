// Runtime services of libhmat_hip: device selection, error text, memory
// helpers and optional HIP-event timing of named kernels.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "hm_common.h"
#include "hm_tsqr.h"

namespace hm {

static thread_local std::string g_last_error;
static std::mutex g_mu;
static int g_inited = 0;
static int g_device = 0;
static hipStream_t g_stream = nullptr;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const char *hip_name(hipError_t e) { return hipGetErrorString(e); }

int ensure_init() {
  if (g_inited) return SUCCESS;
  return hm_hip_init(0);
}

hipStream_t stream_of(void *s) {
  return s ? reinterpret_cast<hipStream_t>(s) : g_stream;
}

static bool timing_on();
int dmalloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  const bool tm = timing_on();
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(p, bytes);
  if (tm)
    timing_host("hipMalloc (host)", std::chrono::duration<double, std::milli>(
                                        std::chrono::steady_clock::now() - t0).count());
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

void dfree(void *p) {
  if (!p) return;
  const bool tm = timing_on();
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipFree(p);
  if (tm)
    timing_host("hipFree (host)", std::chrono::duration<double, std::milli>(
                                      std::chrono::steady_clock::now() - t0).count());
}

// ---------------------------------------------------------------- timing
struct TimingSlot {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
  bool pending = false;
};
struct TimingAcc {
  double total_ms = 0.0;
  long count = 0;
};
static bool g_timing = false;
static bool timing_on() { return g_timing; }
static std::string g_timing_only;  // empty: time every named kernel
static std::vector<TimingSlot> g_slots;
static std::map<std::string, TimingAcc> g_acc;

static void timing_drain_locked() {
  for (auto &sl : g_slots) {
    if (!sl.pending) continue;
    (void)hipEventSynchronize(sl.b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, sl.a, sl.b) == hipSuccess) {
      auto &acc = g_acc[sl.name];
      acc.total_ms += ms;
      acc.count += 1;
    }
    sl.pending = false;
  }
}

Timed timing_begin(const char *name, hipStream_t s) {
  Timed t;
  if (!g_timing) return t;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_timing_only.empty() && g_timing_only != name) return t;
  int slot = -1;
  for (size_t i = 0; i < g_slots.size(); i++)
    if (!g_slots[i].pending) { slot = (int)i; break; }
  if (slot < 0) {
    if (g_slots.size() >= 4096) timing_drain_locked();
    for (size_t i = 0; i < g_slots.size(); i++)
      if (!g_slots[i].pending) { slot = (int)i; break; }
  }
  if (slot < 0) {
    TimingSlot ns;
    if (hipEventCreate(&ns.a) != hipSuccess || hipEventCreate(&ns.b) != hipSuccess)
      return t;
    g_slots.push_back(ns);
    slot = (int)g_slots.size() - 1;
  }
  g_slots[slot].name = name;
  g_slots[slot].pending = true;
  (void)hipEventRecord(g_slots[slot].a, s);
  t.slot = slot;
  return t;
}

void timing_end(const Timed &t, hipStream_t s) {
  if (t.slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  (void)hipEventRecord(g_slots[t.slot].b, s);
}

void timing_host(const char *name, double ms) {
  if (!g_timing) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_timing_only.empty() && g_timing_only != name) return;
  auto &acc = g_acc[name];
  acc.total_ms += ms;
  acc.count += 1;
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_inited && device == g_device) return SUCCESS;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    set_error("hm_hip_init: no HIP device available (%s)",
              e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  if (device < 0 || device >= count) {
    set_error("hm_hip_init: device %d out of range (%d devices)", device, count);
    return INPUT_ERROR;
  }
  HM_HIP(hipSetDevice(device));
  if (!g_stream) HM_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  g_inited = 1;
  return SUCCESS;
}

int hm_hip_release_buffers(void) {
  const int a = release_exec_buffers();
  const int b = release_round_buffer();
  if (a != SUCCESS || b != SUCCESS) {
    set_error("hm_hip_release_buffers: a call on another thread holds a kept buffer");
    return INPUT_ERROR;
  }
  return SUCCESS;
}

int hm_hip_device_count(int *count) {
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return ALLOCATION_FAILURE;
  }
  return SUCCESS;
}

const char *hm_hip_last_error(void) { return g_last_error.c_str(); }

int hm_hip_synchronize(void *stream) {
  HM_OK(ensure_init());
  if (stream) {
    HM_HIP(hipStreamSynchronize(stream_of(stream)));
  } else {
    HM_HIP(hipDeviceSynchronize());
  }
  return SUCCESS;
}

int hm_hip_malloc(void **d_ptr, size_t bytes) {
  HM_OK(ensure_init());
  return dmalloc(d_ptr, bytes);
}

int hm_hip_free(void *d_ptr) {
  if (!d_ptr) return SUCCESS;
  HM_HIP(hipFree(d_ptr));
  return SUCCESS;
}

int hm_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes, void *stream) {
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  HM_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, s));
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

int hm_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes, void *stream) {
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  HM_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, s));
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

int hm_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream) {
  HM_OK(ensure_init());
  HM_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice,
                        stream_of(stream)));
  return SUCCESS;
}

int hm_hip_copy2d_h2d(double *d_dst, long ld_dst, const double *src,
                      long ld_src, int m, int n, void *stream) {
  HM_OK(ensure_init());
  if (m <= 0 || n <= 0) return SUCCESS;
  hipStream_t s = stream_of(stream);
  HM_HIP(hipMemcpy2DAsync(d_dst, ld_dst * sizeof(double), src,
                          ld_src * sizeof(double), m * sizeof(double), n,
                          hipMemcpyHostToDevice, s));
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

int hm_hip_copy2d_d2h(double *dst, long ld_dst, const double *d_src,
                      long ld_src, int m, int n, void *stream) {
  HM_OK(ensure_init());
  if (m <= 0 || n <= 0) return SUCCESS;
  hipStream_t s = stream_of(stream);
  HM_HIP(hipMemcpy2DAsync(dst, ld_dst * sizeof(double), d_src,
                          ld_src * sizeof(double), m * sizeof(double), n,
                          hipMemcpyDeviceToHost, s));
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

int hm_hip_kernel_timing(int enable) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_timing = enable != 0;
  return SUCCESS;
}

int hm_hip_kernel_timing_select(const char *name) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_timing_only = name ? name : "";
  return SUCCESS;
}

int hm_hip_kernel_timing_read(const char *name, double *total_ms, long *count) {
  std::lock_guard<std::mutex> lk(g_mu);
  timing_drain_locked();
  auto it = g_acc.find(name);
  *total_ms = it == g_acc.end() ? 0.0 : it->second.total_ms;
  *count = it == g_acc.end() ? 0 : it->second.count;
  return SUCCESS;
}

int hm_hip_kernel_timing_dump(char *buf, int len) {
  std::lock_guard<std::mutex> lk(g_mu);
  timing_drain_locked();
  std::string out;
  char line[256];
  for (const auto &kv : g_acc) {
    snprintf(line, sizeof(line), "%s\t%.4f\t%ld\n", kv.first.c_str(), kv.second.total_ms,
             kv.second.count);
    out += line;
  }
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s", out.c_str());
  return (int)out.size();
}

int hm_hip_kernel_timing_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  timing_drain_locked();
  g_acc.clear();
  return SUCCESS;
}

}  // extern "C"
