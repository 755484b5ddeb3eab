// Per-launch event timing of the main kernel of each call, and a plain
// device copy (the calibration kernel for hardware byte counters).
//
// The call-level timer (tentaco_timer_start/stop, get_gpu_timer_result)
// brackets whole call sequences, as the reference's GPUTimer does
// (gpu_library.h:949-988). This one brackets only the kernel that does the
// work (slice_kernel, edge_kernel, ttv_kernel, batch_kernel): the launch
// code calls ktimer_begin/ktimer_end around it on the launch stream, so a
// roofline figure can divide algorithmic bytes by that kernel's duration
// alone. Off by default; never records into a stream under capture.

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.h"

namespace tt {
namespace {

constexpr int kMaxPairs = 4096;

struct KTimer {
  std::mutex mu;
  bool on = false;
  bool broken = false;
  int n = 0;
  const char* last = "";
  std::vector<hipEvent_t> ev;  // 2 * kMaxPairs, created on first start
};

KTimer& kt() {
  static KTimer t;
  return t;
}

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus s = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &s) != hipSuccess) return true;
  return s != hipStreamCaptureStatusNone;
}

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, int64_t n4) {
  const int64_t nthr = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += nthr)
    dst[i] = src[i];
}

__global__ void copy_tail_kernel(const unsigned char* __restrict__ src,
                                 unsigned char* __restrict__ dst, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

void ktimer_begin(hipStream_t st) {
  KTimer& t = kt();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.on || t.broken || t.n >= kMaxPairs || capturing(st)) return;
  if (hipEventRecord(t.ev[2 * t.n], st) != hipSuccess) t.broken = true;
}

void ktimer_end(hipStream_t st, const char* kernel) {
  KTimer& t = kt();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.on || t.broken || t.n >= kMaxPairs || capturing(st)) return;
  if (hipEventRecord(t.ev[2 * t.n + 1], st) != hipSuccess) {
    t.broken = true;
    return;
  }
  t.last = kernel;
  ++t.n;
}

}  // namespace tt

using namespace tt;

extern "C" {

int tentaco_ktimer_start(void) {
  KTimer& t = kt();
  std::lock_guard<std::mutex> lk(t.mu);
  if (t.ev.empty()) {
    t.ev.resize(2 * kMaxPairs, nullptr);
    for (auto& e : t.ev) {
      if (hipEventCreate(&e) != hipSuccess) {
        for (auto& f : t.ev)
          if (f) (void)hipEventDestroy(f);
        t.ev.clear();
        return fail(TENTACO_EHIP, "kernel timer: hipEventCreate failed");
      }
    }
  }
  t.on = true;
  t.broken = false;
  t.n = 0;
  t.last = "";
  return TENTACO_OK;
}

int tentaco_ktimer_stop(double* total_ms, int* launches, char* kernel, int kernel_len) {
  KTimer& t = kt();
  std::lock_guard<std::mutex> lk(t.mu);
  t.on = false;
  if (t.broken) return fail(TENTACO_EHIP, "kernel timer: an event record failed");
  double sum = 0.0;
  if (t.n > 0) TT_HIP_CHECK(hipEventSynchronize(t.ev[2 * t.n - 1]));
  for (int i = 0; i < t.n; ++i) {
    float ms = 0.f;
    TT_HIP_CHECK(hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]));
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = t.n;
  if (kernel && kernel_len > 0) {
    int i = 0;
    for (; i + 1 < kernel_len && t.last[i]; ++i) kernel[i] = t.last[i];
    kernel[i] = '\0';
  }
  return TENTACO_OK;
}

int tentaco_dev_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TENTACO_OK;
  if (!dst || !src) return fail(TENTACO_EINVAL, "dev_copy: null pointer");
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
    return fail(TENTACO_EINVAL, "dev_copy: pointers must be 16-byte aligned");
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  if ((s < d + bytes) && (d < s + bytes)) return fail(TENTACO_EINVAL, "dev_copy: ranges overlap");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n4 = int64_t(bytes / 16);
  if (n4 > 0) {
    const int64_t want = (n4 + 255) / 256;
    const int grid = int(want < 16384 ? want : 16384);
    ktimer_begin(st);
    copy_kernel<<<grid, 256, 0, st>>>(reinterpret_cast<const float4*>(src),
                                      reinterpret_cast<float4*>(dst), n4);
    TT_HIP_CHECK(hipGetLastError());
    ktimer_end(st, "copy_kernel");
  }
  const int64_t tail = int64_t(bytes % 16);
  if (tail) {
    copy_tail_kernel<<<1, 64, 0, st>>>(reinterpret_cast<const unsigned char*>(s + n4 * 16),
                                       reinterpret_cast<unsigned char*>(d + n4 * 16), tail);
    TT_HIP_CHECK(hipGetLastError());
  }
  return TENTACO_OK;
}

}  // extern "C"
