// Plain device copy used to calibrate traffic counters and to measure the
// achievable HBM bandwidth on the box (no counterpart in the reference).
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace {
__global__ void copy_probe_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gs) dst[i] = src[i];
}
}  // namespace

extern "C" int dawn_probe_copy(int64_t bytes, int32_t iters, double* gbps) {
  if (bytes < 16 || iters < 1 || !gbps) return DAWN_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return DAWN_ERR_NO_DEVICE;
  }
  const int64_t n = bytes / 16;
  dawn::DevBuf<uint4> a, b;
  DAWN_TRY(a.alloc((size_t)n));
  DAWN_TRY(b.alloc((size_t)n));
  DAWN_HIP(hipMemset(a.p, 1, (size_t)n * 16));
  hipEvent_t e0, e1;
  DAWN_HIP(hipEventCreate(&e0));
  DAWN_HIP(hipEventCreate(&e1));
  const unsigned grid = dawn::grid_for(n, 256, 8192);
  copy_probe_kernel<<<grid, 256>>>(a.p, b.p, n);  // warm-up
  DAWN_HIP(hipEventRecord(e0));
  for (int i = 0; i < iters; i++) copy_probe_kernel<<<grid, 256>>>(a.p, b.p, n);
  DAWN_HIP(hipEventRecord(e1));
  DAWN_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DAWN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 2.0 * (double)n * 16.0 * iters / (ms * 1e-3) / 1e9;  // read + write
  return DAWN_OK;
}

extern "C" int dawn_device_malloc(int64_t bytes, void** out) {
  if (!out || bytes < 0) return DAWN_ERR_INVALID_ARG;
  *out = nullptr;
  const hipError_t e = hipMalloc(out, (size_t)(bytes ? bytes : 1));
  if (e != hipSuccess) {
    *out = nullptr;
    dawn::set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    return e == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY : DAWN_ERR_HIP;
  }
  return DAWN_OK;
}

extern "C" int dawn_device_free(void* p) {
  if (p) DAWN_HIP(hipFree(p));
  return DAWN_OK;
}

extern "C" int dawn_memcpy(void* dst, const void* src, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return DAWN_ERR_INVALID_ARG;
  if (bytes) DAWN_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
  return DAWN_OK;
}

extern "C" int dawn_memset(void* p, int32_t byte_value, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && !p)) return DAWN_ERR_INVALID_ARG;
  if (bytes) {
    DAWN_HIP(hipMemset(p, byte_value, (size_t)bytes));
    DAWN_HIP(hipDeviceSynchronize());
  }
  return DAWN_OK;
}

