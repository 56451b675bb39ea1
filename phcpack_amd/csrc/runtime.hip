// Runtime helpers of the C-ABI: error string, device memory, streams and
// events.  The reference has no such layer (its GPU_* functions allocate,
// copy and free on every call, with no error checks).
#include <hip/hip_runtime.h>
#include <string>
#include "common.hpp"
#include "../../include/phcmd.h"

namespace phc {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace phc

namespace {
// plain streaming copy: 16 bytes per lane per iteration
__global__ void copy_kernel(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" {

int phc_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
  PHC_REQUIRE(bytes % 16 == 0, "phc_device_copy: bytes must be a multiple of 16");
  PHC_REQUIRE(((size_t)dst | (size_t)src) % 16 == 0, "phc_device_copy: pointers must be 16-byte aligned");
  if (bytes == 0) return 0;
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(copy_kernel, dim3(phc::grid_for(n, 256) < 8192 ? phc::grid_for(n, 256) : 8192),
                     dim3(256), 0, (hipStream_t)stream, (double2 *)dst, (const double2 *)src, n);
  PHC_HIP(hipGetLastError());
  return 0;
}

const char *phc_last_error(void) { return phc::g_last_error.c_str(); }

int phc_version(void) { return 100; }

int phc_device_count(int *n) {
  PHC_REQUIRE(n != nullptr, "phc_device_count: null");
  PHC_HIP(hipGetDeviceCount(n));
  return 0;
}

int phc_malloc(void **ptr, size_t bytes) {
  PHC_REQUIRE(ptr != nullptr, "phc_malloc: null");
  *ptr = nullptr;
  if (bytes == 0) bytes = 8;
  PHC_HIP(hipMalloc(ptr, bytes));
  return 0;
}

int phc_free(void *ptr) {
  if (ptr) PHC_HIP(hipFree(ptr));
  return 0;
}

int phc_memcpy_htod(void *dst, const void *src, size_t bytes) {
  if (bytes) PHC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int phc_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  if (bytes) PHC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int phc_memset(void *dst, int value, size_t bytes) {
  if (bytes) PHC_HIP(hipMemset(dst, value, bytes));
  return 0;
}

int phc_synchronize(void) {
  PHC_HIP(hipDeviceSynchronize());
  return 0;
}

int phc_stream_synchronize(void *stream) {
  PHC_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int phc_event_create(void **ev) {
  PHC_REQUIRE(ev != nullptr, "phc_event_create: null");
  hipEvent_t e;
  PHC_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return 0;
}

int phc_event_destroy(void *ev) {
  if (ev) PHC_HIP(hipEventDestroy((hipEvent_t)ev));
  return 0;
}

int phc_event_record(void *ev, void *stream) {
  PHC_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return 0;
}

int phc_event_elapsed_ms(void *start, void *stop, float *ms) {
  PHC_HIP(hipEventSynchronize((hipEvent_t)stop));
  PHC_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return 0;
}

}  // extern "C"
