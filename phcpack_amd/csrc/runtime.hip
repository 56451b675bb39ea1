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

extern "C" {

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
