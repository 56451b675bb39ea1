// Shared host-side helpers for the C-ABI: error reporting and HIP checks.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace phc {

void set_error(const std::string &msg);

#define PHC_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      phc::set_error(std::string(#call) + ": " + hipGetErrorString(e_));     \
      return -1;                                                             \
    }                                                                        \
  } while (0)

#define PHC_REQUIRE(cond, msg)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      phc::set_error(msg);                                                   \
      return -2;                                                             \
    }                                                                        \
  } while (0)

inline int grid_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (1u << 30)) g = (1u << 30);
  return (int)g;
}

}  // namespace phc
