// Butterfly exchanges inside one wavefront (wave64) on the vector ALU.
#pragma once
#include <hip/hip_runtime.h>

namespace wave {

// The value of the butterfly partner lane l ^ M (M = 1, 2, ..., 32) inside a
// butterfly over stages M = 1, 2, 4, ... in increasing order, with every lane
// of the wavefront active.  By stage M the M lanes of every aligned group
// hold the same value, so for M = 4 and 8 the row mirrors (lane i of an 8-
// or 16-lane row to lane 7 - i or 15 - i, in the other half as l ^ M is)
// deliver the partner's value.  DPP moves for M < 16 and the permlane swaps
// for M = 16, 32 (exactly lane l ^ M) run on the vector ALU instead of
// going through LDS as __shfl_xor does.
template <int M> __device__ __forceinline__ int bfly_b32(int x) {
  if constexpr (M == 1) {
    return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (M == 4) {
    return __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false);  // row_half_mirror
  } else if constexpr (M == 8) {
    return __builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false);  // row_mirror
  } else if constexpr (M == 16) {
    // {rows 0,0,2,2 ; rows 1,1,3,3} of x
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (threadIdx.x & 16) ? (int)r[0] : (int)r[1];
  } else {
    // {halves 0,0 ; halves 1,1} of x
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (threadIdx.x & 32) ? (int)r[0] : (int)r[1];
  }
}

template <int M> __device__ __forceinline__ double bfly(double v) {
  const int lo = bfly_b32<M>(__double2loint(v)), hi = bfly_b32<M>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

template <int M> __device__ __forceinline__ float bfly(float v) {
  return __int_as_float(bfly_b32<M>(__float_as_int(v)));
}

// maximum / minimum over the 64 lanes (in every lane)
__device__ __forceinline__ double max64(double v) {
  v = fmax(v, bfly<1>(v));
  v = fmax(v, bfly<2>(v));
  v = fmax(v, bfly<4>(v));
  v = fmax(v, bfly<8>(v));
  v = fmax(v, bfly<16>(v));
  return fmax(v, bfly<32>(v));
}

__device__ __forceinline__ double min64(double v) {
  v = fmin(v, bfly<1>(v));
  v = fmin(v, bfly<2>(v));
  v = fmin(v, bfly<4>(v));
  v = fmin(v, bfly<8>(v));
  v = fmin(v, bfly<16>(v));
  return fmin(v, bfly<32>(v));
}

// (value, index) maximum with the smallest index among equal values, one
// butterfly stage
template <int M> __device__ __forceinline__ void argmax_stage(double &v, int &idx) {
  const double ov = bfly<M>(v);
  const int oi = bfly_b32<M>(idx);
  if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}

}  // namespace wave
