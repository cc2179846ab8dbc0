// lnorm_device.h -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lnorm_internal.h"

namespace lnorm {

// |a - b| + c on unsigned 32-bit values: one v_sad_u32.  With y stored as
// y + 2^31 and b = 2^31 this is |y| + c for the signed y.
__device__ __forceinline__ uint32_t sad32(uint32_t a, uint32_t b, uint32_t c) {
  return (a > b ? a - b : b - a) + c;
}

// |a.lo - b.lo| + |a.hi - b.hi| + c on two unsigned 16-bit halves: one v_sad_u16.
__device__ __forceinline__ uint32_t sad16(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_sad_u16(a, b, c);
}

__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63); }

// A zero the compiler cannot see through.  |y| + 0 written as sad32(y, B, 0)
// is folded into min/max/sub (three instructions); with this accumulator it
// stays one v_sad_u32.  Not volatile: computed once, outside the loops.
__device__ __forceinline__ uint32_t opaque_zero() {
  uint32_t z;
  asm("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// Global wave index of the calling thread.  It is uniform within the wave;
// readfirstlane tells the compiler so, which keeps everything derived from it
// in SGPRs.
__device__ __forceinline__ uint64_t global_wave() {
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)w);
}

// Butterfly reductions over the 64 lanes.  Every lane ends with the same
// value; readfirstlane states that to the compiler, so the result lives in
// an SGPR and an atomic on it is issued once per wave, not once per lane.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    int32_t o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, 64);
    v = v < o ? v : o;
  }
  return uniform_u64(v);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Segment of a lane in the lane-per-segment walks.  Lanes of one wave get
// segments of equal parity, so the Gray-step direction, which depends on the
// parity of the segment index for the top digit (and for every digit when
// d is odd), is uniform across the wave:
//   local = (wave >> 1) * 128 + 2 * lane + (wave & 1).
__device__ __forceinline__ uint64_t lane_segment(uint64_t wave, int lane) {
  return ((wave >> 1) << 7) | ((uint64_t)lane << 1) | (wave & 1);
}

}  // namespace lnorm
