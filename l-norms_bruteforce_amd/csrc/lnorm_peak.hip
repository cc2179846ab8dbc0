// lnorm_peak.hip -- the VALU rate probe behind lnorm_hip_valu_rate: long
// chains of the instructions the walks are built on (v_sad_u16 and
// v_sad_u32 with a wave-uniform SGPR operand, and v_add_u32), with no memory
// traffic in the loop, timed on the device.  Its lane-ops per second are the
// sustained peak the roofline's published 256 CU x 64 lanes x 2.4 GHz
// assumes; divided by 256 x 64 lanes they give the sustained clock.
#include "lnorm_device.h"

namespace lnorm {

constexpr int kRateChains = 8;    // independent chains per lane
constexpr int kRateRounds = 32;   // rounds per loop iteration

// One round: one op per chain, chain j reading chain j + 1's accumulator as
// its first operand, so no operand is loop-invariant: the compiler can
// neither hoist |a - s| out of the loop nor fold rounds together, and each
// round issues exactly kRateChains ops (the last one 7 ops after the chain
// it reads, so the chains stay independent enough to fill the pipeline).
template <int OP>
__global__ __launch_bounds__(256) void k_valu_rate(uint32_t *__restrict__ out, uint32_t s, int iters) {
  uint32_t acc[kRateChains];
#pragma unroll
  for (int j = 0; j < kRateChains; ++j) acc[j] = (threadIdx.x + 1u) * 0x9E3779B9u + (uint32_t)j;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < kRateRounds; ++r) {
#pragma unroll
      for (int j = 0; j < kRateChains; ++j) {
        const uint32_t a = acc[(j + 1) % kRateChains];
        if constexpr (OP == 0) acc[j] = sad16(a, s, acc[j]);
        else if constexpr (OP == 1) acc[j] = sad32(a, s, acc[j]);
        else acc[j] = acc[j] + a;
      }
    }
  }
  uint32_t x = 0;
#pragma unroll
  for (int j = 0; j < kRateChains; ++j) x ^= acc[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

uint64_t valu_rate_lane_ops(int blocks, int iters) {
  return (uint64_t)blocks * 256u * (uint64_t)iters * kRateRounds * kRateChains;
}

hipError_t launch_valu_rate(int op, uint32_t *out, int blocks, int iters, hipStream_t st) {
  if (!out || blocks <= 0 || iters <= 0) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), blk(256);
  const uint32_t s = 0x8000u + 12345u;
  if (op == 0) hipLaunchKernelGGL(k_valu_rate<0>, grid, blk, 0, st, out, s, iters);
  else if (op == 1) hipLaunchKernelGGL(k_valu_rate<1>, grid, blk, 0, st, out, s, iters);
  else if (op == 2) hipLaunchKernelGGL(k_valu_rate<2>, grid, blk, 0, st, out, s, iters);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace lnorm
