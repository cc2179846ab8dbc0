// lnorm_general.hip -- the width-unlimited fallback walk and small helpers.
//
// k_general covers every (kind, d, ny) the register-resident kernels do not
// (ny up to the reference's `length` = 4096, functions.h:2): one wave walks
// one segment, lanes own columns l = lane, lane+64, ..., and the label
// vectors live in a global scratch slice of the wave.  Each step updates the
// lane's columns of the one (L1/LM) or two (L2, Ld) vectors that change and
// reduces their |.| sums across the wave.  Same MODE_MAX / MODE_FIRST
// contract as the plain kernels; segments are d^q ranks starting at
// rank seg*d^q.
#include "lnorm_device.h"

namespace lnorm {

__device__ __forceinline__ int gray_label(uint64_t N, int i, int d) {
  if (d <= 2) return (int)(((N ^ (N >> 1)) >> i) & 1);
  uint64_t x = N;
  for (int k = 0; k < i; ++k) x /= (uint64_t)d;
  const uint32_t r = (uint32_t)(x % (uint64_t)(2 * d));
  return r < (uint32_t)d ? (int)r : (int)(2 * d - 1 - r);
}

template <int MODE>
__global__ __launch_bounds__(64) void k_general(
    const int32_t *__restrict__ A, int R, int C, int kind, int d, uint64_t seglen, uint64_t seg0,
    uint64_t nseg, uint32_t *__restrict__ state, int32_t target, int32_t *__restrict__ unit_best,
    int32_t *__restrict__ gbest, unsigned long long *__restrict__ gfirst, uint32_t B) {
  const int lane = threadIdx.x & 63;
  const bool l1 = (kind == KIND_L1 || kind == KIND_LM);
  const int nvec = l1 ? 1 : d;
  uint32_t *st = state + (size_t)blockIdx.x * (size_t)nvec * (size_t)C;
  uint32_t S[16];

  for (uint64_t sidx = blockIdx.x; sidx < nseg; sidx += gridDim.x) {
    const uint64_t N0 = (seg0 + sidx) * seglen;
    // seed
    for (int l = lane; l < C; l += 64) {
      const uint32_t last = (uint32_t)A[(size_t)(R - 1) * C + l];
      for (int c = 0; c < nvec; ++c) st[(size_t)c * C + l] = (c == 0 ? last : 0u) + kBias32;
    }
    for (int i = 0; i < R - 1; ++i) {
      const int lab = gray_label(N0, i, d < 2 ? 2 : d);
      for (int l = lane; l < C; l += 64) {
        const uint32_t a = (uint32_t)A[(size_t)i * C + l];
        if (l1) st[l] += lab ? a : 0u - a;
        else st[(size_t)lab * C + l] += a;
      }
    }
    auto vec_sum = [&](int c) -> uint32_t {
      uint32_t p = 0;
      for (int l = lane; l < C; l += 64) {
        const uint32_t y = st[(size_t)c * C + l];
        p = (kind == KIND_LM && l == 0) ? p + (y ^ kBias32) : sad32(y, B, p);
      }
      return wave_sum_u32(p);
    };
    for (int c = 0; c < nvec; ++c) S[c] = vec_sum(c);

    int32_t best = INT32_MIN;
    uint64_t first = ~0ull;
    auto eval = [&](uint64_t N) {
      uint32_t acc = 0;
      for (int c = 0; c < nvec; ++c) acc += S[c];
      if (MODE == MODE_MAX) best = (int32_t)acc > best ? (int32_t)acc : best;
      else if ((int32_t)acc == target && first == ~0ull) first = N;
    };
    eval(N0);
    for (uint64_t t = 1; t < seglen; ++t) {
      const uint64_t N = N0 + t;
      int j, from = 0, to = 0, up = 0;
      if (d <= 2) {
        j = __builtin_ctzll(N);
        up = (int)(((N >> (j + 1)) & 1) ^ 1);
        from = up ? 0 : 1;
        to = up ? 1 : 0;
      } else {
        uint64_t x = N;
        j = 0;
        while (x % (uint64_t)d == 0) { x /= (uint64_t)d; ++j; }
        const int v = (int)(x % (uint64_t)d);
        const int p = (int)((x / (uint64_t)d) & 1);
        from = p ? d - v : v - 1;
        to = p ? d - 1 - v : v;
      }
      const int32_t *row = A + (size_t)j * C;
      if (l1) {
        const uint32_t m = up ? 2u : (uint32_t)-2;
        for (int l = lane; l < C; l += 64) st[l] += m * (uint32_t)row[l];
        S[0] = vec_sum(0);
      } else {
        for (int l = lane; l < C; l += 64) {
          const uint32_t a = (uint32_t)row[l];
          st[(size_t)from * C + l] -= a;
          st[(size_t)to * C + l] += a;
        }
        const uint32_t sf = vec_sum(from), stt = vec_sum(to);
        for (int c = 0; c < nvec; ++c) {
          if (c == from) S[c] = sf;
          if (c == to) S[c] = stt;
        }
      }
      eval(N);
    }
    // best and first are wave-uniform (built from wave sums): every lane
    // stores and updates with the same value, so no branch depends on the
    // lane index inside the segment loop
    if (MODE == MODE_MAX) {
      unit_best[sidx] = best;
      atomicMax(gbest, best);
    } else if (first != ~0ull) {
      atomicMin(gfirst, (unsigned long long)first);
    }
  }
}

hipError_t launch_general(Kind kind, Mode mode, const PlainArgs &a, int ny, hipStream_t st) {
  if (!a.A || !a.state || !a.unit_best || !a.gbest || !a.gfirst) return hipErrorInvalidValue;
  // one wave per segment; the state scratch holds grid * nvec * ny words
  uint64_t seglen = 1;
  const int base = a.d < 2 ? 2 : a.d;
  for (int i = 0; i < a.q; ++i) seglen *= (uint64_t)base;
  const uint64_t slices = a.slices ? a.slices : kGeneralSlices;
  const uint64_t grid = a.nseg < slices ? a.nseg : slices;
  if (mode == MODE_MAX)
    hipLaunchKernelGGL(k_general<MODE_MAX>, dim3((unsigned)grid), dim3(64), 0, st, a.A, a.R, ny, (int)kind,
                       a.d, seglen, a.seg0, a.nseg, a.state, a.target, a.unit_best, a.gbest,
                       a.gfirst, kBias32);
  else
    hipLaunchKernelGGL(k_general<MODE_FIRST>, dim3((unsigned)grid), dim3(64), 0, st, a.A, a.R, ny, (int)kind,
                       a.d, seglen, a.seg0, a.nseg, a.state, a.target, a.unit_best, a.gbest,
                       a.gfirst, kBias32);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// helpers: fills (used instead of hipMemsetAsync so the sequence stays
// capturable in a HIP graph) and the first unit attaining the maximum.
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t *p, int32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_first_unit(const int32_t *__restrict__ ub, uint64_t n, const int32_t *__restrict__ gbest,
                             unsigned long long *__restrict__ first) {
  const int32_t target = *gbest;
  unsigned long long mine = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (ub[i] == target) { mine = i; break; }
  const uint64_t m = wave_min_u64(mine);
  if ((threadIdx.x & 63) == 0 && m != ~0ull) atomicMin(first, (unsigned long long)m);
}

// Self-test of the absolute-difference instructions the kernels rely on:
// out[2i] = v_sad_u32(a, b, c), out[2i+1] = v_sad_u16(a, b, c).
__global__ void k_sad_probe(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                            const uint32_t *__restrict__ c, uint32_t *__restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[2 * i] = sad32(a[i], b[i], c[i]);
    out[2 * i + 1] = sad16(a[i], b[i], c[i]);
  }
}

hipError_t launch_sad_probe(const uint32_t *a, const uint32_t *b, const uint32_t *c, uint32_t *out, int n,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_sad_probe, dim3((n + 63) / 64), dim3(64), 0, st, a, b, c, out, n);
  return hipGetLastError();
}

static unsigned grid_for(uint64_t n, int block) {
  uint64_t g = (n + block - 1) / block;
  if (g > 4096) g = 4096;
  return (unsigned)(g ? g : 1);
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, uint64_t n, hipStream_t st) {
  if (!p && n) return hipErrorInvalidValue;   // an unallocated buffer: an error, never a launch
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, uint64_t n, hipStream_t st) {
  if (!p && n) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_first_unit(const int32_t *ub, uint64_t n, const int32_t *gbest, unsigned long long *first,
                             hipStream_t st) {
  if (!ub || !gbest || !first) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_first_unit, dim3(grid_for(n, 256)), dim3(256), 0, st, ub, n, gbest, first);
  return hipGetLastError();
}

}  // namespace lnorm
