// LDS read-bandwidth calibration: the measured peak the subset kernel's LDS
// roofline is quoted against (the kernel reads one 8-byte row word per lane
// per state, so its bound is LDS read bandwidth, not HBM).
//
// Every wave reads whole 64-lane rows (conflict-free: consecutive lanes read
// consecutive words) at fixed offsets from a per-lane base, eight reads in
// flight, and folds them with three-input XORs so the VALU stays well under
// the LDS rate.  An empty asm with a memory clobber makes every iteration
// reload.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "sw_kernels.hpp"

namespace sw {

namespace {

template <int BYTES>
__global__ __launch_bounds__(1024) void k_lds_read(uint64_t iters, uint32_t* out) {
  constexpr int WORDS = 65536 / 4;  // 64 KB of LDS
  __shared__ __align__(16) uint32_t buf[WORDS];
  for (int i = threadIdx.x; i < WORDS; i += blockDim.x) buf[i] = uint32_t(i) * 2654435761u;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  uint32_t a0 = 0, a1 = 0;
  if (BYTES == 16) {
    const uint4* b = reinterpret_cast<const uint4*>(buf) + lane;  // rows of 64 x 16 B
    for (uint64_t i = 0; i < iters; ++i) {
      asm volatile("" ::: "memory");
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint4 v = b[k * 64];
        a0 = __builtin_amdgcn_bitop3_b32(a0, v.x, v.y, 0x96);
        a1 = __builtin_amdgcn_bitop3_b32(a1, v.z, v.w, 0x96);
      }
    }
  } else {
    const uint2* b = reinterpret_cast<const uint2*>(buf) + lane;  // rows of 64 x 8 B
    for (uint64_t i = 0; i < iters; ++i) {
      asm volatile("" ::: "memory");
#pragma unroll
      for (int k = 0; k < 8; k += 2) {
        const uint2 v = b[k * 64];
        const uint2 w = b[(k + 1) * 64];
        a0 = __builtin_amdgcn_bitop3_b32(a0, v.x, v.y, 0x96);
        a1 = __builtin_amdgcn_bitop3_b32(a1, w.x, w.y, 0x96);
      }
    }
  }
  if ((a0 ^ a1) == 0x9e3779b9u) out[blockIdx.x] = 1;  // never true in practice; keeps the reads
}

}  // namespace

hipError_t launch_lds_read(int bytes_per_lane, uint64_t iters, uint32_t blocks, uint32_t* out, hipStream_t st) {
  if (bytes_per_lane == 16)
    hipLaunchKernelGGL(k_lds_read<16>, dim3(blocks), dim3(1024), 0, st, iters, out);
  else if (bytes_per_lane == 8)
    hipLaunchKernelGGL(k_lds_read<8>, dim3(blocks), dim3(1024), 0, st, iters, out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace sw
