// wave.hpp -- wave64 helpers shared by the transport kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ophip {

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Rank of this lane among the set bits of `mask` below it.
__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t first_lane_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ uint32_t popc(uint64_t m) { return __popcll(m); }

// Entry of a block-local queue for the lanes with `want`: one LDS atomic on
// `counter` per wave. Called by all lanes of the wave.
__device__ __forceinline__ uint32_t reserve(uint32_t* counter, bool want) {
  const uint64_t m = __ballot(want);
  uint32_t base = 0;
  if (m != 0) {
    if (lane_id() == 0) base = atomicAdd(counter, popc(m));
    base = __builtin_amdgcn_readfirstlane(base);
  }
  return base + rank_in(m);
}

}  // namespace ophip
