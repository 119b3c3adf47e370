// Wave64 helpers for gfx950 (CDNA4). All of them expect to be called in
// wave-uniform control flow unless stated otherwise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_add_u32(unsigned int);
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_max_u32(unsigned int);
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_min_u32(unsigned int);
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_or_u32(unsigned int);
// v_writelane_b32: the LLVM intrinsic, which clang does not expose as a builtin
extern "C" __device__ __attribute__((const)) unsigned int mce_llvm_writelane(unsigned int v, unsigned int lane,
                                                                             unsigned int old)
    __asm("llvm.amdgcn.writelane.i32");

namespace mce {

constexpr uint32_t kWave = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool b) { return __ballot(b); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// `old` with lane l replaced by the uniform value v (one v_writelane per
// 32 bits, updated in place; l and v must be wave-uniform).
__device__ __forceinline__ uint32_t wrlane(uint32_t old, uint32_t l, uint32_t v) { return mce_llvm_writelane(v, l, old); }
__device__ __forceinline__ uint64_t wrlane64(uint64_t old, uint32_t l, uint64_t v) {
  const uint32_t lo = mce_llvm_writelane((uint32_t)v, l, (uint32_t)old);
  const uint32_t hi = mce_llvm_writelane((uint32_t)(v >> 32), l, (uint32_t)(old >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Bit `lane_id()` of the uniform mask m, with m used as the lane mask itself
// (no per-lane shift and compare).
__device__ __forceinline__ bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Reductions over the active lanes; the result is uniform.
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) { return __ockl_wfred_add_u32(x); }
__device__ __forceinline__ uint32_t wave_max(uint32_t x) { return __ockl_wfred_max_u32(x); }
__device__ __forceinline__ uint32_t wave_min(uint32_t x) { return __ockl_wfred_min_u32(x); }
__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
  const uint32_t lo = __ockl_wfred_or_u32((uint32_t)x), hi = __ockl_wfred_or_u32((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t ctz64(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }
__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__builtin_popcountll(x); }

// Lower bound of `key` in the ascending array a[0, len).
template <typename P>
__device__ __forceinline__ uint32_t lower_bound_u32(P a, uint32_t len, uint32_t key) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Wave-aggregated append: every lane with `take` gets a unique slot of
// *counter; returns the slot (undefined for lanes without take).
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool take) {
  const uint64_t m = ballot(take);
  uint32_t base = 0;
  if (m) {
    const uint32_t leader = ctz64(m);
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)popc64(m));
    base = rdlane(base, leader);
  }
  return base + mbcnt(m);
}

}  // namespace mce
