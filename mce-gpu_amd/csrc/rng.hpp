// Counter-based random numbers and the clique hash, shared by host and device code.
//
// splitmix64 is Vigna's SplitMix64 output function; fmix64 is the MurmurHash3
// 64-bit finalizer. Both are published mixing functions used here as stated.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MCE_HD __host__ __device__ __forceinline__
#else
#define MCE_HD inline
#endif

namespace mce {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

MCE_HD uint64_t splitmix64(uint64_t x) {
  x += kGolden;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

MCE_HD uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Element i of the random stream `stream` under `seed`: a SplitMix64 sequence
// whose starting state is derived from (seed, stream).
MCE_HD uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t key = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
  return splitmix64(key + i * kGolden);
}

// Uniform double in [0, 1) from the top 53 bits.
MCE_HD double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Hash of one clique given the sum of splitmix64 over its vertex ids.
MCE_HD uint64_t clique_hash(uint64_t vertex_mix_sum, uint64_t size) {
  return fmix64(vertex_mix_sum + size * kGolden);
}

}  // namespace mce
