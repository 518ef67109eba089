// Device-side helpers for gfx950 (wave64): counter-based RNG, wave reductions,
// lane broadcasts and the digest mix. Included by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgamd {

constexpr int kWave = 64;
constexpr uint32_t kTile = 64;  // tile edge; one u64 word per tile row

// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014). sm64_at(key, i) is the (i+1)-th output of the
// generator seeded with key, so every draw is addressable by its counter.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t sm64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t sm64_at(uint64_t key, uint64_t i) {
  return sm64_mix(key + (i + 1) * 0x9E3779B97F4A7C15ull);
}
// One independent stream per (seed, purpose tag).
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t tag) {
  return sm64_mix(sm64_mix(seed ^ 0x6A09E667F3BCC909ull) + tag);
}
enum StreamTag : uint64_t { kTagRmat = 1, kTagEr = 2, kTagCommunity = 3, kTagSources = 4 };

// Order-independent digest term for word `value` at position `index`.
__host__ __device__ inline uint64_t digest_term(uint64_t index, uint64_t value) {
  return sm64_mix(index * 0xD1B54A32D192ED03ull ^ sm64_mix(value + 0x8CB92BA72F3D8DD7ull));
}

// Global-memory (address space 1) views of device pointers. A pointer read from memory (a
// descriptor table, say) is generic to the compiler, which then emits flat instructions;
// flat loads also count on the LDS counter, so every wait for a gather would wait for the
// wave's LDS operations too. Loads and stores through these views are global_* ones.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// Value of `v` in lane `src` (src wave-uniform).
__device__ inline uint64_t read_lane(uint64_t v, int src) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), src);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), src);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ inline uint32_t read_first_lane(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}
__device__ inline uint64_t read_first_lane64(uint64_t v) {
  const uint32_t lo = read_first_lane(static_cast<uint32_t>(v));
  const uint32_t hi = read_first_lane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ inline uint32_t shfl_xor32(uint32_t v, int mask) {
  return static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), mask, kWave));
}
__device__ inline uint64_t shfl_xor64(uint64_t v, int mask) {
  const uint32_t lo = shfl_xor32(static_cast<uint32_t>(v), mask);
  const uint32_t hi = shfl_xor32(static_cast<uint32_t>(v >> 32), mask);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
// Value of v in lane src (per-lane src).
__device__ inline uint64_t shfl64(uint64_t v, uint32_t src) {
  const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(src), kWave));
  const uint32_t hi =
      static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(src), kWave));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
// Bitwise OR over the 64 lanes (result in every lane).
__device__ inline uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v |= shfl_xor64(v, m);
  return v;
}
__device__ inline uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += shfl_xor64(v, m);
  return v;
}
// Orders LDS accesses of one wave across lanes (no cross-wave effect).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compact payload of a 64 x 64 bit tile with at most kPayloadMax set bits: entry i =
// (row << 6) | column in bits 12i..12i+11, the entry count in bits 48..50. A tile with more
// bits gets kPayloadDense and its bit count in the low 13 bits. Called by a whole wave with
// lane r holding word r (row r) of the tile; returns the payload in every lane.
constexpr uint32_t kPayloadMax = 4;
constexpr uint64_t kPayloadDense = 1ull << 63;
__device__ inline uint64_t tile_payload(uint64_t word) {
  const uint64_t total = wave_sum(__popcll(word));
  if (total > kPayloadMax) return kPayloadDense | total;
  uint64_t pl = 0;
  uint32_t n = 0;
  for (uint64_t m = ballot(word != 0); m; m &= m - 1) {
    const int r = __ffsll(static_cast<unsigned long long>(m)) - 1;
    for (uint64_t wr = read_lane(word, r); wr; wr &= wr - 1) {
      const uint64_t c = static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(wr)) - 1);
      pl |= ((static_cast<uint64_t>(r) << 6) | c) << (12 * n);
      ++n;
    }
  }
  return pl | (static_cast<uint64_t>(n) << 48);
}

}  // namespace mgamd
