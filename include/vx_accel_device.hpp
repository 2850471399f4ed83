/*
 * vx_accel_device.hpp — device API of the content-addressed set / map
 * (vx_accel_set_* / vx_accel_map_* in vx.h), for HIP kernels of the caller's
 * own. Counterpart of DeviceAccelerationHashSet32<W> / Map32<W>
 * (cuda/hash_table/acceleration_hash/set/common.cuh:40-60,
 * map/common.cuh:45-75): find, insert_unique_unchecked and read_item on
 * items of W u32 words (1..16), each lane working on its own item (no
 * wave-synchronous call contract: inactive lanes simply do not call).
 *
 * Layout: items live in a word pool at item index * W (values, for a map,
 * at the item index); an open-addressing table of u64 slots {fingerprint:32
 * | item index + 1:32}, at most half full, indexes them. A ptr is
 * item index + 1 (0 is the null ptr). Insert takes the next item index from
 * a bump counter (status Overflow once `capacity` items exist), writes the
 * item (and value), then publishes it by claiming an empty slot with a
 * 64-bit CAS; find walks the probe sequence from the item's hash and
 * compares the whole item wherever the fingerprint matches. Equal items
 * inserted unchecked are all stored; find returns the first one on the
 * probe sequence (the same for every caller).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "vx.h"

namespace vx_accel {

enum : uint32_t { kNotFound = 0, kInserted = 1, kOverflow = 4, kFound = 6 };
constexpr uint32_t kMaxWords = 16;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 64-bit hash of the item's words: high half the fingerprint (never 0), low
// bits the first probe position.
__device__ __forceinline__ uint64_t hash_item(const uint32_t (&w)[kMaxWords], uint32_t words) {
  uint64_t h = 0x6A09E667F3BCC909ull ^ words;
#pragma unroll
  for (uint32_t k = 0; k < kMaxWords; ++k)
    if (k < words) h = mix(h ^ w[k]);
  return h;
}

__device__ __forceinline__ uint32_t fingerprint(uint64_t h) {
  const uint32_t f = uint32_t(h >> 32);
  return f ? f : 1u;
}

__device__ __forceinline__ void load_item(const uint32_t* item, uint32_t words, uint32_t (&w)[kMaxWords]) {
#pragma unroll
  for (uint32_t k = 0; k < kMaxWords; ++k) w[k] = k < words ? item[k] : 0u;
}

__device__ __forceinline__ bool equal_at(const vx_accel_view& t, uint32_t index, const uint32_t (&w)[kMaxWords]) {
  const uint32_t* s = t.items + uint64_t(index) * t.words;
  bool eq = true;
#pragma unroll
  for (uint32_t k = 0; k < kMaxWords; ++k)
    if (k < t.words) eq = eq && s[k] == w[k];
  return eq;
}

// find (acceleration_hash/set/device.cuh find): *ptr = item index + 1 and
// status Found, or ptr 0 and NotFound; *value for a map.
__device__ inline uint32_t find(const vx_accel_view& t, const uint32_t* item, uint32_t* ptr, uint32_t* value) {
  uint32_t w[kMaxWords];
  load_item(item, t.words, w);
  const uint64_t h = hash_item(w, t.words);
  const uint32_t fp = fingerprint(h);
  uint64_t pos = h & t.slot_mask;
  for (uint64_t probe = 0; probe <= t.slot_mask; ++probe, pos = (pos + 1) & t.slot_mask) {
    const unsigned long long s = __hip_atomic_load(&t.slots[pos], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (s == 0) break;
    if (uint32_t(s >> 32) != fp) continue;
    const uint32_t index = uint32_t(s) - 1u;
    if (equal_at(t, index, w)) {
      *ptr = index + 1u;
      if (value && t.values) *value = t.values[index];
      return kFound;
    }
  }
  *ptr = 0;
  if (value) *value = 0;
  return kNotFound;
}

// insert_unique_unchecked: the caller guarantees no equal item is present
// or inserted concurrently (equal items would all be stored).
__device__ inline uint32_t insert_unique_unchecked(const vx_accel_view& t, const uint32_t* item, uint32_t value,
                                                   uint32_t* ptr) {
  // One counter update per wave: the lanes in this call take consecutive
  // indices (a counter every lane of the chip hit would serialise).
  const uint64_t active = __ballot(1);
  const uint32_t lane = __lane_id();
  const uint32_t leader = uint32_t(__ffsll((long long)active) - 1);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(active >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(active), 0u));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(t.used, uint32_t(__popcll(active)));
  base = __shfl(base, int(leader), 64);
  const uint32_t index = base + rank;
  if (index >= t.capacity) {
    *ptr = 0;
    return kOverflow;
  }
  uint32_t w[kMaxWords];
  load_item(item, t.words, w);
  uint32_t* dst = t.items + uint64_t(index) * t.words;
#pragma unroll
  for (uint32_t k = 0; k < kMaxWords; ++k)
    if (k < t.words) dst[k] = w[k];
  if (t.values) t.values[index] = value;
  const uint64_t h = hash_item(w, t.words);
  const unsigned long long want = (uint64_t(fingerprint(h)) << 32) | (index + 1u);
  uint64_t pos = h & t.slot_mask;
  for (uint64_t probe = 0; probe <= t.slot_mask; ++probe, pos = (pos + 1) & t.slot_mask) {
    // Release: the item words are visible to any reader that sees the slot.
    unsigned long long expected = 0;
    if (__hip_atomic_compare_exchange_strong(&t.slots[pos], &expected, want, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      *ptr = index + 1u;
      return kInserted;
    }
  }
  *ptr = 0;  // unreachable while the table is at most half full
  return kOverflow;
}

// read_item: the item (and value) behind a ptr from find / insert.
__device__ inline bool read_item(const vx_accel_view& t, uint32_t ptr, uint32_t* out, uint32_t* value) {
  const uint32_t used = __hip_atomic_load(t.used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (ptr == 0 || ptr > t.capacity || ptr > used) return false;
  const uint32_t* s = t.items + uint64_t(ptr - 1u) * t.words;
  for (uint32_t k = 0; k < t.words; ++k) out[k] = s[k];
  if (value && t.values) *value = t.values[ptr - 1u];
  return true;
}

}  // namespace vx_accel
