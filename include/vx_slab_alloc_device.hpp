/*
 * vx_slab_alloc_device.hpp — device API of the reusable slab allocator
 * (vx_slab_alloc_* in vx.h), for HIP kernels of the caller's own.
 * Counterpart of SlabAlloc::allocate / deallocate
 * (cuda/hash_table/slab_alloc.cuh:128-229): ordinals 0..capacity-1 marked in
 * a bitmap, one u64 word per 64 slabs; a freed ordinal can be taken again.
 *
 * The reference scans 1024-slab blocks with a 32-lane warp, one 32-bit
 * bitmap word per lane, from a per-warp resident block. Here each lane
 * allocates on its own: it starts at a word picked by the caller's hint (so
 * concurrent callers spread over the bitmap), claims the lowest free bit of
 * a 64-bit word with one 64-bit atomic OR (re-trying the same word while it
 * still has free bits), and walks on word by word, wrapping around once;
 * an ordinal comes back only if its bit was clear before the OR, and the
 * null ordinal only after every word was seen full.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "vx.h"

namespace vx_slab_alloc_dev {

constexpr uint32_t kNull = 0xffffffffu;

__device__ __forceinline__ uint32_t spread(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  return v ^ (v >> 16);
}

// Bits of word w that name ordinals below the capacity.
__device__ __forceinline__ unsigned long long valid_bits(const vx_slab_alloc_view& a, uint32_t w) {
  const uint32_t tail = a.capacity & 63u;
  return (w + 1 == a.words && tail) ? ((1ull << tail) - 1ull) : ~0ull;
}

__device__ inline uint32_t allocate(const vx_slab_alloc_view& a, uint32_t hint) {
  if (a.capacity == 0 || a.words == 0) return kNull;
  uint32_t w = spread(hint) % a.words;
  for (uint32_t seen = 0; seen < a.words; ++seen) {
    const unsigned long long valid = valid_bits(a, w);
    unsigned long long cur = __hip_atomic_load(&a.bitmap[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (unsigned long long free = ~cur & valid; free; free = ~cur & valid) {
      const unsigned long long bit = free & (~free + 1ull);
      const unsigned long long old = atomicOr(&a.bitmap[w], bit);
      if (!(old & bit)) return w * 64u + uint32_t(__builtin_ctzll(bit));
      cur = old | bit;
    }
    w = w + 1 == a.words ? 0 : w + 1;
  }
  return kNull;
}

// An ordinal outside the capacity is ignored; freeing a free ordinal is a
// no-op (slab_alloc.cuh:150-161).
__device__ inline void deallocate(const vx_slab_alloc_view& a, uint32_t ordinal) {
  if (ordinal >= a.capacity) return;
  atomicAnd(&a.bitmap[ordinal / 64u], ~(1ull << (ordinal % 64u)));
}

}  // namespace vx_slab_alloc_dev
