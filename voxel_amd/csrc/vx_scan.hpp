// Fused single-pass scan: Load(i, payload) -> scan -> Store(i, payload,
// value, exclusive, inclusive), the building block for stream compaction and
// for the DAG pipeline's run/level compactions. Counterpart of the reference's
// decoupled_lookback_fused_kernel (cuda/scan/fused.cuh:102-159), but with
// dynamic tile ids, a wave-wide look-back window, a per-element payload that
// travels from Load to Store in registers, and an optional device-resident
// element count (persistent tiles: blocks claim tiles until the count is
// covered), so dependent stages need no host round trip.
#pragma once

#include <algorithm>

#include "vx_common.hpp"

namespace vx {
namespace scan {

vx_status generic_inplace(uint32_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s);
vx_status compact(const uint32_t* in, const uint32_t* flags, uint32_t n, uint32_t* out, uint32_t* d_count, void* ws,
                  hipStream_t s);

constexpr uint32_t kFusedBlock = 256;
constexpr uint32_t kFusedItems = 8;
constexpr uint32_t kFusedTile = kFusedBlock * kFusedItems;
constexpr uint32_t kFusedMaxBlocks = 2048;

struct NoPayload {};

inline size_t fused_workspace_bytes(uint64_t n) {
  const uint32_t tiles = ceil_div(n, kFusedTile);
  if (tiles <= 1) return 0;
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(tiles);
  return c.size();
}

// Workspace for the device-count form: always carved (the count is unknown).
inline size_t fused_dev_workspace_bytes(uint64_t n_upper) {
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(std::max<uint32_t>(1, ceil_div(n_upper, kFusedTile)));
  return c.size();
}

template <Op OP>
__device__ __forceinline__ uint32_t fcombine(uint32_t a, uint32_t b) {
  if constexpr (OP == kSum) return a + b;
  else return a > b ? a : b;
}

// Decoupled look-back executed by one full wave: lane i reads the record of
// tile (base - i); the nearest inclusive prefix ends the walk, and any
// unpublished record before it makes the wave re-read. Records pack the
// status (bits 32..33: 1 aggregate, 2 inclusive prefix) over a 32-bit value
// so one 64-bit agent-scope load observes both.
template <Op OP>
__device__ uint32_t lookback_window(const uint64_t* rec, uint32_t tile) {
  const uint32_t lane = lane_id();
  uint32_t excl = 0;
  int64_t base = int64_t(tile) - 1;
  while (base >= 0) {
    const int64_t idx = base - lane;
    const uint64_t r = idx >= 0 ? ld_agent(&rec[idx]) : (uint64_t(2) << 32);
    const uint32_t st = uint32_t(r >> 32);
    const uint64_t pref = ballot(st == 2);
    const uint64_t inval = ballot(st == 0);
    const uint32_t first = pref ? uint32_t(__builtin_ctzll(pref)) : 64u;
    const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1);
    if (inval & need) continue;
    uint32_t contrib = lane <= first ? uint32_t(r) : 0u;
    if constexpr (OP == kSum) contrib = wave_inclusive_sum(contrib);
    else contrib = wave_inclusive_max(contrib);
    excl = fcombine<OP>(excl, shfl(contrib, 63));
    if (first < 64) break;
    base -= 64;
  }
  return excl;
}

// One tile: kFusedBlock threads x kFusedItems elements, wave-segment order
// (lane-contiguous per round, so every round of loads is coalesced).
template <Op OP, class Load, class Store>
__device__ __forceinline__ void fused_tile(uint32_t tile, uint32_t n, bool multi, const Load& ld, const Store& st,
                                           uint64_t* rec, uint32_t* s_wave, uint32_t* s_prefix) {
  constexpr uint32_t kWaves = kFusedBlock / kWave;
  using P = typename Load::Payload;
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t seg = tile * kFusedTile + wave * (kWave * kFusedItems);
  uint32_t v[kFusedItems], ex[kFusedItems];
  P pl[kFusedItems];
  uint32_t carry = 0;
#pragma unroll
  for (uint32_t k = 0; k < kFusedItems; ++k) {
    const uint32_t i = seg + k * kWave + lane;
    v[k] = i < n ? ld(i, pl[k]) : 0u;
    uint32_t inc;
    if constexpr (OP == kSum) inc = wave_inclusive_sum(v[k]);
    else inc = wave_inclusive_max(v[k]);
    uint32_t e = __shfl_up(inc, 1, kWave);
    if (lane == 0) e = 0;
    ex[k] = fcombine<OP>(carry, e);
    carry = fcombine<OP>(carry, shfl(inc, 63));
  }
  if (lane == 0) s_wave[wave] = carry;
  __syncthreads();
  uint32_t before = 0, agg = 0;
#pragma unroll
  for (uint32_t w = 0; w < kWaves; ++w) {
    const uint32_t t = s_wave[w];
    if (w < wave) before = fcombine<OP>(before, t);
    agg = fcombine<OP>(agg, t);
  }
  if (wave == 0) {
    uint32_t prefix = 0;
    if (multi) {
      if (tile == 0) {
        if (lane == 0) st_agent(&rec[0], (uint64_t(2) << 32) | agg);
      } else {
        if (lane == 0) st_agent(&rec[tile], (uint64_t(1) << 32) | agg);
        prefix = lookback_window<OP>(rec, tile);
        if (lane == 0) st_agent(&rec[tile], (uint64_t(2) << 32) | fcombine<OP>(prefix, agg));
      }
    }
    if (lane == 0) *s_prefix = prefix;
  }
  __syncthreads();
  const uint32_t prefix = fcombine<OP>(*s_prefix, before);
#pragma unroll
  for (uint32_t k = 0; k < kFusedItems; ++k) {
    const uint32_t i = seg + k * kWave + lane;
    if (i < n) {
      const uint32_t e = fcombine<OP>(prefix, ex[k]);
      st(i, pl[k], v[k], e, fcombine<OP>(e, v[k]));
    }
  }
}

template <Op OP, class Load, class Store>
__global__ __launch_bounds__(kFusedBlock) void fused_kernel(uint32_t n, Load ld, Store st, uint32_t* counter,
                                                            uint64_t* rec) {
  __shared__ uint32_t s_wave[kFusedBlock / kWave];
  __shared__ uint32_t s_tile, s_prefix;
  const bool multi = gridDim.x > 1;
  if (threadIdx.x == 0) s_tile = multi ? atomicAdd(counter, 1u) : 0u;
  __syncthreads();
  fused_tile<OP>(s_tile, n, multi, ld, st, rec, s_wave, &s_prefix);
}

// Persistent form: the element count is read from device memory; blocks claim
// tiles until it is covered. Every block exits once the claimed tile lies past
// the count, so no block waits on a tile nobody will process.
template <Op OP, class Load, class Store>
__global__ __launch_bounds__(kFusedBlock) void fused_dev_kernel(const uint32_t* n_ptr, uint32_t n_cap, Load ld,
                                                                Store st, uint32_t* counter, uint64_t* rec) {
  __shared__ uint32_t s_wave[kFusedBlock / kWave];
  __shared__ uint32_t s_tile, s_prefix;
  const uint32_t n = min(*n_ptr, n_cap);
  const uint32_t tiles = (n + kFusedTile - 1) / kFusedTile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    __syncthreads();
    if (tile >= tiles) break;
    fused_tile<OP>(tile, n, tiles > 1, ld, st, rec, s_wave, &s_prefix);
    __syncthreads();
  }
}

// Host launcher. `ws` must hold fused_workspace_bytes(n) bytes when n spans
// more than one tile.
template <Op OP, class Load, class Store>
vx_status fused(uint32_t n, Load ld, Store st, void* ws, hipStream_t s) {
  if (n == 0) return VX_OK;
  const uint32_t tiles = ceil_div(n, kFusedTile);
  uint32_t* counter = nullptr;
  uint64_t* rec = nullptr;
  if (tiles > 1) {
    Carver c(ws);
    counter = c.take<uint32_t>(64);
    rec = c.take<uint64_t>(tiles);
    const vx_status r = fill_u32(counter, (reinterpret_cast<char*>(rec + tiles) - reinterpret_cast<char*>(counter)) / 4, 0, s);
    if (r != VX_OK) return r;
  }
  fused_kernel<OP, Load, Store><<<tiles, kFusedBlock, 0, s>>>(n, ld, st, counter, rec);
  return last_launch();
}

// Host launcher for a device-resident count bounded by n_upper. `ws` must hold
// fused_dev_workspace_bytes(n_upper) bytes.
template <Op OP, class Load, class Store>
vx_status fused_dev(const uint32_t* n_ptr, uint64_t n_upper, Load ld, Store st, void* ws, hipStream_t s) {
  const uint32_t tiles = std::max<uint32_t>(1, ceil_div(n_upper, kFusedTile));
  Carver c(ws);
  uint32_t* counter = c.take<uint32_t>(64);
  uint64_t* rec = c.take<uint64_t>(tiles);
  const vx_status r = fill_u32(counter, (reinterpret_cast<char*>(rec + tiles) - reinterpret_cast<char*>(counter)) / 4, 0, s);
  if (r != VX_OK) return r;
  const uint32_t blocks = std::min(tiles, kFusedMaxBlocks);
  const uint32_t n_cap = uint32_t(std::min<uint64_t>(n_upper, 0xFFFFFFFFu));
  fused_dev_kernel<OP, Load, Store><<<blocks, kFusedBlock, 0, s>>>(n_ptr, n_cap, ld, st, counter, rec);
  return last_launch();
}

}  // namespace scan
}  // namespace vx
