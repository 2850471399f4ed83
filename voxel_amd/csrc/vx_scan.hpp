// Fused single-pass scan: Load(i) -> scan -> Store(i, value, exclusive,
// inclusive), the building block for stream compaction and for the DAG
// pipeline's run/level compactions. Counterpart of the reference's
// decoupled_lookback_fused_kernel (cuda/scan/fused.cuh:102-159), but with a
// dynamic tile id and a wave-wide look-back window.
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

inline size_t fused_workspace_bytes(uint32_t n) {
  const uint32_t tiles = ceil_div(n, kFusedTile);
  if (tiles <= 1) return 0;
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(tiles);
  return c.size();
}

template <Op OP>
__device__ __forceinline__ uint32_t fcombine(uint32_t a, uint32_t b) {
  if constexpr (OP == kSum) return a + b;
  else return a > b ? a : b;
}

// Decoupled look-back executed by one full wave: lane i reads the record of
// tile (base - i); the nearest inclusive prefix ends the walk, and any
// unpublished record before it makes the wave re-read.
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

template <Op OP, class Load, class Store>
__global__ __launch_bounds__(kFusedBlock) void fused_kernel(uint32_t n, Load ld, Store st, uint32_t* counter,
                                                            uint64_t* rec) {
  constexpr uint32_t kWaves = kFusedBlock / kWave;
  __shared__ uint32_t s_wave[kWaves];
  __shared__ uint32_t s_tile, s_prefix;
  const bool multi = gridDim.x > 1;
  if (threadIdx.x == 0) s_tile = multi ? atomicAdd(counter, 1u) : 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t seg = tile * kFusedTile + wave * (kWave * kFusedItems);

  uint32_t v[kFusedItems], ex[kFusedItems];
  uint32_t carry = 0;
#pragma unroll
  for (uint32_t k = 0; k < kFusedItems; ++k) {
    const uint32_t i = seg + k * kWave + lane;
    v[k] = i < n ? ld(i) : 0u;
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
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const uint32_t prefix = fcombine<OP>(s_prefix, before);
#pragma unroll
  for (uint32_t k = 0; k < kFusedItems; ++k) {
    const uint32_t i = seg + k * kWave + lane;
    if (i < n) {
      const uint32_t e = fcombine<OP>(prefix, ex[k]);
      st(i, v[k], e, fcombine<OP>(e, v[k]));
    }
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

}  // namespace scan
}  // namespace vx
