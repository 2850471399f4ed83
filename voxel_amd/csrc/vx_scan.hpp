// Fused single-pass scan: Load(i, payload) -> scan -> Store(i, payload,
// value, exclusive, inclusive, n), the building block for stream compaction and
// for the DAG pipeline's run/level compactions. Counterpart of the reference's
// decoupled_lookback_fused_kernel (cuda/scan/fused.cuh:102-159), but with
// dynamic tile ids, a wave-wide look-back window, a per-element payload that
// travels from Load to Store in registers, an optional device-resident element
// count (persistent tiles: blocks claim tiles until the count is covered, so
// dependent stages need no host round trip), and a segmented monoid.
#pragma once

#include <algorithm>

#include "vx_common.hpp"

namespace vx {
namespace scan {

size_t workspace_bytes_u64(uint32_t n);
vx_status run_u64(uint64_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s);
vx_status generic_inplace(uint32_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s);
vx_status compact(const uint32_t* in, const uint32_t* flags, uint32_t n, uint32_t* out, uint32_t* d_count, void* ws,
                  hipStream_t s);

// 1024 x 8 = 8192 elements per tile: the tile rate of a decoupled look-back
// is bounded by the claim atomic (one same-address device atomic per tile)
// and by each tile's publish -> look-back round trip, so on this part a
// 4x larger tile than 256 x 8 roughly halves the time of every fused pass.
// Persistent grid: two 1024-thread blocks per CU.
constexpr uint32_t kFusedBlock = 1024;
constexpr uint32_t kFusedItems = 8;
constexpr uint32_t kFusedTile = kFusedBlock * kFusedItems;
constexpr uint32_t kFusedMaxBlocks = 512;

struct NoPayload {};

// Monoids. A look-back record packs {status:2 at kStatusShift, value below}.
struct SumOp {
  using T = uint32_t;
  static constexpr uint32_t kStatusShift = 32;
  __host__ __device__ static T identity() { return 0u; }
  __device__ static T combine(T a, T b) { return a + b; }
};
struct MaxOp {
  using T = uint32_t;
  static constexpr uint32_t kStatusShift = 32;
  __host__ __device__ static T identity() { return 0u; }
  __device__ static T combine(T a, T b) { return a > b ? a : b; }
};
// Segmented sum: bit 32 marks "a segment head is inside", bits 0..31 the sum
// since the last head. combine((f1,v1),(f2,v2)) = (f1|f2, f2 ? v2 : v1+v2).
struct SegSumOp {
  using T = uint64_t;
  static constexpr uint32_t kStatusShift = 40;
  static constexpr uint64_t kHead = 1ull << 32;
  __host__ __device__ static T identity() { return 0ull; }
  __device__ static T combine(T a, T b) {
    const uint64_t f = (a | b) & kHead;
    return (b & kHead) ? (f | uint32_t(b)) : (f | uint32_t(uint32_t(a) + uint32_t(b)));
  }
  __host__ __device__ static T make(uint32_t v, bool head) { return (head ? kHead : 0ull) | v; }
};

// Two counts in one word: bits 0..29 and 30..59 (each count < 2^30), so u64
// addition adds them independently; status at bit 62.
struct PairSumOp {
  using T = uint64_t;
  static constexpr uint32_t kStatusShift = 62;
  static constexpr uint32_t kBits = 30;
  __host__ __device__ static T identity() { return 0ull; }
  __device__ static T combine(T a, T b) { return a + b; }
  __host__ __device__ static T make(uint32_t lo, uint32_t hi) { return uint64_t(lo) | (uint64_t(hi) << kBits); }
  __host__ __device__ static uint32_t lo(T v) { return uint32_t(v) & ((1u << kBits) - 1); }
  __host__ __device__ static uint32_t hi(T v) { return uint32_t(v >> kBits) & ((1u << kBits) - 1); }
};

template <Op OP>
struct OpOf;
template <>
struct OpOf<kSum> {
  using type = SumOp;
};
template <>
struct OpOf<kMax> {
  using type = MaxOp;
};

inline size_t fused_workspace_bytes(uint64_t n) {
  const uint32_t tiles = ceil_div(n, kFusedTile);
  if (tiles <= 1) return 0;
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(tiles);
  return c.size();
}

// Items per thread of the device-count form: kFusedItems for passes over
// millions of elements (fewer tiles, fewer claims and look-backs),
// kFusedItemsSmall for the per-height passes whose counts are small (half the
// registers, so two 1024-thread blocks fit a CU and a tile's chain of
// dependent loads is shorter: measured 20% faster there, 15% slower on the
// big passes).
constexpr uint32_t kFusedItemsSmall = 4;

// Workspace for the device-count form: always carved (the count is unknown),
// sized for the smaller tile.
inline size_t fused_dev_workspace_bytes(uint64_t n_upper) {
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(std::max<uint32_t>(1, ceil_div(n_upper, kFusedBlock * kFusedItemsSmall)));
  return c.size();
}

template <class M>
__device__ __forceinline__ typename M::T wave_inclusive(typename M::T v) {
  return wave_inclusive_dpp(v, [](typename M::T a, typename M::T b) { return M::combine(a, b); });
}

// Decoupled look-back executed by one full wave: lane i reads the record of
// tile (base - i); the nearest inclusive prefix ends the walk, and any
// unpublished record before it makes the wave re-read. The window is folded
// oldest-first, so non-commutative (segmented) monoids are exact.
template <class M>
__device__ typename M::T lookback_window(const uint64_t* rec, uint32_t tile) {
  using T = typename M::T;
  constexpr uint64_t kValMask = (uint64_t(1) << M::kStatusShift) - 1;
  const uint32_t lane = lane_id();
  T excl = M::identity();
  int64_t base = int64_t(tile) - 1;
  while (base >= 0) {
    const int64_t idx = base - lane;
    const uint64_t r = idx >= 0 ? ld_agent(&rec[idx]) : (uint64_t(2) << M::kStatusShift);
    const uint32_t st = uint32_t(r >> M::kStatusShift);
    const uint64_t pref = ballot(st == 2);
    const uint64_t inval = ballot(st == 0);
    const uint32_t first = pref ? uint32_t(__builtin_ctzll(pref)) : 64u;
    const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1);
    if (inval & need) continue;
    const T mine = lane <= first ? T(r & kValMask) : M::identity();
    // Reverse lane order (oldest first), then an inclusive scan: lane 63 holds
    // rec[base-63] (+) ... (+) rec[base].
    const T rev = __shfl(mine, int(63 - lane), kWave);
    const T win = wave_last(wave_inclusive<M>(rev));
    excl = M::combine(win, excl);
    if (first < 64) break;
    base -= 64;
  }
  return excl;
}

// A Load may split its reads from its arithmetic: `Raw`, a
// `fetch<N>(idx, n, raw)` that reads the N items of a thread (indices clamped
// to the count n, so it reads without guards, round by round for all items
// at once), and `finish(i, raw, payload)` that does the rest without memory
// reads. A load written as one call per item waits for each item's reads
// before the next item's are issued whenever the call branches.
template <class L>
concept SplitLoad = requires { typename L::Raw; };

// One tile: kFusedBlock threads x ITEMS elements, wave-segment order
// (lane-contiguous per round, so every round of loads is coalesced). Grid
// kernels use kFusedItems; a single-workgroup pass over at most kFusedBlock
// elements uses one item per thread (a quarter of the registers).
template <class M, uint32_t ITEMS = kFusedItems, class Load, class Store>
__device__ __forceinline__ void fused_tile(uint32_t tile, uint32_t n, bool multi, const Load& ld, const Store& st,
                                           uint64_t* rec, typename M::T* s_wave, typename M::T* s_prefix) {
  using T = typename M::T;
  constexpr uint32_t kWaves = kFusedBlock / kWave;
  using P = typename Load::Payload;
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t seg = tile * (kFusedBlock * ITEMS) + wave * (kWave * ITEMS);
  T v[ITEMS], ex[ITEMS];
  P pl[ITEMS];
  T carry = M::identity();
  Load lc = ld;  // per-thread copy: a Load may keep state from call to call
  // Every item's loads first, then the scans: a wave scan between two items'
  // loads would make each item wait for its own round trip.
  if constexpr (SplitLoad<Load>) {
    uint32_t idx[ITEMS];
    typename Load::Raw raw[ITEMS];
    const uint32_t last = n ? n - 1 : 0u;
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) idx[k] = min(seg + k * kWave + lane, last);
    if (n) lc.template fetch<ITEMS>(idx, n, raw);
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) {
      const uint32_t i = seg + k * kWave + lane;
      v[k] = i < n ? T(lc.finish(i, raw[k], pl[k])) : M::identity();
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) {
      const uint32_t i = seg + k * kWave + lane;
      v[k] = i < n ? T(lc(i, pl[k])) : M::identity();
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < ITEMS; ++k) {
    const T inc = wave_inclusive<M>(v[k]);
    ex[k] = M::combine(carry, wave_shift_up1(inc));
    carry = M::combine(carry, wave_last(inc));
  }
  if (lane == 0) s_wave[wave] = carry;
  __syncthreads();
  T before = M::identity(), agg = M::identity();
#pragma unroll
  for (uint32_t w = 0; w < kWaves; ++w) {
    const T t = s_wave[w];
    if (w < wave) before = M::combine(before, t);
    agg = M::combine(agg, t);
  }
  if (wave == 0) {
    T prefix = M::identity();
    if (multi) {
      if (tile == 0) {
        if (lane == 0) st_agent(&rec[0], (uint64_t(2) << M::kStatusShift) | uint64_t(agg));
      } else {
        if (lane == 0) st_agent(&rec[tile], (uint64_t(1) << M::kStatusShift) | uint64_t(agg));
        prefix = lookback_window<M>(rec, tile);
        if (lane == 0) st_agent(&rec[tile], (uint64_t(2) << M::kStatusShift) | uint64_t(M::combine(prefix, agg)));
      }
    }
    if (lane == 0) *s_prefix = prefix;
  }
  __syncthreads();
  const T prefix = M::combine(*s_prefix, before);
#pragma unroll
  for (uint32_t k = 0; k < ITEMS; ++k) {
    const uint32_t i = seg + k * kWave + lane;
    if (i < n) {
      const T e = M::combine(prefix, ex[k]);
      st(i, pl[k], v[k], e, M::combine(e, v[k]), n);
    }
  }
}

template <class M, class Load, class Store>
__global__ __launch_bounds__(kFusedBlock) void fused_kernel(uint32_t n, Load ld, Store st, uint32_t* counter,
                                                            uint64_t* rec) {
  __shared__ typename M::T s_wave[kFusedBlock / kWave];
  __shared__ typename M::T s_prefix;
  __shared__ uint32_t s_tile;
  const bool multi = gridDim.x > 1;
  if (threadIdx.x == 0) s_tile = multi ? atomicAdd(counter, 1u) : 0u;
  __syncthreads();
  fused_tile<M>(s_tile, n, multi, ld, st, rec, s_wave, &s_prefix);
}

// Persistent form: the element count is read from device memory; blocks claim
// tiles until it is covered. Every block exits once the claimed tile lies past
// the count, so no block waits on a tile nobody will process.
template <class M, uint32_t ITEMS, class Load, class Store>
__global__ __launch_bounds__(kFusedBlock) void fused_dev_kernel(const uint32_t* n_ptr, uint32_t n_cap, Load ld,
                                                                Store st, uint32_t* counter, uint64_t* rec) {
  constexpr uint32_t kTile = kFusedBlock * ITEMS;
  __shared__ typename M::T s_wave[kFusedBlock / kWave];
  __shared__ typename M::T s_prefix;
  __shared__ uint32_t s_tile;
  const uint32_t n = min(*n_ptr, n_cap);
  const uint32_t tiles = (n + kTile - 1) / kTile;
  if (tiles <= 1) {  // one tile: block 0 takes it without a claim or look-back
    if (tiles == 1 && blockIdx.x == 0) fused_tile<M, ITEMS>(0u, n, false, ld, st, rec, s_wave, &s_prefix);
    return;
  }
  // Blocks beyond the tile count leave at once: they would only add
  // contended claims on the counter (the grid is sized for the upper bound).
  if (blockIdx.x >= tiles) return;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    __syncthreads();
    if (tile >= tiles) break;
    fused_tile<M, ITEMS>(tile, n, tiles > 1, ld, st, rec, s_wave, &s_prefix);
    __syncthreads();
  }
}

// Host launcher. `ws` must hold fused_workspace_bytes(n) bytes when n spans
// more than one tile.
template <class M, class Load, class Store>
vx_status fused_m(uint32_t n, Load ld, Store st, void* ws, hipStream_t s) {
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
  fused_kernel<M, Load, Store><<<tiles, kFusedBlock, 0, s>>>(n, ld, st, counter, rec);
  return last_launch();
}

template <Op OP, class Load, class Store>
vx_status fused(uint32_t n, Load ld, Store st, void* ws, hipStream_t s) {
  return fused_m<typename OpOf<OP>::type>(n, ld, st, ws, s);
}

// Host launcher for a device-resident count bounded by n_upper. `ws` must hold
// fused_dev_workspace_bytes(n_upper) bytes; with `clear` false the caller
// guarantees they are zero (a pipeline clears all its scans' workspaces with
// one fill instead of one per scan).
template <class M, uint32_t ITEMS = kFusedItems, class Load, class Store>
vx_status fused_dev_m(const uint32_t* n_ptr, uint64_t n_upper, Load ld, Store st, void* ws, hipStream_t s,
                      bool clear = true) {
  static_assert(ITEMS >= kFusedItemsSmall, "the workspace holds records for tiles of kFusedItemsSmall items");
  const uint32_t tiles = std::max<uint32_t>(1, ceil_div(n_upper, kFusedBlock * ITEMS));
  Carver c(ws);
  uint32_t* counter = c.take<uint32_t>(64);
  uint64_t* rec = c.take<uint64_t>(tiles);
  if (clear) {
    const vx_status r =
        fill_u32(counter, (reinterpret_cast<char*>(rec + tiles) - reinterpret_cast<char*>(counter)) / 4, 0, s);
    if (r != VX_OK) return r;
  }
  const uint32_t blocks = std::min(tiles, kFusedMaxBlocks);
  const uint32_t n_cap = uint32_t(std::min<uint64_t>(n_upper, 0xFFFFFFFFu));
  fused_dev_kernel<M, ITEMS, Load, Store><<<blocks, kFusedBlock, 0, s>>>(n_ptr, n_cap, ld, st, counter, rec);
  return last_launch();
}

template <Op OP, uint32_t ITEMS = kFusedItems, class Load, class Store>
vx_status fused_dev(const uint32_t* n_ptr, uint64_t n_upper, Load ld, Store st, void* ws, hipStream_t s,
                    bool clear = true) {
  return fused_dev_m<typename OpOf<OP>::type, ITEMS>(n_ptr, n_upper, ld, st, ws, s, clear);
}

}  // namespace scan
}  // namespace vx
