// Onesweep LSD radix sort for gfx950: stable, u32/u64 keys with an optional
// u32 payload.
//
// Replaces algo::cuda::sort::sort_keys / sort_by_key (cuda/sort/sort.cuh:100-167)
// and follows the onesweep structure of cuda/sort/radix/onesweep/host.cuh:12-205:
//   1. one read of the keys builds the digit histograms of every pass
//      (global_offsets.cuh:52-99 — here LDS atomics + one global flush per block);
//   2. per pass, one partition kernel: dynamic tile id, wave64 multisplit ranks
//      (RB ballots over 64-bit lane masks instead of 32-bit warp masks,
//      warp_rank/multi_split.cuh:8-48), per-digit decoupled look-back
//      (lookback.cuh:7-79), and an LDS-staged scatter so consecutive output
//      positions of one digit are written by consecutive lanes.
// Look-back records are u32 {status:2, value:30}, hence n < 2^30
// (onesweep/host.cuh:85-86). Record rows of the next pass are cleared by the
// current pass, so no memset runs between passes. Partition blocks are
// persistent (they claim tiles until the key count is covered), which also
// lets the DAG pipeline sort a device-resident count without a host sync.
#include <algorithm>
#include <mutex>

#include "vx_scan.hpp"
#include "vx_sort.hpp"

namespace vx {
namespace sort {

namespace {

constexpr uint32_t kStatusAgg = 1u << 30;
constexpr uint32_t kStatusPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kMaxLds = 160 * 1024;  // gfx950 LDS per CU; checked against the device at run time

template <class K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
  return uint32_t(key >> shift) & mask;
}

template <class K, bool VALS, uint32_t RB, uint32_t BLOCK, uint32_t KPT>
struct Cfg {
  static constexpr uint32_t R = 1u << RB;
  static constexpr uint32_t WAVES = BLOCK / kWave;
  static constexpr uint32_t TILE = BLOCK * KPT;
  // The staged tile (keys, values) and the per-wave digit histograms share
  // one LDS region: the histograms are dead once every key's position is in
  // registers, before the first key is staged.
  static constexpr size_t KEYS_BYTES = size_t(TILE) * (sizeof(K) + (VALS ? 4 : 0));
  static constexpr size_t HIST_BYTES = size_t(WAVES) * R * 4;
  static constexpr size_t SHARED = ((KEYS_BYTES > HIST_BYTES ? KEYS_BYTES : HIST_BYTES) + 15) / 16 * 16;
  static constexpr size_t LDS = SHARED + size_t(3) * R * 4 + 64;
  static constexpr bool kFeasible = LDS <= kMaxLds && R <= BLOCK;
};

__device__ __forceinline__ uint32_t count_of(const uint32_t* n_ptr, uint32_t n_arg) {
  return n_ptr ? min(*n_ptr, n_arg) : n_arg;
}

// ---- pass 0: histograms of every digit position in one read ----------------
// Lanes of the wave whose digit equals this lane's, among the lanes of
// `init`: one ballot per digit bit, each narrowing the match with one
// three-input bit operation per 32-lane half (m &= ~(ballot ^ my bit)).
template <uint32_t RB>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t init) {
  uint32_t lo = uint32_t(init), hi = uint32_t(init >> 32);
#pragma unroll
  for (uint32_t b = 0; b < RB; ++b) {
    const uint32_t bm = uint32_t(__builtin_amdgcn_sbfe(int(d), b, 1));  // all ones iff bit b is set
    const uint64_t bal = ballot(bm != 0u);
    lo &= ~(uint32_t(bal) ^ bm);
    hi &= ~(uint32_t(bal >> 32) ^ bm);
  }
  return (uint64_t(hi) << 32) | lo;
}

// The same match over a digit of `bits` bits (a wave-uniform run-time value).
__device__ __forceinline__ uint64_t match_digit_n(uint32_t d, uint32_t bits, uint64_t init) {
  uint32_t lo = uint32_t(init), hi = uint32_t(init >> 32);
  for (uint32_t b = 0; b < bits; ++b) {
    const uint32_t bm = uint32_t(__builtin_amdgcn_sbfe(int(d), b, 1));
    const uint64_t bal = ballot(bm != 0u);
    lo &= ~(uint32_t(bal) ^ bm);
    hi &= ~(uint32_t(bal >> 32) ^ bm);
  }
  return (uint64_t(hi) << 32) | lo;
}

// A digit of at most this many bits is counted once per distinct value in
// the wave: with 64 lanes over 2^bits values, same-address LDS atomics from
// one instruction (which serialise) would otherwise dominate, as for the
// 2-bit top digit of 42-bit keys in 8-bit digits.
constexpr uint32_t kMatchCountBits = 6;

// Adds the wave's digits of one pass (lanes at and above `valid_lanes` hold
// no key) into the block's LDS histogram h.
// COHERENT (keys that arrive in spatially coherent runs, like the edit's
// brick records in brush order): equal digits then come in runs over
// consecutive lanes, and same-address LDS atomics from one instruction
// serialise, so the first lane of each run adds the run's length; a wave
// with many runs (the low digits) adds one per lane instead. Random keys add
// one per lane, or, for narrow digits, one per distinct value.
template <bool COHERENT, bool NARROW = false>
__device__ __forceinline__ void count_digits(uint32_t* h, uint32_t d, uint32_t bits, uint32_t lane,
                                             uint32_t valid_lanes) {
  const bool valid = lane < valid_lanes;
  if constexpr (COHERENT) {
    const uint32_t key = valid ? d : 0xffffffffu;  // a run of its own, never added
    // The previous lane's digit (lane 0 sees one that differs from its own).
    const uint32_t prev = uint32_t(__builtin_amdgcn_update_dpp(int(~key), int(key), 0x138 /* wave_shr:1 */, 0xf, 0xf,
                                                               false));
    const uint64_t heads = ballot(key != prev);
    if (__popcll(heads) <= 16) {
      if (valid && key != prev) {
        const uint64_t above = (heads >> lane) >> 1;  // heads after this lane
        const uint32_t len = above ? uint32_t(__builtin_ctzll(above)) + 1u : kWave - lane;
        atomicAdd(&h[d], len);
      }
      return;
    }
  }
  if (NARROW && bits <= kMatchCountBits) {  // wave-uniform
    const uint64_t live = valid_lanes >= kWave ? ~0ull : ((1ull << valid_lanes) - 1ull);
    const uint64_t peers = match_digit_n(d, bits, live);
    if (valid && lanes_below(peers) == 0) atomicAdd(&h[d], uint32_t(__popcll(peers)));
    return;
  }
  if (valid) atomicAdd(&h[d], 1u);
}

// Each wave reads runs of kHistKpt * 64 consecutive keys, all loads of a run
// issued before any is counted; two 1024-thread blocks per CU (eight waves
// per SIMD) keep the final flush of the per-block histograms into the global
// ones to a few hundred atomics per counter. The block keeps kHistCopies
// copies of its histograms (wave w counts into copy w % kHistCopies), so at
// most four waves contend for one LDS counter; the flush sums them.
constexpr uint32_t kHistKpt = 8;
constexpr uint32_t kHistBlock = 1024;
constexpr uint32_t kHistCopies = 4;

template <class K, uint32_t RB, bool COHERENT>
__global__ __launch_bounds__(kHistBlock, 8) void histogram_kernel(const K* __restrict__ keys, const uint32_t* n_ptr,
                                                               uint32_t n_arg, uint32_t begin_bit, uint32_t key_bits,
                                                               uint32_t* __restrict__ hist) {
  constexpr uint32_t R = 1u << RB;
  constexpr uint32_t MAXP = (sizeof(K) * 8 + RB - 1) / RB;
  constexpr uint32_t CHUNK = kHistBlock * kHistKpt;
  constexpr uint32_t COPY = MAXP * R;
  static_assert(size_t(kHistCopies) * COPY * 4 <= 64 * 1024, "histogram copies exceed static LDS");
  __shared__ uint32_t s_all[kHistCopies * COPY];
  const uint32_t n = count_of(n_ptr, n_arg);
  const uint32_t passes = (key_bits + RB - 1) / RB;
  const uint32_t full = key_bits / RB;  // passes of all RB bits
  const uint32_t lane = lane_id();
  for (uint32_t i = threadIdx.x; i < kHistCopies * COPY; i += kHistBlock) s_all[i] = 0;
  __syncthreads();
  uint32_t* s = s_all + (wave_id() % kHistCopies) * COPY;
  for (uint64_t c0 = uint64_t(blockIdx.x) * CHUNK; c0 < n; c0 += uint64_t(gridDim.x) * CHUNK) {
    const uint32_t first = wave_id() * (kWave * kHistKpt);  // the wave's first key in the chunk
    const K* run = keys + c0 + first;
    const uint32_t left = uint32_t(min<uint64_t>(CHUNK, n - c0));
    const uint32_t wave_keys = left > first ? left - first : 0u;
    K key[kHistKpt];
#pragma unroll
    for (uint32_t j = 0; j < kHistKpt; ++j) key[j] = j * kWave + lane < wave_keys ? run[j * kWave + lane] : K(0);
#pragma unroll
    for (uint32_t j = 0; j < kHistKpt; ++j) {
      const uint32_t valid_lanes = wave_keys > j * kWave ? min(wave_keys - j * kWave, kWave) : 0u;
      if (valid_lanes == 0) break;  // uniform across the wave
      // Passes of RB bits, then the narrower last pass, if any (only it can
      // take the per-distinct-value count).
#pragma unroll
      for (uint32_t p = 0; p < MAXP; ++p) {
        if (p < full)
          count_digits<COHERENT>(s + p * R, digit_of(key[j], begin_bit + p * RB, R - 1), RB, lane, valid_lanes);
      }
      if (passes > full) {
        const uint32_t bits = key_bits - full * RB;
        count_digits<COHERENT, true>(s + full * R, digit_of(key[j], begin_bit + full * RB, (1u << bits) - 1), bits,
                                     lane, valid_lanes);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * R; i += kHistBlock) {
    uint32_t v = 0;
#pragma unroll
    for (uint32_t c = 0; c < kHistCopies; ++c) v += s_all[c * COPY + i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// Loads of a wave's KPT keys (lane-contiguous rounds) and their ranks among
// the wave's earlier keys of the same digit, kept in the wave's LDS
// histogram `wh`. FULL: the whole tile lies below n (no per-key guards).
template <bool FULL, class K, bool VALS, uint32_t RB, uint32_t KPT>
__device__ __forceinline__ void load_and_rank(const K* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t base,
                                              uint32_t n, uint32_t lane, uint32_t shift, uint32_t dmask, uint32_t* wh,
                                              K (&key)[KPT], uint32_t (&val)[KPT], uint32_t (&dig)[KPT],
                                              uint32_t (&rank)[KPT]) {
#pragma unroll
  for (uint32_t j = 0; j < KPT; ++j) {
    const uint64_t i = base + j * kWave + lane;
    const bool valid = FULL || i < n;
    key[j] = valid ? kin[i] : K(0);
    if constexpr (VALS) val[j] = valid ? vin[i] : 0u;
    dig[j] = digit_of(key[j], shift, dmask);
  }
#pragma unroll
  for (uint32_t j = 0; j < KPT; ++j) {
    const bool valid = FULL || base + j * kWave + lane < n;
    const uint64_t peers = match_digit<RB>(dig[j], FULL ? ~0ull : ballot(valid));
    const uint32_t below = lanes_below(peers);
    uint32_t cnt = 0;
    if (valid) cnt = wh[dig[j]];
    rank[j] = cnt + below;
    if (valid && below == 0) wh[dig[j]] = cnt + uint32_t(__popcll(peers));
  }
}

// ---- partition pass -------------------------------------------------------
// One tile per claim: load, rank (wave64 multisplit), per-digit wave scan and
// publication of the tile aggregate, then the first look-back window is
// issued before the digit scan and the LDS staging, so its round trip to the
// predecessors' records overlaps them; the walk is finished afterwards and
// the staged keys are scattered to their digits' output runs.
template <class K, bool VALS, uint32_t RB, uint32_t BLOCK, uint32_t KPT>
__global__ __launch_bounds__(BLOCK) void partition_kernel(const K* __restrict__ kin, K* __restrict__ kout,
                                                          const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                          const uint32_t* n_ptr, uint32_t n_arg, uint32_t shift,
                                                          uint32_t dmask, const uint32_t* __restrict__ pass_counts,
                                                          uint32_t* tile_counter, uint32_t* rec, uint32_t* rec_next) {
  using C = Cfg<K, VALS, RB, BLOCK, KPT>;
  constexpr uint32_t R = C::R, WAVES = C::WAVES, TILE = C::TILE;
  constexpr int kLookWindow = 8;
  extern __shared__ __align__(16) unsigned char smem[];
  K* s_keys = reinterpret_cast<K*>(smem);
  uint32_t* s_vals = reinterpret_cast<uint32_t*>(smem + size_t(TILE) * sizeof(K));
  uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);  // [WAVES][R], aliases the staged tile
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(smem + C::SHARED);  // tile count per digit
  uint32_t* s_start = s_cnt + R;             // tile-local start per digit
  uint32_t* s_global = s_start + R;          // global base - tile start per digit
  uint32_t* s_misc = s_global + R;

  const uint32_t n = count_of(n_ptr, n_arg);
  const uint32_t tiles = (n + TILE - 1) / TILE;
  const uint32_t t = threadIdx.x, lane = lane_id(), wave = wave_id();
  uint32_t* wh = s_hist + wave * R;
  // The pass's digit offsets, the exclusive scan of its histogram, computed
  // by every block (2 KB read) instead of by a launch of its own; thread t
  // keeps digit t's.
  uint32_t digit_base = 0;
  {
    const uint32_t c = t < R ? pass_counts[t] : 0u;
    const uint32_t inc = wave_inclusive_sum(c);
    if (lane == kWave - 1) s_cnt[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += s_cnt[w];
    digit_base = before + inc - c;
    __syncthreads();
  }

  for (;;) {
    for (uint32_t i = t; i < WAVES * R; i += BLOCK) s_hist[i] = 0;
    if (t == 0) s_misc[0] = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = s_misc[0];
    if (tile >= tiles) break;
    const uint64_t tile_base = uint64_t(tile) * TILE;
    const uint64_t base = tile_base + uint64_t(wave) * kWave * KPT;

    K key[KPT];
    uint32_t val[KPT], dig[KPT], rank[KPT];
    const bool full = tile_base + TILE <= n;
    if (full)
      load_and_rank<true, K, VALS, RB, KPT>(kin, vin, base, n, lane, shift, dmask, wh, key, val, dig, rank);
    else
      load_and_rank<false, K, VALS, RB, KPT>(kin, vin, base, n, lane, shift, dmask, wh, key, val, dig, rank);
    __syncthreads();
    // Per digit (one per thread, R <= BLOCK): exclusive scan over waves (wave
    // order = input order), publication of the tile aggregate, and the first
    // look-back window over the predecessors' records.
    uint32_t look[kLookWindow];
    if (t < R) {
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t w = 0; w < WAVES; ++w) {
        const uint32_t c = s_hist[w * R + t];
        s_hist[w * R + t] = sum;
        sum += c;
      }
      s_cnt[t] = sum;
      st_agent(&rec[size_t(tile) * R + t], (tile == 0 ? kStatusPrefix : kStatusAgg) | sum);
#pragma unroll
      for (int q = 0; q < kLookWindow; ++q) {
        const int64_t jq = int64_t(tile) - 1 - q;
        look[q] = jq >= 0 ? ld_agent(&rec[size_t(jq) * R + t]) : kStatusPrefix;
      }
    }
    __syncthreads();
    // Exclusive scan over digits of the tile counts (wave 0, R <= 512).
    if (wave == 0) {
      constexpr uint32_t PER = (R + kWave - 1) / kWave;
      uint32_t loc[PER];
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t d = lane * PER + q;
        loc[q] = d < R ? s_cnt[d] : 0u;
        sum += loc[q];
      }
      const uint32_t inc = wave_inclusive_sum(sum);
      uint32_t run = inc - sum;
#pragma unroll
      for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t d = lane * PER + q;
        if (d < R) s_start[d] = run;
        run += loc[q];
      }
    }
    __syncthreads();
    // Tile positions into registers, then (the histograms dead) the staging.
#pragma unroll
    for (uint32_t j = 0; j < KPT; ++j) rank[j] += s_start[dig[j]] + wh[dig[j]];
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < KPT; ++j) {
      if (full || base + j * kWave + lane < n) {
        s_keys[rank[j]] = key[j];
        if constexpr (VALS) s_vals[rank[j]] = val[j];
      }
    }
    // Finish the per-digit look-back, kLookWindow records per round trip: the
    // loads of one window are independent, so a walk across tiles that have
    // only published aggregates costs one memory latency per window instead
    // of one per tile; an unpublished record makes the walk re-read from it.
    if (t < R) {
      uint32_t excl = 0;
      if (tile > 0) {
        int64_t j = int64_t(tile) - 1;
        bool done = false, first = true;
        while (!done) {
          uint32_t r[kLookWindow];
#pragma unroll
          for (int q = 0; q < kLookWindow; ++q)
            r[q] = first ? look[q] : (j - q >= 0 ? ld_agent(&rec[size_t(j - q) * R + t]) : kStatusPrefix);
          first = false;
          int consumed = 0;
#pragma unroll
          for (int q = 0; q < kLookWindow; ++q) {
            const uint32_t st = r[q] & ~kValueMask;
            if (done || consumed != q || st == 0) continue;  // stop at the first unpublished record
            excl += r[q] & kValueMask;
            consumed = q + 1;
            if (st == kStatusPrefix) done = true;
          }
          j -= consumed;
        }
        st_agent(&rec[size_t(tile) * R + t], kStatusPrefix | (excl + s_cnt[t]));
      }
      s_global[t] = digit_base + excl - s_start[t];
      rec_next[size_t(tile) * R + t] = 0u;
    }
    __syncthreads();
    const uint32_t tile_n = uint32_t(min<uint64_t>(TILE, n - tile_base));
    for (uint32_t p = t; p < tile_n; p += BLOCK) {
      const K k = s_keys[p];
      const uint32_t g = s_global[digit_of(k, shift, dmask)] + p;
      kout[g] = k;
      if constexpr (VALS) vout[g] = s_vals[p];
    }
    __syncthreads();
  }
}

template <class T>
__global__ void copy_kernel(const T* __restrict__ in, T* __restrict__ out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[i];
}

// ---- tuning table -----------------------------------------------------------
std::mutex g_mu;
// Rows: u32 pairs, u64 pairs, u32 keys only, u64 keys only.
Tuning g_tuning[4] = {{8, 15, 256}, {8, 15, 256}, {8, 15, 256}, {9, 8, 1024}};

uint32_t tuning_row(size_t key_bytes, bool vals) { return (key_bytes == 8 ? 1u : 0u) + (vals ? 0u : 2u); }

template <class K, bool VALS>
struct Launcher {
  size_t lds = 0;
  uint32_t tile = 0;
  void (*kernel)(const K*, K*, const uint32_t*, uint32_t*, const uint32_t*, uint32_t, uint32_t, uint32_t,
                 const uint32_t*, uint32_t*, uint32_t*, uint32_t*) = nullptr;
  uint32_t block = 0;
};

template <class K, bool VALS, uint32_t RB, uint32_t BLOCK, uint32_t KPT>
bool pick3(Launcher<K, VALS>& L) {
  using C = Cfg<K, VALS, RB, BLOCK, KPT>;
  if constexpr (C::kFeasible) {
    L.kernel = partition_kernel<K, VALS, RB, BLOCK, KPT>;
    L.lds = C::LDS;
    L.tile = C::TILE;
    L.block = BLOCK;
    return true;
  } else {
    return false;
  }
}

template <class K, bool VALS, uint32_t RB, uint32_t BLOCK>
bool pick2(Launcher<K, VALS>& L, uint32_t kpt) {
  switch (kpt) {
    case 8: return pick3<K, VALS, RB, BLOCK, 8>(L);
    case 15: return pick3<K, VALS, RB, BLOCK, 15>(L);
    case 23: return pick3<K, VALS, RB, BLOCK, 23>(L);
  }
  return false;
}

template <class K, bool VALS, uint32_t RB>
bool pick1(Launcher<K, VALS>& L, uint32_t block, uint32_t kpt) {
  switch (block) {
    case 256: return pick2<K, VALS, RB, 256>(L, kpt);
    case 512: return pick2<K, VALS, RB, 512>(L, kpt);
    case 1024: return pick2<K, VALS, RB, 1024>(L, kpt);
  }
  return false;
}

struct DeviceLimits {
  size_t max_lds = 64 * 1024;
  uint32_t cus = 256;
};

const DeviceLimits& limits() {
  static DeviceLimits cached;
  static bool done = false;
  if (!done) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        cached.max_lds = size_t(v);
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
        cached.cus = uint32_t(v);
    }
    done = true;
  }
  return cached;
}

// 9-bit digits (4 passes instead of 5 for 33..36-bit keys, 5 instead of 6 for
// 41..45): 512 digits need a block of at least 512 threads.
template <class K, bool VALS>
bool pick9(Launcher<K, VALS>& L, uint32_t block, uint32_t kpt) {
  if (block == 512 && kpt == 8) return pick3<K, VALS, 9, 512, 8>(L);
  if (block == 1024 && kpt == 8) return pick3<K, VALS, 9, 1024, 8>(L);
  if (block == 512 && kpt == 15) return pick3<K, VALS, 9, 512, 15>(L);
  if (block == 512 && kpt == 23) return pick3<K, VALS, 9, 512, 23>(L);
  if (block == 1024 && kpt == 15) return pick3<K, VALS, 9, 1024, 15>(L);
  return false;
}

template <class K, bool VALS>
bool pick(Launcher<K, VALS>& L, const Tuning& tn) {
  bool ok = false;
  if (tn.rb == 8) ok = pick1<K, VALS, 8>(L, tn.block, tn.kpt);
  else if (tn.rb == 4) ok = pick1<K, VALS, 4>(L, tn.block, tn.kpt);
  else if (tn.rb == 9) ok = pick9<K, VALS>(L, tn.block, tn.kpt);
  return ok && L.lds <= limits().max_lds;
}

Tuning tuning_for(size_t key_bytes, bool vals = true) {
  std::lock_guard<std::mutex> g(g_mu);
  return g_tuning[tuning_row(key_bytes, vals)];
}

struct Layout {
  uint32_t passes, tiles;
  uint32_t *hist, *counters, *rec0, *rec1, *alt_vals;
  void* alt_keys;
  size_t bytes;
};

Layout layout(void* ws, uint64_t n, size_t key_bytes, uint32_t key_bits, bool vals, const Tuning& tn, uint32_t tile,
              bool carve_alt) {
  Layout L{};
  const uint32_t R = 1u << tn.rb;
  L.passes = (key_bits + tn.rb - 1) / tn.rb;
  L.tiles = std::max<uint32_t>(1, ceil_div(n, tile));
  Carver c(ws);
  L.hist = c.take<uint32_t>(size_t(L.passes) * R);
  L.counters = c.take<uint32_t>(64);
  L.rec0 = c.take<uint32_t>(size_t(L.tiles) * R);
  L.rec1 = c.take<uint32_t>(size_t(L.tiles) * R);
  if (carve_alt) {
    L.alt_keys = c.take<char>(size_t(n) * key_bytes);
    L.alt_vals = vals ? c.take<uint32_t>(n) : nullptr;
  }
  L.bytes = c.size();
  return L;
}

// Core driver. Sorts (keys, vals) using (alt_keys, alt_vals) as the second
// buffer; *in_alt reports where the result ended up.
template <class K, bool VALS, bool COHERENT = false>
vx_status drive(K* keys, uint32_t* vals, K* alt_keys, uint32_t* alt_vals, const uint32_t* n_ptr, uint32_t n,
                uint32_t begin_bit, uint32_t key_bits, const Layout& lay, const Launcher<K, VALS>& L, const Tuning& tn,
                bool* in_alt, hipStream_t s) {
  const uint32_t R = 1u << tn.rb;
  // hist, counters and rec0 are contiguous: clear them with one fill.
  const size_t clear_words = (reinterpret_cast<char*>(lay.rec0 + size_t(lay.tiles) * R) - reinterpret_cast<char*>(lay.hist)) / 4;
  vx_status st = fill_u32(lay.hist, clear_words, 0, s);
  if (st != VX_OK) return st;
  const uint32_t hblocks =
      std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, kHistBlock * kHistKpt), 2 * limits().cus));
  {
    TimedLaunch tl(VX_K_SORT_HISTOGRAM, s);
    if (tn.rb == 8)
      histogram_kernel<K, 8, COHERENT><<<hblocks, kHistBlock, 0, s>>>(keys, n_ptr, n, begin_bit, key_bits, lay.hist);
    else if (tn.rb == 9)
      histogram_kernel<K, 9, COHERENT><<<hblocks, kHistBlock, 0, s>>>(keys, n_ptr, n, begin_bit, key_bits, lay.hist);
    else
      histogram_kernel<K, 4, COHERENT><<<hblocks, kHistBlock, 0, s>>>(keys, n_ptr, n, begin_bit, key_bits, lay.hist);
  }
  if ((st = last_launch()) != VX_OK) return st;

  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(L.kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(L.lds));
  // Persistent blocks: as many as can be resident at once (LDS, registers
  // and waves per CU all count).
  int resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(L.kernel), int(L.block),
                                                   L.lds) != hipSuccess || resident < 1)
    resident = 1;
  const uint32_t per_cu = uint32_t(resident);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(lay.tiles, limits().cus * per_cu));

  K* src = keys;
  K* dst = alt_keys;
  uint32_t* vsrc = vals;
  uint32_t* vdst = alt_vals;
  for (uint32_t p = 0; p < lay.passes; ++p) {
    const uint32_t shift = begin_bit + p * tn.rb;
    const uint32_t bits = std::min(tn.rb, key_bits - p * tn.rb);
    uint32_t* rec = (p & 1) ? lay.rec1 : lay.rec0;
    uint32_t* rec_next = (p & 1) ? lay.rec0 : lay.rec1;
    {
      TimedLaunch tl(VX_K_SORT_PARTITION, s);
      L.kernel<<<grid, L.block, L.lds, s>>>(src, dst, vsrc, vdst, n_ptr, n, shift, (1u << bits) - 1,
                                            lay.hist + size_t(p) * R, lay.counters + p, rec, rec_next);
    }
    if ((st = last_launch()) != VX_OK) return st;
    std::swap(src, dst);
    std::swap(vsrc, vdst);
  }
  *in_alt = src != keys;
  return VX_OK;
}

template <class K, bool VALS>
size_t ws_bytes(uint32_t n, uint32_t key_bits) {
  if (n <= 1) return 0;
  const Tuning tn = tuning_for(sizeof(K), VALS);
  Launcher<K, VALS> L;
  if (!pick(L, tn)) return 0;
  return layout(nullptr, n, sizeof(K), key_bits, VALS, tn, L.tile, true).bytes;
}

template <class K, bool VALS>
vx_status run(K* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws, size_t ws_size, hipStream_t s) {
  if (n <= 1) return VX_OK;
  const Tuning tn = tuning_for(sizeof(K), VALS);
  Launcher<K, VALS> L;
  if (!pick(L, tn)) return VX_UNSUPPORTED;
  const Layout lay = layout(ws, n, sizeof(K), key_bits, VALS, tn, L.tile, true);
  if (!ws || ws_size < lay.bytes) return VX_INVALID_VALUE;
  bool in_alt = false;
  vx_status st = drive<K, VALS>(keys, vals, static_cast<K*>(lay.alt_keys), lay.alt_vals, nullptr, n, 0, key_bits, lay,
                                L, tn, &in_alt, s);
  if (st != VX_OK) return st;
  if (in_alt) {
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 8192);
    copy_kernel<K><<<blocks, 256, 0, s>>>(static_cast<K*>(lay.alt_keys), keys, n);
    if ((st = last_launch()) != VX_OK) return st;
    if constexpr (VALS) {
      copy_kernel<uint32_t><<<blocks, 256, 0, s>>>(lay.alt_vals, vals, n);
      if ((st = last_launch()) != VX_OK) return st;
    }
  }
  return VX_OK;
}

}  // namespace

size_t workspace_bytes(uint32_t n, uint32_t key_bytes, uint32_t key_bits, bool vals) {
  if (key_bytes == 8) return vals ? ws_bytes<uint64_t, true>(n, key_bits) : ws_bytes<uint64_t, false>(n, key_bits);
  return vals ? ws_bytes<uint32_t, true>(n, key_bits) : ws_bytes<uint32_t, false>(n, key_bits);
}

vx_status sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws, size_t ws_size, hipStream_t s) {
  return vals ? run<uint32_t, true>(keys, vals, n, key_bits, ws, ws_size, s)
              : run<uint32_t, false>(keys, nullptr, n, key_bits, ws, ws_size, s);
}

vx_status sort_u64(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws, size_t ws_size, hipStream_t s) {
  return vals ? run<uint64_t, true>(keys, vals, n, key_bits, ws, ws_size, s)
              : run<uint64_t, false>(keys, nullptr, n, key_bits, ws, ws_size, s);
}

uint32_t passes(uint32_t key_bits, const Tuning& tn) { return (key_bits + tn.rb - 1) / tn.rb; }

Tuning pairs_u64_dev_tuning() {
  Tuning tn = tuning_for(8);
  Launcher<uint64_t, true> L;
  if (!pick(L, tn)) tn = Tuning{8, 8, 1024};
  return tn;
}

size_t pairs_u64_dev_workspace_bytes(uint64_t n_upper, uint32_t key_bits, const Tuning& tn) {
  Launcher<uint64_t, true> L;
  if (!pick(L, tn)) return 0;
  return layout(nullptr, n_upper, 8, key_bits, true, tn, L.tile, false).bytes;
}

vx_status pairs_u64_dev(uint64_t* keys, uint32_t* vals, uint64_t* alt_keys, uint32_t* alt_vals, const uint32_t* n_ptr,
                        uint64_t n_upper, uint32_t key_bits, void* ws, const Tuning& tn, bool* in_alt,
                        hipStream_t s) {
  Launcher<uint64_t, true> L;
  if (!pick(L, tn)) return VX_UNSUPPORTED;
  if (n_upper >= (1u << 30)) return VX_INVALID_VALUE;
  const Layout lay = layout(ws, n_upper, 8, key_bits, true, tn, L.tile, false);
  return drive<uint64_t, true>(keys, vals, alt_keys, alt_vals, n_ptr, uint32_t(n_upper), 0, key_bits, lay, L, tn,
                               in_alt, s);
}

// Packed u64 keys sorted by bits [begin_bit, begin_bit + key_bits): the
// u64 keys-only tuning row; 8-bit digits where 9-bit ones save no pass.
Tuning keys_u64_dev_tuning(uint32_t key_bits) {
  Tuning tn = tuning_for(8, false);
  if (tn.rb == 9 && (key_bits + 8) / 9 == (key_bits + 7) / 8) {
    Launcher<uint64_t, false> L;
    const Tuning t8{8, tn.kpt, tn.block};
    if (pick(L, t8)) tn = t8;
  }
  Launcher<uint64_t, false> L;
  if (!pick(L, tn)) tn = Tuning{8, 8, 1024};
  return tn;
}

size_t keys_u64_dev_workspace_bytes(uint64_t n_upper, uint32_t key_bits, const Tuning& tn) {
  Launcher<uint64_t, false> L;
  if (!pick(L, tn)) return 0;
  return layout(nullptr, n_upper, 8, key_bits, false, tn, L.tile, false).bytes;
}

vx_status keys_u64_dev(uint64_t* keys, uint64_t* alt_keys, const uint32_t* n_ptr, uint64_t n_upper, uint32_t begin_bit,
                       uint32_t key_bits, void* ws, const Tuning& tn, bool* in_alt, hipStream_t s) {
  Launcher<uint64_t, false> L;
  if (!pick(L, tn)) return VX_UNSUPPORTED;
  if (n_upper >= (1u << 30) || begin_bit + key_bits > 64) return VX_INVALID_VALUE;
  const Layout lay = layout(ws, n_upper, 8, key_bits, false, tn, L.tile, false);
  return drive<uint64_t, false, true>(keys, nullptr, alt_keys, nullptr, n_ptr, uint32_t(n_upper), begin_bit, key_bits,
                                      lay, L, tn, in_alt, s);
}

vx_status set_tuning(uint32_t key_bytes, uint32_t rb, uint32_t kpt, uint32_t block) {
  const Tuning tn{rb, kpt, block};
  const bool vals = !(key_bytes & VX_SORT_KEYS_ONLY);
  const uint32_t kb = key_bytes & ~VX_SORT_KEYS_ONLY;
  bool ok;
  if (kb == 8) {
    Launcher<uint64_t, true> Lv;
    Launcher<uint64_t, false> Lk;
    ok = vals ? pick(Lv, tn) : pick(Lk, tn);
  } else if (kb == 4) {
    Launcher<uint32_t, true> Lv;
    Launcher<uint32_t, false> Lk;
    ok = vals ? pick(Lv, tn) : pick(Lk, tn);
  } else {
    return VX_INVALID_VALUE;
  }
  if (!ok) return VX_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_mu);
  g_tuning[tuning_row(kb, vals)] = tn;
  return VX_OK;
}

vx_status get_tuning(uint32_t key_bytes, uint32_t* rb, uint32_t* kpt, uint32_t* block) {
  const uint32_t kb = key_bytes & ~VX_SORT_KEYS_ONLY;
  if ((kb != 4 && kb != 8) || !rb || !kpt || !block) return VX_INVALID_VALUE;
  const Tuning tn = tuning_for(kb, !(key_bytes & VX_SORT_KEYS_ONLY));
  *rb = tn.rb;
  *kpt = tn.kpt;
  *block = tn.block;
  return VX_OK;
}

}  // namespace sort
}  // namespace vx

// ---- C ABI -------------------------------------------------------------------
extern "C" vx_status vx_sort_workspace_size(uint32_t n, uint32_t key_bytes, uint32_t key_bits, int32_t with_values,
                                            size_t* bytes) {
  if (!bytes || (key_bytes != 4 && key_bytes != 8) || key_bits == 0 || key_bits > key_bytes * 8) return VX_INVALID_VALUE;
  *bytes = vx::sort::workspace_bytes(n, key_bytes, key_bits, with_values != 0);
  return VX_OK;
}

static vx_status check_sort_args(const void* keys, uint32_t n, uint32_t key_bits, uint32_t max_bits) {
  if (!keys && n) return VX_INVALID_VALUE;
  if (key_bits == 0 || key_bits > max_bits) return VX_INVALID_VALUE;
  if (n >= (1u << 30)) return VX_INVALID_VALUE;
  return VX_OK;
}

extern "C" vx_status vx_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws,
                                       size_t ws_bytes, void* stream) {
  const vx_status st = check_sort_args(keys, n, key_bits, 32);
  if (st != VX_OK) return st;
  return vx::sort::sort_u32(keys, vals, n, key_bits, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* ws,
                                       size_t ws_bytes, void* stream) {
  const vx_status st = check_sort_args(keys, n, key_bits, 64);
  if (st != VX_OK) return st;
  return vx::sort::sort_u64(keys, vals, n, key_bits, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_sort_set_tuning(uint32_t key_bytes, uint32_t rb, uint32_t kpt, uint32_t block) {
  return vx::sort::set_tuning(key_bytes, rb, kpt, block);
}

extern "C" vx_status vx_sort_get_tuning(uint32_t key_bytes, uint32_t* rb, uint32_t* kpt, uint32_t* block) {
  return vx::sort::get_tuning(key_bytes, rb, kpt, block);
}
