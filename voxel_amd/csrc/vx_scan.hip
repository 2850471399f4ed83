// Single-pass prefix scan with decoupled look-back, written for wave64.
//
// Replaces algo::cuda::scan's default BlockBased<DecoupledLookback<
// WarpShuffleBlock<512,4>>> (cuda/scan/scan.cuh:22-23,
// cuda/scan/block_based/decoupled_lookback.cuh:116-222) and the fused
// transform -> scan -> post-scan kernel used for stream compaction
// (cuda/scan/fused.cuh:102-159).
//
// Layout: a tile is kBlock threads x 16 elements. Each wave owns a contiguous
// 1024-element segment that it reads as 4 rounds of 64 lanes x uint4, so every
// load instruction moves 1 KiB of consecutive bytes; loads and stores are
// streaming (nontemporal). A tile publishes its
// aggregate, then wave 0 inspects up to 64 predecessor records per step
// (ballot over the 64 lanes) instead of walking back one tile at a time.
#include <mutex>

#include "vx_scan.hpp"

namespace vx {

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) p[i] = v;
}

vx_status fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return VX_OK;
  const uint32_t blocks = uint32_t(std::min<uint64_t>(ceil_div(n, 256), 4096));
  TimedLaunch tl(VX_K_FILL, s);
  fill_u32_kernel<<<blocks, 256, 0, s>>>(p, n, v);
  return last_launch();
}

KernelTimer*& current_timer() {
  static thread_local KernelTimer* t = nullptr;
  return t;
}

namespace scan {

namespace {

template <Op OP>
__device__ __forceinline__ uint32_t combine(uint32_t a, uint32_t b) {
  if constexpr (OP == kSum) return a + b;
  else return a > b ? a : b;
}

template <Op OP>
__device__ __forceinline__ uint32_t wave_incl(uint32_t v) {
  if constexpr (OP == kSum) return wave_inclusive_sum(v);
  else return wave_inclusive_max(v);
}

// Scan of a tile whose kBlock*4K elements were loaded as K uint4 per lane in
// wave-segment order; returns, per lane and round, the exclusive prefix of the
// first element of the vector within the tile, and the tile aggregate.
template <Op OP, uint32_t kBlock, int K>
struct TileScan {
  static constexpr uint32_t kWaves = kBlock / kWave;
  uint32_t* s_wave;  // [kWaves] LDS

  __device__ uint32_t run(const uint4 (&v)[K], uint32_t (&excl)[K]) {
    const uint32_t lane = lane_id(), wave = wave_id();
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t s = combine<OP>(combine<OP>(v[k].x, v[k].y), combine<OP>(v[k].z, v[k].w));
      const uint32_t inc = wave_incl<OP>(s);
      excl[k] = combine<OP>(carry, wave_shift_up1(inc));
      carry = combine<OP>(carry, wave_last(inc));
    }
    if (lane == 0) s_wave[wave] = carry;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
      const uint32_t t = s_wave[w];
      if (w < wave) before = combine<OP>(before, t);
      total = combine<OP>(total, t);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) excl[k] = combine<OP>(before, excl[k]);
    return total;
  }
};

}  // namespace

// Decoupled look-back (lookback_window in vx_scan.hpp): records pack status
// (bits 32..33: 1 aggregate, 2 inclusive prefix) over a 32-bit value, so one
// 64-bit agent-scope load observes both, and wave 0 inspects 64 predecessors
// per step.

namespace {

// Each wave owns 64*4*K consecutive elements and moves them as K rounds of
// 64 lanes x uint4 (1 KiB per load instruction).
template <Op OP, bool INCL, uint32_t kBlock, int K>
__global__ __launch_bounds__(kBlock) void scan_vec_kernel(uint32_t* __restrict__ data, uint32_t n,
                                                         uint32_t* counter, uint64_t* rec) {
  constexpr uint32_t kSeg = kWave * 4 * K;
  constexpr uint32_t kTile = kBlock * 4 * K;
  __shared__ uint32_t s_wave[kBlock / kWave];
  __shared__ uint32_t s_tile, s_prefix;
  const bool multi = gridDim.x > 1;
  if (threadIdx.x == 0) s_tile = multi ? atomicAdd(counter, 1u) : 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint64_t seg = uint64_t(tile) * kTile + uint64_t(wave) * kSeg;

  uint4 v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t i = seg + k * 256u + lane * 4u;
    if (i + 3 < n) {
      v[k] = ld_stream(data + i);
    } else {
      v[k].x = i < n ? data[i] : 0u;
      v[k].y = i + 1 < n ? data[i + 1] : 0u;
      v[k].z = i + 2 < n ? data[i + 2] : 0u;
      v[k].w = 0u;
    }
  }
  uint32_t excl[K];
  TileScan<OP, kBlock, K> ts{s_wave};
  const uint32_t agg = ts.run(v, excl);

  if (wave == 0) {
    uint32_t prefix = 0;
    if (multi) {
      if (tile == 0) {
        if (lane == 0) st_agent(&rec[0], (uint64_t(2) << 32) | agg);
      } else {
        if (lane == 0) st_agent(&rec[tile], (uint64_t(1) << 32) | agg);
        prefix = lookback_window<typename OpOf<OP>::type>(rec, tile);
        if (lane == 0) st_agent(&rec[tile], (uint64_t(2) << 32) | combine<OP>(prefix, agg));
      }
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const uint32_t prefix = s_prefix;

#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t i = seg + k * 256u + lane * 4u;
    uint32_t a = combine<OP>(prefix, excl[k]);
    uint4 o;
    if constexpr (INCL) {
      a = combine<OP>(a, v[k].x); o.x = a;
      a = combine<OP>(a, v[k].y); o.y = a;
      a = combine<OP>(a, v[k].z); o.z = a;
      a = combine<OP>(a, v[k].w); o.w = a;
    } else {
      o.x = a; a = combine<OP>(a, v[k].x);
      o.y = a; a = combine<OP>(a, v[k].y);
      o.z = a; a = combine<OP>(a, v[k].z);
      o.w = a;
    }
    if (i + 3 < n) {
      st_stream(data + i, o);
    } else {
      if (i < n) data[i] = o.x;
      if (i + 1 < n) data[i + 1] = o.y;
      if (i + 2 < n) data[i + 2] = o.z;
    }
  }
}

// The smallest tile any configuration uses; the workspace is sized for it so
// that every tuning fits one workspace query.
constexpr uint32_t kMinScanTile = 256 * 16;

struct ScanTuning {
  uint32_t block = 256;
  uint32_t k = 8;
};
ScanTuning g_scan_tuning;

template <Op OP, bool INCL, uint32_t kBlock, int K>
vx_status launch_cfg(uint32_t* data, uint32_t n, void* ws, hipStream_t s) {
  constexpr uint32_t kTile = kBlock * 4 * K;
  const uint32_t tiles = ceil_div(n, kTile);
  uint32_t* counter = nullptr;
  uint64_t* rec = nullptr;
  if (tiles > 1) {
    Carver c(ws);
    counter = c.take<uint32_t>(64);
    rec = c.take<uint64_t>(tiles);
    const vx_status st = fill_u32(counter, (reinterpret_cast<char*>(rec + tiles) - reinterpret_cast<char*>(counter)) / 4, 0, s);
    if (st != VX_OK) return st;
  }
  scan_vec_kernel<OP, INCL, kBlock, K><<<tiles, kBlock, 0, s>>>(data, n, counter, rec);
  return last_launch();
}

template <Op OP, bool INCL>
vx_status launch(uint32_t* data, uint32_t n, void* ws, hipStream_t s) {
  const ScanTuning tn = g_scan_tuning;
  // A single tile never needs the workspace: use the smallest configuration
  // that covers n when the caller passed none.
  if (n <= kMinScanTile) return launch_cfg<OP, INCL, 256, 4>(data, n, ws, s);
  if (tn.block == 256 && tn.k == 4) return launch_cfg<OP, INCL, 256, 4>(data, n, ws, s);
  if (tn.block == 256 && tn.k == 16) return launch_cfg<OP, INCL, 256, 16>(data, n, ws, s);
  if (tn.block == 512 && tn.k == 4) return launch_cfg<OP, INCL, 512, 4>(data, n, ws, s);
  if (tn.block == 512 && tn.k == 8) return launch_cfg<OP, INCL, 512, 8>(data, n, ws, s);
  if (tn.block == 1024 && tn.k == 4) return launch_cfg<OP, INCL, 1024, 4>(data, n, ws, s);
  return launch_cfg<OP, INCL, 256, 8>(data, n, ws, s);
}

}  // namespace

size_t workspace_bytes(uint32_t n) {
  const uint32_t tiles = ceil_div(n, kMinScanTile);
  if (tiles <= 1) return 0;
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint64_t>(tiles);
  return c.size();
}

vx_status set_tuning(uint32_t block, uint32_t k) {
  const bool ok = (block == 256 && (k == 4 || k == 8 || k == 16)) || (block == 512 && (k == 4 || k == 8)) ||
                  (block == 1024 && k == 4);
  if (!ok) return VX_UNSUPPORTED;
  g_scan_tuning = ScanTuning{block, k};
  return VX_OK;
}

vx_status get_tuning(uint32_t* block, uint32_t* k) {
  if (!block || !k) return VX_INVALID_VALUE;
  *block = g_scan_tuning.block;
  *k = g_scan_tuning.k;
  return VX_OK;
}

vx_status run(uint32_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s) {
  if (n == 0) return VX_OK;
  if (reinterpret_cast<uintptr_t>(data) & 15) return generic_inplace(data, n, inclusive, op, ws, s);
  if (op == kSum) return inclusive ? launch<kSum, true>(data, n, ws, s) : launch<kSum, false>(data, n, ws, s);
  return inclusive ? launch<kMax, true>(data, n, ws, s) : launch<kMax, false>(data, n, ws, s);
}

// ---- u64 scans -----------------------------------------------------------------
// Same structure as the u32 scan (wave-contiguous segments, DPP wave scans,
// a wave-wide look-back window), for 64-bit elements. A u64 value leaves no
// room for status bits in the record, so a tile publishes {value, flag}: the
// value first, then the flag with release semantics; a reader loads the flag
// with acquire semantics before the value it selects (vals[2t] aggregate,
// vals[2t + 1] inclusive prefix).
namespace {

constexpr uint32_t kBlock64 = 256;
constexpr uint32_t kVec64 = 8;                            // uint4 (two elements) per lane and tile
constexpr uint32_t kTile64 = kBlock64 * 2 * kVec64;       // 4096 elements, 32 KiB
constexpr uint32_t kFlagAgg = 1, kFlagPrefix = 2;

template <Op OP>
__device__ __forceinline__ uint64_t combine64(uint64_t a, uint64_t b) {
  if constexpr (OP == kSum) return a + b;
  else return a > b ? a : b;
}

template <Op OP>
__device__ __forceinline__ uint64_t wave_incl64(uint64_t v) {
  return wave_inclusive_dpp(v, [](uint64_t a, uint64_t b) { return combine64<OP>(a, b); });
}

template <Op OP>
__device__ uint64_t lookback64(const uint32_t* flags, const uint64_t* vals, uint32_t tile) {
  const uint32_t lane = lane_id();
  uint64_t excl = 0;  // identity of sum and of max over u64
  int64_t base = int64_t(tile) - 1;
  while (base >= 0) {
    const int64_t idx = base - lane;
    const uint32_t st =
        idx >= 0 ? __hip_atomic_load(&flags[idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : kFlagPrefix;
    const uint64_t pref = ballot(st == kFlagPrefix);
    const uint64_t inval = ballot(st == 0);
    const uint32_t first = pref ? uint32_t(__builtin_ctzll(pref)) : 64u;
    const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1);
    if (inval & need) continue;
    uint64_t mine = 0;
    if (lane <= first && idx >= 0) mine = ld_agent(&vals[2 * idx + (st == kFlagPrefix ? 1 : 0)]);
    const uint64_t rev = __shfl(mine, int(63 - lane), kWave);
    excl = combine64<OP>(wave_last(wave_incl64<OP>(rev)), excl);
    if (first < 64) break;
    base -= 64;
  }
  return excl;
}

__device__ __forceinline__ void publish64(uint32_t* flag, uint64_t* val, uint64_t v, uint32_t f) {
  st_agent(val, v);
  __hip_atomic_store(flag, f, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <Op OP, bool INCL, bool ALIGNED>
__global__ __launch_bounds__(kBlock64) void scan64_kernel(uint64_t* __restrict__ data, uint32_t n, uint32_t* counter,
                                                          uint32_t* flags, uint64_t* vals) {
  constexpr uint32_t kSeg = kWave * 2 * kVec64;
  __shared__ uint64_t s_wave[kBlock64 / kWave];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const bool multi = gridDim.x > 1;
  if (threadIdx.x == 0) s_tile = multi ? atomicAdd(counter, 1u) : 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint64_t seg = uint64_t(tile) * kTile64 + uint64_t(wave) * kSeg;

  uint64_t a[kVec64], b[kVec64], excl[kVec64];
  uint64_t carry = 0;
#pragma unroll
  for (uint32_t k = 0; k < kVec64; ++k) {
    const uint64_t i = seg + k * (2 * kWave) + lane * 2;
    if (ALIGNED && i + 1 < n) {
      const uint4 v = ld_stream(reinterpret_cast<const uint32_t*>(data + i));
      a[k] = uint64_t(v.x) | (uint64_t(v.y) << 32);
      b[k] = uint64_t(v.z) | (uint64_t(v.w) << 32);
    } else {
      a[k] = i < n ? data[i] : 0ull;
      b[k] = i + 1 < n ? data[i + 1] : 0ull;
    }
    const uint64_t inc = wave_incl64<OP>(combine64<OP>(a[k], b[k]));
    excl[k] = combine64<OP>(carry, wave_shift_up1(inc));
    carry = combine64<OP>(carry, wave_last(inc));
  }
  if (lane == 0) s_wave[wave] = carry;
  __syncthreads();
  uint64_t before = 0, agg = 0;
#pragma unroll
  for (uint32_t w = 0; w < kBlock64 / kWave; ++w) {
    const uint64_t t = s_wave[w];
    if (w < wave) before = combine64<OP>(before, t);
    agg = combine64<OP>(agg, t);
  }
  if (wave == 0) {
    uint64_t prefix = 0;
    if (multi) {
      if (tile == 0) {
        if (lane == 0) publish64(&flags[0], &vals[1], agg, kFlagPrefix);
      } else {
        if (lane == 0) publish64(&flags[tile], &vals[2 * size_t(tile)], agg, kFlagAgg);
        prefix = lookback64<OP>(flags, vals, tile);
        if (lane == 0) publish64(&flags[tile], &vals[2 * size_t(tile) + 1], combine64<OP>(prefix, agg), kFlagPrefix);
      }
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const uint64_t prefix = combine64<OP>(s_prefix, before);
#pragma unroll
  for (uint32_t k = 0; k < kVec64; ++k) {
    const uint64_t i = seg + k * (2 * kWave) + lane * 2;
    const uint64_t e = combine64<OP>(prefix, excl[k]);
    const uint64_t oa = INCL ? combine64<OP>(e, a[k]) : e;
    const uint64_t ob = INCL ? combine64<OP>(combine64<OP>(e, a[k]), b[k]) : combine64<OP>(e, a[k]);
    if (ALIGNED && i + 1 < n) {
      st_stream(reinterpret_cast<uint32_t*>(data + i),
                make_uint4(uint32_t(oa), uint32_t(oa >> 32), uint32_t(ob), uint32_t(ob >> 32)));
    } else {
      if (i < n) data[i] = oa;
      if (i + 1 < n) data[i + 1] = ob;
    }
  }
}

template <Op OP, bool INCL>
vx_status launch64(uint64_t* data, uint32_t n, void* ws, hipStream_t s) {
  const uint32_t tiles = ceil_div(n, kTile64);
  uint32_t* counter = nullptr;
  uint32_t* flags = nullptr;
  uint64_t* vals = nullptr;
  if (tiles > 1) {
    Carver c(ws);
    counter = c.take<uint32_t>(64);
    flags = c.take<uint32_t>(tiles);
    vals = c.take<uint64_t>(2 * size_t(tiles));
    const vx_status st =
        fill_u32(counter, (reinterpret_cast<char*>(flags + tiles) - reinterpret_cast<char*>(counter)) / 4, 0, s);
    if (st != VX_OK) return st;
  }
  if (reinterpret_cast<uintptr_t>(data) & 15)
    scan64_kernel<OP, INCL, false><<<tiles, kBlock64, 0, s>>>(data, n, counter, flags, vals);
  else
    scan64_kernel<OP, INCL, true><<<tiles, kBlock64, 0, s>>>(data, n, counter, flags, vals);
  return last_launch();
}

}  // namespace

size_t workspace_bytes_u64(uint32_t n) {
  const uint32_t tiles = ceil_div(n, kTile64);
  if (tiles <= 1) return 0;
  Carver c(nullptr);
  c.take<uint32_t>(64);
  c.take<uint32_t>(tiles);
  c.take<uint64_t>(2 * size_t(tiles));
  return c.size();
}

vx_status run_u64(uint64_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s) {
  if (n == 0) return VX_OK;
  if (op == kSum) return inclusive ? launch64<kSum, true>(data, n, ws, s) : launch64<kSum, false>(data, n, ws, s);
  return inclusive ? launch64<kMax, true>(data, n, ws, s) : launch64<kMax, false>(data, n, ws, s);
}

// ---- generic fused scan (transform -> scan -> post-scan) ---------------------
namespace {

struct InplaceLoad {
  using Payload = NoPayload;
  const uint32_t* p;
  __device__ uint32_t operator()(uint32_t i, Payload&) const { return p[i]; }
};

template <Op OP>
struct InplaceStoreOp {
  uint32_t* p;
  bool inclusive;
  __device__ void operator()(uint32_t i, const NoPayload&, uint32_t, uint32_t excl, uint32_t incl, uint32_t) const {
    p[i] = inclusive ? incl : excl;
  }
};

struct FlagLoad {
  using Payload = NoPayload;
  const uint32_t* flags;
  __device__ uint32_t operator()(uint32_t i, Payload&) const { return flags[i] != 0u; }
};
struct CompactStore {
  const uint32_t* in;
  uint32_t* out;
  uint32_t* count;
  __device__ void operator()(uint32_t i, const NoPayload&, uint32_t v, uint32_t excl, uint32_t incl, uint32_t n) const {
    if (v) out[excl] = in[i];
    if (i == n - 1) *count = incl;
  }
};

}  // namespace

vx_status generic_inplace(uint32_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s) {
  if (op == kSum) return fused<kSum>(n, InplaceLoad{data}, InplaceStoreOp<kSum>{data, inclusive}, ws, s);
  return fused<kMax>(n, InplaceLoad{data}, InplaceStoreOp<kMax>{data, inclusive}, ws, s);
}

vx_status compact(const uint32_t* in, const uint32_t* flags, uint32_t n, uint32_t* out, uint32_t* d_count, void* ws,
                  hipStream_t s) {
  if (n == 0) return fill_u32(d_count, 1, 0, s);
  return fused<kSum>(n, FlagLoad{flags}, CompactStore{in, out, d_count}, ws, s);
}

}  // namespace scan
}  // namespace vx

// ---- scratch for the allocation-free entry point ------------------------------
// vx_exclusive_sum_u32_noalloc keeps its look-back state (claim counter and
// one record per tile) in this file-scope device array instead of a caller
// workspace: 256 bytes of counters and one 8-byte record per tile of the
// smallest tile (4096 elements), for up to 2^28 elements — 512 KiB per
// device, held until the process ends. Host code never passes `scratch`
// itself: its device address, per device, comes from hipGetSymbolAddress.
namespace {
constexpr uint32_t kScratchMaxN = 1u << 28;
constexpr size_t kScratchBytes = 256 + size_t(kScratchMaxN / (256 * 16)) * 8;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
vx_status scratch_for_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return VX_DEVICE_ERROR;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VX_DEVICE_ERROR;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return VX_OK;
}
}  // namespace

// ---- C ABI -------------------------------------------------------------------
using vx::scan::Op;

extern "C" vx_status vx_exclusive_sum_u32_noalloc(uint32_t* data, uint32_t n, void* stream) {
  if (!data && n) return VX_INVALID_VALUE;
  if (n > kScratchMaxN) return VX_INVALID_VALUE;
  const size_t need = std::max(vx::scan::workspace_bytes(n), vx::scan::fused_workspace_bytes(n));
  if (need > kScratchBytes) return VX_INVALID_VALUE;
  void* ws = nullptr;
  const vx_status st = scratch_for_current_device(&ws);
  if (st != VX_OK) return st;
  // The scan clears the counter and look-back records it uses with a fill
  // kernel before its own launch.
  return vx::scan::run(data, n, false, vx::scan::kSum, ws, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_scan_workspace_size(uint32_t n, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  *bytes = std::max(vx::scan::workspace_bytes(n), vx::scan::fused_workspace_bytes(n));
  return VX_OK;
}

static vx_status scan_entry(uint32_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream, bool incl, Op op) {
  if (!data && n) return VX_INVALID_VALUE;
  size_t need = 0;
  vx_scan_workspace_size(n, &need);
  if (need && (!ws || ws_bytes < need)) return VX_INVALID_VALUE;
  return vx::scan::run(data, n, incl, op, ws, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_exclusive_sum_u32(uint32_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry(d, n, ws, b, s, false, vx::scan::kSum);
}
extern "C" vx_status vx_inclusive_sum_u32(uint32_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry(d, n, ws, b, s, true, vx::scan::kSum);
}
extern "C" vx_status vx_exclusive_max_u32(uint32_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry(d, n, ws, b, s, false, vx::scan::kMax);
}
extern "C" vx_status vx_inclusive_max_u32(uint32_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry(d, n, ws, b, s, true, vx::scan::kMax);
}

extern "C" vx_status vx_scan_workspace_size_u64(uint32_t n, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  *bytes = vx::scan::workspace_bytes_u64(n);
  return VX_OK;
}

static vx_status scan_entry_u64(uint64_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream, bool incl,
                                Op op) {
  if (!data && n) return VX_INVALID_VALUE;
  if ((reinterpret_cast<uintptr_t>(data) & 7) != 0) return VX_INVALID_VALUE;
  const size_t need = vx::scan::workspace_bytes_u64(n);
  if (need && (!ws || ws_bytes < need)) return VX_INVALID_VALUE;
  return vx::scan::run_u64(data, n, incl, op, ws, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_exclusive_sum_u64(uint64_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry_u64(d, n, ws, b, s, false, vx::scan::kSum);
}
extern "C" vx_status vx_inclusive_sum_u64(uint64_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry_u64(d, n, ws, b, s, true, vx::scan::kSum);
}
extern "C" vx_status vx_exclusive_max_u64(uint64_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry_u64(d, n, ws, b, s, false, vx::scan::kMax);
}
extern "C" vx_status vx_inclusive_max_u64(uint64_t* d, uint32_t n, void* ws, size_t b, void* s) {
  return scan_entry_u64(d, n, ws, b, s, true, vx::scan::kMax);
}

extern "C" vx_status vx_scan_set_tuning(uint32_t block, uint32_t vectors_per_lane) {
  return vx::scan::set_tuning(block, vectors_per_lane);
}

extern "C" vx_status vx_scan_get_tuning(uint32_t* block, uint32_t* vectors_per_lane) {
  return vx::scan::get_tuning(block, vectors_per_lane);
}

extern "C" vx_status vx_compact_u32(const uint32_t* in, const uint32_t* flags, uint32_t n, uint32_t* out,
                                    uint32_t* d_count, void* ws, size_t ws_bytes, void* stream) {
  if (!d_count || (n && (!in || !flags || !out))) return VX_INVALID_VALUE;
  size_t need = 0;
  vx_scan_workspace_size(n, &need);
  if (need && (!ws || ws_bytes < need)) return VX_INVALID_VALUE;
  return vx::scan::compact(in, flags, n, out, d_count, ws, static_cast<hipStream_t>(stream));
}
