// Shared device helpers for the gfx950 kernels (wave64, CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/vx.h"

namespace vx {

constexpr uint32_t kWave = 64;

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

inline vx_status from_hip(hipError_t e) { return e == hipSuccess ? VX_OK : VX_DEVICE_ERROR; }
inline vx_status last_launch() { return from_hip(hipGetLastError()); }

// Carves consecutive 256-byte aligned regions out of a caller workspace.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
  size_t size() const { return align_up(off, 256); }
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / kWave; }

// Number of set bits of `mask` below this lane (mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <class T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, kWave); }

// ---- wave64 scans on DPP --------------------------------------------------
// Cross-lane moves through DPP (data-parallel primitives: the VALU reads a
// neighbour lane's register in the same instruction) instead of ds_bpermute
// round trips through LDS. Controls (GFX9 encoding): row_shr:k shifts within
// each 16-lane row, row_bcast:15 / row_bcast:31 carry lane 15 of a row (lane
// 31 of the wave) into the next row(s), wave_shr:1 shifts the whole wave by
// one lane. Lanes without a source read 0, the identity of every monoid
// scanned here (sum, max, segmented sum, pair sum).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, ROW_MASK, 0xf, true));
}
template <int CTRL, int ROW_MASK = 0xf, class T>
__device__ __forceinline__ T dpp_move(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
  if constexpr (sizeof(T) == 4) {
    return T(dpp_u32<CTRL, ROW_MASK>(uint32_t(v)));
  } else {
    const uint64_t u = uint64_t(v);
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>(uint32_t(u));
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>(uint32_t(u >> 32));
    return T((uint64_t(hi) << 32) | lo);
  }
}

// Inclusive scan over the 64 lanes of a wave for a monoid whose identity is
// zero: combine(earlier, later).
template <class T, class F>
__device__ __forceinline__ T wave_inclusive_dpp(T v, F combine) {
  v = combine(dpp_move<kDppRowShr1>(v), v);
  v = combine(dpp_move<kDppRowShr2>(v), v);
  v = combine(dpp_move<kDppRowShr4>(v), v);
  v = combine(dpp_move<kDppRowShr8>(v), v);
  v = combine(dpp_move<kDppRowBcast15, 0xa>(v), v);
  v = combine(dpp_move<kDppRowBcast31, 0xc>(v), v);
  return v;
}

// The value of the lane below (0 in lane 0): inclusive -> exclusive.
template <class T>
__device__ __forceinline__ T wave_shift_up1(T v) {
  return dpp_move<kDppWaveShr1>(v);
}

// Lane 63's value in every lane.
template <class T>
__device__ __forceinline__ T wave_last(T v) {
  if constexpr (sizeof(T) == 4) {
    return T(__builtin_amdgcn_readlane(int(v), 63));
  } else {
    const uint64_t u = uint64_t(v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), 63));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), 63));
    return T((uint64_t(hi) << 32) | lo);
  }
}

// Inclusive wave scan (sum) over 64 lanes.
template <class T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
  return wave_inclusive_dpp(v, [](T a, T b) { return T(a + b); });
}

template <class T>
__device__ __forceinline__ T wave_inclusive_max(T v) {
  return wave_inclusive_dpp(v, [](T a, T b) { return a > b ? a : b; });
}

// Streaming (nontemporal) 16-byte loads and stores for data touched once:
// the accesses carry the nontemporal hint, so a pass over gigabytes does not
// evict the caches' working set (measured: an in-place scan of 1 GiB runs
// 10% faster than with plain loads and stores).
typedef uint32_t u32x4_nt __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ld_stream(const uint32_t* p) {
  const u32x4_nt t = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt*>(p));
  return make_uint4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void st_stream(uint32_t* p, uint4 v) {
  u32x4_nt t;
  t.x = v.x;
  t.y = v.y;
  t.z = v.z;
  t.w = v.w;
  __builtin_nontemporal_store(t, reinterpret_cast<u32x4_nt*>(p));
}
template <class T>
__device__ __forceinline__ T ld_stream1(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_stream1(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Plain load from device (global) memory through a pointer whose address
// space the compiler cannot see (one read out of a descriptor): a generic
// pointer becomes a flat access, which also occupies the LDS wait counter,
// so every wait for it drains both counters.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ uint4 gld4(const uint32_t* p) {  // 16-byte aligned
  typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
  const v4u32 v = *(const __attribute__((address_space(1))) v4u32*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Device-coherent (agent scope) relaxed atomics: every XCD sees the value,
// which plain loads through a per-XCD L2 would not guarantee.
template <class T>
__device__ __forceinline__ T ld_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Counter-based splitmix64 stream shared by all generators.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed, uint64_t stream) {
  return mix64(seed ^ mix64(stream));
}
__host__ __device__ __forceinline__ uint64_t stream_at(uint64_t base, uint64_t i) {
  return mix64(base + i * 0x9E3779B97F4A7C15ull);
}
__host__ __device__ __forceinline__ uint32_t uniform_u32(uint64_t r, uint32_t lo, uint32_t hi) {
  const uint64_t span = uint64_t(hi) - lo + 1;
  return lo + uint32_t(((r >> 32) * span) >> 32);
}

// Optional per-launch timing. While a KernelTimer is installed on the calling
// host thread, every launch bracketed by a TimedLaunch gets a hipEvent pair
// on its stream, tagged with its kernel kind (VX_K_*). Scopes do not nest:
// an inner TimedLaunch inside an outer one is a no-op.
// With period p > 1, of the launches of a timed kind within one call only
// the one at index (phase mod p) gets events, phase advancing by one per
// call: over p calls every position is timed once, with one event pair per
// call instead of one per launch.
struct KernelTimer {
  static constexpr uint32_t kEvents = 1024;
  hipEvent_t ev[kEvents] = {};
  uint32_t kind[kEvents / 2] = {};
  uint32_t used = 0;
  uint32_t depth = 0;
  uint64_t mask = ~0ull;  // kinds that are timed
  uint32_t period = 1;
  uint32_t phase = 0;
  uint32_t seen[64] = {};  // launches of each kind in the current call
};
KernelTimer*& current_timer();

struct TimedLaunch {
  KernelTimer* t = nullptr;
  uint32_t slot = 0;
  hipStream_t s;
  TimedLaunch(uint32_t kind, hipStream_t stream) : s(stream) {
    KernelTimer* k = current_timer();
    if (!k || k->depth++ > 0 || k->used + 2 > KernelTimer::kEvents || !((k->mask >> kind) & 1u)) return;
    if (k->period > 1 && (k->seen[kind & 63]++ % k->period) != k->phase % k->period) return;
    t = k;
    slot = k->used;
    k->kind[slot / 2] = kind;
    if (hipEventRecord(k->ev[slot], s) != hipSuccess) t = nullptr;
  }
  ~TimedLaunch() {
    KernelTimer* k = current_timer();
    if (k) --k->depth;
    if (t && hipEventRecord(t->ev[slot + 1], s) == hipSuccess) t->used = slot + 2;
  }
};

// Zero a u32 range (graph-capturable replacement for hipMemsetAsync).
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v);
vx_status fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

// Scan building blocks used by other translation units.
namespace scan {
// Returns bytes of workspace the single-pass scan needs for n elements.
size_t workspace_bytes(uint32_t n);
enum Op : uint32_t { kSum = 0, kMax = 1 };
vx_status run(uint32_t* data, uint32_t n, bool inclusive, Op op, void* ws, hipStream_t s);
}  // namespace scan

}  // namespace vx
