// Device-side view and helpers of the gfx950 Hash DAG.
//
// Storage: two u32 word pools (nodes, leaves) and two open-addressing hash
// tables of u64 slots {fingerprint:32 | value:32}. A slot value is either a
// pool offset (>= kPoolStart) of an interned item, kTomb (a dead slot that
// never matches), or kTag | i: item i of the level currently being interned
// (only while that level's probe/commit kernels run).
//
// Refs (hash_dag_gpu/config.cuh:30-37): 0 EMPTY, 1 FULL, otherwise a pool
// offset; the level being walked tells which pool. Node items are
// [child_mask | fill_mask << 8, refs of children in child_mask, in index
// order] (detail/node.cuh:266-289). Leaf items are the occupancy u64 (2 words)
// or 64 u8 materials (16 words).
#pragma once

#include "vx_common.hpp"

namespace vx {
namespace dag {

constexpr uint32_t kEmpty = 0;
constexpr uint32_t kFull = 1;
constexpr uint32_t kTomb = 1;        // pool offset 1 never holds an item
constexpr uint32_t kPoolStart = 16;  // first item offset (keeps 16-word leaves 64-byte aligned)
constexpr uint32_t kTag = 0x80000000u;
constexpr uint32_t kNodeStride = 9;
constexpr uint32_t kMaxProbe = 1u << 20;

// Per-item interning result: kind in the high word.
enum : uint32_t { kResFinal = 0, kResDup = 1, kResRep = 2, kResPending = 3 };

enum : uint32_t { kStatusOk = 0, kStatusOverflow = 1 };

// Control words (device): pool usage, status, root and per-level snapshots.
enum : uint32_t {
  kCtlNodeUsed = 0,
  kCtlLeafUsed = 1,
  kCtlStatus = 2,
  kCtlRoot = 3,
  kCtlNodeBase = 4,
  kCtlLeafBase = 5,
  kCtlWords = 16
};

struct Dev {
  uint32_t* node_pool;
  uint32_t* leaf_pool;
  uint64_t* node_slots;
  uint64_t* leaf_slots;
  uint64_t node_mask, leaf_mask;
  uint64_t node_cap, leaf_cap;  // words
  uint32_t* ctl;
  // First-error word of the edit in flight: the scene's own ctl[kCtlStatus]
  // for direct calls; an executor points it into its workspace, so two
  // executors editing disjoint scene sets never share a status word.
  uint32_t* status;
  // Pool usage snapshots taken when a level's probe starts ([0] nodes, [1]
  // leaves; the commit places new items from there): ctl + kCtlNodeBase for
  // direct calls, the executor's workspace for an executor, for the same
  // reason as `status`.
  uint32_t* base;
  uint32_t depth;       // inner levels D; leaves sit at level D
  uint32_t leaf_words;  // 2 or 16
};

__host__ __device__ __forceinline__ uint64_t res_make(uint32_t kind, uint32_t v) { return (uint64_t(kind) << 32) | v; }
__host__ __device__ __forceinline__ uint32_t res_kind(uint64_t r) { return uint32_t(r >> 32); }

// Ref of item k of a level, following a duplicate to its representative.
__device__ __forceinline__ uint32_t resolve(const uint64_t* res, uint32_t k) {
  uint64_t r = res[k];
  if (res_kind(r) == kResDup) r = res[uint32_t(r)];
  return uint32_t(r);
}

__device__ __forceinline__ void set_status(const Dev& d, uint32_t v) { atomicCAS(d.status, kStatusOk, v); }

// child_ref (detail/node.cuh:148-175): sentinels propagate; otherwise decode
// the dense ref list with popcount.
__device__ __forceinline__ uint32_t child_of(const uint32_t* pool, uint32_t ref, uint32_t c) {
  if (ref <= kFull) return ref;
  const uint32_t h = pool[ref];
  const uint32_t cm = h & 0xffu, fm = (h >> 8) & 0xffu, bit = 1u << c;
  if (cm & bit) return pool[ref + 1 + __popc(cm & (bit - 1))];
  return (fm & bit) ? kFull : kEmpty;
}

__device__ __forceinline__ void children_of(const uint32_t* pool, uint32_t ref, uint32_t (&ch)[8]) {
  if (ref <= kFull) {
#pragma unroll
    for (int c = 0; c < 8; ++c) ch[c] = ref;
    return;
  }
  const uint32_t h = pool[ref];
  const uint32_t cm = h & 0xffu, fm = (h >> 8) & 0xffu;
  uint32_t k = ref + 1;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t bit = 1u << c;
    if (cm & bit) ch[c] = pool[k++];
    else ch[c] = (fm & bit) ? kFull : kEmpty;
  }
}

// build_item_from_children + the uniform collapse of canonicalize_item
// (detail/node.cuh:210-222, 266-289). Returns the item width, or 0 with *ref
// set when the node collapses to a sentinel.
__device__ __forceinline__ uint32_t make_node_item(const uint32_t (&ch)[8], uint32_t* item, uint32_t* ref) {
  uint32_t cm = 0, fm = 0, w = 1;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (ch[c] == kFull) fm |= 1u << c;
    else if (ch[c] != kEmpty) {
      cm |= 1u << c;
      item[w++] = ch[c];
    }
  }
  if (cm == 0 && fm == 0) {
    *ref = kEmpty;
    return 0;
  }
  if (cm == 0 && fm == 0xffu) {
    *ref = kFull;
    return 0;
  }
  item[0] = cm | (fm << 8);
  return w;
}

// make_node_item into a register array: every index is a compile-time
// constant (word q of the item is the child c whose rank among the table-
// backed children is q - 1), so the array stays in registers instead of
// going to scratch. Words past the width are zero.
__device__ __forceinline__ uint32_t make_node_item_regs(const uint32_t (&ch)[8], uint32_t (&item)[9], uint32_t* ref) {
  uint32_t cm = 0, fm = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (ch[c] == kFull) fm |= 1u << c;
    else if (ch[c] != kEmpty) cm |= 1u << c;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) item[q] = 0;
  if (cm == 0 && (fm == 0 || fm == 0xffu)) {
    *ref = fm ? kFull : kEmpty;
    return 0;
  }
  item[0] = cm | (fm << 8);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t rank = 1 + __popc(cm & ((1u << c) - 1));
    const bool on = (cm >> c) & 1u;
#pragma unroll
    for (int q = 1; q < 9; ++q)
      if (on && rank == uint32_t(q)) item[q] = ch[c];
  }
  return 1 + __popc(cm);
}

// Item hash (slot index from the low bits, fingerprint from the high 32):
// the words two at a time as 64-bit chunks, each folded in with one
// multiplication (h = (h ^ chunk) * K, then h ^= h >> 32: a bijection of h per
// chunk), and MurmurHash3's fmix64 avalanche at the end. 64-bit
// multiplications are the cost here (each is several quarter-rate 32-bit
// multiplies on this part): 10 for a 16-word leaf instead of 32 with a full
// mix per word.
__host__ __device__ __forceinline__ uint64_t hash_chunk(uint64_t h, uint64_t c) {
  h = (h ^ c) * 0x9E3779B97F4A7C15ull;
  return h ^ (h >> 32);
}
__host__ __device__ __forceinline__ uint64_t hash_final(uint64_t h) {
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDull;
  h ^= h >> 33;
  h *= 0xC4CEB9FE1A85EC53ull;
  return h ^ (h >> 33);
}

__device__ __forceinline__ uint64_t hash_words(const uint32_t* w, uint32_t n) {
  uint64_t h = 0x243F6A8885A308D3ull ^ n;
  uint32_t k = 0;
  for (; k + 1 < n; k += 2) h = hash_chunk(h, w[k] | (uint64_t(w[k + 1]) << 32));
  if (k < n) h = hash_chunk(h, w[k]);
  return hash_final(h);
}

// hash_words of a register array of up to kN words (constant indices).
template <uint32_t kN>
__device__ __forceinline__ uint64_t hash_words_regs(const uint32_t (&w)[kN], uint32_t n) {
  uint64_t h = 0x243F6A8885A308D3ull ^ n;
#pragma unroll
  for (uint32_t k = 0; k < kN; k += 2) {
    const uint64_t hi = k + 1 < kN ? uint64_t(w[k + 1 < kN ? k + 1 : k]) << 32 : 0ull;
    if (k + 1 < n) h = hash_chunk(h, w[k] | hi);
    else if (k < n) h = hash_chunk(h, w[k]);
  }
  return hash_final(h);
}

__device__ __forceinline__ bool words_equal(const uint32_t* a, const uint32_t* b, uint32_t n) {
  for (uint32_t k = 0; k < n; ++k)
    if (a[k] != b[k]) return false;
  return true;
}

// Leaf payload helpers. Material leaves: byte i of the 64 is voxel bit i
// (bit = x + 4y + 16z, detail/node.cuh:87-94), word i/4, byte i%4.
// (Bit b of the nibble lands on bit 8b of nib * 0x204081, no two partial
// products overlapping; x * 255 = (x << 8) - x turns each 0x01 byte into 0xff.)
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t nib) {
  const uint32_t x = (nib * 0x204081u) & 0x01010101u;
  return (x << 8) - x;
}

__host__ __device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__host__ __device__ __forceinline__ uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x | (x >> 4)) & 0x100f00f00f00f00full;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x | (x >> 16)) & 0x1f00000000ffffull;
  x = (x | (x >> 32)) & 0x1fffffull;
  return uint32_t(x);
}
__host__ __device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

inline uint32_t grid_for(uint64_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, 256), 4096)); }

// Smallest power of two >= v (slot counts).
uint64_t pow2_at_least(uint64_t v);

// A level's items: fixed stride; width from the header for node items.
struct ItemSet {
  const uint32_t* items;
  uint32_t stride;
  uint32_t fixed_width;  // 0 for node items
  __device__ __forceinline__ uint32_t width(const uint32_t* it) const {
    return fixed_width ? fixed_width : 1 + __popc(it[0] & 0xffu);
  }
};

// Scratch device buffer owned by a host scope (builders and maintenance calls
// that size their temporaries on the host).
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* get() const {
    return static_cast<T*>(p);
  }
};

inline vx_status dev_alloc(DevBuf& b, size_t bytes) {
  return hipMalloc(&b.p, bytes > 256 ? bytes : 256) == hipSuccess ? VX_OK : VX_OUT_OF_CAPACITY;
}

inline uint32_t read_u32(const uint32_t* p, hipStream_t s) {
  uint32_t v = 0;
  (void)hipMemcpyAsync(&v, p, 4, hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  return v;
}

// Probe + commit one level's items into the node (leaves=false) or leaf
// table. res[i] must be FINAL (sentinel) or PENDING (hashes[i] set) on entry;
// every entry is FINAL or DUP afterwards, and hashes[] is overwritten. fused_ws holds
// scan::fused_dev_workspace_bytes(n_ub).
vx_status intern(const Dev& d, bool leaves, ItemSet set, uint64_t* hashes, const uint32_t* n_ptr, uint64_t n_ub,
                 uint64_t* res, void* fused_ws, hipStream_t s);

// Pinned host memory for an edit's uploads (counts, scene table, brush jobs)
// and its status read-back, grown on demand and reused: every edit call
// synchronises its stream before it returns, so no copy from the buffer is in
// flight when the next call writes it.
struct HostStage {
  void* p = nullptr;
  size_t bytes = 0;
  size_t readback = 0;  // offset of the read-back words of the current call
  vx_status reserve(size_t want);
  HostStage() = default;
  HostStage(const HostStage&) = delete;
  HostStage& operator=(const HostStage&) = delete;
  ~HostStage();
};

}  // namespace dag
}  // namespace vx

// Host owner of one scene's device storage.
struct vx_dag {
  vx::dag::Dev dev{};
  uint32_t depth = 0;
  bool materials = false;
  vx_dag_capacity cap{};
  uint32_t* node_pool = nullptr;
  uint32_t* leaf_pool = nullptr;
  uint64_t* node_slots = nullptr;
  uint64_t* leaf_slots = nullptr;
  uint32_t* ctl = nullptr;
  // Bumped whenever the scene's buffers move (vx_dag_grow): an executor that
  // holds a copy of `dev` refreshes it, and re-captures its graph, when the
  // generation it copied is older.
  uint64_t generation = 0;
  // Optional per-stage and per-kernel timing of edits (hipEvents on the
  // edit's stream).
  bool profile = false;
  bool stage_events = true;  // per-stage events while profiled
  hipEvent_t ev[VX_DAG_STAGES + 1] = {};
  vx::KernelTimer* timer = nullptr;
  vx_dag_profile prof{};
  vx::dag::HostStage stage;
};
