// Slab hash set (u64 keys), slab hash map (u32 -> u32, packed 64-bit slots)
// and cuckoo hash set (u64 keys) for gfx950.
//
// Slab tables follow algo::cuda slab_hash (slab_hash/set/device.cuh:32-316,
// slab_hash/map/device.cuh:49-393): a bucket is a chain of slabs, keys come
// from a ballot work queue, and ballots find matches and empty slots. A slab
// is 64 u64 words = 512 bytes, words 0..62 hold slots and word 63 the
// next-slab link (the reference's 32-lane slab holds 31 u32 slots + next);
// a wave walks four keys at once, 16 lanes per key. Status codes are the reference's
// (slab_hash/common.cuh:18-25). Slabs beyond the bucket heads come from a
// bump allocator (hash_table/bump_allocator.cuh:84-121).
//
// The cuckoo set follows cuckoo_hash.cuh:73-120 (four salted hash functions,
// atomicExch eviction chains), storing the u64 key itself in the slot so one
// 64-bit exchange moves a whole entry.
#include <vector>

#include "../../include/vx_slab_alloc_device.hpp"
#include "vx_common.hpp"

namespace vx {
namespace hash {

constexpr uint64_t kEmpty64 = ~0ull;
constexpr uint64_t kTomb64 = ~0ull - 1;
constexpr uint32_t kNull = 0xffffffffu;
constexpr uint32_t kSlotLanes = 63;

enum : uint32_t {
  kNotFound = 0,
  kInserted = 1,
  kReplaced = 2,
  kErased = 3,
  kOverflow = 4,
  kInvalidKey = 5,
  kFound = 6,
  kAlreadyPresent = 7
};

struct SlabTable {
  uint64_t* slabs;     // [slab][64]
  uint32_t buckets;
  uint32_t capacity;   // total slabs (buckets + overflow)
  vx_slab_alloc_view alloc;  // overflow slabs: slab buckets + ordinal
};

__host__ __device__ __forceinline__ uint32_t bucket_of(uint64_t key, uint32_t buckets) {
  return uint32_t(mix64(key) % buckets);
}

// An overflow slab from the table's slab allocator (SlabAlloc, the
// reference's default, slab_hash/set/host.cuh:107): overflow slabs are
// pre-cleared to ~0, and the hint spreads concurrent callers over the bitmap.
__device__ __forceinline__ uint32_t alloc_slab(const SlabTable& t, uint32_t hint) {
  const uint32_t o = vx_slab_alloc_dev::allocate(t.alloc, hint);
  return o == vx_slab_alloc_dev::kNull ? kNull : t.buckets + o;
}

// Link word values (low 32 bits of lane 63's word) besides slab ids.
constexpr uint32_t kLinkLock = 0xfffffffeu;  // a wave is allocating the next slab
constexpr uint32_t kLinkFull = 0xfffffffdu;  // the allocator was exhausted here

// ---- wave groups -------------------------------------------------------------
// A wave works on four keys at once: a group of 16 lanes per key, each lane
// holding words gl, gl+16, gl+32, gl+48 of the key's current slab (four
// independent loads, one 128-byte line per load instruction across the
// group). A slab walk is a chain of dependent memory round trips, so four
// walks in flight per wave quadruple the keys per round trip of the
// reference's one-key-per-warp queue (slab_hash/set/device.cuh:32-316).
constexpr uint32_t kGroupLanes = 16;
constexpr uint32_t kGroupsPerWave = kWave / kGroupLanes;  // keys in flight per wave
constexpr uint32_t kLaneWords = kWave / kGroupLanes;      // slab words per lane

struct Group {
  uint32_t g, gl, first;  // group of this lane, lane within the group, group's first lane
};

__device__ __forceinline__ Group my_group() {
  const uint32_t l = lane_id();
  return Group{l / kGroupLanes, l % kGroupLanes, (l / kGroupLanes) * kGroupLanes};
}

// Lowest slab word index (0..62) for which the group's predicate holds, or 64.
// Word (lane gl, part q) is gl + 16 q. Executed by every lane of the wave.
__device__ __forceinline__ uint32_t group_first(const Group& G, const bool (&p)[kLaneWords]) {
  uint32_t best = 64;
#pragma unroll
  for (int q = kLaneWords - 1; q >= 0; --q) {
    const uint32_t bits = uint32_t(ballot(p[q]) >> G.first) & ((1u << kGroupLanes) - 1u);
    if (bits) best = kGroupLanes * uint32_t(q) + uint32_t(__builtin_ctz(bits));
  }
  return best;
}

// Lane src_j of the wave's j-th pending key (64 if none): the same for every
// lane, computed from the wave-uniform queue.
__device__ __forceinline__ uint32_t pending_lane(uint64_t queue, uint32_t j) {
  uint32_t src = 64;
#pragma unroll
  for (uint32_t x = 0; x < kGroupsPerWave; ++x) {
    const uint32_t c = queue ? uint32_t(__builtin_ctzll(queue)) : 64u;
    if (x == j) src = c;
    if (queue) queue &= queue - 1;
  }
  return src;
}

// Called by a whole group at a slab whose link is not a slab id. The group's
// first lane locks the link, allocates and publishes the id in one step, so no
// slab is ever allocated and lost to a race. A group that finds the link
// locked gets kLinkLock back and re-reads the slab instead of waiting (the
// holder may be another group of the same wave, which can only proceed once
// this one yields). Returns a slab id, kLinkLock, or kNull when the allocator
// is exhausted.
__device__ uint32_t group_link(const SlabTable& t, uint64_t* s, const Group& G) {
  uint32_t link = kNull;
  if (G.gl == 0) {
    const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&s[63]), (unsigned long long)kEmpty64,
                                    (unsigned long long)(0xffffffff00000000ull | kLinkLock));
    if (prev == kEmpty64) {
      const uint32_t id = alloc_slab(t, blockIdx.x * blockDim.x + threadIdx.x);
      link = id == kNull ? kLinkFull : id;
      st_agent(&s[63], id == kNull ? uint64_t(0xffffffff00000000ull | kLinkFull) : uint64_t(id));
    } else {
      link = uint32_t(prev);
    }
  }
  link = __shfl(link, int(G.first), kWave);
  return (link == kLinkFull || link == kNull) ? kNull : link;
}

// Bound on the rounds of one walk. A round either moves to the next slab,
// ends the walk, or re-reads the slab because another walker changed it (won
// a slot, linked a slab): every such retry is progress of some other key, and
// at most 63 slots per slab can be won, so the walks of one launch end well
// within this bound; it only guards against a corrupted chain.
__device__ __forceinline__ uint64_t walk_guard(const SlabTable& t) { return (uint64_t(t.capacity) + 1) * 256 + 1024; }

// The group's words of slab `s` (kEmpty64 for a group with nothing to do).
__device__ __forceinline__ void group_load(const uint64_t* s, const Group& G, bool idle, uint64_t (&v)[kLaneWords]) {
#pragma unroll
  for (uint32_t q = 0; q < kLaneWords; ++q) v[q] = idle ? kEmpty64 : ld_agent(&s[G.gl + kGroupLanes * q]);
}

// This lane's copy of word part q = w / kGroupLanes (the owner lane of word w
// holds word w there).
__device__ __forceinline__ uint64_t own_word(const uint64_t (&v)[kLaneWords], uint32_t w) {
  const uint32_t q = w / kGroupLanes;
  uint64_t r = v[0];
#pragma unroll
  for (uint32_t x = 1; x < kLaneWords; ++x)
    if (q == x) r = v[x];
  return r;
}

// ---- set --------------------------------------------------------------------
enum class SetOp { kInsert, kContains, kErase };

// Keys of K bits (u64, or u32 stored zero-extended); the two largest values
// of K are reserved (empty and deleted slots, slab_hash/common.cuh:10-30).
template <class K>
__device__ __forceinline__ bool reserved_key(uint64_t key) {
  return key == uint64_t(K(~K(0))) || key == uint64_t(K(~K(0)) - 1);
}

template <SetOp OP, class K = uint64_t>
__global__ void slab_set_kernel(SlabTable t, const K* __restrict__ keys, uint32_t n, uint8_t* found,
                                uint32_t* status) {
  const uint32_t lane = lane_id();
  const Group G = my_group();
  const uint32_t waves = (gridDim.x * blockDim.x) / kWave;
  const uint32_t wave_global = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  for (uint32_t base = wave_global * kWave; base < n; base += waves * kWave) {
    const uint32_t i = base + lane;
    uint64_t key = i < n ? uint64_t(keys[i]) : 0;
    bool active = i < n;
    uint32_t st = kNotFound;
    if (active && reserved_key<K>(key)) {
      st = kInvalidKey;
      active = false;
    }
    uint64_t queue = ballot(active);
    while (queue) {
      const uint32_t src = pending_lane(queue, G.g);
      bool done = src == 64;
      const uint64_t k = shfl(key, done ? 0 : int(src));
      uint32_t slab = done ? 0u : bucket_of(k, t.buckets);
      uint32_t result = kOverflow;
      for (uint64_t guard = 0, limit = walk_guard(t); guard < limit && ballot(!done); ++guard) {
        uint64_t* s = t.slabs + size_t(slab) * kWave;
        uint64_t v[kLaneWords];
        group_load(s, G, done, v);
        bool pm[kLaneWords], pe[kLaneWords];
#pragma unroll
        for (uint32_t q = 0; q < kLaneWords; ++q) {
          const bool slot = !done && G.gl + kGroupLanes * q < kSlotLanes;
          pm[q] = slot && v[q] == k;
          pe[q] = slot && v[q] == kEmpty64;
        }
        const uint32_t m = group_first(G, pm);
        const uint32_t e = OP == SetOp::kInsert ? group_first(G, pe) : 64u;
        const uint32_t link = uint32_t(__shfl(v[kLaneWords - 1], int(G.first + kGroupLanes - 1), kWave));
        if (done) continue;
        if (m < 64) {
          if constexpr (OP == SetOp::kInsert) {
            result = kAlreadyPresent;
            done = true;
          } else if constexpr (OP == SetOp::kContains) {
            result = kFound;
            done = true;
          } else {
            const uint32_t owner = G.first + m % kGroupLanes;
            uint32_t ok = 0;
            if (lane == owner)
              ok = atomicCAS(reinterpret_cast<unsigned long long*>(&s[m]), (unsigned long long)k,
                             (unsigned long long)kTomb64) == k;
            if (__shfl(ok, int(owner), kWave)) {
              result = kErased;
              done = true;
            }  // else raced: re-read this slab
          }
          continue;
        }
        if constexpr (OP == SetOp::kInsert) {
          if (e < 64) {
            const uint32_t owner = G.first + e % kGroupLanes;
            uint32_t ok = 0;
            if (lane == owner)
              ok = atomicCAS(reinterpret_cast<unsigned long long*>(&s[e]), (unsigned long long)kEmpty64,
                             (unsigned long long)k) == kEmpty64;
            if (__shfl(ok, int(owner), kWave)) {
              result = kInserted;
              done = true;
            }  // else lost the slot: re-read (the winner may hold the same key)
            continue;
          }
        }
        uint32_t next = link;
        if (next >= kLinkFull) {  // null, being linked, or exhausted
          if constexpr (OP != SetOp::kInsert) {
            result = kNotFound;
            done = true;
            continue;
          } else {
            next = next == kLinkFull ? kNull : group_link(t, s, G);
            if (next == kNull) {
              result = kOverflow;
              done = true;
              continue;
            }
            if (next == kLinkLock) continue;  // being linked elsewhere: re-read
          }
        }
        slab = next;
      }
#pragma unroll
      for (uint32_t j = 0; j < kGroupsPerWave; ++j) {
        const uint32_t r = __shfl(result, int(j * kGroupLanes), kWave);
        if (lane == pending_lane(queue, j)) {
          st = r;
          active = false;
        }
      }
      queue = ballot(active);
    }
    if (i < n) {
      if (status) status[i] = st;
      if constexpr (OP == SetOp::kContains)
        if (found) found[i] = st == kFound;
    }
  }
}

// ---- map (u32 -> u32 packed in one u64 slot) ----------------------------------
enum class MapOp { kInsertOrReplace, kFind, kErase };

__device__ __forceinline__ uint32_t slot_key(uint64_t v) { return uint32_t(v >> 32); }

template <MapOp OP>
__global__ void slab_map_kernel(SlabTable t, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals_in,
                                uint32_t* vals_out, uint32_t n, uint32_t* status) {
  const uint32_t lane = lane_id();
  const Group G = my_group();
  const uint32_t waves = (gridDim.x * blockDim.x) / kWave;
  const uint32_t wave_global = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  constexpr uint32_t kEmptyKey = 0xffffffffu, kTombKey = 0xfffffffeu;
  for (uint32_t base = wave_global * kWave; base < n; base += waves * kWave) {
    const uint32_t i = base + lane;
    const uint32_t key = i < n ? keys[i] : 0u;
    const uint32_t val = (OP == MapOp::kInsertOrReplace && i < n) ? vals_in[i] : 0u;
    bool active = i < n;
    uint32_t st = kNotFound, outv = 0xffffffffu;
    if (active && (key == kEmptyKey || key == kTombKey)) {
      st = kInvalidKey;
      active = false;
    }
    uint64_t queue = ballot(active);
    while (queue) {
      const uint32_t src = pending_lane(queue, G.g);
      bool done = src == 64;
      const uint32_t k = shfl(key, done ? 0 : int(src));
      const uint32_t kv = shfl(val, done ? 0 : int(src));
      const uint64_t packed = (uint64_t(k) << 32) | kv;
      uint32_t slab = done ? 0u : bucket_of(k, t.buckets);
      uint32_t result = kOverflow, found_val = 0xffffffffu;
      for (uint64_t guard = 0, limit = walk_guard(t); guard < limit && ballot(!done); ++guard) {
        uint64_t* s = t.slabs + size_t(slab) * kWave;
        uint64_t v[kLaneWords];
        group_load(s, G, done, v);
        bool pm[kLaneWords], pe[kLaneWords];
#pragma unroll
        for (uint32_t q = 0; q < kLaneWords; ++q) {
          const bool slot = !done && G.gl + kGroupLanes * q < kSlotLanes;
          pm[q] = slot && slot_key(v[q]) == k;
          pe[q] = slot && v[q] == kEmpty64;
        }
        const uint32_t m = group_first(G, pm);
        const uint32_t e = OP == MapOp::kInsertOrReplace ? group_first(G, pe) : 64u;
        const uint32_t link = uint32_t(__shfl(v[kLaneWords - 1], int(G.first + kGroupLanes - 1), kWave));
        if (done) continue;
        if (m < 64) {
          const uint32_t owner = G.first + m % kGroupLanes;
          const uint64_t mine = own_word(v, m);
          if constexpr (OP == MapOp::kFind) {
            found_val = uint32_t(__shfl(mine, int(owner), kWave));
            result = kFound;
            done = true;
          } else {
            const uint64_t repl = OP == MapOp::kErase ? ((uint64_t(kTombKey) << 32) | 0xffffffffu) : packed;
            uint32_t ok = 0;
            if (lane == owner)
              ok = atomicCAS(reinterpret_cast<unsigned long long*>(&s[m]), (unsigned long long)mine,
                             (unsigned long long)repl) == mine;
            if (__shfl(ok, int(owner), kWave)) {
              result = OP == MapOp::kErase ? kErased : kReplaced;
              done = true;
            }
          }
          continue;
        }
        if constexpr (OP == MapOp::kInsertOrReplace) {
          if (e < 64) {
            const uint32_t owner = G.first + e % kGroupLanes;
            uint32_t ok = 0;
            if (lane == owner)
              ok = atomicCAS(reinterpret_cast<unsigned long long*>(&s[e]), (unsigned long long)kEmpty64,
                             (unsigned long long)packed) == kEmpty64;
            if (__shfl(ok, int(owner), kWave)) {
              result = kInserted;
              done = true;
            }
            continue;
          }
        }
        uint32_t next = link;
        if (next >= kLinkFull) {  // null, being linked, or exhausted
          if constexpr (OP != MapOp::kInsertOrReplace) {
            result = kNotFound;
            done = true;
            continue;
          } else {
            next = next == kLinkFull ? kNull : group_link(t, s, G);
            if (next == kNull) {
              result = kOverflow;
              done = true;
              continue;
            }
            if (next == kLinkLock) continue;
          }
        }
        slab = next;
      }
#pragma unroll
      for (uint32_t j = 0; j < kGroupsPerWave; ++j) {
        const uint32_t r = __shfl(result, int(j * kGroupLanes), kWave);
        const uint32_t fv = __shfl(found_val, int(j * kGroupLanes), kWave);
        if (lane == pending_lane(queue, j)) {
          st = r;
          outv = fv;
          active = false;
        }
      }
      queue = ballot(active);
    }
    if (i < n) {
      if (status) status[i] = st;
      if constexpr (OP == MapOp::kFind) vals_out[i] = outv;
    }
  }
}

// ---- cuckoo -------------------------------------------------------------------
constexpr uint32_t kCuckooHashes = 4;

__device__ __forceinline__ uint64_t cuckoo_slot(uint64_t key, uint32_t h, uint64_t slots) {
  constexpr uint64_t salts[kCuckooHashes] = {0x9e3779b97f4a7c15ull, 0x85ebca6bc2b2ae35ull, 0x27d4eb2f165667b1ull,
                                             0xd6e8feb86659fd93ull};
  return __umul64hi(mix64(key ^ salts[h]), slots);
}

__global__ void cuckoo_insert_kernel(uint64_t* table, uint64_t slots, const uint64_t* __restrict__ keys, uint32_t n,
                                     uint32_t max_evictions, uint32_t* failed) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint64_t cur = keys[i];
    if (cur == kEmpty64) {
      atomicAdd(failed, 1u);
      continue;
    }
    uint32_t h = 0;
    bool placed = false;
    for (uint32_t a = 0; a < max_evictions; ++a) {
      const uint64_t slot = cuckoo_slot(cur, h, slots);
      const uint64_t prev = atomicExch(reinterpret_cast<unsigned long long*>(&table[slot]), (unsigned long long)cur);
      if (prev == kEmpty64) {
        placed = true;
        break;
      }
      // The evicted key sat at one of its own hash slots: continue it with
      // the next function after that one (cuckoo_hash.cuh:104-110).
      cur = prev;
      uint32_t j = 0;
      while (j < kCuckooHashes && cuckoo_slot(cur, j, slots) != slot) ++j;
      h = (j + 1) % kCuckooHashes;
    }
    if (!placed) atomicAdd(failed, 1u);
  }
}

__global__ void cuckoo_find_kernel(const uint64_t* table, uint64_t slots, const uint64_t* __restrict__ keys,
                                   uint32_t n, uint8_t* found) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    uint8_t f = 0;
#pragma unroll
    for (uint32_t h = 0; h < kCuckooHashes; ++h)
      if (table[cuckoo_slot(k, h, slots)] == k) f = 1;
    found[i] = k == kEmpty64 ? 0 : f;
  }
}

inline uint32_t grid_for(uint64_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, 256), 8192)); }

}  // namespace hash
}  // namespace vx

using namespace vx;
using namespace vx::hash;

struct vx_slab_table {
  SlabTable t{};
  uint64_t* mem = nullptr;
  unsigned long long* bitmap = nullptr;
};
struct vx_slab_set : vx_slab_table {};
struct vx_slab_set_u32 : vx_slab_table {};
struct vx_slab_map : vx_slab_table {};
struct vx_slab_alloc {
  vx_slab_alloc_view v{};
};
struct vx_cuckoo_set {
  uint64_t* table = nullptr;
  uint64_t slots = 0;
  uint32_t max_evictions = 0;
};

namespace {

uint32_t bitmap_words(uint32_t capacity) { return (capacity + 63) / 64; }

vx_status slab_create(vx_slab_table* h, uint32_t buckets, uint32_t overflow) {
  if (buckets == 0) return VX_INVALID_VALUE;
  const uint64_t cap = uint64_t(buckets) + overflow;
  if (cap >= kNull) return VX_INVALID_VALUE;
  if (hipMalloc(&h->mem, cap * kWave * 8) != hipSuccess) return VX_OUT_OF_CAPACITY;
  if (hipMalloc(&h->bitmap, std::max<size_t>(256, size_t(bitmap_words(overflow)) * 8)) != hipSuccess)
    return VX_OUT_OF_CAPACITY;
  h->t = SlabTable{h->mem, buckets, uint32_t(cap), vx_slab_alloc_view{h->bitmap, overflow, bitmap_words(overflow)}};
  return VX_OK;
}

// Every slab empty and every overflow slab free (the allocator reset with
// the table, as SlabHashSetU32::clear does).
vx_status slab_clear(vx_slab_table* h, hipStream_t s) {
  vx_status st = fill_u32(reinterpret_cast<uint32_t*>(h->mem), uint64_t(h->t.capacity) * kWave * 2, 0xffffffffu, s);
  if (st != VX_OK) return st;
  return fill_u32(reinterpret_cast<uint32_t*>(h->bitmap), uint64_t(h->t.alloc.words) * 2, 0, s);
}

void slab_free(vx_slab_table* h) {
  if (h->mem) hipFree(h->mem);
  if (h->bitmap) hipFree(h->bitmap);
}

// Allocated ordinals of a bitmap (host popcount over a device copy).
vx_status bitmap_used(const vx_slab_alloc_view& a, uint32_t* used) {
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  std::vector<unsigned long long> words(a.words);
  if (a.words && hipMemcpy(words.data(), a.bitmap, size_t(a.words) * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return VX_DEVICE_ERROR;
  uint32_t n = 0;
  for (unsigned long long w : words) n += uint32_t(__builtin_popcountll(w));
  *used = n;
  return VX_OK;
}

uint32_t slab_grid(uint32_t n) {
  return std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, 256), 4096));
}

}  // namespace

// ---- C ABI: slab set ---------------------------------------------------------------
extern "C" vx_status vx_slab_set_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_set** out) {
  if (!out) return VX_INVALID_VALUE;
  auto* h = new vx_slab_set();
  vx_status st = slab_create(h, buckets, overflow_slabs);
  if (st == VX_OK) st = slab_clear(h, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    slab_free(h);
    delete h;
    return st;
  }
  *out = h;
  return VX_OK;
}

extern "C" vx_status vx_slab_set_destroy(vx_slab_set* h) {
  if (!h) return VX_INVALID_VALUE;
  slab_free(h);
  delete h;
  return VX_OK;
}

extern "C" vx_status vx_slab_set_clear(vx_slab_set* h, void* stream) {
  if (!h) return VX_INVALID_VALUE;
  return slab_clear(h, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_slab_set_insert(vx_slab_set* h, const uint64_t* keys, uint32_t n, uint32_t* status,
                                        void* stream) {
  if (!h || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kInsert><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->t, keys, n, nullptr,
                                                                                              status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_contains(const vx_slab_set* h, const uint64_t* keys, uint32_t n, uint8_t* found,
                                          uint32_t* status, void* stream) {
  if (!h || (!keys && n) || (!found && !status && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kContains><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->t, keys, n, found,
                                                                                                status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_erase(vx_slab_set* h, const uint64_t* keys, uint32_t n, uint32_t* status,
                                       void* stream) {
  if (!h || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kErase><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->t, keys, n, nullptr,
                                                                                             status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_slabs_used(const vx_slab_set* h, uint32_t* used) {
  if (!h || !used) return VX_INVALID_VALUE;
  uint32_t overflow = 0;
  const vx_status st = bitmap_used(h->t.alloc, &overflow);
  if (st != VX_OK) return st;
  *used = h->t.buckets + overflow;
  return VX_OK;
}

// ---- C ABI: slab set of u32 keys (SlabHashSetU32) ----------------------------------
extern "C" vx_status vx_slab_set_u32_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_set_u32** out) {
  if (!out) return VX_INVALID_VALUE;
  auto* h = new vx_slab_set_u32();
  vx_status st = slab_create(h, buckets, overflow_slabs);
  if (st == VX_OK) st = slab_clear(h, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    slab_free(h);
    delete h;
    return st;
  }
  *out = h;
  return VX_OK;
}

extern "C" vx_status vx_slab_set_u32_destroy(vx_slab_set_u32* h) {
  if (!h) return VX_INVALID_VALUE;
  slab_free(h);
  delete h;
  return VX_OK;
}

extern "C" vx_status vx_slab_set_u32_clear(vx_slab_set_u32* h, void* stream) {
  if (!h) return VX_INVALID_VALUE;
  return slab_clear(h, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_slab_set_u32_insert(vx_slab_set_u32* h, const uint32_t* keys, uint32_t n, uint32_t* status,
                                            void* stream) {
  if (!h || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kInsert, uint32_t><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(
      h->t, keys, n, nullptr, status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_u32_contains(const vx_slab_set_u32* h, const uint32_t* keys, uint32_t n,
                                              uint8_t* found, uint32_t* status, void* stream) {
  if (!h || (!keys && n) || (!found && !status && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kContains, uint32_t><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(
      h->t, keys, n, found, status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_u32_erase(vx_slab_set_u32* h, const uint32_t* keys, uint32_t n, uint32_t* status,
                                           void* stream) {
  if (!h || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_set_kernel<SetOp::kErase, uint32_t><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(
      h->t, keys, n, nullptr, status);
  return last_launch();
}

extern "C" vx_status vx_slab_set_u32_slabs_used(const vx_slab_set_u32* h, uint32_t* used) {
  if (!h || !used) return VX_INVALID_VALUE;
  uint32_t overflow = 0;
  const vx_status st = bitmap_used(h->t.alloc, &overflow);
  if (st != VX_OK) return st;
  *used = h->t.buckets + overflow;
  return VX_OK;
}

// ---- C ABI: the slab allocator itself (SlabAlloc) ----------------------------------
namespace {
__global__ void slab_alloc_kernel(vx_slab_alloc_view a, uint32_t n, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = vx_slab_alloc_dev::allocate(a, i);
}
__global__ void slab_dealloc_kernel(vx_slab_alloc_view a, const uint32_t* ordinals, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    vx_slab_alloc_dev::deallocate(a, ordinals[i]);
}
}  // namespace

extern "C" vx_status vx_slab_alloc_create(uint32_t capacity, vx_slab_alloc** out) {
  if (!out) return VX_INVALID_VALUE;
  auto* h = new vx_slab_alloc();
  h->v.capacity = capacity;
  h->v.words = bitmap_words(capacity);
  if (hipMalloc(&h->v.bitmap, std::max<size_t>(256, size_t(h->v.words) * 8)) != hipSuccess) {
    delete h;
    return VX_OUT_OF_CAPACITY;
  }
  vx_status st = fill_u32(reinterpret_cast<uint32_t*>(h->v.bitmap), uint64_t(h->v.words) * 2, 0, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    hipFree(h->v.bitmap);
    delete h;
    return st;
  }
  *out = h;
  return VX_OK;
}

extern "C" vx_status vx_slab_alloc_destroy(vx_slab_alloc* h) {
  if (!h) return VX_INVALID_VALUE;
  if (h->v.bitmap) hipFree(h->v.bitmap);
  delete h;
  return VX_OK;
}

extern "C" vx_status vx_slab_alloc_clear(vx_slab_alloc* h, void* stream) {
  if (!h) return VX_INVALID_VALUE;
  return fill_u32(reinterpret_cast<uint32_t*>(h->v.bitmap), uint64_t(h->v.words) * 2, 0,
                  static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_slab_alloc_allocate(vx_slab_alloc* h, uint32_t n, uint32_t* d_ordinals, void* stream) {
  if (!h || (n && !d_ordinals)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_alloc_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->v, n, d_ordinals);
  return last_launch();
}

extern "C" vx_status vx_slab_alloc_deallocate(vx_slab_alloc* h, const uint32_t* d_ordinals, uint32_t n,
                                              void* stream) {
  if (!h || (n && !d_ordinals)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_dealloc_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->v, d_ordinals, n);
  return last_launch();
}

extern "C" vx_status vx_slab_alloc_used(const vx_slab_alloc* h, uint32_t* used) {
  if (!h || !used) return VX_INVALID_VALUE;
  return bitmap_used(h->v, used);
}

extern "C" vx_status vx_slab_alloc_view_get(const vx_slab_alloc* h, vx_slab_alloc_view* out) {
  if (!h || !out) return VX_INVALID_VALUE;
  *out = h->v;
  return VX_OK;
}

// ---- C ABI: slab map -----------------------------------------------------------------
extern "C" vx_status vx_slab_map_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_map** out) {
  if (!out) return VX_INVALID_VALUE;
  auto* h = new vx_slab_map();
  vx_status st = slab_create(h, buckets, overflow_slabs);
  if (st == VX_OK) st = slab_clear(h, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    slab_free(h);
    delete h;
    return st;
  }
  *out = h;
  return VX_OK;
}

extern "C" vx_status vx_slab_map_destroy(vx_slab_map* h) {
  if (!h) return VX_INVALID_VALUE;
  slab_free(h);
  delete h;
  return VX_OK;
}

extern "C" vx_status vx_slab_map_clear(vx_slab_map* h, void* stream) {
  if (!h) return VX_INVALID_VALUE;
  return slab_clear(h, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_slab_map_insert_or_replace(vx_slab_map* h, const uint32_t* keys, const uint32_t* values,
                                                   uint32_t n, uint32_t* status, void* stream) {
  if (!h || (n && (!keys || !values))) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_map_kernel<MapOp::kInsertOrReplace><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(
      h->t, keys, values, nullptr, n, status);
  return last_launch();
}

extern "C" vx_status vx_slab_map_find(const vx_slab_map* h, const uint32_t* keys, uint32_t* values, uint32_t n,
                                      uint32_t* status, void* stream) {
  if (!h || (n && (!keys || !values))) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_map_kernel<MapOp::kFind><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->t, keys, nullptr,
                                                                                            values, n, status);
  return last_launch();
}

extern "C" vx_status vx_slab_map_erase(vx_slab_map* h, const uint32_t* keys, uint32_t n, uint32_t* status,
                                       void* stream) {
  if (!h || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  slab_map_kernel<MapOp::kErase><<<slab_grid(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->t, keys, nullptr,
                                                                                             nullptr, n, status);
  return last_launch();
}

// ---- C ABI: cuckoo set ------------------------------------------------------------------
extern "C" vx_status vx_cuckoo_set_create(uint64_t slots, uint32_t max_evictions, vx_cuckoo_set** out) {
  if (!out || slots < 8) return VX_INVALID_VALUE;
  auto* h = new vx_cuckoo_set();
  h->slots = slots;
  h->max_evictions = max_evictions ? max_evictions : 512;
  if (hipMalloc(&h->table, slots * 8) != hipSuccess) {
    delete h;
    return VX_OUT_OF_CAPACITY;
  }
  vx_status st = fill_u32(reinterpret_cast<uint32_t*>(h->table), slots * 2, 0xffffffffu, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    hipFree(h->table);
    delete h;
    return st;
  }
  *out = h;
  return VX_OK;
}

extern "C" vx_status vx_cuckoo_set_destroy(vx_cuckoo_set* h) {
  if (!h) return VX_INVALID_VALUE;
  if (h->table) hipFree(h->table);
  delete h;
  return VX_OK;
}

extern "C" vx_status vx_cuckoo_set_clear(vx_cuckoo_set* h, void* stream) {
  if (!h) return VX_INVALID_VALUE;
  return fill_u32(reinterpret_cast<uint32_t*>(h->table), h->slots * 2, 0xffffffffu, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_cuckoo_set_insert(vx_cuckoo_set* h, const uint64_t* keys, uint32_t n, uint32_t* d_failed,
                                          void* stream) {
  if (!h || !d_failed || (!keys && n)) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  cuckoo_insert_kernel<<<hash::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->table, h->slots, keys, n,
                                                                                        h->max_evictions, d_failed);
  return last_launch();
}

extern "C" vx_status vx_cuckoo_set_contains(const vx_cuckoo_set* h, const uint64_t* keys, uint32_t n, uint8_t* found,
                                            void* stream) {
  if (!h || (n && (!keys || !found))) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  cuckoo_find_kernel<<<hash::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(h->table, h->slots, keys, n,
                                                                                      found);
  return last_launch();
}
