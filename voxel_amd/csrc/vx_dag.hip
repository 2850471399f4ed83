// Hash DAG edit for gfx950 — one pipeline for one scene or many.
//
// Replaces the reference's hash_dag_gpu edit pipeline (detail/edit_pipeline.cuh:31-97)
// with a level-compacted design that also batches many scenes into one launch
// sequence (the reference edits one 1024^3 DAG per call):
//   records   brushes (or voxels) -> one record per touched 4^3 leaf brick
//             {(scene, leaf Morton) key, voxel mask, material}, compacted in
//             brush order by the fused scan (replaces
//             pack_leaf_edit_records_kernel, build_leaf_masks/pipeline.cuh:16-36);
//   sort      onesweep sort of (key, record index): a segmented sort, the scene
//             id sits above the Morton bits; stable, so records of one leaf
//             stay in brush order (last writer wins);
//   heights   run compaction -> distinct leaves (height 0); then per height the
//             distinct parents with child ranges (replaces the per-depth request
//             emission + binary-search parent link, detail/edit/requests.cuh:51-173);
//             a scene whose root is at this height contributes its root entry
//             and stops — every later kernel runs over exactly one height's
//             entries, not over 9N threads;
//   old refs  one top-down pass per height: each entry reads one parent item
//             (replaces the per-request root walk, detail/edit/snapshots.cuh:14-53);
//   rebuild   leaves then heights 1..H: build item -> probe -> commit
//             (replaces apply_leaf_masks / process_depth,
//             detail/edit/leaf_update.cuh:52-109, detail/edit/rebuild.cuh:30-89).
// Interning is a true insert-if-absent on a u64-slot table: the probe kernel
// claims an empty slot by CAS with {fingerprint, kTag | item index}; an equal
// item that meets the tagged slot becomes a duplicate of that representative
// (the in-batch dedup the reference does with a temporary map,
// detail/edit/temp_dedup.cuh:12-64); the commit kernel — a segmented scan over
// the representatives' widths, one segment per scene — copies each
// representative into its scene's pool at the scanned offset and rewrites its
// slot with the pool offset. No pool space is ever allocated for a duplicate,
// and no thread waits on another.
// Status: first error wins per scene (detail/edit_common.cuh:96-99); a scene's
// root is published only if that scene saw no error (detail/edit/root.cuh:13-25).
#include <cmath>
#include <type_traits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "vx_dag.hpp"
#include "vx_brick_mask.hpp"
#include "vx_scan.hpp"
#include "vx_sort.hpp"

using namespace vx;
using namespace vx::dag;

namespace vx {
namespace dag {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kMaxHeights = 16;

// Level links: an entry's parent index (at the height above) and its child
// slot in the parent (the low 3 bits of its key), packed as index << 3 | slot
// so that the passes that follow a link read one word instead of the link
// and the 8-byte key; kNone for a root. Entries per height stay below
// kMaxEntries (the record count bounds them).
constexpr uint32_t kMaxEntries = 1u << 29;
__device__ __forceinline__ uint32_t link_make(uint32_t parent, uint64_t key) { return (parent << 3) | uint32_t(key & 7u); }
__device__ __forceinline__ uint32_t link_parent(uint32_t link) { return link >> 3; }
__device__ __forceinline__ uint32_t link_slot(uint32_t link) { return link & 7u; }

// Composite keys: scene id above `shift` bits of (leaf Morton >> 3h).
struct Codec {
  uint32_t shift;
  __host__ __device__ __forceinline__ uint32_t scene(uint64_t k) const { return uint32_t(k >> shift); }
  __host__ __device__ __forceinline__ uint64_t parent(uint64_t k) const {
    const uint64_t m = (uint64_t(1) << shift) - 1;
    return (k & ~m) | ((k & m) >> 3);
  }
};

// The scene an entry belongs to: a device table indexed by the key's scene id,
// or a single scene passed by value.
struct Scenes {
  const Dev* table;
  const uint64_t* keys;
  Codec codec;
  Dev single;
  __device__ __forceinline__ uint32_t id(uint32_t i) const { return table ? codec.scene(keys[i]) : 0u; }
  __device__ __forceinline__ const Dev& of(uint32_t i) const { return table ? table[codec.scene(keys[i])] : single; }
  __device__ __forceinline__ const Dev& at(uint32_t s) const { return table ? table[s] : single; }
  // The same with the level's keys given (no copy of the Scenes, whose
  // embedded Dev would otherwise go to scratch once a reference to it is taken).
  __device__ __forceinline__ uint32_t id(const uint64_t* k, uint32_t i) const { return table ? codec.scene(k[i]) : 0u; }
  __device__ __forceinline__ const Dev& of(const uint64_t* k, uint32_t i) const {
    return table ? table[codec.scene(k[i])] : single;
  }
};

// ---- interning --------------------------------------------------------------
struct Table {
  uint64_t* slots;
  uint64_t mask;
  uint32_t* pool;
  uint64_t cap;
  uint32_t* used;
  uint32_t* base;
};

__device__ __forceinline__ Table table_of(const Dev& d, bool leaves) {
  return leaves ? Table{d.leaf_slots, d.leaf_mask, d.leaf_pool, d.leaf_cap, d.ctl + kCtlLeafUsed, d.base + 1}
                : Table{d.node_slots, d.node_mask, d.node_pool, d.node_cap, d.ctl + kCtlNodeUsed, d.base};
}

// Item words in registers, every load issued before any use: a word-by-word
// copy or compare loop would wait one memory latency per word. kMax is the
// widest item of the set (9 for nodes, the leaf width for leaves); aligned
// full-width items move as 16-byte vectors.
template <uint32_t kMax>
__device__ __forceinline__ void load_item(const uint32_t* p, uint32_t n, uint32_t (&w)[kMax]) {
  if constexpr (kMax % 4 == 0) {
    if (n == kMax && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
#pragma unroll
      for (uint32_t k = 0; k < kMax; k += 4) {
        const uint4 v = gld4(p + k);
        w[k] = v.x;
        w[k + 1] = v.y;
        w[k + 2] = v.z;
        w[k + 3] = v.w;
      }
      return;
    }
  }
  if constexpr (kMax % 4 == 0) {
    // (Not taken by aligned full-width items; kept small in registers.)
#pragma unroll
    for (uint32_t k = 0; k < kMax; ++k) w[k] = k < n ? gld(p + k) : 0u;
  } else {
    // Words past the item's width read its first word again (no branch on
    // n, which would make each load wait for the one before; no line past
    // the item), then read as 0.
#pragma unroll
    for (uint32_t k = 0; k < kMax; ++k) w[k] = gld(p + (k < n ? k : 0u));
#pragma unroll
    for (uint32_t k = 0; k < kMax; ++k) w[k] &= k < n ? 0xffffffffu : 0u;
  }
}

template <uint32_t kMax>
__device__ __forceinline__ bool item_equal(const uint32_t (&a)[kMax], const uint32_t* p, uint32_t n) {
  uint32_t b[kMax];
  load_item<kMax>(p, n, b);
  uint32_t diff = 0;
#pragma unroll
  for (uint32_t k = 0; k < kMax; ++k) diff |= a[k] ^ b[k];
  return diff == 0;
}

template <uint32_t kMax>
__device__ __forceinline__ void store_item(uint32_t* p, uint32_t n, const uint32_t (&w)[kMax]) {
  if constexpr (kMax % 4 == 0) {
    if (n == kMax && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
#pragma unroll
      for (uint32_t k = 0; k < kMax; k += 4)
        *reinterpret_cast<uint4*>(p + k) = make_uint4(w[k], w[k + 1], w[k + 2], w[k + 3]);
      return;
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < kMax; ++k)
    if (k < n) p[k] = w[k];
}

// Probe of entry i (PENDING entries only): find an equal item in the table or
// in this batch, or claim an empty slot as the representative.
template <uint32_t kMax>
__device__ __forceinline__ void probe_entry(const ItemSet& set, const uint64_t* __restrict__ hashes, uint64_t* res,
                                            const Scenes& sc, bool leaves, uint32_t i) {
  // One round of reads: the entry's result, its hash and its whole item slot
  // (a slot holds set.stride == kMax words), the width taken from the loaded
  // header and the words past it cleared afterwards.
  const uint64_t r0 = res[i];
  const uint64_t h = hashes[i];
  uint32_t mine[kMax];
  load_item<kMax>(set.items + size_t(i) * set.stride, kMax, mine);
  if (res_kind(r0) != kResPending) return;
  {
    const Dev& d = sc.of(i);
    const Table t = table_of(d, leaves);
    const uint32_t w = set.fixed_width ? set.fixed_width : 1u + __popc(mine[0] & 0xffu);
#pragma unroll
    for (uint32_t k = 0; k < kMax; ++k) mine[k] &= k < w ? 0xffffffffu : 0u;
    const uint32_t fp = uint32_t(h >> 32);
    uint64_t idx = h & t.mask;
    const uint64_t limit = t.mask < kMaxProbe ? t.mask + 1 : kMaxProbe;
    uint64_t out = res_make(kResFinal, kEmpty);
    bool done = false;
    for (uint64_t probe = 0; probe < limit && !done; ++probe) {
      // Plain (cacheable) read: during a probe launch a slot only goes from
      // empty to tagged, so a stale read can only show "empty", and the CAS
      // below then returns the slot's real value.
      uint64_t s = t.slots[idx];
      if (s == 0) {
        const uint64_t tagged = (uint64_t(fp) << 32) | kTag | i;
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&t.slots[idx]), 0ull,
                                        static_cast<unsigned long long>(tagged));
        if (prev == 0) {
          out = res_make(kResRep, uint32_t(idx));
          done = true;
          break;
        }
        s = prev;
      }
      if (uint32_t(s >> 32) == fp) {
        const uint32_t v = uint32_t(s);
        if (v & kTag) {
          const uint32_t j = v & ~kTag;
          if (item_equal<kMax>(mine, set.items + size_t(j) * set.stride, w)) {
            out = res_make(kResDup, j);
            done = true;
            break;
          }
        } else if (v >= kPoolStart) {
          // An item of width w that would run past the pool's end cannot be
          // the one stored at v (equal headers mean equal widths), and must
          // not be read.
          if (uint64_t(v) + w <= t.cap && item_equal<kMax>(mine, t.pool + v, w)) {
            out = res_make(kResFinal, v);
            done = true;
            break;
          }
        }
      }
      idx = (idx + 1) & t.mask;
    }
    if (!done) set_status(d, kStatusOverflow);
    res[i] = out;
  }
}

// Read-only lookup of an item among the items interned by earlier edits, for
// a build step: before a level's probe the table holds no tagged slot, so a
// walk that ends at an empty slot proves the item is not interned yet. An
// entry found here becomes FINAL at once and its item is never written; 0
// (walk ended at an empty or tagged slot, or after kFindProbes slots) leaves
// the decision to the probe.
constexpr uint32_t kFindProbes = 64;

// *empty_at: the empty slot that ended the walk (else the home slot): every
// slot before it on the probe sequence holds an item that is not this one,
// and stays so while the level's probe only tags empty slots, so the probe
// may start there.
template <uint32_t kMax>
__device__ __forceinline__ uint32_t find_interned(const Table& t, const uint32_t (&w)[kMax], uint32_t n, uint64_t h,
                                                  uint64_t* empty_at) {
  const uint32_t fp = uint32_t(h >> 32);
  uint64_t idx = h & t.mask;
  *empty_at = idx;
  for (uint32_t p = 0; p < kFindProbes; ++p) {
    const uint64_t s = gld(t.slots + idx);
    if (s == 0) {
      *empty_at = idx;
      return 0;
    }
    if (uint32_t(s >> 32) == fp) {
      const uint32_t v = uint32_t(s);
      if (v & kTag) return 0;
      if (v >= kPoolStart && uint64_t(v) + n <= t.cap && item_equal<kMax>(w, t.pool + v, n)) return v;
    }
    idx = (idx + 1) & t.mask;
  }
  return 0;
}

template <uint32_t kMax>
__global__ void probe_kernel(ItemSet set, const uint64_t* __restrict__ hashes, const uint32_t* n_ptr, uint64_t* res,
                             Scenes sc, uint32_t n_scenes, bool leaves) {
  if (blockIdx.x == 0)
    for (uint32_t s = threadIdx.x; s < n_scenes; s += blockDim.x) {
      const Table t = table_of(sc.at(s), leaves);
      *t.base = *t.used;
    }
  const uint32_t n = *n_ptr;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    probe_entry<kMax>(set, hashes, res, sc, leaves, i);
}

// Commit, part 1: segmented scan (one segment per scene) of representative
// widths. For each representative it stores the final slot value
// {fingerprint, pool offset} in hashes[i] (the fingerprint rides in the scan
// payload, so nothing is read back); a representative that does not fit is
// tombstoned here and resolves to EMPTY. Part 2 moves the items.
// Everything the store half needs is loaded by the load half (whose loads
// are all in flight together), so the store half only writes.
struct CommitPayload {
  uint64_t r;
  uint32_t i;     // the entry
  uint32_t fp;
  uint32_t base;  // the scene's pool words in use before this level
  bool last;      // last entry of its scene's segment
};

// Position t of the scan is entry list[t] (entry t without a list); a list
// is in entry order, so scene segments stay contiguous. A split load: the
// reads of all of a thread's items go out together, round by round (the
// entries; their results, fingerprints, node headers and scene ids; the
// scenes' pool words in use), none behind a branch on a loaded value.
struct CommitLoad {
  using Payload = CommitPayload;
  const uint64_t* res;
  const uint64_t* hashes;
  const uint32_t* n_ptr;
  const uint32_t* list;
  ItemSet set;
  Scenes sc;
  bool leaves;
  struct Raw {
    uint64_t r;
    uint32_t i, fp, w, base;
    bool head, last;
  };
  template <uint32_t N>
  __device__ void fetch(const uint32_t (&t)[N], uint32_t n, Raw (&x)[N]) const {
    uint32_t ip[N], in[N], hdr[N];
    if (list) {
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) {
        x[k].i = gld(list + t[k]);
        ip[k] = gld(list + (t[k] ? t[k] - 1 : 0u));
        in[k] = gld(list + min(t[k] + 1, n - 1));
      }
    } else {
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) {
        x[k].i = t[k];
        ip[k] = t[k] ? t[k] - 1 : 0u;
        in[k] = min(t[k] + 1, n - 1);
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) {
      x[k].r = gld(res + x[k].i);
      x[k].fp = gld(reinterpret_cast<const uint32_t*>(hashes + x[k].i) + 1);  // the hash's high word
      hdr[k] = 0;
    }
    if (!set.fixed_width) {
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) hdr[k] = gld(set.items + size_t(x[k].i) * set.stride);
    }
    if (sc.table) {
      uint32_t id[N];
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) {  // every load unconditional (no short-circuit)
        id[k] = sc.codec.scene(gld(sc.keys + x[k].i));
        const uint32_t before = sc.codec.scene(gld(sc.keys + ip[k])), after = sc.codec.scene(gld(sc.keys + in[k]));
        x[k].head = (t[k] == 0) | (before != id[k]);
        x[k].last = (t[k] == n - 1) | (after != id[k]);
      }
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) x[k].base = gld(gld(&sc.table[id[k]].base) + (leaves ? 1 : 0));
    } else {
      const uint32_t b = gld(table_of(sc.single, leaves).base);
#pragma unroll
      for (uint32_t k = 0; k < N; ++k) {
        x[k].head = t[k] == 0;
        x[k].last = t[k] == n - 1;
        x[k].base = b;
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) {
      const bool rep = res_kind(x[k].r) == kResRep;
      x[k].w = rep ? (set.fixed_width ? set.fixed_width : 1u + __popc(hdr[k] & 0xffu)) : 0u;
      if (!rep) x[k].fp = 0;
    }
  }
  __device__ uint64_t finish(uint32_t, const Raw& x, CommitPayload& p) const {
    p.i = x.i;
    p.r = x.r;
    p.fp = x.fp;
    p.base = x.base;
    p.last = x.last;
    return scan::SegSumOp::make(x.w, x.head);
  }
};

// With slot_idx (final mode) the store half also makes each representative
// FINAL at once (its ref is its pool offset) and notes, per position t of
// the scan, the slot its place step publishes (kNone: nothing to place), so
// that the next height's build can run beside the place.
struct CommitStore {
  uint64_t* res;
  uint64_t* slot_vals;  // the hashes array, reused
  Scenes sc;
  bool leaves;
  uint32_t* slot_idx = nullptr;
  __device__ void operator()(uint32_t pos, const CommitPayload& p, uint64_t v, uint64_t, uint64_t incl,
                             uint32_t) const {
    const uint32_t i = p.i;
    const Dev& d = sc.of(i);
    const Table t = table_of(d, leaves);
    const uint32_t base = p.base;
    const uint32_t w = uint32_t(v), seg_incl = uint32_t(incl);
    uint32_t place = kNone;
    if (w) {
      const uint64_t fp_hi = uint64_t(p.fp) << 32;
      const uint64_t off = uint64_t(base) + seg_incl - w;
      if (off + w > t.cap || off >= kTag) {
        set_status(d, kStatusOverflow);
        st_agent(&t.slots[uint32_t(p.r)], fp_hi | kTomb);
        res[i] = res_make(kResFinal, kEmpty);
      } else {
        slot_vals[i] = fp_hi | uint32_t(off);
        if (slot_idx) {
          res[i] = res_make(kResFinal, uint32_t(off));
          place = uint32_t(p.r);
        }
      }
    }
    if (slot_idx) slot_idx[pos] = place;
    if (p.last) {
      const uint64_t used = uint64_t(base) + seg_incl;
      *t.used = used > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(used);
    }
  }
};

// Commit, part 2: one thread per entry; each representative copies its item
// to the pool, publishes its slot and becomes FINAL. Independent entries, so
// the item loads of many entries are in flight at once.
template <uint32_t kMax>
__device__ __forceinline__ void place_entry(const ItemSet& set, const uint64_t* __restrict__ slot_vals, uint64_t* res,
                                            const Scenes& sc, bool leaves, uint32_t i) {
  // Node items (up to 9 words): one round of reads (result, slot value, the
  // whole item slot: set.stride == kMax words), as in probe_entry. Leaf
  // items (16-byte vectors): the result first, so that the batch's
  // duplicates, which are not copied, read no item.
  const uint64_t r = res[i];
  const uint64_t sv = slot_vals[i];
  uint32_t words[kMax];
  constexpr bool kVec = kMax % 4 == 0;
  if (!kVec) load_item<kMax>(set.items + size_t(i) * set.stride, kMax, words);
  if (res_kind(r) != kResRep) return;
  if (kVec) load_item<kMax>(set.items + size_t(i) * set.stride, kMax, words);
  const Table t = table_of(sc.of(i), leaves);
  const uint32_t off = uint32_t(sv);
  const uint32_t w = set.fixed_width ? set.fixed_width : 1u + __popc(words[0] & 0xffu);
  store_item<kMax>(t.pool + off, w, words);
  st_agent(&t.slots[uint32_t(r)], sv);
  res[i] = res_make(kResFinal, off);
}

template <uint32_t kMax>
__global__ void place_kernel(ItemSet set, const uint64_t* __restrict__ slot_vals, const uint32_t* n_ptr,
                             const uint32_t* __restrict__ list, uint64_t* res, Scenes sc, bool leaves) {
  const uint32_t n = *n_ptr;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    place_entry<kMax>(set, slot_vals, res, sc, leaves, list ? list[t] : t);
}

// Place after a final-mode commit: position t of the commit scan copies its
// entry's item to the pool and publishes the slot the commit noted (results
// are already FINAL).
struct PlaceArgs {
  ItemSet set;
  const uint64_t* slot_vals;
  const uint32_t* slot_idx;
  const uint32_t* n_ptr;
  const uint32_t* list;
  Scenes sc;
  bool leaves;
};

template <uint32_t kMax>
__device__ __forceinline__ void place_noted(const PlaceArgs& a, uint32_t block, uint32_t blocks) {
  const uint32_t n = *a.n_ptr;
  for (uint32_t t = block * blockDim.x + threadIdx.x; t < n; t += blocks * blockDim.x) {
    // One round of reads: the noted slot, the entry, its slot value and item
    // slot (set.stride == kMax words); the entry is t without a list.
    const uint32_t slot = a.slot_idx[t];
    const uint32_t i = a.list ? a.list[t] : t;
    if (slot == kNone) continue;
    const uint64_t sv = a.slot_vals[i];
    uint32_t words[kMax];
    load_item<kMax>(a.set.items + size_t(i) * a.set.stride, kMax, words);
    const Table tb = table_of(a.sc.of(i), a.leaves);
    const uint32_t w = a.set.fixed_width ? a.set.fixed_width : 1u + __popc(words[0] & 0xffu);
    store_item<kMax>(tb.pool + uint32_t(sv), w, words);
    st_agent(&tb.slots[slot], sv);
  }
}

template <uint32_t kMax>
__global__ void place_noted_kernel(PlaceArgs a) {
  place_noted<kMax>(a, blockIdx.x, gridDim.x);
}

// The commit scan alone, in final mode (the place runs later, in
// place_noted_kernel or beside the next height's build).
template <uint32_t kMax>
vx_status commit_final(const Scenes& sc, uint32_t n_scenes, bool leaves, ItemSet set, uint64_t* hashes,
                       const uint32_t* n_ptr, const uint32_t* list, uint64_t n_ub, uint64_t* res, uint32_t* slot_idx,
                       void* fused_ws, hipStream_t s, bool clear_ws) {
  TimedLaunch tl(leaves ? VX_K_LEAF_COMMIT : VX_K_NODE_COMMIT, s);
  const CommitLoad ld{res, hashes, n_ptr, list, set, sc, leaves};
  const CommitStore sto{res, hashes, sc, leaves, slot_idx};
  constexpr uint32_t kBig = scan::kFusedItems, kSmall = scan::kFusedItemsSmall;
  if (n_scenes == 1)
    return leaves ? scan::fused_dev_m<scan::SumOp, kBig>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws)
                  : scan::fused_dev_m<scan::SumOp, kSmall>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws);
  return leaves ? scan::fused_dev_m<scan::SegSumOp, kBig>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws)
                : scan::fused_dev_m<scan::SegSumOp, kSmall>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws);
}

// Commit and place of the entries list[0 .. *n_ptr) (0 .. *n_ptr without a
// list) after their probe.
template <uint32_t kMax>
vx_status commit_place(const Scenes& sc, uint32_t n_scenes, bool leaves, ItemSet set, uint64_t* hashes,
                       const uint32_t* n_ptr, const uint32_t* list, uint64_t n_ub, uint64_t* res, void* fused_ws,
                       hipStream_t s, bool clear_ws) {
  vx_status st;
  {
    // One scene has one segment: a plain 32-bit sum (CommitLoad's value
    // truncated to its low word, the width) replaces the 64-bit segmented sum,
    // halving the scan's cross-lane work and registers.
    TimedLaunch tl(leaves ? VX_K_LEAF_COMMIT : VX_K_NODE_COMMIT, s);
    const CommitLoad ld{res, hashes, n_ptr, list, set, sc, leaves};
    const CommitStore sto{res, hashes, sc, leaves};
    // Leaves: one pass over many entries; node heights: small counts.
    constexpr uint32_t kBig = scan::kFusedItems, kSmall = scan::kFusedItemsSmall;
    if (n_scenes == 1)
      st = leaves ? scan::fused_dev_m<scan::SumOp, kBig>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws)
                  : scan::fused_dev_m<scan::SumOp, kSmall>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws);
    else
      st = leaves ? scan::fused_dev_m<scan::SegSumOp, kBig>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws)
                  : scan::fused_dev_m<scan::SegSumOp, kSmall>(n_ptr, n_ub, ld, sto, fused_ws, s, clear_ws);
  }
  if (st != VX_OK) return st;
  {
    TimedLaunch tl(leaves ? VX_K_LEAF_PLACE : VX_K_NODE_PLACE, s);
    place_kernel<kMax><<<grid_for(n_ub), 256, 0, s>>>(set, hashes, n_ptr, list, res, sc, leaves);
  }
  return last_launch();
}

template <uint32_t kMax>
vx_status intern_width(const Scenes& sc, uint32_t n_scenes, bool leaves, ItemSet set, uint64_t* hashes,
                       const uint32_t* n_ptr, uint64_t n_ub, uint64_t* res, void* fused_ws, hipStream_t s,
                       bool clear_ws) {
  {
    TimedLaunch tl(leaves ? VX_K_LEAF_PROBE : VX_K_NODE_PROBE, s);
    probe_kernel<kMax><<<grid_for(n_ub), 256, 0, s>>>(set, hashes, n_ptr, res, sc, n_scenes, leaves);
  }
  const vx_status st = last_launch();
  if (st != VX_OK) return st;
  return commit_place<kMax>(sc, n_scenes, leaves, set, hashes, n_ptr, nullptr, n_ub, res, fused_ws, s, clear_ws);
}

vx_status intern_scenes(const Scenes& sc, uint32_t n_scenes, bool leaves, ItemSet set, uint64_t* hashes,
                        const uint32_t* n_ptr, uint64_t n_ub, uint64_t* res, void* fused_ws, hipStream_t s,
                        bool clear_ws = true) {
  if (!leaves)
    return intern_width<kNodeStride>(sc, n_scenes, leaves, set, hashes, n_ptr, n_ub, res, fused_ws, s, clear_ws);
  if (set.fixed_width == 16)
    return intern_width<16>(sc, n_scenes, leaves, set, hashes, n_ptr, n_ub, res, fused_ws, s, clear_ws);
  if (set.fixed_width == 2)
    return intern_width<2>(sc, n_scenes, leaves, set, hashes, n_ptr, n_ub, res, fused_ws, s, clear_ws);
  return VX_INVALID_VALUE;
}

vx_status intern(const Dev& d, bool leaves, ItemSet set, uint64_t* hashes, const uint32_t* n_ptr, uint64_t n_ub,
                 uint64_t* res, void* fused_ws, hipStream_t s) {
  const Scenes sc{nullptr, nullptr, Codec{63}, d};
  return intern_scenes(sc, 1, leaves, set, hashes, n_ptr, n_ub, res, fused_ws, s);
}

// ---- records ----------------------------------------------------------------
// 80 bytes, 16-byte aligned: a job is read as five 16-byte loads.
struct alignas(16) BrushJob {
  vx_brush b;
  int32_t bx0, by0, bz0;
  uint32_t nx, ny, nz;
  uint32_t start;
  uint32_t scene;
  float inv_nx, inv_ny;  // 1/nx, 1/ny rounded to nearest (see divmod_small)
  uint32_t row_start;    // first brick row (ny * nz rows per job)
  uint32_t pad;
};
static_assert(sizeof(BrushJob) == 80, "BrushJob layout");
static_assert(offsetof(BrushJob, bx0) == 32 && offsetof(BrushJob, scene) == 60, "BrushJob words read by the records");

// a / d and a % d for a < 2^22 and d >= 1 from inv = float(1/d): a is exact
// in float and the product carries a relative error below 2^-23, so it is
// within 1/2 of a / d (< 2^22) and one correction step either way makes the
// truncated quotient exact. Larger a take the integer division.
__device__ __forceinline__ void divmod_small(uint32_t a, uint32_t d, float inv, uint32_t& q, uint32_t& r) {
  if (a < (1u << 22)) {
    q = uint32_t(float(a) * inv);
    int32_t rem = int32_t(a) - int32_t(q * d);
    if (rem < 0) {
      --q;
      rem += int32_t(d);
    } else if (rem >= int32_t(d)) {
      ++q;
      rem -= int32_t(d);
    }
    r = uint32_t(rem);
  } else {
    q = a / d;
    r = a - q * d;
  }
}

struct Record {
  uint64_t key;
  uint64_t mask;
  uint32_t mat;
};


// Record outputs. With index_bits > 0 the sort key is packed above the
// record index (key << index_bits | index) and sorted by its upper bits, so
// the sort moves 8 bytes per record instead of a 12-byte (key, index) pair.
struct RecordOut {
  uint64_t* key;
  uint32_t* val;
  uint4* pay;  // {mask low, mask high, material, 0}: one 16-byte read per record later
  uint32_t* n_rec;
  uint32_t index_bits;
  __device__ __forceinline__ void put_payload(uint32_t pos, uint64_t mask, uint32_t mat) const {
    pay[pos] = make_uint4(uint32_t(mask), uint32_t(mask >> 32), mat, 0u);
  }
  __device__ __forceinline__ void put_key(uint32_t pos, uint64_t k) const {
    if (index_bits) {
      key[pos] = (k << index_bits) | pos;
    } else {
      key[pos] = k;
      val[pos] = pos;
    }
  }
};

// Sorted records, in either layout of RecordOut.
struct SortedRecs {
  const uint64_t* keys;
  const uint32_t* vals;
  uint32_t index_bits;
  __device__ __forceinline__ uint64_t key(uint32_t i) const { return index_bits ? keys[i] >> index_bits : keys[i]; }
  __device__ __forceinline__ uint32_t index(uint32_t i) const {
    // One load in either layout, no branch (a branch would make the loads
    // issued before it wait at the join): the packed key's low word holds
    // the index (index_bits <= 32), else the value array does.
    const uint32_t* src = index_bits ? reinterpret_cast<const uint32_t*>(keys) + 2 * size_t(i) : vals + i;
    const uint32_t m = index_bits ? uint32_t((uint64_t(1) << index_bits) - 1) : 0xffffffffu;
    return gld(src) & m;
  }
};

// Brush jobs on the device: job_of finds the job of brick row i (ROWS) or of
// brick i of the concatenated clipped brick boxes.
struct BrushCands {
  const BrushJob* jobs;
  const uint32_t* n_jobs;  // device count
  bool materials;
  Codec codec;
  // The job of candidate i: the last job whose start is <= i. The lanes of a
  // wave hold consecutive candidates: one wave-uniform binary search (scalar
  // loads) for the lowest active lane, then each lane steps forward over the
  // few jobs that start inside its wave's window. With a hint (the job of an
  // earlier candidate of the same lane) the search is skipped.
  template <bool ROWS = false>
  __device__ __forceinline__ uint32_t job_of(uint32_t i, uint32_t nj, uint32_t hint) const {
    auto first = [&](uint32_t k) { return ROWS ? jobs[k].row_start : jobs[k].start; };
    uint32_t lo = hint;
    if (hint == kNone) {
      const uint32_t i0 = __builtin_amdgcn_readfirstlane(i);
      uint32_t hi = nj;
      lo = 0;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first(mid) <= i0) lo = mid;
        else hi = mid;
      }
    }
    while (lo + 1 < nj && first(lo + 1) <= i) ++lo;
    return lo;
  }
};

// The brush path's input: its jobs (the records come from the brick rows
// below).
struct BrushLoad {
  BrushCands c;
};

// ---- brush records by brick rows ---------------------------------------------
// A job's clipped brick box is walked by rows (fixed brick y and z, the box's
// x run). The touched bricks of a row are one run: the whole row for a box;
// for a sphere the bricks whose voxel nearest to the centre lies inside, an
// interval in closed form (integer square root): a brick is touched iff its
// voxel nearest to the centre (clamped axis by axis, which minimises each
// squared term) is inside. A scan over the rows counts them and drops empty rows; then
// one thread per record writes it, in the same order as a walk over every
// brick of the box (rows in order, x within a row). Only touched bricks are
// enumerated: about 37% of a sphere batch's box bricks are not.
struct RowRef {
  uint32_t job;
  uint32_t yz;  // by | bz << 16 (job brick indices)
  uint32_t x0;  // first touched brick of the row (job brick index)
};

// The touched run [x0, x1) of row (by, bz) of job j, in job brick indices.
__device__ __forceinline__ void row_run(const BrushJob& j, uint32_t by, uint32_t bz, uint32_t& x0, uint32_t& x1) {
  const vx_brush& b = j.b;
  x0 = 0;
  x1 = j.nx;
  if (b.kind != 0) return;  // every brick of a box's clipped range is touched
  const int64_t gy = int64_t(j.by0) + by, gz = int64_t(j.bz0) + bz;
  const int64_t dy = max(int64_t(0), max(4 * gy - b.cy, int64_t(b.cy) - (4 * gy + 3)));
  const int64_t dz = max(int64_t(0), max(4 * gz - b.cz, int64_t(b.cz) - (4 * gz + 3)));
  const int64_t r2 = int64_t(b.hx) * b.hx - dy * dy - dz * dz;
  if (r2 < 0) {
    x1 = 0;
    return;
  }
  int64_t q = int64_t(sqrt(double(r2)));  // floor(sqrt(r2)), corrected to exact
  while (q * q > r2) --q;
  while ((q + 1) * (q + 1) <= r2) ++q;
  // Bricks g with voxels [4g, 4g + 3] meeting [cx - q, cx + q].
  auto floor4 = [](int64_t a) { return a >= 0 ? a / 4 : -((3 - a) / 4); };
  const int64_t lo = floor4(int64_t(b.cx) - q) - j.bx0, hi = floor4(int64_t(b.cx) + q) - j.bx0 + 1;
  const int64_t nx = j.nx;
  x0 = uint32_t(min(max(lo, int64_t(0)), nx));
  x1 = uint32_t(min(max(hi, int64_t(x0)), nx));
}

struct RowLoad {
  using Payload = RowRef;
  BrushCands c;
  uint32_t hint = kNone;
  uint32_t nj = 0;
  __device__ uint64_t operator()(uint32_t r, RowRef& p) {
    if (hint == kNone) nj = *c.n_jobs;
    p.job = c.job_of<true>(r, nj, hint);
    hint = p.job;
    const BrushJob j = c.jobs[p.job];
    uint32_t by, bz;
    divmod_small(r - j.row_start, j.ny, j.inv_ny, bz, by);
    uint32_t x0, x1;
    row_run(j, by, bz, x0, x1);
    p.yz = by | (bz << 16);
    p.x0 = x0;
    return scan::PairSumOp::make(x1 > x0 ? 1u : 0u, x1 - x0);
  }
};

// Records per chunk of the record kernel (a multiple of the wave width).
constexpr uint32_t kRecChunk = 1024;

// The touched rows, compacted: record start, job, brick y|z, first brick;
// and, per chunk of kRecChunk records, the row that holds its first record.
struct TouchedRows {
  uint32_t* rec;
  uint32_t* job;
  uint32_t* yz;
  uint32_t* x0;
  uint32_t* n_rows;
  uint32_t* chunk_row;
};

struct RowStore {
  TouchedRows rows;
  uint32_t* n_rec;
  __device__ void operator()(uint32_t r, const RowRef& p, uint64_t v, uint64_t excl, uint64_t incl,
                             uint32_t n) const {
    using PS = scan::PairSumOp;
    if (PS::lo(v)) {
      const uint32_t k = PS::lo(excl);
      const uint64_t e = PS::hi(excl), len = PS::hi(v);
      rows.rec[k] = uint32_t(e);
      rows.job[k] = p.job;
      rows.yz[k] = p.yz;
      rows.x0[k] = p.x0;
      // The chunk starts among this row's records [e, e + len).
      for (uint64_t ch = (e + kRecChunk - 1) / kRecChunk; ch * kRecChunk < e + len; ++ch)
        rows.chunk_row[ch] = k;
    }
    if (r == n - 1) {
      *rows.n_rows = PS::lo(incl);
      *n_rec = PS::hi(incl);
    }
  }
};

// What a record needs of its row, kept in LDS for the 64 rows of a window
// (four 16-byte words per row): the brick x of the batch's record 0 (record
// i has gx = gx0 + i), the key bits of its scene, y and z (key = those |
// spread3(gx)), the brush; and w: for a sphere of radius up to
// kSphere16MaxR, the row's sphere_row16 terms (what the row's bricks share
// of their masks), otherwise the row's brick y and z and the brush's
// cy, cz, hy, hz.
struct alignas(16) RowFields {
  uint32_t gx0, key_lo, key_hi, kind;
  int32_t cx, hx;
  uint32_t material, pad;
  uint32_t w[8];
};
static_assert(sizeof(RowFields) == 64, "RowFields layout");

constexpr uint32_t kRecBlock = 256;

// One thread per record. A wave takes chunks of kRecChunk consecutive records
// (grid stride); the row scan noted the row holding each chunk's first record.
// The wave loads a window of 64 consecutive rows (lane l: row w + l, and its
// brush), puts each row's fields in LDS and keeps the rows' first records in
// its lanes, then walks the window's records in steps of at most 64: each
// lane finds its record's row among the lanes (a 6-step search; the first
// records ascend over the lanes) and reads that row's fields from LDS. The
// memory round trips (rows, then brushes) come once per window, not once per
// step; the next window starts at the record where this one ends.
__global__ __launch_bounds__(kRecBlock) void brush_records_kernel(BrushCands c, TouchedRows rows, RecordOut out) {
  __shared__ RowFields s_rows[kRecBlock / kWave][kWave];
  RowFields* const wrows = s_rows[wave_id()];
  const uint32_t n = *out.n_rec, nr = *rows.n_rows;
  const uint32_t lane = lane_id();
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t chunks = (n + kRecChunk - 1) / kRecChunk;
  for (uint32_t ch = blockIdx.x * (blockDim.x / kWave) + wave_id(); ch < chunks; ch += waves) {
    const uint32_t r1 = uint32_t(min<uint64_t>(n, uint64_t(ch + 1) * kRecChunk));
    uint32_t i0 = ch * kRecChunk;
    uint32_t wrow = rows.chunk_row[ch];
    while (i0 < r1) {
      const uint32_t rr = wrow + lane;
      const bool have = rr < nr;
      const uint32_t start = have ? rows.rec[rr] : 0xffffffffu;  // ascending over the lanes
      const uint32_t job = have ? rows.job[rr] : 0u;
      const uint32_t yz = have ? rows.yz[rr] : 0u;
      const uint32_t x0 = have ? rows.x0[rr] : 0u;
      const uint32_t wend = wrow + kWave < nr ? rows.rec[wrow + kWave] : n;  // the window's end
      {
        const uint32_t* jp = reinterpret_cast<const uint32_t*>(c.jobs + job);
        const uint4 q0 = gld4(jp), q1 = gld4(jp + 4), q2 = gld4(jp + 8), q3 = gld4(jp + 12);
        RowFields f;
        f.gx0 = q2.x + x0 - start;  // BrushJob: bx0, by0, bz0 at words 8..10, scene at word 15
        const uint32_t gy = q2.y + (yz & 0xffffu), gz = q2.z + (yz >> 16);
        const uint64_t key = (uint64_t(q3.w) << c.codec.shift) | (spread3(gy) << 1) | (spread3(gz) << 2);
        f.key_lo = uint32_t(key);
        f.key_hi = uint32_t(key >> 32);
        f.kind = q1.z;
        f.cx = int32_t(q0.x);
        f.hx = int32_t(q0.w);
        f.material = c.materials ? q1.w : (q1.w ? 1u : 0u);
        f.pad = 0u;
        if (f.kind == 0 && f.hx <= kSphere16MaxR) {  // brick_mask's choice for this brush
          sphere_row16(int32_t(gy * 4) - int32_t(q0.y), int32_t(gz * 4) - int32_t(q0.z), f.hx * f.hx, f.w);
        } else {
          f.w[0] = gy;
          f.w[1] = gz;
          f.w[2] = q0.y;
          f.w[3] = q0.z;
          f.w[4] = q1.x;
          f.w[5] = q1.y;
          f.w[6] = f.w[7] = 0u;
        }
        wrows[lane] = f;  // read by other lanes of this wave only (LDS is in order within a wave)
      }
      const uint32_t stop = min(wend, r1);
      for (uint32_t s0 = i0; s0 < stop; s0 += kWave) {
        const uint32_t i = s0 + lane;
        // k = the number of lanes whose row starts at or before record i, minus one.
        uint32_t k = 0;
#pragma unroll
        for (uint32_t step = 32; step >= 1; step >>= 1) {
          const uint32_t probe = uint32_t(__shfl(start, int(k + step), kWave));
          if (k + step < kWave && probe <= i) k += step;
        }
        const RowFields g = wrows[k];
        if (i < stop) {
          const uint32_t gx = g.gx0 + i;
          out.put_key(i, ((uint64_t(g.key_hi) << 32) | g.key_lo) | spread3(gx));
          uint64_t mask;
          if (g.kind == 0 && g.hx <= kSphere16MaxR) {
            mask = sphere_mask16_row(int32_t(gx * 4) - g.cx, g.w);
          } else {
            const vx_brush b{g.cx, int32_t(g.w[2]), int32_t(g.w[3]), g.hx, int32_t(g.w[4]), int32_t(g.w[5]), g.kind,
                             g.material};
            mask = brick_mask(b, int32_t(gx * 4), int32_t(g.w[0] * 4), int32_t(g.w[1] * 4));
          }
          out.put_payload(i, mask, g.material);
        }
      }
      i0 = max(i0, stop);  // [i0, stop) done
      wrow += kWave;
    }
  }
}

struct VoxelLoad {
  using Payload = Record;
  const vx_voxel* edits;
  uint32_t world;
  uint32_t mat;
  __device__ uint32_t operator()(uint32_t i, Record& rec) const {
    const vx_voxel e = edits[i];
    if (e.x >= world || e.y >= world || e.z >= world) return 0u;  // ignored (build_leaf_masks/pipeline.cuh:25-31)
    rec.key = morton3(e.x >> 2, e.y >> 2, e.z >> 2);
    rec.mask = 1ull << ((e.x & 3u) + 4u * (e.y & 3u) + 16u * (e.z & 3u));
    rec.mat = mat;
    return 1u;
  }
};

struct RecordStore {
  RecordOut out;
  __device__ void operator()(uint32_t i, const Record& rec, uint32_t v, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    if (v) {
      out.put_key(excl, rec.key);
      out.put_payload(excl, rec.mask, rec.mat);
    }
    if (i == n - 1) *out.n_rec = incl;
  }
};

// ---- height structure -------------------------------------------------------
// Sorted records -> height 0 (distinct leaves with record ranges) and height
// 1 (their distinct parents with leaf ranges) in one pass: a pair count (new
// leaf, new parent) per record. A record starts a new parent when it starts a
// new leaf whose parent differs from the previous record's; entries of a
// scene whose root is at height 0 get no parent.
struct RunUpPayload {
  uint64_t key;
  bool leaf;
  bool inc;
};

// The depth of the scene of each key (reads of the scene table, all issued
// together, or the single scene's).
template <uint32_t N, class Raw>
__device__ __forceinline__ void fetch_depths(const Scenes& sc, Raw (&r)[N]) {
  if (sc.table) {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) r[k].depth = gld(&sc.table[sc.codec.scene(r[k].key)].depth);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) r[k].depth = sc.single.depth;
  }
}

// Split load (the reads of all of a thread's items first): the record's key
// and its predecessor's (index clamped, not branched on), then the scenes'
// depths.
struct RunUpLoad {
  using Payload = RunUpPayload;
  SortedRecs recs;
  Scenes sc;  // keys unused: scene from the record key
  struct Raw {
    uint64_t key, prev;
    uint32_t depth;
  };
  template <uint32_t N>
  __device__ void fetch(const uint32_t (&i)[N], uint32_t, Raw (&r)[N]) const {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) {
      r[k].key = recs.key(i[k]);
      r[k].prev = recs.key(i[k] ? i[k] - 1 : 0u);
    }
    fetch_depths<N>(sc, r);
  }
  __device__ uint64_t finish(uint32_t i, const Raw& r, RunUpPayload& p) const {
    p.key = r.key;
    p.leaf = i == 0 || r.prev != p.key;
    p.inc = r.depth > 0;
    const bool par = p.leaf && p.inc && (i == 0 || sc.codec.parent(r.prev) != sc.codec.parent(p.key));
    return scan::PairSumOp::make(p.leaf, par);
  }
};

struct RunUpStore {
  Codec codec;
  uint64_t* keys0;
  uint32_t* first0;
  uint32_t* parent0;
  uint64_t* keys1;
  uint32_t* first1;
  uint32_t* cnt;  // cnt[1] leaves, cnt[2] height-1 entries
  uint32_t* root_pos;
  __device__ void operator()(uint32_t i, const RunUpPayload& p, uint64_t v, uint64_t excl, uint64_t incl,
                             uint32_t n) const {
    using PS = scan::PairSumOp;
    const uint32_t leaves = PS::lo(incl), parents = PS::hi(incl);
    if (p.leaf) {
      const uint32_t l = leaves - 1;
      keys0[l] = p.key;
      first0[l] = i;
      parent0[l] = p.inc ? link_make(parents - 1, p.key) : kNone;
      if (!p.inc) root_pos[codec.scene(p.key)] = l;
      if (PS::hi(v)) {
        keys1[PS::hi(excl)] = codec.parent(p.key);
        first1[PS::hi(excl)] = l;
      }
    }
    if (i == n - 1) {
      cnt[1] = leaves;
      first0[leaves] = i + 1;
      cnt[2] = parents;
      first1[parents] = leaves;
    }
  }
};

// Height h -> h + 1: distinct parents; entries of a scene whose root sits at
// height h are its root and get no parent.
struct UpPayload {
  uint64_t key;
  uint32_t scene;
  bool inc;
};

struct UpLoad {  // split load, as RunUpLoad
  using Payload = UpPayload;
  const uint64_t* keys;
  Scenes sc;
  uint32_t h;
  struct Raw {
    uint64_t key, prev;
    uint32_t depth;
  };
  template <uint32_t N>
  __device__ void fetch(const uint32_t (&i)[N], uint32_t, Raw (&r)[N]) const {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) {
      r[k].key = gld(keys + i[k]);
      r[k].prev = gld(keys + (i[k] ? i[k] - 1 : 0u));
    }
    fetch_depths<N>(sc, r);
  }
  __device__ uint32_t finish(uint32_t i, const Raw& r, UpPayload& p) const {
    p.key = r.key;
    p.scene = sc.codec.scene(p.key);
    p.inc = r.depth > h;
    if (!p.inc) return 0u;
    return i == 0 || sc.codec.parent(r.prev) != sc.codec.parent(p.key);
  }
};

struct UpStore {
  Codec codec;
  uint64_t* up_keys;
  uint32_t* up_first;
  uint32_t* parent;
  uint32_t* up_n;
  uint32_t* root_pos;  // per scene: index of its root entry at its root height
  __device__ void operator()(uint32_t i, const UpPayload& p, uint32_t f, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    parent[i] = p.inc ? link_make(incl - 1, p.key) : kNone;
    if (!p.inc) root_pos[p.scene] = i;
    if (f) {
      up_keys[excl] = codec.parent(p.key);
      up_first[excl] = i;
    }
    if (i == n - 1) {
      *up_n = incl;
      up_first[incl] = i + 1;
    }
  }
};

// The ref entry j of height h had before the edit: its scene's root at the
// root height, else the matching child of its parent's old node.
__device__ __forceinline__ void topdown_entry(const Scenes& sc, const uint32_t* __restrict__ parent,
                                              const uint32_t* __restrict__ old_up, uint32_t* __restrict__ old,
                                              uint32_t h, uint32_t j) {
  const Dev& d = sc.of(j);
  const uint32_t link = parent[j];
  old[j] = d.depth == h ? d.ctl[kCtlRoot] : child_of(d.node_pool, old_up[link_parent(link)], link_slot(link));
}

__global__ void topdown_kernel(Scenes sc, const uint32_t* n_ptr, const uint32_t* __restrict__ parent,
                               const uint32_t* __restrict__ old_up, uint32_t* __restrict__ old, uint32_t h) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    topdown_entry(sc, parent, old_up, old, h, j);
}

// Old refs of the entries of heights h_lo..h_hi in one launch (one launch per
// height would cost a launch each for a few thousand entries): each entry
// walks down from its scene's root along its own key, through nodes that
// the other walks keep in L2 near the root. Same result as the height-by-
// height pass: the child of the parent's old node, a sentinel staying a
// sentinel.
struct OldRefHeights {
  const uint64_t* keys[kMaxHeights + 1];
  uint32_t* old[kMaxHeights + 1];
};

__global__ void old_refs_kernel(Scenes sc, const uint32_t* __restrict__ counts, OldRefHeights hs, uint32_t h_lo,
                                uint32_t h_hi) {
  __shared__ uint32_t start[kMaxHeights + 2];  // flat index of each height's first entry
  if (threadIdx.x == 0) {
    start[h_lo] = 0;
    for (uint32_t h = h_lo; h <= h_hi; ++h) start[h + 1] = start[h] + counts[1 + h];
  }
  __syncthreads();
  const uint32_t total = start[h_hi + 1];
  const uint64_t mask = (uint64_t(1) << sc.codec.shift) - 1;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    uint32_t h = h_lo;
    while (g >= start[h + 1]) ++h;
    const uint32_t j = g - start[h];
    const uint64_t* keys = hs.keys[h];
    const Dev& d = sc.of(keys, j);
    const uint64_t m = keys[j] & mask;
    uint32_t ref = d.ctl[kCtlRoot];
    for (uint32_t l = d.depth - h; l-- > 0 && ref > kFull;) ref = child_of(d.node_pool, ref, uint32_t(m >> (3 * l)) & 7u);
    hs.old[h][j] = ref;
  }
}

// ---- rebuild ---------------------------------------------------------------
// The leaf's old ref is found here (its parent's old node, one child read)
// instead of in a separate top-down pass over the leaves: this kernel is
// bandwidth-bound with many entries in flight, so the dependent reads hide
// behind the streaming work.
// The leaves the build leaves PENDING, listed in entry order without global
// atomics (one counter hit by every wave of the chip serialises) and without
// giving up the build's grid-stride order (the whole chip works through one
// window of entries, whose parents and old nodes then stay in L2): slice s =
// entries [64 s, 64 s + 64), the span one wave covers in one step, keeps its
// pending entries, in order, in pend[64 s ...] and their number in count[s];
// a fused scan over the slice counts gives each slice's base in the list,
// and the leaf probe walks the slices, writing the list as it goes.
constexpr uint32_t kLeafBlock = 256;

struct LeafSlices {
  uint32_t* pend;      // 64 per slice
  uint32_t* count;     // per slice
  uint32_t* base;      // per slice: exclusive prefix of count
  uint32_t* list;      // the pending entries, contiguous, in entry order
  uint32_t* n_slices;  // ceil(leaves / 64), written by the build
  uint32_t* total;     // pending entries
};

struct SliceLoad {
  using Payload = scan::NoPayload;
  const uint32_t* count;
  __device__ uint32_t operator()(uint32_t s, Payload&) const { return count[s]; }
};
struct SliceStore {
  uint32_t* base;
  uint32_t* total;
  __device__ void operator()(uint32_t s, const scan::NoPayload&, uint32_t, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    base[s] = excl;
    if (s == n - 1) *total = incl;
  }
};

// child_of with the node's header and its eight possible child words read
// together (one memory round trip instead of two; the reads past the node's
// end are clamped to the pool and unused). The child is picked with masks,
// not a select, so the compiler cannot fold the nine loads back into a
// header read followed by one dependent child read.
__device__ __forceinline__ uint32_t child_of_wide(const uint32_t* pool, uint64_t cap, uint32_t ref, uint32_t c) {
  const uint64_t base = ref > kFull ? ref : 0u;
  uint32_t v[9];
#pragma unroll
  for (uint32_t k = 0; k < 9; ++k) v[k] = gld(pool + min<uint64_t>(base + k, cap - 1));
  const uint32_t cm = v[0] & 0xffu, fm = (v[0] >> 8) & 0xffu, bit = 1u << c;
  const uint32_t at = 1u + __popc(cm & (bit - 1));
  uint32_t r = 0;
#pragma unroll
  for (uint32_t k = 1; k < 9; ++k) r |= v[k] & (0u - uint32_t(at == k));
  if (ref <= kFull) return ref;
  if (cm & bit) return r;
  return (fm & bit) ? kFull : kEmpty;
}

// One leaf entry j of the build: its old leaf (the parent's old node, one
// child), the entry's records applied in order, then the lookup among the
// interned leaves. The reads run in three rounds of independent loads
// (indices are clamped instead of branched on, so no round waits for a
// conditional region): the record range, parent link and scene root; the
// first record's index and the parent's old ref; the first record's mask and
// material (one 16-byte read) with the parent's old node. Most leaves have
// one record, so the records cost no round trip of their own. d is the
// scene's descriptor by value, so its fields are plain loads (kernel
// arguments in the single-scene case). Returns whether the entry is left
// PENDING (its item and hash written).
template <uint32_t LW>
__device__ __forceinline__ bool build_leaf(const Dev d, uint32_t j, const uint32_t* __restrict__ parent,
                                           const uint32_t* __restrict__ old_up, const uint32_t* __restrict__ first,
                                           const SortedRecs& recs, const uint4* __restrict__ rpay,
                                           uint32_t* __restrict__ items,
                                           uint64_t* __restrict__ hashes, uint64_t* __restrict__ res) {
  const bool top = d.depth == 0;  // the leaf is the scene's root
  const uint32_t r0 = first[j], r1 = first[j + 1];
  const uint32_t link = parent[j];
  const uint32_t root = gld(d.ctl + kCtlRoot);
  // n > 0 implies at least one record: record 0 stands in, unused, for an
  // entry without records; entry 0 of the level above stands in for the
  // root's missing parent.
  const uint32_t up = old_up[top ? 0u : link_parent(link)];
  const uint32_t idx0 = recs.index(r0 < r1 ? r0 : 0u);
  const uint4 pay0 = rpay[idx0];
  const uint32_t child = child_of_wide(d.node_pool, d.node_cap, up, link_slot(link));
  const uint32_t ref = top ? root : child;
  // A first record whose mask covers the whole brick (a brush interior)
  // overwrites every voxel, so the old leaf is not read.
  const bool covered = r0 < r1 && (pay0.x & pay0.y) == 0xffffffffu;
  uint32_t w[LW];
  if (ref == kEmpty || covered) {
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) w[k] = 0;
  } else if (ref == kFull) {
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) w[k] = 0xffffffffu;
  } else {
    load_item<LW>(d.leaf_pool + ref, LW, w);
  }
  // Records in order: the first from the registers loaded above, the rest
  // (rare) read here.
  auto apply = [&w](uint64_t m, uint32_t mat) {
    if constexpr (LW == 2) {
      const uint32_t lo = uint32_t(m), hi = uint32_t(m >> 32);
      if (mat) {
        w[0] |= lo;
        w[1] |= hi;
      } else {
        w[0] &= ~lo;
        w[1] &= ~hi;
      }
    } else {
      const uint32_t fill = mat * 0x01010101u;
#pragma unroll
      for (uint32_t k = 0; k < LW; ++k) {
        const uint32_t bm = nibble_bytes(uint32_t(m >> (4 * k)) & 15u);
        w[k] = (w[k] & ~bm) | (fill & bm);
      }
    }
  };
  auto mask_of = [](const uint4& q) { return (uint64_t(q.y) << 32) | q.x; };
  if (r0 < r1) apply(mask_of(pay0), pay0.z);
  for (uint32_t r = r0 + 1; r < r1; ++r) {
    const uint4 q = rpay[recs.index(r)];
    apply(mask_of(q), q.z);
  }
  uint32_t any = 0, all = 0xffffffffu;
#pragma unroll
  for (uint32_t k = 0; k < LW; ++k) {
    any |= w[k];
    all &= w[k];
  }
  if (!any) {
    res[j] = res_make(kResFinal, kEmpty);
    return false;
  }
  if (LW == 2 && all == 0xffffffffu) {
    res[j] = res_make(kResFinal, kFull);
    return false;
  }
  const uint64_t h = hash_words(w, LW);
  uint64_t start = 0;
  const uint32_t v = find_interned<LW>(table_of(d, true), w, LW, h, &start);
  if (v) {
    res[j] = res_make(kResFinal, v);
    return false;
  }
  uint32_t* it = items + size_t(j) * LW;
#pragma unroll
  for (uint32_t k = 0; k < LW; ++k) it[k] = w[k];
  // The probe starts where the lookup's walk ended (the slot index in the
  // low word, which the probe masks; the fingerprint stays in the high word).
  hashes[j] = (h & 0xFFFFFFFF00000000ull) | uint32_t(start);
  res[j] = res_make(kResPending, 0);
  return true;
}

template <uint32_t LW>
__global__ __launch_bounds__(kLeafBlock, 8) void leaf_build_kernel(
    Scenes sc, const uint32_t* n_ptr, const uint32_t* __restrict__ parent, const uint32_t* __restrict__ old_up,
    const uint32_t* __restrict__ first, SortedRecs recs, const uint4* __restrict__ rpay,
    uint32_t* __restrict__ items, uint64_t* __restrict__ hashes,
    uint64_t* __restrict__ res, LeafSlices ls) {
  const uint32_t n = *n_ptr;
  const uint32_t slices = (n + kWave - 1) / kWave;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ls.n_slices = slices;
  const uint32_t waves = gridDim.x * (kLeafBlock / kWave);
  for (uint32_t sl = blockIdx.x * (kLeafBlock / kWave) + wave_id(); sl < slices; sl += waves) {
    const uint32_t j = sl * kWave + lane_id();
    bool pending = false;
    if (j < n) {
      if (sc.table)
        pending = build_leaf<LW>(sc.table[sc.codec.scene(sc.keys[j])], j, parent, old_up, first, recs, rpay, items,
                                 hashes, res);
      else
        pending = build_leaf<LW>(sc.single, j, parent, old_up, first, recs, rpay, items, hashes, res);
    }
    const uint64_t pm = ballot(pending);
    if (pending) ls.pend[sl * kWave + lanes_below(pm)] = j;
    if (lane_id() == 0) ls.count[sl] = uint32_t(__popcll(pm));
  }
}

// Leaf probe over the slices: a wave takes kProbeSlices slices at a time and
// spreads their pending entries over its lanes (a slice holds a few pending
// entries on average, so a wave per slice would leave most lanes idle). The
// bases are an exclusive scan in slice order, so entry e of the group (in
// order) goes to list position base[first slice] + e.
constexpr uint32_t kProbeSlices = 16;

template <uint32_t kMax>
__global__ void probe_slices_kernel(ItemSet set, const uint64_t* __restrict__ hashes, LeafSlices ls, uint64_t* res,
                                    Scenes sc, uint32_t n_scenes) {
  if (blockIdx.x == 0)
    for (uint32_t s = threadIdx.x; s < n_scenes; s += blockDim.x) {
      const Table t = table_of(sc.at(s), true);
      *t.base = *t.used;
    }
  const uint32_t slices = *ls.n_slices;
  const uint32_t lane = lane_id();
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t s0 = (blockIdx.x * (blockDim.x / kWave) + wave_id()) * kProbeSlices; s0 < slices;
       s0 += waves * kProbeSlices) {
    const uint32_t c = lane < kProbeSlices && s0 + lane < slices ? ls.count[s0 + lane] : 0u;
    const uint32_t incl = wave_inclusive_sum(c);  // lanes >= kProbeSlices carry the group total
    const uint32_t total = wave_last(incl);
    const uint32_t b0 = ls.base[s0];
    for (uint32_t e = lane; e < total; e += kWave) {
      // Slice q of the group holding entry e: the number of slices whose
      // inclusive count is <= e; ex = the entries of the slices before it.
      uint32_t q = 0, ex = 0;
#pragma unroll
      for (uint32_t k = 0; k < kProbeSlices; ++k) {
        const uint32_t ik = uint32_t(__builtin_amdgcn_readlane(int(incl), int(k)));
        if (ik <= e) {
          q = k + 1;
          ex = ik;
        }
      }
      const uint32_t i = ls.pend[(s0 + q) * kWave + (e - ex)];
      ls.list[b0 + e] = i;
      probe_entry<kMax>(set, hashes, res, sc, true, i);
    }
  }
}

// The children of old node `ref` in two rounds of reads: its header, then one
// load per child slot (a slot without a child, and a sentinel ref, read the
// header's word or pool word 0 again: no branch on a loaded value, which
// would make each load wait for the one before, and no line outside the
// node). children_of reads the same words one after another.
__device__ __forceinline__ uint32_t old_header(const uint32_t* pool, uint32_t ref) {
  return gld(pool + (ref > kFull ? ref : 0u));
}
__device__ __forceinline__ void old_children_of_header(const uint32_t* pool, uint32_t ref, uint32_t hdr,
                                                       uint32_t (&ch)[8]) {
  const uint32_t cm = ref > kFull ? hdr & 0xffu : 0u, fm = (hdr >> 8) & 0xffu;
  uint32_t w[8];
#pragma unroll
  for (uint32_t c = 0; c < 8; ++c) {
    const uint32_t bit = 1u << c;
    w[c] = gld(pool + (ref > kFull ? ref + ((cm & bit) ? 1u + __popc(cm & (bit - 1u)) : 0u) : 0u));
  }
#pragma unroll
  for (uint32_t c = 0; c < 8; ++c) {
    const uint32_t bit = 1u << c;
    ch[c] = ref <= kFull ? ref : (cm & bit) ? w[c] : (fm & bit) ? kFull : kEmpty;
  }
}
__device__ __forceinline__ void old_children(const uint32_t* pool, uint32_t ref, uint32_t (&ch)[8]) {
  old_children_of_header(pool, ref, old_header(pool, ref), ch);
}

// Grid form: a block takes 256 consecutive parents; their child entries are
// one contiguous range of the height below, read coalesced by the whole block
// into per-parent LDS slots (a per-thread walk over each parent's children
// touches ~6 scattered entries per lane per step, and the lines it drags in
// are evicted before the next step), then each thread merges its slots into
// the old node's children. Every read goes out in rounds without branches on
// loaded values (a load under such a branch makes the next one wait for it):
// the old ref and the old node's header before the child entries; the child
// entries 8 per thread at a time (links and results, then the results of
// duplicates); the old node's child words, one load per child slot (a slot
// without a child reads the header's word again, so no line outside the
// node is touched).
constexpr uint32_t kBuildBlock = 256;
constexpr uint32_t kBuildBatch = 8;

struct BuildArgs {
  Scenes sc;
  const uint32_t* n_ptr;
  const uint32_t* old;
  const uint32_t* first;
  const uint32_t* cparent;
  const uint64_t* cres;
  uint32_t* items;
  uint64_t* hashes;
  uint64_t* res;
};

// Blocks `block` of `blocks` (grid stride) of a height's build.
__device__ __forceinline__ void build_nodes(const BuildArgs& a, uint32_t block, uint32_t blocks) {
  __shared__ uint32_t s_ch[kBuildBlock * 8];
  const Scenes& sc = a.sc;
  const uint32_t* __restrict__ old = a.old;
  const uint32_t* __restrict__ first = a.first;
  const uint32_t* __restrict__ cparent = a.cparent;
  const uint64_t* __restrict__ cres = a.cres;
  uint32_t* __restrict__ items = a.items;
  uint64_t* __restrict__ hashes = a.hashes;
  uint64_t* __restrict__ res = a.res;
  const uint32_t n = *a.n_ptr;
  const uint32_t t = threadIdx.x;
  for (uint32_t j0 = block * kBuildBlock; j0 < n; j0 += blocks * kBuildBlock) {
    const uint32_t jn = min(j0 + kBuildBlock, n);
    const uint32_t j = j0 + t;
    const bool live = j < n;
    const uint32_t* pool = sc.of(live ? j : j0).node_pool;
    const uint32_t o = live ? old[j] : kEmpty;
    const uint32_t hdr = old_header(pool, o);
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) s_ch[c * kBuildBlock + t] = kNone;
    __syncthreads();
    const uint32_t kb = first[j0], k1 = first[jn];
    for (uint32_t k0 = kb; k0 < k1; k0 += kBuildBatch * kBuildBlock) {
      uint32_t link[kBuildBatch];
      uint64_t r[kBuildBatch];
#pragma unroll
      for (uint32_t u = 0; u < kBuildBatch; ++u) {
        const uint32_t k = k0 + u * kBuildBlock + t;
        const uint32_t kk = k < k1 ? k : kb;
        link[u] = cparent[kk];
        r[u] = cres[kk];
      }
      uint64_t rep[kBuildBatch];  // a duplicate's representative (the same line otherwise)
#pragma unroll
      for (uint32_t u = 0; u < kBuildBatch; ++u) {
        const uint32_t k = k0 + u * kBuildBlock + t;
        rep[u] = cres[res_kind(r[u]) == kResDup ? uint32_t(r[u]) : (k < k1 ? k : kb)];
      }
#pragma unroll
      for (uint32_t u = 0; u < kBuildBatch; ++u) r[u] = res_kind(r[u]) == kResDup ? rep[u] : r[u];
#pragma unroll
      for (uint32_t u = 0; u < kBuildBatch; ++u) {
        const uint32_t k = k0 + u * kBuildBlock + t;
        const uint32_t pj = link_parent(link[u]);
        if (k < k1 && pj >= j0 && pj < jn) s_ch[link_slot(link[u]) * kBuildBlock + (pj - j0)] = uint32_t(r[u]);
      }
    }
    __syncthreads();
    if (live) {
      uint32_t ch[8];
      old_children_of_header(pool, o, hdr, ch);
#pragma unroll
      for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t v = s_ch[c * kBuildBlock + t];
        if (v != kNone) ch[c] = v;
      }
      // The item in registers (hashed there, not read back from memory).
      uint32_t item[kNodeStride];
      uint32_t ref = 0;
      const uint32_t w = make_node_item_regs(ch, item, &ref);
      if (!w) {
        res[j] = res_make(kResFinal, ref);
      } else {
        uint32_t* it = items + size_t(j) * kNodeStride;
#pragma unroll
        for (uint32_t q = 0; q < kNodeStride; ++q) it[q] = item[q];
        hashes[j] = hash_words_regs(item, w);
        res[j] = res_make(kResPending, 0);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBuildBlock) void node_build_kernel(BuildArgs a) { build_nodes(a, blockIdx.x, gridDim.x); }

// One launch for two independent steps: the place of height h (its first
// place_blocks blocks; results are already FINAL after the final-mode
// commit) and the build of height h + 1 (the other blocks), which reads
// only those results. The heights' items and hashes sit in separate
// buffers (even and odd heights).
template <uint32_t kMax>
__global__ __launch_bounds__(kBuildBlock) void build_and_place_kernel(PlaceArgs pa, uint32_t place_blocks,
                                                                      BuildArgs ba) {
  if (blockIdx.x < place_blocks)
    place_noted<kMax>(pa, blockIdx.x, place_blocks);
  else
    build_nodes(ba, blockIdx.x - place_blocks, gridDim.x - place_blocks);
}

struct ResByHeight {
  const uint64_t* p[kMaxHeights];
};

__global__ void publish_kernel(Scenes sc, uint32_t n_scenes, const uint32_t* __restrict__ root_pos, ResByHeight rh) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_scenes; s += gridDim.x * blockDim.x) {
    const uint32_t pos = root_pos[s];
    const Dev& d = sc.at(s);
    if (pos == kNone || *d.status != kStatusOk) continue;
    d.ctl[kCtlRoot] = resolve(rh.p[d.depth], pos);
  }
}

// Start of an edit, one launch (a fill per buffer would cost a launch each):
// every scene's status OK, the first `n_counts` count words zero, no root
// position yet, and the fused scans' claim counters and look-back rows zero.
__global__ void begin_kernel(Scenes sc, uint32_t n_scenes, uint32_t* counts, uint32_t n_counts, uint32_t* root_pos,
                             uint32_t* fused, uint64_t fused_words) {
  const uint64_t t0 = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x, stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t s = t0; s < n_scenes; s += stride) {
    *sc.at(uint32_t(s)).status = kStatusOk;
    root_pos[s] = kNone;
  }
  for (uint64_t i = t0; i < n_counts; i += stride) counts[i] = 0;
  for (uint64_t i = t0; i < fused_words; i += stride) fused[i] = 0;
}

// ---- workspace ---------------------------------------------------------------
struct Heights {
  uint64_t* keys[kMaxHeights + 1];
  uint32_t* first[kMaxHeights + 1];
  uint32_t* parent[kMaxHeights + 1];
  uint32_t* old[kMaxHeights + 1];
  uint64_t* res[kMaxHeights + 1];
  uint64_t cap[kMaxHeights + 1];
};

constexpr uint32_t kCntRows = 58, kCntSlices = 59, kCntPending = 60;

struct EditWs {
  uint32_t* counts;  // [0] records, [1 + h] entries at height h, [kCntPending] pending leaves
  uint32_t* root_pos;
  uint32_t* status;  // per-scene first-error words of an executor's edits
  uint32_t* base;    // per-scene pool snapshots of an executor's edits (2 words each)
  uint64_t *rkey, *rkey_alt;
  uint4* rpay;  // per record: brick mask and material (RecordOut::pay)
  uint32_t *rval, *rval_alt;
  void* sort_ws;
  sort::Tuning sort_tn;  // the tuning sort_ws was sized for
  uint32_t index_bits;  // > 0: records sort as packed u64 keys (RecordOut)
  Heights hs;
  uint32_t* items;       // even heights (leaves included)
  uint32_t* items_odd;   // odd heights
  uint64_t* hashes;      // even heights
  uint64_t* hashes_odd;  // odd heights
  uint32_t* slot_idx;    // per commit position: the slot its place publishes (CommitStore)
  uint32_t* items_at(uint32_t h) const { return (h & 1) ? items_odd : items; }
  uint64_t* hashes_at(uint32_t h) const { return (h & 1) ? hashes_odd : hashes; }
  // Workspaces of the pipeline's fused scans, one per scan so that a single
  // fill at the start of an edit clears them all.
  uint32_t* fused_all;
  size_t fused_bytes;
  void* rec_ws;
  void* run_ws;
  void* up_ws[kMaxHeights + 1];
  void* commit_ws[kMaxHeights + 1];
  uint32_t *pend, *pend_list, *slice_count, *slice_base;  // pending leaves (LeafSlices)
  uint32_t *row_rec, *row_job, *row_yz, *row_x0;          // touched brick rows (TouchedRows)
  uint32_t* chunk_row;                                     // per kRecChunk records (TouchedRows)
  void* slice_ws;
  BrushJob* jobs;
  Dev* scenes;
  size_t scenes_off, jobs_off;  // byte offsets from counts (the workspace base)
  size_t bytes;
};

// ---- tail: the small top heights in one workgroup ---------------------------
// Heights with at most one entry per thread (capacity <= 1024) are processed
// by a single 1024-thread workgroup that walks them in order with a barrier
// between steps, instead of one launch per step and height: near the root
// every launch would handle a handful of entries and its fixed cost (launch,
// first memory round trip, drain) would dominate. (With several entries per
// thread the dependent loads of each entry would run back to back, so larger
// heights stay on the grid kernels.) Global writes of one step are visible
// to the whole workgroup after the barrier.
constexpr uint64_t kTailCap = scan::kFusedBlock;

struct TailArgs {
  Heights hs;
  Scenes sc;  // keys set per height
  Codec codec;
  uint32_t* cnt;
  uint32_t* root_pos;
  uint32_t* items;
  uint64_t* hashes;
  uint32_t n_scenes;
  uint32_t h_lo, h_hi;  // heights [h_lo, h_hi]
};

__device__ __forceinline__ uint32_t tail_count(const TailArgs& a, uint32_t h) {
  return min(a.cnt[1 + h], uint32_t(kTailCap));
}

// Parents of heights h_lo..h_hi (the UpLoad / UpStore scan, one tile each).
__global__ __launch_bounds__(scan::kFusedBlock) void tail_up_kernel(TailArgs a) {
  __shared__ uint32_t s_wave[scan::kFusedBlock / kWave];
  __shared__ uint32_t s_prefix;
  for (uint32_t h = a.h_lo; h <= a.h_hi; ++h) {
    Scenes sh = a.sc;
    sh.keys = a.hs.keys[h];
    scan::fused_tile<scan::SumOp, 1>(0u, tail_count(a, h), false, UpLoad{a.hs.keys[h], sh, h},
                                  UpStore{a.codec, a.hs.keys[h + 1], a.hs.first[h + 1], a.hs.parent[h], &a.cnt[2 + h],
                                          a.root_pos},
                                  nullptr, s_wave, &s_prefix);
    __syncthreads();
  }
}

// Old refs of heights h_hi down to h_lo.

// Node items of heights h_lo..h_hi: build, probe, commit, place. Each entry
// stays with its thread from build to place (its item, hash and result in
// registers), the height's items sit in LDS for the tagged-slot compares, and
// the results of one height stay in LDS as the child refs of the next: the
// only global round trips left per height are the old node, the level links,
// the slot probe and the stores.
struct TailRegLoad {
  using Payload = scan::NoPayload;
  uint64_t v;
  __device__ uint64_t operator()(uint32_t, scan::NoPayload&) const { return v; }
};
struct TailRegStore {
  uint32_t* incl;
  __device__ void operator()(uint32_t, const scan::NoPayload&, uint64_t, uint64_t, uint64_t in, uint32_t) const {
    *incl = uint32_t(in);
  }
};

__global__ __launch_bounds__(scan::kFusedBlock) void tail_nodes_kernel(TailArgs a) {
  __shared__ uint64_t s_wave[scan::kFusedBlock / kWave];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_item[kTailCap * kNodeStride];
  __shared__ uint64_t s_res[2][kTailCap];
  const uint32_t t = threadIdx.x;
  for (uint32_t h = a.h_lo; h <= a.h_hi; ++h) {
    const Scenes& sh = a.sc;
    const uint64_t* keys = a.hs.keys[h];
    const uint32_t n = tail_count(a, h);
    const bool act = t < n;
    const bool lds_children = h > a.h_lo;  // the height below was built here
    const uint64_t* cres_lds = s_res[(h - 1) & 1];
    const uint64_t* cres_glb = a.hs.res[h - 1];
    // Build.
    uint32_t it[kNodeStride] = {};
    uint32_t w = 0;
    uint64_t r = 0, hash = 0;
    if (act) {
      // Reads in rounds without branches on loaded values (as in
      // node_build_kernel): the old ref, the range of child entries (at most
      // 8), the old node's header; its child words and the entries' links
      // and results; the representatives of duplicates.
      uint32_t ch[8];
      const uint32_t* pool = sh.of(keys, t).node_pool;
      const uint32_t o = a.hs.old[h][t];
      const uint32_t k0 = a.hs.first[h][t], k1 = a.hs.first[h][t + 1];
      old_children(pool, o, ch);
      uint32_t link[8];
      uint64_t cr[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t k = k0 + u < k1 ? k0 + u : k0;
        link[u] = a.hs.parent[h - 1][k];
        cr[u] = lds_children ? cres_lds[k] : cres_glb[k];
      }
      uint64_t rep[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        const uint32_t k = res_kind(cr[u]) == kResDup ? uint32_t(cr[u]) : (k0 + u < k1 ? k0 + u : k0);
        rep[u] = lds_children ? cres_lds[k] : cres_glb[k];
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) {
        if (k0 + u >= k1 || link_parent(link[u]) != t) continue;
        const uint32_t v = uint32_t(res_kind(cr[u]) == kResDup ? rep[u] : cr[u]);
        const uint32_t slot = link_slot(link[u]);
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c)
          if (c == slot) ch[c] = v;  // constant indices: ch stays in registers
      }
      uint32_t ref = 0;
      w = make_node_item_regs(ch, it, &ref);
      if (!w) {
        r = res_make(kResFinal, ref);
      } else {
        hash = hash_words_regs(it, w);
        r = res_make(kResPending, 0);
      }
#pragma unroll
      for (uint32_t q = 0; q < kNodeStride; ++q) s_item[t * kNodeStride + q] = it[q];
    }
    for (uint32_t sidx = t; sidx < a.n_scenes; sidx += blockDim.x) {
      const Table tb = table_of(sh.at(sidx), false);
      *tb.base = *tb.used;
    }
    __syncthreads();
    // Probe: an equal item in the table or earlier in this height, else claim
    // an empty slot (the same protocol as probe_entry, with the item held in
    // registers and this height's tagged items read from LDS).
    if (act && res_kind(r) == kResPending) {
      const Dev& d = sh.of(keys, t);
      const Table tb = table_of(d, false);
      const uint32_t fp = uint32_t(hash >> 32);
      uint64_t idx = hash & tb.mask;
      const uint64_t limit = tb.mask < kMaxProbe ? tb.mask + 1 : kMaxProbe;
      uint64_t out = res_make(kResFinal, kEmpty);
      bool done = false;
      for (uint64_t probe = 0; probe < limit && !done; ++probe) {
        uint64_t sv = tb.slots[idx];
        if (sv == 0) {
          const uint64_t tagged = (uint64_t(fp) << 32) | kTag | t;
          const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&tb.slots[idx]), 0ull,
                                          static_cast<unsigned long long>(tagged));
          if (prev == 0) {
            out = res_make(kResRep, uint32_t(idx));
            done = true;
            break;
          }
          sv = prev;
        }
        if (uint32_t(sv >> 32) == fp) {
          const uint32_t v = uint32_t(sv);
          if (v & kTag) {
            const uint32_t j = v & ~kTag;
            uint32_t diff = 0;  // words past the width are zero in both
#pragma unroll
            for (uint32_t q = 0; q < kNodeStride; ++q) diff |= it[q] ^ s_item[j * kNodeStride + q];
            if (!diff) {
              out = res_make(kResDup, j);
              done = true;
              break;
            }
          } else if (v >= kPoolStart) {
            if (uint64_t(v) + w <= tb.cap && item_equal<kNodeStride>(it, tb.pool + v, w)) {
              out = res_make(kResFinal, v);
              done = true;
              break;
            }
          }
        }
        idx = (idx + 1) & tb.mask;
      }
      if (!done) set_status(d, kStatusOverflow);
      r = out;
    }
    __syncthreads();  // every claim of this height is in the table before pool offsets are handed out
    // Commit: pool offsets of the representatives (a segmented sum of their
    // widths per scene), then place: copy, publish the slot, become final.
    const uint32_t wr = (act && res_kind(r) == kResRep) ? w : 0u;
    const uint32_t id = act ? sh.id(keys, t) : 0u;
    const bool head = act && (t == 0 || sh.id(keys, t - 1) != id);
    uint32_t incl = 0;
    scan::fused_tile<scan::SegSumOp, 1>(0u, n, false, TailRegLoad{scan::SegSumOp::make(wr, head)},
                                        TailRegStore{&incl}, nullptr, s_wave, &s_prefix);
    if (act) {
      const Dev& d = sh.of(keys, t);
      const Table tb = table_of(d, false);
      const uint32_t base = *tb.base;
      if (wr) {
        const uint64_t off = uint64_t(base) + incl - wr;
        const uint64_t fp_hi = uint64_t(uint32_t(hash >> 32)) << 32;
        if (off + wr > tb.cap || off >= kTag) {
          set_status(d, kStatusOverflow);
          st_agent(&tb.slots[uint32_t(r)], fp_hi | kTomb);
          r = res_make(kResFinal, kEmpty);
        } else {
          store_item<kNodeStride>(tb.pool + off, wr, it);
          st_agent(&tb.slots[uint32_t(r)], fp_hi | uint32_t(off));
          r = res_make(kResFinal, uint32_t(off));
        }
      }
      if (t == n - 1 || sh.id(keys, t + 1) != id) {
        const uint64_t used = uint64_t(base) + incl;
        *tb.used = used > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(used);
      }
      s_res[h & 1][t] = r;
      a.hs.res[h][t] = r;
    }
    __syncthreads();
  }
}

struct Plan {
  uint32_t n_scenes = 1;
  uint32_t max_depth = 0;
  uint32_t leaf_words = 2;
  uint32_t key_bits = 0;
  Codec codec{0};
};

Plan make_plan(const std::vector<const vx_dag*>& dags) {
  Plan p;
  p.n_scenes = uint32_t(dags.size());
  for (const vx_dag* g : dags) p.max_depth = std::max(p.max_depth, g->depth);
  p.leaf_words = dags[0]->dev.leaf_words;
  p.codec.shift = 3 * p.max_depth;
  uint32_t sbits = 0;
  while ((1u << sbits) < p.n_scenes) ++sbits;
  p.key_bits = p.codec.shift + sbits;
  return p;
}

EditWs carve(void* base, const Plan& p, uint64_t r_ub, uint32_t n_jobs, uint64_t candidates) {
  EditWs w{};
  Carver c(base);
  // The host uploads [counts, scene table, brush jobs] as one block.
  w.counts = c.take<uint32_t>(64);
  w.scenes_off = align_up(c.off, 256);
  w.scenes = c.take<Dev>(p.n_scenes);
  w.jobs_off = align_up(c.off, 256);
  w.jobs = c.take<BrushJob>(std::max<uint32_t>(1, n_jobs));
  w.root_pos = c.take<uint32_t>(p.n_scenes);
  w.status = c.take<uint32_t>(p.n_scenes);
  w.base = c.take<uint32_t>(2 * size_t(p.n_scenes));
  const uint64_t R = std::max<uint64_t>(1, r_ub);
  // Packed records when the index of the last record fits below the key
  // (VOXEL_AMD_SORT_PAIRS=1 forces the pair layout, for testing).
  uint32_t ib = 1;
  while (ib < 32 && (uint64_t(1) << ib) < R) ++ib;
  const char* force_pairs = std::getenv("VOXEL_AMD_SORT_PAIRS");
  const bool pairs = force_pairs && force_pairs[0] == '1';
  w.index_bits = !pairs && std::max(p.key_bits, 1u) + ib <= 64 ? ib : 0u;
  w.rkey = c.take<uint64_t>(R);
  w.rkey_alt = c.take<uint64_t>(R);
  w.rpay = c.take<uint4>(R);
  w.rval = w.index_bits ? nullptr : c.take<uint32_t>(R);
  w.rval_alt = w.index_bits ? nullptr : c.take<uint32_t>(R);
  w.sort_tn = w.index_bits ? sort::keys_u64_dev_tuning(std::max(p.key_bits, 1u)) : sort::pairs_u64_dev_tuning();
  w.sort_ws = c.take<char>(w.index_bits ? sort::keys_u64_dev_workspace_bytes(R, std::max(p.key_bits, 1u), w.sort_tn)
                                        : sort::pairs_u64_dev_workspace_bytes(R, p.key_bits, w.sort_tn));
  // Items and hashes of even heights (the leaves included) and of odd
  // heights in separate buffers: a height's place runs beside the next
  // height's build.
  uint64_t max_cap[2] = {1, 1}, max_items[2] = {1, 1}, max_cap_all = 1;
  const uint32_t H = p.max_depth;
  for (uint32_t h = 0; h <= H + 1; ++h) {
    uint64_t cap = R;
    if (h > 0) {
      const uint32_t lv = h <= H ? H - h : 0;
      if (3 * lv < 50) cap = std::min<uint64_t>(R, uint64_t(p.n_scenes) << (3 * lv));
    }
    w.hs.cap[h] = cap;
    w.hs.keys[h] = c.take<uint64_t>(cap);
    w.hs.first[h] = c.take<uint32_t>(cap + 1);
    w.hs.parent[h] = c.take<uint32_t>(cap);
    w.hs.old[h] = c.take<uint32_t>(cap);
    w.hs.res[h] = c.take<uint64_t>(cap);
    max_cap[h & 1] = std::max(max_cap[h & 1], cap);
    max_cap_all = std::max(max_cap_all, cap);
    max_items[h & 1] = std::max<uint64_t>(max_items[h & 1], cap * (h == 0 ? p.leaf_words : kNodeStride));
  }
  w.items = c.take<uint32_t>(max_items[0]);
  w.items_odd = c.take<uint32_t>(max_items[1]);
  w.slot_idx = c.take<uint32_t>(max_cap_all);
  const uint64_t slices = (w.hs.cap[0] + kWave - 1) / kWave;
  w.pend = c.take<uint32_t>(slices * kWave);
  w.pend_list = c.take<uint32_t>(w.hs.cap[0]);
  // Touched brick rows of a brush batch: at most one per candidate brick.
  const uint64_t rows_ub = std::max<uint64_t>(candidates, 1);
  w.row_rec = c.take<uint32_t>(rows_ub);
  w.row_job = c.take<uint32_t>(rows_ub);
  w.row_yz = c.take<uint32_t>(rows_ub);
  w.row_x0 = c.take<uint32_t>(rows_ub);
  w.chunk_row = c.take<uint32_t>(R / kRecChunk + 2);
  w.slice_count = c.take<uint32_t>(slices);
  w.slice_base = c.take<uint32_t>(slices);
  w.hashes = c.take<uint64_t>(max_cap[0]);
  w.hashes_odd = c.take<uint64_t>(max_cap[1]);
  {
    char* start = c.take<char>(0);
    w.rec_ws = c.take<char>(scan::fused_dev_workspace_bytes(std::max<uint64_t>(candidates, 1)));
    w.run_ws = c.take<char>(scan::fused_dev_workspace_bytes(R));
    w.slice_ws = c.take<char>(scan::fused_dev_workspace_bytes((w.hs.cap[0] + kWave - 1) / kWave));
    for (uint32_t h = 0; h <= H; ++h) {
      w.up_ws[h] = c.take<char>(scan::fused_dev_workspace_bytes(w.hs.cap[h]));
      w.commit_ws[h] = c.take<char>(scan::fused_dev_workspace_bytes(w.hs.cap[h]));
    }
    char* end = c.take<char>(0);
    w.fused_all = reinterpret_cast<uint32_t*>(start);
    w.fused_bytes = size_t(end - start);
  }
  w.bytes = c.size();
  return w;
}

inline void mark(vx_dag* g, uint32_t i, hipStream_t s) {
  if (g && g->profile && g->stage_events) hipEventRecord(g->ev[i], s);
}

// Everything after record generation: sort, heights, old refs, rebuild,
// publish. `g` (may be null) receives the stage timings.
vx_status rebuild(vx_dag* g, const Plan& p, const Scenes& sc, EditWs& w, uint64_t r_ub, hipStream_t s) {
  const uint32_t H = p.max_depth;
  uint32_t* cnt = w.counts;
  Heights& L = w.hs;
  vx_status st;
  bool in_alt = false;
  mark(g, 1, s);
  if (w.index_bits)
    st = sort::keys_u64_dev(w.rkey, w.rkey_alt, &cnt[0], std::max<uint64_t>(r_ub, 1), w.index_bits,
                            std::max(p.key_bits, 1u), w.sort_ws, w.sort_tn, &in_alt, s);
  else
    st = sort::pairs_u64_dev(w.rkey, w.rval, w.rkey_alt, w.rval_alt, &cnt[0], std::max<uint64_t>(r_ub, 1),
                             std::max(p.key_bits, 1u), w.sort_ws, w.sort_tn, &in_alt, s);
  if (st != VX_OK) return st;
  mark(g, 2, s);
  const SortedRecs recs{in_alt ? w.rkey_alt : w.rkey, in_alt ? w.rval_alt : w.rval, w.index_bits};
  {
    TimedLaunch tl(VX_K_RUNS, s);
    Scenes sr = sc;
    sr.keys = nullptr;
    st = scan::fused_dev_m<scan::PairSumOp>(
        &cnt[0], r_ub, RunUpLoad{recs, sr},
        RunUpStore{p.codec, L.keys[0], L.first[0], L.parent[0], L.keys[1], L.first[1], cnt, w.root_pos}, w.run_ws, s,
        false);
  }
  if (st != VX_OK) return st;
  // Heights from hu up go to the single-workgroup tail kernels.
  uint32_t hu = H + 1;
  for (uint32_t h = 1; h <= H; ++h)
    if (L.cap[h] <= kTailCap) {
      hu = h;
      break;
    }
  const TailArgs ta{L, sc, p.codec, cnt, w.root_pos, w.items, w.hashes, p.n_scenes, hu, H};
  for (uint32_t h = 1; h < hu; ++h) {
    Scenes sh = sc;
    sh.keys = L.keys[h];
    {
      TimedLaunch tl(VX_K_UP, s);
      st = scan::fused_dev<scan::kSum, scan::kFusedItemsSmall>(&cnt[1 + h], L.cap[h], UpLoad{L.keys[h], sh, h},
                                       UpStore{p.codec, L.keys[h + 1], L.first[h + 1], L.parent[h], &cnt[2 + h],
                                               w.root_pos},
                                       w.up_ws[h], s, false);
    }
    if (st != VX_OK) return st;
  }
  if (hu <= H) {
    TimedLaunch tl(VX_K_TAIL_UP, s);
    tail_up_kernel<<<1, scan::kFusedBlock, 0, s>>>(ta);
  }
  if ((st = last_launch()) != VX_OK) return st;
  mark(g, 3, s);
  // Heights 2..H walk from the root in one launch; height 1 (millions of
  // entries) reads its parent's old node; the leaves' old refs are read by
  // the leaf build.
  if (H >= 2) {
    OldRefHeights oh{};
    uint64_t walkers = 0;
    for (uint32_t h = 2; h <= H; ++h) {
      oh.keys[h] = L.keys[h];
      oh.old[h] = L.old[h];
      walkers += L.cap[h];
    }
    TimedLaunch tl(VX_K_TOPDOWN, s);
    old_refs_kernel<<<grid_for(walkers), 256, 0, s>>>(sc, cnt, oh, 2, H);
  }
  if ((st = last_launch()) != VX_OK) return st;
  if (H >= 1) {
    Scenes sh = sc;
    sh.keys = L.keys[1];
    TimedLaunch tl(VX_K_TOPDOWN, s);
    topdown_kernel<<<grid_for(L.cap[1]), 256, 0, s>>>(sh, &cnt[1 + 1], L.parent[1], L.old[2], L.old[1], 1);
  }
  if ((st = last_launch()) != VX_OK) return st;
  mark(g, 4, s);
  PlaceArgs pending{};  // the place step not launched yet (results already FINAL)
  {
    Scenes sh = sc;
    sh.keys = L.keys[0];
    const LeafSlices ls{w.pend, w.slice_count, w.slice_base, w.pend_list, &cnt[kCntSlices], &cnt[kCntPending]};
    {
      TimedLaunch tl(VX_K_LEAF_BUILD, s);
      if (p.leaf_words == 2)
        leaf_build_kernel<2><<<grid_for(L.cap[0]), kLeafBlock, 0, s>>>(sh, &cnt[1], L.parent[0], L.old[1],
                                                                       L.first[0], recs, w.rpay, w.items,
                                                                       w.hashes, L.res[0], ls);
      else
        leaf_build_kernel<16><<<grid_for(L.cap[0]), kLeafBlock, 0, s>>>(sh, &cnt[1], L.parent[0], L.old[1],
                                                                        L.first[0], recs, w.rpay, w.items,
                                                                        w.hashes, L.res[0], ls);
    }
    if ((st = last_launch()) != VX_OK) return st;
    {
      // The slice scan is timed with the probe (one kind).
      TimedLaunch tl(VX_K_LEAF_PROBE, s);
      const uint64_t slices_ub = (L.cap[0] + kWave - 1) / kWave;
      if ((st = scan::fused_dev_m<scan::SumOp>(ls.n_slices, slices_ub, SliceLoad{ls.count},
                                               SliceStore{ls.base, ls.total}, w.slice_ws, s, false)) != VX_OK)
        return st;
      if (p.leaf_words == 2)
        probe_slices_kernel<2><<<grid_for(L.cap[0]), 256, 0, s>>>(ItemSet{w.items, 2, 2}, w.hashes, ls, L.res[0], sh,
                                                                  p.n_scenes);
      else
        probe_slices_kernel<16><<<grid_for(L.cap[0]), 256, 0, s>>>(ItemSet{w.items, 16, 16}, w.hashes, ls, L.res[0],
                                                                   sh, p.n_scenes);
    }
    if ((st = last_launch()) != VX_OK) return st;
    const ItemSet leaf_set{w.items, p.leaf_words, p.leaf_words};
    st = p.leaf_words == 2 ? commit_final<2>(sh, p.n_scenes, true, leaf_set, w.hashes, ls.total, ls.list, L.cap[0],
                                             L.res[0], w.slot_idx, w.commit_ws[0], s, false)
                           : commit_final<16>(sh, p.n_scenes, true, leaf_set, w.hashes, ls.total, ls.list,
                                              L.cap[0], L.res[0], w.slot_idx, w.commit_ws[0], s, false);
    if (st != VX_OK) return st;
    pending = PlaceArgs{leaf_set, w.hashes, w.slot_idx, ls.total, ls.list, sh, true};
  }
  mark(g, 5, s);
  // Each height's place runs in one launch with the next height's build
  // (build_and_place_kernel); the last grid height's place runs alone.
  uint32_t pending_kmax = p.leaf_words;
  uint64_t pending_ub = L.cap[0];
  auto place_alone = [&]() {
    TimedLaunch tl(pending.leaves ? VX_K_LEAF_PLACE : VX_K_NODE_PLACE, s);
    const uint32_t grid = grid_for(pending_ub);
    if (pending_kmax == 2) place_noted_kernel<2><<<grid, 256, 0, s>>>(pending);
    else if (pending_kmax == 16) place_noted_kernel<16><<<grid, 256, 0, s>>>(pending);
    else place_noted_kernel<kNodeStride><<<grid, 256, 0, s>>>(pending);
    return last_launch();
  };
  for (uint32_t h = 1; h < hu; ++h) {
    Scenes sh = sc;
    sh.keys = L.keys[h];
    {
      TimedLaunch tl(VX_K_NODE_BUILD, s);
      const BuildArgs ba{sh, &cnt[1 + h], L.old[h], L.first[h], L.parent[h - 1], L.res[h - 1], w.items_at(h),
                         w.hashes_at(h), L.res[h]};
      const uint32_t pb = grid_for(pending_ub), grid = pb + grid_for(L.cap[h]);
      if (pending_kmax == 2) build_and_place_kernel<2><<<grid, kBuildBlock, 0, s>>>(pending, pb, ba);
      else if (pending_kmax == 16) build_and_place_kernel<16><<<grid, kBuildBlock, 0, s>>>(pending, pb, ba);
      else build_and_place_kernel<kNodeStride><<<grid, kBuildBlock, 0, s>>>(pending, pb, ba);
    }
    if ((st = last_launch()) != VX_OK) return st;
    const ItemSet set{w.items_at(h), kNodeStride, 0};
    {
      TimedLaunch tl(VX_K_NODE_PROBE, s);
      probe_kernel<kNodeStride><<<grid_for(L.cap[h]), 256, 0, s>>>(set, w.hashes_at(h), &cnt[1 + h], L.res[h], sh,
                                                                    p.n_scenes, false);
    }
    if ((st = last_launch()) != VX_OK) return st;
    if ((st = commit_final<kNodeStride>(sh, p.n_scenes, false, set, w.hashes_at(h), &cnt[1 + h], nullptr, L.cap[h],
                                        L.res[h], w.slot_idx, w.commit_ws[h], s, false)) != VX_OK)
      return st;
    pending = PlaceArgs{set, w.hashes_at(h), w.slot_idx, &cnt[1 + h], nullptr, sh, false};
    pending_kmax = kNodeStride;
    pending_ub = L.cap[h];
  }
  if ((st = place_alone()) != VX_OK) return st;
  if (hu <= H) {
    TimedLaunch tl(VX_K_TAIL_NODES, s);
    tail_nodes_kernel<<<1, scan::kFusedBlock, 0, s>>>(ta);
  }
  if ((st = last_launch()) != VX_OK) return st;
  mark(g, 6, s);
  ResByHeight rh{};
  for (uint32_t h = 0; h <= H && h < kMaxHeights; ++h) rh.p[h] = L.res[h];
  {
    TimedLaunch tl(VX_K_PUBLISH, s);
    publish_kernel<<<grid_for(p.n_scenes), 256, 0, s>>>(sc, p.n_scenes, w.root_pos, rh);
  }
  st = last_launch();
  mark(g, 7, s);
  return st;
}

// Brush jobs: brick bounding boxes clipped to the world; empty ones dropped.
// Brush jobs; *rows (may be null) is the running count of brick rows (each
// job's rows start there).
// Brush fields the header documents: kind 0 (sphere) or 1 (box), material
// 0 (remove) or 1..255.
bool brushes_valid(const vx_brush* b, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i)
    if (b[i].kind > 1 || b[i].material > 255) return false;
  return true;
}

uint64_t make_jobs(const vx_brush* b, uint32_t n, uint32_t world, uint32_t scene, uint64_t start,
                   std::vector<BrushJob>* jobs, uint64_t* rows = nullptr) {
  uint64_t total = start;
  for (uint32_t i = 0; i < n; ++i) {
    const vx_brush& br = b[i];
    const int64_t c[3] = {br.cx, br.cy, br.cz};
    const int64_t e[3] = {br.hx, br.kind == 0 ? br.hx : br.hy, br.kind == 0 ? br.hx : br.hz};
    int64_t lo[3], cnt[3];
    bool empty = br.hx < 0 || (br.kind != 0 && (br.hy < 0 || br.hz < 0));
    for (int a = 0; a < 3 && !empty; ++a) {
      const int64_t v0 = std::max<int64_t>(0, c[a] - e[a]);
      const int64_t v1 = std::min<int64_t>(int64_t(world) - 1, c[a] + e[a]);
      if (v0 > v1) empty = true;
      lo[a] = v0 >> 2;
      cnt[a] = (v1 >> 2) - lo[a] + 1;
    }
    if (empty) continue;
    BrushJob j{};
    j.b = br;
    j.bx0 = int32_t(lo[0]);
    j.by0 = int32_t(lo[1]);
    j.bz0 = int32_t(lo[2]);
    j.nx = uint32_t(cnt[0]);
    j.ny = uint32_t(cnt[1]);
    j.nz = uint32_t(cnt[2]);
    j.start = uint32_t(total);
    j.scene = scene;
    j.inv_nx = float(1.0 / double(j.nx));
    j.inv_ny = float(1.0 / double(j.ny));
    if (rows) {
      j.row_start = uint32_t(*rows);
      *rows += uint64_t(cnt[1]) * cnt[2];
    }
    total += uint64_t(cnt[0]) * cnt[1] * cnt[2];
    if (jobs) jobs->push_back(j);
  }
  return total - start;
}

// Reads back every scene's status (and the counts when profiling) into the
// pinned read-back words `upload` reserved, then synchronises the stream.
vx_status finish(vx_dag* g, const std::vector<vx_dag*>& dags, const EditWs& w, const Plan& p, HostStage& hs,
                 hipStream_t s) {
  uint32_t* const rb = reinterpret_cast<uint32_t*>(static_cast<char*>(hs.p) + hs.readback);
  uint32_t* const status = rb + 64;
  uint32_t* const counts = rb;
  for (size_t i = 0; i < dags.size(); ++i)
    if (hipMemcpyAsync(&status[i], dags[i]->ctl + kCtlStatus, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
      return VX_DEVICE_ERROR;
  const bool prof = g && g->profile;
  if (prof && hipMemcpyAsync(counts, w.counts, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  if (prof) {
    for (uint32_t i = 0; i < VX_DAG_STAGES && g->stage_events; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, g->ev[i], g->ev[i + 1]) == hipSuccess) g->prof.ms[i] += ms;
    }
    g->prof.edits += 1;
    g->prof.records += counts[0];
    g->prof.leaves += counts[1];
    g->prof.pending_leaves += counts[kCntPending];
    for (uint32_t h = 0; h <= p.max_depth && h < 16; ++h) g->prof.height_entries[h] += counts[1 + h];
    for (uint32_t h = 1; h <= p.max_depth; ++h) g->prof.nodes += counts[1 + h];
    const uint32_t kb = std::max(p.key_bits, 1u);
    g->prof.sort_passes = sort::passes(kb, w.sort_tn);
    g->prof.sort_record_bytes = w.index_bits ? 8 : 12;
    if (g->timer) {
      KernelTimer& t = *g->timer;
      for (uint32_t k = 0; k + 1 < t.used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]) != hipSuccess) continue;
        g->prof.kernel_ms[t.kind[k / 2]] += ms;
        g->prof.kernel_launches[t.kind[k / 2]] += 1;
      }
      t.used = 0;
    }
  }
  for (size_t i = 0; i < dags.size(); ++i)
    if (status[i] != kStatusOk) return VX_OUT_OF_CAPACITY;
  return VX_OK;
}

vx_status HostStage::reserve(size_t want) {
  if (want <= bytes) return VX_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  bytes = 0;
  const size_t size = align_up(std::max<size_t>(want, size_t(1) << 16), size_t(1) << 16);
  if (hipHostMalloc(&p, size, hipHostMallocDefault) != hipSuccess) {
    p = nullptr;
    return VX_DEVICE_ERROR;
  }
  bytes = size;
  return VX_OK;
}

HostStage::~HostStage() {
  if (p) (void)hipHostFree(p);
}

// Counts words the host uploads before the device pipeline runs.
constexpr uint32_t kCntCandidates = 62;
constexpr uint32_t kCntJobs = 63;

// Host -> device inputs of one edit (scene table, brush jobs, candidate and
// job counts) in one copy from pinned memory laid out like the workspace's
// upload block; reserves the read-back words `finish` uses after it.
vx_status upload(EditWs& w, const std::vector<Dev>& devs, const std::vector<BrushJob>* jobs, uint32_t n_cand,
                 HostStage& hs, hipStream_t s) {
  const size_t n_jobs = jobs ? jobs->size() : 0;
  const size_t bytes = w.jobs_off + n_jobs * sizeof(BrushJob);
  hs.readback = align_up(bytes, 256);
  if (hs.reserve(hs.readback + (4 * devs.size() + 64) * sizeof(uint32_t)) != VX_OK) return VX_DEVICE_ERROR;
  char* h = static_cast<char*>(hs.p);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(h);
  std::memset(cnt, 0, 64 * sizeof(uint32_t));
  cnt[kCntCandidates] = n_cand;
  cnt[kCntJobs] = uint32_t(n_jobs);
  std::memcpy(h + w.scenes_off, devs.data(), devs.size() * sizeof(Dev));
  if (n_jobs) std::memcpy(h + w.jobs_off, jobs->data(), n_jobs * sizeof(BrushJob));
  return from_hip(hipMemcpyAsync(w.counts, h, bytes, hipMemcpyHostToDevice, s));
}

// The record writer that matches each candidate kind.
inline RecordOut record_out(EditWs& w) {
  return RecordOut{w.rkey, w.rval, w.rpay, &w.counts[0], w.index_bits};
}
inline RecordStore store_for(const VoxelLoad&, EditWs& w) {
  return RecordStore{record_out(w)};
}


// All device work of an edit (no host synchronisation, no memset nodes), so a
// stream capture of it replays as a hipGraph. Grids are sized by the
// capacities in `w`; the actual counts are read on the device.
template <class Load>
vx_status device_pipeline(vx_dag* prof_dag, const Plan& p, const Scenes& sc, EditWs& w, uint64_t cand_cap, Load load,
                          hipStream_t s) {
  vx_status st;
  {
    TimedLaunch tl(VX_K_OTHER, s);
    const uint64_t fused_words = w.fused_bytes / 4;
    begin_kernel<<<grid_for(std::max<uint64_t>(fused_words, p.n_scenes)), 256, 0, s>>>(
        sc, p.n_scenes, w.counts, kCntCandidates, w.root_pos, w.fused_all, fused_words);
  }
  if ((st = last_launch()) != VX_OK) return st;
  mark(prof_dag, 0, s);
  {
    TimedLaunch tl(VX_K_RECORDS, s);
    if constexpr (std::is_same_v<Load, BrushLoad>) {
      // Brushes: a scan over the brick rows (counts[kCntCandidates] of them),
      // then one thread per touched brick.
      const TouchedRows rows{w.row_rec, w.row_job, w.row_yz, w.row_x0, &w.counts[kCntRows], w.chunk_row};
      st = scan::fused_dev_m<scan::PairSumOp>(&w.counts[kCntCandidates], cand_cap, RowLoad{load.c},
                                              RowStore{rows, &w.counts[0]}, w.rec_ws, s, false);
      if (st == VX_OK) {
        // About one wave per chunk of records (cand_cap bounds the records).
        brush_records_kernel<<<grid_for(cand_cap / (kRecChunk / kWave)), kRecBlock, 0, s>>>(load.c, rows,
                                                                                             record_out(w));
        st = last_launch();
      }
    } else {
      st = scan::fused_dev<scan::kSum>(&w.counts[kCntCandidates], cand_cap, load, store_for(load, w), w.rec_ws, s,
                                       false);
    }
  }
  if (st != VX_OK) return st;
  return rebuild(prof_dag, p, sc, w, cand_cap, s);
}

// Shared driver of the direct (uncaptured) calls.
template <class Load>
// n_first: the first scan's element count (voxels, or brush brick rows);
// cap: the upper bound of the records (the grids and workspace are sized by it).
vx_status run_edit(const std::vector<vx_dag*>& dags, vx_dag* prof_dag, const Plan& p, EditWs& w,
                   const std::vector<BrushJob>* jobs, uint32_t n_first, uint64_t cap, Load load, hipStream_t s) {
  vx_status st;
  std::vector<Dev> devs;
  for (vx_dag* g : dags) devs.push_back(g->dev);
  HostStage& hs = dags[0]->stage;
  if ((st = upload(w, devs, jobs, n_first, hs, s)) != VX_OK) return st;
  const Scenes sc{p.n_scenes > 1 ? w.scenes : nullptr, nullptr, p.codec, devs[0]};
  {
    // Per-launch timing of this call's kernels while the dag is profiled.
    struct Install {
      KernelTimer* prev;
      Install(KernelTimer* t, uint64_t call) : prev(current_timer()) {
        if (t) {
          t->used = 0;
          t->depth = 0;
          t->phase = uint32_t(call % t->period);
          for (uint32_t& c : t->seen) c = 0;
          current_timer() = t;
        }
      }
      ~Install() { current_timer() = prev; }
    } install(prof_dag && prof_dag->profile ? prof_dag->timer : nullptr, prof_dag ? prof_dag->prof.edits : 0);
    if ((st = device_pipeline(prof_dag, p, sc, w, cap, load, s)) != VX_OK) return st;
  }
  return finish(prof_dag, dags, w, p, hs, s);
}

// ---- queries -------------------------------------------------------------------
__global__ void get_voxels_kernel(Dev d, const vx_voxel* __restrict__ q, uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t world = 4u << d.depth;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const vx_voxel v = q[i];
    uint32_t r = 0;
    if (v.x < world && v.y < world && v.z < world) {
      uint32_t ref = d.ctl[kCtlRoot];
      for (uint32_t lvl = 0; lvl < d.depth && ref > kFull; ++lvl) {
        const uint32_t sh = d.depth + 1 - lvl;  // log2(world) - 1 - level
        const uint32_t c = ((v.x >> sh) & 1u) | (((v.y >> sh) & 1u) << 1) | (((v.z >> sh) & 1u) << 2);
        ref = child_of(d.node_pool, ref, c);
      }
      const uint32_t bit = (v.x & 3u) + 4u * (v.y & 3u) + 16u * (v.z & 3u);
      if (ref == kFull) r = 1;
      else if (ref != kEmpty) {
        if (d.leaf_words == 2) r = (d.leaf_pool[ref + (bit >> 5)] >> (bit & 31)) & 1u;
        else r = (d.leaf_pool[ref + (bit >> 2)] >> ((bit & 3) * 8)) & 0xffu;
      }
    }
    out[i] = r;
  }
}

__global__ void reset_ctl_kernel(uint32_t* ctl) {
  for (uint32_t i = threadIdx.x; i < kCtlWords; i += blockDim.x) ctl[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl[kCtlNodeUsed] = kPoolStart;
    ctl[kCtlLeafUsed] = kPoolStart;
    ctl[kCtlRoot] = kEmpty;
  }
}

uint64_t pow2_at_least(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace dag
}  // namespace vx

// ---- C ABI -----------------------------------------------------------------------
extern "C" vx_status vx_dag_reset(vx_dag* g, void* stream) {
  if (!g) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  vx_status st;
  if ((st = fill_u32(reinterpret_cast<uint32_t*>(g->node_slots), (g->dev.node_mask + 1) * 2, 0, s)) != VX_OK) return st;
  if ((st = fill_u32(reinterpret_cast<uint32_t*>(g->leaf_slots), (g->dev.leaf_mask + 1) * 2, 0, s)) != VX_OK) return st;
  reset_ctl_kernel<<<1, 64, 0, s>>>(g->ctl);
  return last_launch();
}

extern "C" vx_status vx_dag_create(uint32_t inner_depth, int32_t materials, const vx_dag_capacity* cap, vx_dag** out) {
  if (!out || inner_depth == 0 || inner_depth > 13) return VX_INVALID_VALUE;
  vx_dag_capacity c = cap ? *cap : vx_dag_capacity{1ull << 24, 1ull << 24, 1ull << 21, 1ull << 21};
  if (c.node_words < 64 || c.leaf_words < 64 || c.node_slots < 64 || c.leaf_slots < 64) return VX_INVALID_VALUE;
  c.node_words = std::min<uint64_t>(c.node_words, kTag);
  c.leaf_words = std::min<uint64_t>(c.leaf_words, kTag);
  c.node_slots = pow2_at_least(c.node_slots);
  c.leaf_slots = pow2_at_least(c.leaf_slots);
  if (c.node_slots > (1ull << 32) || c.leaf_slots > (1ull << 32)) return VX_INVALID_VALUE;
  vx_dag* g = new (std::nothrow) vx_dag();
  if (!g) return VX_DEVICE_ERROR;
  g->depth = inner_depth;
  g->materials = materials != 0;
  g->cap = c;
  bool ok = hipMalloc(&g->node_pool, c.node_words * 4) == hipSuccess &&
            hipMalloc(&g->leaf_pool, c.leaf_words * 4) == hipSuccess &&
            hipMalloc(&g->node_slots, c.node_slots * 8) == hipSuccess &&
            hipMalloc(&g->leaf_slots, c.leaf_slots * 8) == hipSuccess &&
            hipMalloc(&g->ctl, kCtlWords * 4) == hipSuccess;
  if (!ok) {
    vx_dag_destroy(g);
    return VX_OUT_OF_CAPACITY;
  }
  Dev& d = g->dev;
  d.node_pool = g->node_pool;
  d.leaf_pool = g->leaf_pool;
  d.node_slots = g->node_slots;
  d.leaf_slots = g->leaf_slots;
  d.node_mask = c.node_slots - 1;
  d.leaf_mask = c.leaf_slots - 1;
  d.node_cap = c.node_words;
  d.leaf_cap = c.leaf_words;
  d.ctl = g->ctl;
  d.status = g->ctl + kCtlStatus;
  d.base = g->ctl + kCtlNodeBase;
  d.depth = inner_depth;
  d.leaf_words = g->materials ? 16 : 2;
  vx_status st = vx_dag_reset(g, nullptr);
  if (st == VX_OK) st = from_hip(hipDeviceSynchronize());
  if (st != VX_OK) {
    vx_dag_destroy(g);
    return st;
  }
  *out = g;
  return VX_OK;
}

extern "C" vx_status vx_dag_destroy(vx_dag* g) {
  if (!g) return VX_INVALID_VALUE;
  if (g->node_pool) hipFree(g->node_pool);
  if (g->leaf_pool) hipFree(g->leaf_pool);
  if (g->node_slots) hipFree(g->node_slots);
  if (g->leaf_slots) hipFree(g->leaf_slots);
  if (g->ctl) hipFree(g->ctl);
  for (auto& e : g->ev)
    if (e) hipEventDestroy(e);
  if (g->timer) {
    for (auto& e : g->timer->ev)
      if (e) hipEventDestroy(e);
    delete g->timer;
  }
  delete g;
  return VX_OK;
}

extern "C" vx_status vx_dag_stats_get(const vx_dag* g, vx_dag_stats* out) {
  if (!g || !out) return VX_INVALID_VALUE;
  uint32_t ctl[kCtlWords];
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  if (hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  out->node_words_used = ctl[kCtlNodeUsed];
  out->leaf_words_used = ctl[kCtlLeafUsed];
  out->node_slots = g->dev.node_mask + 1;
  out->leaf_slots = g->dev.leaf_mask + 1;
  out->root = ctl[kCtlRoot];
  out->inner_depth = g->depth;
  out->leaf_words = g->dev.leaf_words;
  out->last_status = ctl[kCtlStatus];
  return VX_OK;
}

extern "C" vx_status vx_dag_voxel_edits_workspace_size(const vx_dag* g, uint32_t count, size_t* bytes) {
  if (!g || !bytes) return VX_INVALID_VALUE;
  *bytes = count ? carve(nullptr, make_plan({g}), count, 0, count).bytes : 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_apply_voxel_edits(vx_dag* g, const vx_voxel* edits, uint32_t count, uint32_t op,
                                              uint32_t material, void* ws, size_t ws_bytes, void* stream) {
  if (!g || (!edits && count) || op > 1) return VX_INVALID_VALUE;  // edit.cuh:28-37
  if (count == 0) return VX_OK;
  if (count >= kMaxEntries) return VX_INVALID_VALUE;
  const Plan p = make_plan({g});
  EditWs w = carve(ws, p, count, 0, count);
  if (!ws || ws_bytes < w.bytes) return VX_INVALID_VALUE;
  const uint32_t mat = op == 1 ? 0u : (g->materials ? (material & 0xffu) : 1u);
  if (op == 0 && g->materials && mat == 0) return VX_INVALID_VALUE;
  return run_edit({g}, g, p, w, nullptr, count, count, VoxelLoad{edits, 4u << g->depth, mat},
                  static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_dag_brush_workspace_size(const vx_dag* g, const vx_brush* b, uint32_t n, size_t* bytes) {
  if (!g || !bytes || (!b && n) || (n && !brushes_valid(b, n))) return VX_INVALID_VALUE;
  std::vector<BrushJob> jobs;
  const uint64_t cand = make_jobs(b, n, 4u << g->depth, 0, 0, &jobs);
  *bytes = cand ? carve(nullptr, make_plan({g}), cand, uint32_t(jobs.size()), cand).bytes : 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_apply_brushes(vx_dag* g, const vx_brush* b, uint32_t n, void* ws, size_t ws_bytes,
                                          void* stream) {
  if (!g || (!b && n) || (n && !brushes_valid(b, n))) return VX_INVALID_VALUE;
  std::vector<BrushJob> jobs;
  uint64_t rows = 0;
  const uint64_t cand = make_jobs(b, n, 4u << g->depth, 0, 0, &jobs, &rows);
  if (cand == 0) return VX_OK;
  if (cand >= kMaxEntries) return VX_INVALID_VALUE;
  const Plan p = make_plan({g});
  EditWs w = carve(ws, p, cand, uint32_t(jobs.size()), cand);
  if (!ws || ws_bytes < w.bytes) return VX_INVALID_VALUE;
  return run_edit({g}, g, p, w, &jobs, uint32_t(rows), cand,
                  BrushLoad{BrushCands{w.jobs, &w.counts[kCntJobs], g->materials, p.codec}},
                  static_cast<hipStream_t>(stream));
}

// Batched: batches[i] applies to dags[i]; batches of one scene apply in array
// order, scenes are independent; all of it runs as one pipeline.
static vx_status batch_setup(vx_dag* const* dags, const vx_brush* const* brushes, const uint32_t* counts,
                             uint32_t n_batches, std::vector<vx_dag*>* scenes, std::vector<BrushJob>* jobs,
                             uint64_t* cand, uint64_t* rows = nullptr) {
  if (!dags || !brushes || !counts || n_batches == 0) return VX_INVALID_VALUE;
  std::unordered_map<const vx_dag*, uint32_t> index;
  for (uint32_t i = 0; i < n_batches; ++i) {
    if (!dags[i] || (!brushes[i] && counts[i])) return VX_INVALID_VALUE;
    if (counts[i] && !brushes_valid(brushes[i], counts[i])) return VX_INVALID_VALUE;
    if (index.emplace(dags[i], uint32_t(scenes->size())).second) scenes->push_back(dags[i]);
  }
  const vx_dag* first = (*scenes)[0];
  for (vx_dag* g : *scenes)
    if (g->materials != first->materials) return VX_INVALID_VALUE;
  *cand = 0;
  // Jobs grouped by scene (scene order = first appearance), batch order kept.
  for (uint32_t s = 0; s < scenes->size(); ++s)
    for (uint32_t i = 0; i < n_batches; ++i)
      if (index[dags[i]] == s)
        *cand += make_jobs(brushes[i], counts[i], 4u << dags[i]->depth, s, *cand, jobs, rows);
  return VX_OK;
}

extern "C" vx_status vx_dag_batch_workspace_size(vx_dag* const* dags, const vx_brush* const* brushes,
                                                 const uint32_t* counts, uint32_t n_batches, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  std::vector<vx_dag*> scenes;
  std::vector<BrushJob> jobs;
  uint64_t cand = 0;
  const vx_status st = batch_setup(dags, brushes, counts, n_batches, &scenes, &jobs, &cand);
  if (st != VX_OK) return st;
  const Plan p = make_plan(std::vector<const vx_dag*>(scenes.begin(), scenes.end()));
  *bytes = cand ? carve(nullptr, p, cand, uint32_t(jobs.size()), cand).bytes : 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_apply_batches(vx_dag* const* dags, const vx_brush* const* brushes, const uint32_t* counts,
                                          uint32_t n_batches, void* ws, size_t ws_bytes, void* stream) {
  std::vector<vx_dag*> scenes;
  std::vector<BrushJob> jobs;
  uint64_t cand = 0, rows = 0;
  vx_status st = batch_setup(dags, brushes, counts, n_batches, &scenes, &jobs, &cand, &rows);
  if (st != VX_OK) return st;
  if (cand == 0) return VX_OK;
  if (cand >= kMaxEntries || scenes.size() > (1u << 20)) return VX_INVALID_VALUE;
  const Plan p = make_plan(std::vector<const vx_dag*>(scenes.begin(), scenes.end()));
  if (p.key_bits > 64) return VX_INVALID_VALUE;
  EditWs w = carve(ws, p, cand, uint32_t(jobs.size()), cand);
  if (!ws || ws_bytes < w.bytes) return VX_INVALID_VALUE;
  return run_edit(scenes, scenes[0], p, w, &jobs, uint32_t(rows), cand,
                  BrushLoad{BrushCands{w.jobs, &w.counts[kCntJobs], scenes[0]->materials, p.codec}},
                  static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_dag_get_voxels(const vx_dag* g, const vx_voxel* q, uint32_t n, uint32_t* out, void* stream) {
  if (!g || (n && (!q || !out))) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  get_voxels_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(g->dev, q, n, out);
  return last_launch();
}

extern "C" vx_status vx_dag_profile_enable(vx_dag* g, int32_t enable) {
  if (!g) return VX_INVALID_VALUE;
  if (enable && !g->profile) {
    for (auto& e : g->ev)
      if (!e && hipEventCreate(&e) != hipSuccess) return VX_DEVICE_ERROR;
    if (!g->timer) {
      g->timer = new (std::nothrow) KernelTimer();
      if (!g->timer) return VX_DEVICE_ERROR;
      for (auto& e : g->timer->ev)
        if (hipEventCreate(&e) != hipSuccess) return VX_DEVICE_ERROR;
    }
  }
  g->profile = (enable & VX_PROFILE_ON) != 0;
  g->stage_events = (enable & VX_PROFILE_NO_STAGES) == 0;
  g->prof = vx_dag_profile{};
  if (g->timer) g->timer->period = 1;
  return VX_OK;
}

extern "C" vx_status vx_dag_profile_kernel_mask(vx_dag* g, uint64_t mask) {
  if (!g || !g->timer) return VX_INVALID_VALUE;  // enable profiling first
  g->timer->mask = mask;
  return VX_OK;
}

extern "C" vx_status vx_dag_profile_kernel_period(vx_dag* g, uint32_t period) {
  if (!g || !g->timer || period == 0) return VX_INVALID_VALUE;  // enable profiling first
  g->timer->period = period;
  return VX_OK;
}

extern "C" vx_status vx_dag_profile_read(const vx_dag* g, vx_dag_profile* out) {
  if (!g || !out) return VX_INVALID_VALUE;
  *out = g->prof;
  return VX_OK;
}

// Exact voxel count of each brush inside the world: spheres row by row
// (integer square root of r^2 - dy^2 - dz^2), boxes as clipped extents.
extern "C" vx_status vx_brush_voxel_ops(const vx_brush* b, uint32_t n, uint32_t world, uint64_t* ops) {
  if (!ops || (!b && n)) return VX_INVALID_VALUE;
  uint64_t total = 0;
  auto clip_len = [&](int64_t lo, int64_t hi) -> int64_t {
    lo = std::max<int64_t>(lo, 0);
    hi = std::min<int64_t>(hi, int64_t(world) - 1);
    return hi >= lo ? hi - lo + 1 : 0;
  };
  for (uint32_t i = 0; i < n; ++i) {
    const vx_brush& br = b[i];
    if (br.kind != 0) {
      total += uint64_t(clip_len(br.cx - int64_t(br.hx), br.cx + int64_t(br.hx))) *
               clip_len(br.cy - int64_t(br.hy), br.cy + int64_t(br.hy)) *
               clip_len(br.cz - int64_t(br.hz), br.cz + int64_t(br.hz));
      continue;
    }
    const int64_t r = br.hx;
    if (r < 0) continue;
    for (int64_t dz = -r; dz <= r; ++dz) {
      const int64_t z = br.cz + dz;
      if (z < 0 || z >= int64_t(world)) continue;
      for (int64_t dy = -r; dy <= r; ++dy) {
        const int64_t y = br.cy + dy;
        const int64_t rem = r * r - dy * dy - dz * dz;
        if (rem < 0 || y < 0 || y >= int64_t(world)) continue;
        int64_t s = int64_t(std::sqrt(double(rem)));
        while (s * s > rem) --s;
        while ((s + 1) * (s + 1) <= rem) ++s;
        total += uint64_t(clip_len(br.cx - s, br.cx + s));
      }
    }
  }
  *ops = total;
  return VX_OK;
}

// ---- persistent batched executor ------------------------------------------------
// Owns a workspace sized for a candidate capacity and its own stream; the whole
// device pipeline of a batched edit is captured into a hipGraph on the first
// run and replayed for every later batch set within the capacity (the counts
// it works on are uploaded to device memory before each replay).
struct vx_batcher {
  std::vector<vx_dag*> scenes;
  std::unordered_map<const vx_dag*, uint32_t> index;
  Plan plan;
  uint64_t cap_cand = 0;
  uint32_t cap_jobs = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  EditWs w{};
  hipStream_t stream = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<Dev> devs;
  std::vector<uint64_t> generations;  // of each scene when its `devs` entry was copied
  std::vector<BrushJob> jobs;
  HostStage stage;
  // Asynchronous submissions: one in flight at a time, completed by query.
  hipEvent_t done = nullptr;
  bool submitted = false;  // `done` has been recorded at least once
  bool inflight = false;
  bool empty = false;      // the submission in flight had no candidates
  vx_status last = VX_OK;  // outcome of the last completed submission
};

extern "C" vx_status vx_batcher_destroy(vx_batcher* b) {
  if (!b) return VX_INVALID_VALUE;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->done) hipEventDestroy(b->done);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->ws) hipFree(b->ws);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return VX_OK;
}

extern "C" vx_status vx_batcher_create(vx_dag* const* scenes, uint32_t n_scenes, uint64_t max_candidates,
                                       uint32_t max_jobs, int32_t use_graph, vx_batcher** out) {
  if (!scenes || !n_scenes || !out || !max_candidates || !max_jobs || max_candidates >= kMaxEntries)
    return VX_INVALID_VALUE;
  auto* b = new (std::nothrow) vx_batcher();
  if (!b) return VX_DEVICE_ERROR;
  for (uint32_t i = 0; i < n_scenes; ++i) {
    if (!scenes[i] || scenes[i]->materials != scenes[0]->materials || !b->index.emplace(scenes[i], i).second) {
      delete b;
      return VX_INVALID_VALUE;
    }
    b->scenes.push_back(scenes[i]);
    b->devs.push_back(scenes[i]->dev);
    b->generations.push_back(scenes[i]->generation);
  }
  b->plan = make_plan(std::vector<const vx_dag*>(b->scenes.begin(), b->scenes.end()));
  if (b->plan.key_bits > 64) {
    delete b;
    return VX_INVALID_VALUE;
  }
  b->cap_cand = max_candidates;
  b->cap_jobs = max_jobs;
  b->use_graph = use_graph != 0;
  b->ws_bytes = carve(nullptr, b->plan, max_candidates, max_jobs, max_candidates).bytes;
  if (hipMalloc(&b->ws, b->ws_bytes) != hipSuccess) {
    b->ws = nullptr;
    vx_batcher_destroy(b);
    return VX_OUT_OF_CAPACITY;
  }
  b->w = carve(b->ws, b->plan, max_candidates, max_jobs, max_candidates);
  for (uint32_t i = 0; i < n_scenes; ++i) {
    b->devs[i].status = b->w.status + i;
    b->devs[i].base = b->w.base + 2 * i;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    b->stream = nullptr;
    vx_batcher_destroy(b);
    return VX_DEVICE_ERROR;
  }
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    b->done = nullptr;
    vx_batcher_destroy(b);
    return VX_DEVICE_ERROR;
  }
  *out = b;
  return VX_OK;
}

extern "C" vx_status vx_batcher_info(const vx_batcher* b, uint64_t* workspace_bytes, int32_t* graph_ready) {
  if (!b) return VX_INVALID_VALUE;
  if (workspace_bytes) *workspace_bytes = b->ws_bytes;
  if (graph_ready) *graph_ready = b->exec != nullptr;
  return VX_OK;
}

extern "C" vx_status vx_batcher_layout(const vx_batcher* b, uint32_t* key_bits, uint32_t* index_bits) {
  if (!b) return VX_INVALID_VALUE;
  if (key_bits) *key_bits = b->plan.key_bits;
  if (index_bits) *index_bits = b->w.index_bits;
  return VX_OK;
}

// Read-back block of a submission (pinned host memory after the upload):
// per scene ctl words [node used, leaf used] and the executor's status word.
constexpr uint32_t kReadbackWords = 4;

static vx_status batcher_submit(vx_batcher* b, const uint32_t* scene_of_batch, const vx_brush* const* brushes,
                                const uint32_t* counts, uint32_t n_batches, uint64_t* candidates,
                                const vx_batcher* after) {
  if (!b || (n_batches && (!scene_of_batch || !brushes || !counts))) return VX_INVALID_VALUE;
  if (b->inflight || after == b) return VX_INVALID_VALUE;  // complete the last submission first
  for (uint32_t i = 0; i < n_batches; ++i)
    if (scene_of_batch[i] >= b->scenes.size() || (!brushes[i] && counts[i]) ||
        (counts[i] && !brushes_valid(brushes[i], counts[i])))
      return VX_INVALID_VALUE;
  b->jobs.clear();
  uint64_t cand = 0, rows = 0;
  for (uint32_t s = 0; s < b->scenes.size(); ++s)
    for (uint32_t i = 0; i < n_batches; ++i)
      if (scene_of_batch[i] == s)
        cand += make_jobs(brushes[i], counts[i], 4u << b->scenes[s]->depth, s, cand, &b->jobs, &rows);
  if (candidates) *candidates = cand;
  if (cand > b->cap_cand || b->jobs.size() > b->cap_jobs) return VX_INVALID_VALUE;
  // A scene whose buffers moved since it was copied (vx_dag_grow): take its
  // new descriptor, and capture the pipeline again (the graph holds the old
  // one in its kernel arguments).
  for (uint32_t i = 0; i < b->scenes.size(); ++i) {
    if (b->generations[i] == b->scenes[i]->generation) continue;
    Dev d = b->scenes[i]->dev;
    d.status = b->w.status + i;
    d.base = b->w.base + 2 * i;
    b->devs[i] = d;
    b->generations[i] = b->scenes[i]->generation;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  hipStream_t s = b->stream;
  vx_status st;
  const uint32_t n_scenes = uint32_t(b->scenes.size());
  b->empty = cand == 0;
  if (!b->empty) {
    // The brush path's first scan runs over brick rows. The upload goes to this
    // executor's own workspace, so it may overlap another executor's edit.
    if ((st = upload(b->w, b->devs, &b->jobs, uint32_t(rows), b->stage, s)) != VX_OK) return st;
    if (after && after->submitted && hipStreamWaitEvent(s, after->done, 0) != hipSuccess) return VX_DEVICE_ERROR;
    const Scenes sc{b->plan.n_scenes > 1 ? b->w.scenes : nullptr, nullptr, b->plan.codec, b->devs[0]};
    const BrushLoad load{BrushCands{b->w.jobs, &b->w.counts[kCntJobs], b->scenes[0]->materials, b->plan.codec}};
    if (b->use_graph) {
      if (!b->exec) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) return VX_DEVICE_ERROR;
        const vx_status cap_st = device_pipeline(nullptr, b->plan, sc, b->w, b->cap_cand, load, s);
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(s, &g);
        if (cap_st != VX_OK || end != hipSuccess || !g) {
          if (g) hipGraphDestroy(g);
          return cap_st != VX_OK ? cap_st : VX_DEVICE_ERROR;
        }
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
          hipGraphDestroy(g);
          b->exec = nullptr;
          return VX_DEVICE_ERROR;
        }
        b->graph = g;
      }
      if (hipGraphLaunch(b->exec, s) != hipSuccess) return VX_DEVICE_ERROR;
    } else {
      if ((st = device_pipeline(nullptr, b->plan, sc, b->w, b->cap_cand, load, s)) != VX_OK) return st;
    }
  } else {
    b->stage.readback = 0;
    if (b->stage.reserve(size_t(n_scenes) * kReadbackWords * 4) != VX_OK) return VX_DEVICE_ERROR;
    if (after && after->submitted && hipStreamWaitEvent(s, after->done, 0) != hipSuccess) return VX_DEVICE_ERROR;
  }
  // Read-back of every scene's pool usage and status, then the completion event.
  uint32_t* rb = reinterpret_cast<uint32_t*>(static_cast<char*>(b->stage.p) + b->stage.readback);
  for (uint32_t i = 0; i < n_scenes; ++i) {
    if (hipMemcpyAsync(rb + kReadbackWords * i, b->scenes[i]->ctl, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        (!b->empty && hipMemcpyAsync(rb + kReadbackWords * i + 2, b->w.status + i, sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s) != hipSuccess))
      return VX_DEVICE_ERROR;
  }
  if (hipEventRecord(b->done, s) != hipSuccess) return VX_DEVICE_ERROR;
  b->submitted = true;
  b->inflight = true;
  return VX_OK;
}

static vx_status batcher_query(vx_batcher* b, int32_t wait, int32_t* done, uint32_t* scene_status,
                               uint64_t* scene_used) {
  if (!b) return VX_INVALID_VALUE;
  const uint32_t n_scenes = uint32_t(b->scenes.size());
  if (b->inflight) {
    const hipError_t e = wait ? hipEventSynchronize(b->done) : hipEventQuery(b->done);
    if (e == hipErrorNotReady) {
      if (done) *done = 0;
      return VX_OK;
    }
    b->inflight = false;
    if (e != hipSuccess) return b->last = VX_DEVICE_ERROR;
    const uint32_t* rb = reinterpret_cast<const uint32_t*>(static_cast<const char*>(b->stage.p) + b->stage.readback);
    b->last = VX_OK;
    for (uint32_t i = 0; i < n_scenes; ++i)
      if (!b->empty && rb[kReadbackWords * i + 2] != kStatusOk) b->last = VX_OUT_OF_CAPACITY;
  }
  if (done) *done = 1;
  if (b->submitted && (scene_status || scene_used)) {
    const uint32_t* rb = reinterpret_cast<const uint32_t*>(static_cast<const char*>(b->stage.p) + b->stage.readback);
    for (uint32_t i = 0; i < n_scenes; ++i) {
      if (scene_status) scene_status[i] = b->empty ? kStatusOk : rb[kReadbackWords * i + 2];
      if (scene_used) {
        scene_used[2 * i] = rb[kReadbackWords * i + kCtlNodeUsed];
        scene_used[2 * i + 1] = rb[kReadbackWords * i + kCtlLeafUsed];
      }
    }
  }
  return b->last;
}

extern "C" vx_status vx_batcher_submit(vx_batcher* b, const uint32_t* scene_of_batch, const vx_brush* const* brushes,
                                       const uint32_t* counts, uint32_t n_batches, uint64_t* candidates,
                                       const vx_batcher* after) {
  return batcher_submit(b, scene_of_batch, brushes, counts, n_batches, candidates, after);
}

extern "C" vx_status vx_batcher_query(vx_batcher* b, int32_t wait, int32_t* done, uint32_t* scene_status,
                                      uint64_t* scene_used) {
  return batcher_query(b, wait, done, scene_status, scene_used);
}

extern "C" vx_status vx_batcher_run(vx_batcher* b, const uint32_t* scene_of_batch, const vx_brush* const* brushes,
                                    const uint32_t* counts, uint32_t n_batches, uint64_t* candidates) {
  const vx_status st = batcher_submit(b, scene_of_batch, brushes, counts, n_batches, candidates, nullptr);
  if (st != VX_OK) return st;
  return batcher_query(b, 1, nullptr, nullptr, nullptr);
}
