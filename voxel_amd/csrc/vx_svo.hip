// Pointer-SVO voxel edits on gfx950: the algo::svt::cuda API
// (svt/cuda/voxel/edit.cuh:14-92, storage.cuh:12-122, query.cuh:12-39)
// re-hosted on this library's sort and scan.
//
// Storage (device_view.cuh:10-46): nodes of eight u32 slots {child index:30 |
// pointer bit 30 | filled bit 31}, leaves of two u32 masks (voxel bit
// x + 4y + 16z), counters {node_count, leaf_count, free_node_count,
// free_leaf_count} and two LIFO free-index stacks. World 4096^3, ten inner
// levels, root at index 0.
//
// One edit batch, in the reference's stage order (voxel/detail/pipeline.cuh:
// 37-88), with its own kernels:
//   records  fused compaction: valid edits -> packed (leaf key << 6 | leaf bit)
//   sort     onesweep over the 36 key bits (4 passes of 9 bits)
//   masks    fused scan over the sorted records: a leaf starts where the key
//            changes; each wave ORs its lanes' bits per leaf with a segmented
//            shuffle scan and emits one store per leaf piece (an atomic OR only
//            for pieces of a leaf that crosses a 64-record boundary)
//   walk     pair-count scan over the masks: each walks the existing tree to
//            its first missing slot (recording the path), and counts the new
//            nodes / leaf it must take — those of its path below the missing
//            slot that the previous mask's path does not already create
//   (host)   one read-back of the counts: capacity is checked before anything
//            of the tree changes
//   build    one thread per mask takes its ranks (free stack top first, then
//            the bump counter, batch_kernels.cuh:12-79), writes its new nodes
//            (slots uniform with the replaced slot's fill) and leaf, links
//            them, and sets or clears the leaf bits
//   link     the first new node of a mask whose parent was created by an
//            earlier mask is linked into it (the parent's creator is found by a
//            galloping search for the start of the shared-prefix run)
//   collapse depth 10 .. 1, a fused scan each: the first mask of every
//            depth-d prefix checks its depth-d leaf / node for uniformity,
//            pushes it on the free stack in mask order and sets the parent slot
//            (collapse/kernels.cuh:12-92, batch_kernels.cuh:81-139); a level
//            that frees nothing ends the walk up (nothing above can become
//            uniform), via a device-side count.
#include <algorithm>
#include <new>
#include <vector>

#include "vx_dag.hpp"
#include "vx_scan.hpp"
#include "vx_sort.hpp"

using namespace vx;

struct vx_svo {
  uint32_t* nodes = nullptr;
  uint32_t* leaves = nullptr;
  uint32_t* counters = nullptr;
  uint32_t* free_nodes = nullptr;
  uint32_t* free_leaves = nullptr;
  uint32_t max_nodes = 0, max_leaves = 0;
};

namespace {

constexpr uint32_t kDepth = 10;
constexpr uint32_t kWorld = 4096;
constexpr uint32_t kIndexMask = 0x3fffffffu;
constexpr uint32_t kChild = 1u << 30;
constexpr uint32_t kFilled = 1u << 31;
constexpr uint32_t kKeyBits = 36;  // 30-bit leaf key + 6-bit leaf offset

// Control words of a batch (device). kError is set by any index check that
// fails before the host read-back.
enum : uint32_t { kNValid = 0, kNMasks = 1, kTotNodes = 2, kTotLeaves = 3, kError = 4, kActive = 16 };

struct Svo {
  uint32_t* nodes;
  uint32_t* leaves;
  uint32_t* counters;
  uint32_t* free_nodes;
  uint32_t* free_leaves;
  uint32_t max_nodes, max_leaves;
};

Svo dev_of(const vx_svo* g) {
  return Svo{g->nodes, g->leaves, g->counters, g->free_nodes, g->free_leaves, g->max_nodes, g->max_leaves};
}

__device__ __forceinline__ uint32_t child_at(uint32_t key, uint32_t d) { return (key >> ((kDepth - 1 - d) * 3)) & 7u; }

// Leaf key of make_leaf_key (detail/geometry.cuh:12-25): child index
// x | y << 1 | z << 2 per level, top level highest = the Morton code of the
// 10-bit leaf coordinates.
__device__ __forceinline__ uint32_t leaf_key(uint32_t x, uint32_t y, uint32_t z) {
  return uint32_t(dag::morton3(x >> 2, y >> 2, z >> 2));
}

// Number of leading child indices two distinct keys share.
__device__ __forceinline__ uint32_t common_depth(uint32_t a, uint32_t b) { return (__clz(a ^ b) - 2) / 3; }

// ---- 1. records -------------------------------------------------------------------
struct RecLoad {
  using Payload = uint64_t;
  const vx_voxel* edits;
  __device__ uint32_t operator()(uint32_t i, uint64_t& rec) const {
    const vx_voxel v = edits[i];
    rec = (uint64_t(leaf_key(v.x, v.y, v.z)) << 6) | ((v.x & 3u) | ((v.y & 3u) << 2) | ((v.z & 3u) << 4));
    return v.x < kWorld && v.y < kWorld && v.z < kWorld;
  }
};

struct RecStore {
  uint64_t* out;
  uint32_t* n_out;
  __device__ void operator()(uint32_t i, const uint64_t& rec, uint32_t v, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    if (v) out[excl] = rec;
    if (i == n - 1) *n_out = incl;
  }
};

// ---- 3. masks ---------------------------------------------------------------------
struct MaskPayload {
  uint64_t bits;
  uint32_t key;
  uint32_t flags;  // 1 first record of its leaf, 2 emits, 4 emits with an atomic OR
};

struct MaskLoad {
  using Payload = MaskPayload;
  const uint64_t* recs;
  const uint32_t* n_ptr;
  // Called with the lanes of a wave on 64 consecutive records (only the
  // trailing lanes of the last chunk are inactive, and shuffles read only
  // lower lanes).
  __device__ uint32_t operator()(uint32_t i, MaskPayload& p) const {
    const uint32_t n = *n_ptr;
    const uint64_t r = recs[i];
    const uint64_t prev = i ? recs[i - 1] : ~0ull;
    const uint64_t next = i + 1 < n ? recs[i + 1] : ~0ull;
    const bool head = i == 0 || (prev >> 6) != (r >> 6);
    uint64_t v = 1ull << (r & 63u);  // OR is idempotent: repeated voxels need no care
    uint32_t h = head;
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t tv = __shfl_up(v, d, kWave);
      const uint32_t th = __shfl_up(h, d, kWave);
      if (lane >= uint32_t(d) && !h) {
        v |= tv;
        h |= th;
      }
    }
    const bool tail = (next >> 6) != (r >> 6);
    const bool emit = tail || lane == kWave - 1;
    p.bits = v;
    p.key = uint32_t(r >> 6);
    p.flags = (head ? 1u : 0u) | (emit ? 2u : 0u) | (emit && !(tail && h) ? 4u : 0u);
    return head;
  }
};

struct MaskStore {
  uint32_t* mkey;
  uint64_t* mbits;
  uint32_t* n_masks;
  __device__ void operator()(uint32_t i, const MaskPayload& p, uint32_t, uint32_t, uint32_t incl, uint32_t n) const {
    const uint32_t rank = incl - 1;
    if (p.flags & 1u) mkey[rank] = p.key;
    if (p.flags & 4u) atomicOr(reinterpret_cast<unsigned long long*>(&mbits[rank]), (unsigned long long)p.bits);
    else if (p.flags & 2u) mbits[rank] = p.bits;
    if (i == n - 1) *n_masks = incl;
  }
};

__global__ void masks_out_kernel(const uint32_t* n_ptr, const uint32_t* __restrict__ mkey,
                                 const uint64_t* __restrict__ mbits, vx_leaf_mask* __restrict__ out,
                                 uint32_t* leaf_count) {
  const uint32_t n = *n_ptr;
  if (blockIdx.x == 0 && threadIdx.x == 0) *leaf_count = n;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    out[j] = vx_leaf_mask{mkey[j], uint32_t(mbits[j]), uint32_t(mbits[j] >> 32)};
}

// ---- 4. walk and count ----------------------------------------------------------
// info word per mask: start depth of its own new items (4 bits) | first
// missing depth (4 bits, 10 = none) << 4 | inherited fill << 8 | depth shared
// with the previous mask << 12.
__device__ __forceinline__ uint32_t info_start(uint32_t f) { return f & 15u; }
__device__ __forceinline__ uint32_t info_missing(uint32_t f) { return (f >> 4) & 15u; }
__device__ __forceinline__ bool info_filled(uint32_t f) { return (f >> 8) & 1u; }
__device__ __forceinline__ uint32_t info_common(uint32_t f) { return (f >> 12) & 15u; }

// Path entries are stored depth-major: path[d * stride + j].
struct WalkLoad {
  using Payload = uint32_t;
  const uint32_t* mkey;
  Svo v;
  uint32_t* path;
  uint32_t stride;
  uint32_t* error;
  __device__ uint64_t operator()(uint32_t j, uint32_t& info) const {
    const uint32_t key = mkey[j];
    const uint32_t c = j ? common_depth(key, mkey[j - 1]) : 0u;
    uint32_t node = 0, fm = kDepth, inh = 0;
    path[j] = 0;
    for (uint32_t d = 0; d < kDepth; ++d) {
      const uint32_t s = v.nodes[size_t(node) * 8 + child_at(key, d)];
      if (!(s & kChild)) {
        fm = d;
        inh = s >> 31;
        break;
      }
      node = s & kIndexMask;
      if (node >= (d + 1 < kDepth ? v.max_nodes : v.max_leaves)) {  // a corrupt tree: stop here
        *error = 1;
        fm = kDepth;
        break;
      }
      path[size_t(d + 1) * stride + j] = node;
    }
    uint32_t start = 0, nn = 0, nl = 0;
    if (fm < kDepth) {
      start = max(fm + 1, c + 1);
      nn = start < kDepth ? kDepth - start : 0u;
      nl = 1;
    }
    info = start | (fm << 4) | (inh << 8) | (c << 12);
    return scan::PairSumOp::make(nn, nl);
  }
};

struct WalkStore {
  uint32_t* info;
  uint32_t* base_node;
  uint32_t* leaf_rank;
  uint32_t* totals;
  __device__ void operator()(uint32_t j, const uint32_t& f, uint64_t, uint64_t excl, uint64_t incl, uint32_t n) const {
    info[j] = f;
    base_node[j] = scan::PairSumOp::lo(excl);
    leaf_rank[j] = scan::PairSumOp::hi(excl);
    if (j == n - 1) {
      totals[0] = scan::PairSumOp::lo(incl);
      totals[1] = scan::PairSumOp::hi(incl);
    }
  }
};

// ---- 5. build and link ---------------------------------------------------------
struct Alloc {
  uint32_t node_count, leaf_count, free_nodes, free_leaves;  // before the batch
  uint32_t new_nodes, new_leaves;                            // taken by the batch
};

__device__ __forceinline__ uint32_t node_at(const Svo& v, const Alloc& a, uint32_t r) {
  return r < a.free_nodes ? v.free_nodes[a.free_nodes - 1 - r] : a.node_count + (r - a.free_nodes);
}
__device__ __forceinline__ uint32_t leaf_at(const Svo& v, const Alloc& a, uint32_t r) {
  return r < a.free_leaves ? v.free_leaves[a.free_leaves - 1 - r] : a.leaf_count + (r - a.free_leaves);
}

__device__ __forceinline__ void apply_bits(uint32_t* w, uint64_t bits, uint32_t op) {
  if (op == 0) {
    w[0] |= uint32_t(bits);
    w[1] |= uint32_t(bits >> 32);
  } else {
    w[0] &= ~uint32_t(bits);
    w[1] &= ~uint32_t(bits >> 32);
  }
}

__global__ void build_kernel(Svo v, Alloc a, uint32_t n, const uint32_t* __restrict__ mkey,
                             const uint64_t* __restrict__ mbits, const uint32_t* __restrict__ info,
                             const uint32_t* __restrict__ base_node, const uint32_t* __restrict__ leaf_rank,
                             uint32_t* path, uint32_t stride, uint32_t op) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint32_t rn = min(a.new_nodes, a.free_nodes), rl = min(a.new_leaves, a.free_leaves);
    v.counters[0] = a.node_count + (a.new_nodes - rn);
    v.counters[1] = a.leaf_count + (a.new_leaves - rl);
    v.counters[2] = a.free_nodes - rn;
    v.counters[3] = a.free_leaves - rl;
  }
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t key = mkey[j], f = info[j];
    const uint64_t bits = mbits[j];
    const uint32_t fm = info_missing(f);
    if (fm == kDepth) {
      uint32_t* w = v.leaves + size_t(path[size_t(kDepth) * stride + j]) * 2;
      apply_bits(w, bits, op);
      continue;
    }
    const uint32_t start = info_start(f);
    const uint32_t slot_fill = info_filled(f) ? kFilled : 0u;
    uint32_t parent = start == fm + 1 ? path[size_t(fm) * stride + j] : ~0u;
    uint32_t rank = base_node[j];
    for (uint32_t e = start; e <= kDepth; ++e) {
      uint32_t idx;
      if (e < kDepth) {
        idx = node_at(v, a, rank++);
        uint4* q = reinterpret_cast<uint4*>(v.nodes + size_t(idx) * 8);
        const uint4 u = make_uint4(slot_fill, slot_fill, slot_fill, slot_fill);
        q[0] = u;
        q[1] = u;
      } else {
        idx = leaf_at(v, a, leaf_rank[j]);
        uint32_t w[2] = {slot_fill ? 0xffffffffu : 0u, slot_fill ? 0xffffffffu : 0u};
        apply_bits(w, bits, op);
        *reinterpret_cast<uint2*>(v.leaves + size_t(idx) * 2) = make_uint2(w[0], w[1]);
      }
      path[size_t(e) * stride + j] = idx;
      if (parent != ~0u) v.nodes[size_t(parent) * 8 + child_at(key, e - 1)] = idx | kChild | slot_fill;
      parent = idx;
    }
  }
}

// First index i <= j whose key shares the top `shift`-free bits with key j
// (keys are sorted, so the run is contiguous): gallop back, then bisect.
__device__ __forceinline__ uint32_t run_start(const uint32_t* __restrict__ mkey, uint32_t j, uint32_t shift) {
  const uint32_t prefix = mkey[j] >> shift;
  uint32_t hi = j, step = 1, lo = 0;
  for (;;) {
    if (hi < step) {
      lo = 0;
      break;
    }
    const uint32_t t = hi - step;
    if ((mkey[t] >> shift) == prefix) {
      hi = t;
      step <<= 1;
    } else {
      lo = t + 1;
      break;
    }
  }
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    if ((mkey[m] >> shift) == prefix) hi = m;
    else lo = m + 1;
  }
  return lo;
}

__global__ void link_kernel(Svo v, uint32_t n, const uint32_t* __restrict__ mkey, const uint32_t* __restrict__ info,
                            uint32_t* path, uint32_t stride) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t f = info[j];
    const uint32_t fm = info_missing(f), start = info_start(f);
    if (fm == kDepth || start == fm + 1) continue;
    const uint32_t p = start - 1;  // the depth shared with the previous mask
    const uint32_t key = mkey[j];
    const uint32_t creator = run_start(mkey, j, (kDepth - p) * 3);
    const uint32_t parent = path[size_t(p) * stride + creator];
    if (parent >= v.max_nodes) continue;
    v.nodes[size_t(parent) * 8 + child_at(key, p)] =
        path[size_t(start) * stride + j] | kChild | (info_filled(f) ? kFilled : 0u);
    path[size_t(p) * stride + j] = parent;
  }
}

// ---- 6. collapse -------------------------------------------------------------------
struct CollapsePayload {
  uint32_t idx, parent_slot, full, base;
};

struct CollapseLoad {
  using Payload = CollapsePayload;
  Svo v;
  const uint32_t* mkey;
  const uint32_t* info;
  const uint32_t* path;
  uint32_t stride;
  uint32_t d;
  __device__ uint32_t operator()(uint32_t j, CollapsePayload& p) const {
    if (info_common(info[j]) >= d) return 0;  // not the first mask of its depth-d prefix
    const uint32_t idx = path[size_t(d) * stride + j];
    if (idx >= (d == kDepth ? v.max_leaves : v.max_nodes)) return 0;
    bool uniform, full;
    if (d == kDepth) {
      const uint2 w = *reinterpret_cast<const uint2*>(v.leaves + size_t(idx) * 2);
      full = w.x == 0xffffffffu && w.y == 0xffffffffu;
      uniform = full || (w.x == 0 && w.y == 0);
    } else {
      const uint4* q = reinterpret_cast<const uint4*>(v.nodes + size_t(idx) * 8);
      const uint4 a = q[0], b = q[1];
      const uint32_t any_child = (a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w) & kChild;
      const uint32_t all = a.x & a.y & a.z & a.w & b.x & b.y & b.z & b.w;
      const uint32_t some = a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w;
      full = !any_child && (all & kFilled);
      uniform = !any_child && (full || !(some & kFilled));
    }
    if (!uniform) return 0;
    const uint32_t parent = path[size_t(d - 1) * stride + j];
    if (parent >= v.max_nodes) return 0;
    p.idx = idx;
    p.parent_slot = parent * 8 + child_at(mkey[j], d - 1);
    p.full = full;
    p.base = v.counters[d == kDepth ? 3 : 2];
    return 1;
  }
};

struct CollapseStore {
  Svo v;
  uint32_t d;
  const uint32_t* n_masks;
  uint32_t* active_next;
  __device__ void operator()(uint32_t j, const CollapsePayload& p, uint32_t f, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    if (f && p.base + excl < (d == kDepth ? v.max_leaves : v.max_nodes)) {
      (d == kDepth ? v.free_leaves : v.free_nodes)[p.base + excl] = p.idx;
      v.nodes[p.parent_slot] = p.full ? kFilled : 0u;
    }
    if (j == n - 1) {
      v.counters[d == kDepth ? 3 : 2] += incl;
      *active_next = incl ? *n_masks : 0u;
    }
  }
};

__global__ void set_active_kernel(uint32_t* cnt) { cnt[kActive + kDepth] = cnt[kNMasks]; }

__global__ void reset_kernel(Svo v) {
  if (threadIdx.x < 8) v.nodes[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    v.counters[0] = 1;
    v.counters[1] = 0;
    v.counters[2] = 0;
    v.counters[3] = 0;
  }
}

__global__ void get_kernel(Svo v, const vx_voxel* __restrict__ q, uint32_t n, uint32_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const vx_voxel p = q[i];
    uint32_t r = 0;
    if (p.x < kWorld && p.y < kWorld && p.z < kWorld) {
      const uint32_t key = leaf_key(p.x, p.y, p.z);
      uint32_t node = 0, d = 0;
      for (; d < kDepth; ++d) {
        const uint32_t s = v.nodes[size_t(node) * 8 + child_at(key, d)];
        if (!(s & kChild)) {
          r = s >> 31;
          break;
        }
        node = s & kIndexMask;
        if (node >= (d + 1 < kDepth ? v.max_nodes : v.max_leaves)) break;  // corrupt tree: read as empty
      }
      if (d == kDepth) {
        const uint32_t bit = (p.x & 3u) | ((p.y & 3u) << 2) | ((p.z & 3u) << 4);
        r = (v.leaves[size_t(node) * 2 + (bit >> 5)] >> (bit & 31u)) & 1u;
      }
    }
    out[i] = r;
  }
}

// Workspace of one batch of `count` edits.
struct Ws {
  uint32_t* cnt;
  uint64_t* recs;
  uint64_t* alt;
  void* sort_ws;
  sort::Tuning sort_tn;
  void* fused_ws;
  uint32_t* mkey;
  uint64_t* mbits;
  uint32_t* info;
  uint32_t* base_node;
  uint32_t* leaf_rank;
  uint32_t* path;
  size_t bytes;
};

Ws carve(void* p, uint32_t count, bool edit) {
  const uint64_t n = std::max<uint32_t>(count, 1);
  Carver c(p);
  Ws w{};
  w.cnt = c.take<uint32_t>(64);
  w.recs = c.take<uint64_t>(n);
  w.alt = c.take<uint64_t>(n);
  w.sort_tn = sort::keys_u64_dev_tuning(kKeyBits);
  w.sort_ws = c.take<char>(sort::keys_u64_dev_workspace_bytes(n, kKeyBits, w.sort_tn));
  w.fused_ws = c.take<char>(scan::fused_dev_workspace_bytes(n));
  w.mkey = c.take<uint32_t>(n);
  w.mbits = c.take<uint64_t>(n);
  if (edit) {
    w.info = c.take<uint32_t>(n);
    w.base_node = c.take<uint32_t>(n);
    w.leaf_rank = c.take<uint32_t>(n);
    w.path = c.take<uint32_t>(n * (kDepth + 1));
  }
  w.bytes = c.size();
  return w;
}

// Stages 1-3: leaf masks (keys, bits) of the valid edits, count in
// w.cnt[kNMasks]. Sorted by key.
vx_status build_masks(const vx_voxel* edits, uint32_t count, const Ws& w, hipStream_t s) {
  vx_status st;
  if ((st = fill_u32(w.cnt, 64, 0, s)) != VX_OK) return st;
  if ((st = fill_u32(reinterpret_cast<uint32_t*>(w.mbits), uint64_t(count) * 2, 0, s)) != VX_OK) return st;
  if ((st = scan::fused<scan::kSum>(count, RecLoad{edits}, RecStore{w.recs, w.cnt + kNValid}, w.fused_ws, s)) != VX_OK)
    return st;
  bool in_alt = false;
  if ((st = sort::keys_u64_dev(w.recs, w.alt, w.cnt + kNValid, count, 0, kKeyBits, w.sort_ws, w.sort_tn, &in_alt, s)) != VX_OK)
    return st;
  const uint64_t* sorted = in_alt ? w.alt : w.recs;
  return scan::fused_dev<scan::kSum>(w.cnt + kNValid, count, MaskLoad{sorted, w.cnt + kNValid},
                                     MaskStore{w.mkey, w.mbits, w.cnt + kNMasks}, w.fused_ws, s);
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------------
extern "C" vx_status vx_svo_create(uint32_t max_node_count, uint32_t max_leaf_count, vx_svo** out) {
  if (!out || max_node_count == 0 || max_leaf_count == 0 || max_node_count > kIndexMask + 1 ||
      max_leaf_count > kIndexMask + 1)
    return VX_INVALID_VALUE;
  vx_svo* g = new (std::nothrow) vx_svo();
  if (!g) return VX_DEVICE_ERROR;
  g->max_nodes = max_node_count;
  g->max_leaves = max_leaf_count;
  const bool ok = hipMalloc(&g->nodes, size_t(max_node_count) * 32) == hipSuccess &&
                  hipMalloc(&g->leaves, size_t(max_leaf_count) * 8) == hipSuccess &&
                  hipMalloc(&g->counters, 16) == hipSuccess &&
                  hipMalloc(&g->free_nodes, size_t(max_node_count) * 4) == hipSuccess &&
                  hipMalloc(&g->free_leaves, size_t(max_leaf_count) * 4) == hipSuccess;
  if (!ok) {
    vx_svo_destroy(g);
    return VX_OUT_OF_CAPACITY;
  }
  // The tree starts reset (empty root, counters {1, 0, 0, 0}): the
  // reference's constructor only allocates and leaves the counters and free
  // stacks as the allocation found them until reset_svo, and an edit on such
  // a tree pops whatever indices those words hold.
  reset_kernel<<<1, 64>>>(dev_of(g));
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    vx_svo_destroy(g);
    return VX_DEVICE_ERROR;
  }
  *out = g;
  return VX_OK;
}

extern "C" vx_status vx_svo_destroy(vx_svo* g) {
  if (!g) return VX_INVALID_VALUE;
  for (void* p : {static_cast<void*>(g->nodes), static_cast<void*>(g->leaves), static_cast<void*>(g->counters),
                  static_cast<void*>(g->free_nodes), static_cast<void*>(g->free_leaves)})
    if (p) (void)hipFree(p);
  delete g;
  return VX_OK;
}

extern "C" vx_status vx_svo_reset(vx_svo* g, void* stream) {
  if (!g) return VX_INVALID_VALUE;
  reset_kernel<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(dev_of(g));
  return last_launch();
}

extern "C" vx_status vx_svo_view_get(const vx_svo* g, vx_svo_view* out) {
  if (!g || !out) return VX_INVALID_VALUE;
  *out = vx_svo_view{g->nodes, g->leaves, reinterpret_cast<vx_svo_counters*>(g->counters), g->free_nodes,
                     g->free_leaves, g->max_nodes, g->max_leaves};
  return VX_OK;
}

extern "C" vx_status vx_svo_counters_get(const vx_svo* g, vx_svo_counters* out, void* stream) {
  if (!g || !out) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(out, g->counters, 16, hipMemcpyDeviceToHost, s) != hipSuccess) return VX_DEVICE_ERROR;
  return from_hip(hipStreamSynchronize(s));
}

extern "C" vx_status vx_voxel_edits_to_leaf_masks_workspace_size(uint32_t count, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  *bytes = carve(nullptr, count, false).bytes;
  return VX_OK;
}

extern "C" vx_status vx_voxel_edits_to_leaf_masks(const vx_voxel* edits, uint32_t count, vx_leaf_mask* leaf_masks,
                                                  uint32_t* leaf_count, void* ws, size_t ws_bytes, void* stream) {
  if ((!edits && count) || !leaf_masks || !leaf_count) return VX_INVALID_VALUE;
  if (count >= (1u << 30)) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (count == 0) return fill_u32(leaf_count, 1, 0, s);
  const Ws w = carve(ws, count, false);
  if (ws_bytes < w.bytes || !ws) return VX_INVALID_VALUE;
  vx_status st = build_masks(edits, count, w, s);
  if (st != VX_OK) return st;
  masks_out_kernel<<<dag::grid_for(count), 256, 0, s>>>(w.cnt + kNMasks, w.mkey, w.mbits, leaf_masks, leaf_count);
  return last_launch();
}

extern "C" vx_status vx_svo_voxel_edits_workspace_size(uint32_t count, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  *bytes = carve(nullptr, count, true).bytes;
  return VX_OK;
}

extern "C" vx_status vx_svo_apply_voxel_edits(vx_svo* g, const vx_voxel* edits, uint32_t count, uint32_t op, void* ws,
                                              size_t ws_bytes, void* stream) {
  if (!g || (!edits && count) || op > 1 || !ws) return VX_INVALID_VALUE;
  if (count >= (1u << 30)) return VX_INVALID_VALUE;
  const Ws w = carve(ws, count, true);
  if (ws_bytes < w.bytes) return VX_INVALID_VALUE;
  if (count == 0) return VX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Svo v = dev_of(g);
  vx_status st = build_masks(edits, count, w, s);
  if (st != VX_OK) return st;
  const uint32_t stride = count;
  if ((st = scan::fused_dev_m<scan::PairSumOp>(
           w.cnt + kNMasks, count, WalkLoad{w.mkey, v, w.path, stride, w.cnt + kError},
           WalkStore{w.info, w.base_node, w.leaf_rank, w.cnt + kTotNodes}, w.fused_ws, s)) != VX_OK)
    return st;
  // One read-back: mask count, totals and the counters.
  uint32_t h[12];
  if (hipMemcpyAsync(h, w.cnt, 32, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(h + 8, g->counters, 16, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (h[kError]) return VX_DEVICE_ERROR;
  const uint32_t n_masks = h[kNMasks];
  if (n_masks == 0) return VX_OK;
  Alloc a{h[8], h[9], h[10], h[11], h[kTotNodes], h[kTotLeaves]};
  const uint64_t take_n = a.new_nodes > a.free_nodes ? a.new_nodes - a.free_nodes : 0;
  const uint64_t take_l = a.new_leaves > a.free_leaves ? a.new_leaves - a.free_leaves : 0;
  if (a.node_count + take_n > g->max_nodes || a.leaf_count + take_l > g->max_leaves) return VX_OUT_OF_CAPACITY;

  const uint32_t grid = dag::grid_for(n_masks);
  build_kernel<<<grid, 256, 0, s>>>(v, a, n_masks, w.mkey, w.mbits, w.info, w.base_node, w.leaf_rank, w.path, stride,
                                    op);
  if ((st = last_launch()) != VX_OK) return st;
  link_kernel<<<grid, 256, 0, s>>>(v, n_masks, w.mkey, w.info, w.path, stride);
  if ((st = last_launch()) != VX_OK) return st;
  set_active_kernel<<<1, 1, 0, s>>>(w.cnt);
  if ((st = last_launch()) != VX_OK) return st;
  for (uint32_t d = kDepth; d >= 1; --d) {
    if ((st = scan::fused_dev<scan::kSum>(w.cnt + kActive + d, n_masks,
                                          CollapseLoad{v, w.mkey, w.info, w.path, stride, d},
                                          CollapseStore{v, d, w.cnt + kNMasks, w.cnt + kActive + d - 1}, w.fused_ws,
                                          s)) != VX_OK)
      return st;
  }
  return VX_OK;
}

extern "C" vx_status vx_svo_get_voxels(const vx_svo* g, const vx_voxel* q, uint32_t n, uint32_t* out, void* stream) {
  if (!g || ((!q || !out) && n)) return VX_INVALID_VALUE;
  if (n == 0) return VX_OK;
  get_kernel<<<dag::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(dev_of(g), q, n, out);
  return last_launch();
}

// Canonical export: nodes in pre-order from the root (children 0..7), leaves
// in the order the walk reaches them; a pointer slot carries the child's
// canonical id with the slot's own flag bits.
extern "C" vx_status vx_svo_export_canonical(const vx_svo* g, uint32_t* out, uint64_t cap, uint64_t* n_words) {
  if (!g || !n_words) return VX_INVALID_VALUE;
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  uint32_t c[4];
  if (hipMemcpy(c, g->counters, 16, hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nn = std::min(c[0], g->max_nodes), nl = std::min(c[1], g->max_leaves);
  std::vector<uint32_t> nodes(nn * 8), leaves(nl * 2);
  if (nn && hipMemcpy(nodes.data(), g->nodes, nn * 32, hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  if (nl && hipMemcpy(leaves.data(), g->leaves, nl * 8, hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  std::vector<uint32_t> node_recs(8), leaf_recs;
  struct Frame {
    uint32_t node, depth, id, c;
  };
  std::vector<Frame> stack{Frame{0, 0, 0, 0}};
  uint32_t next_node = 1, next_leaf = 0;
  while (!stack.empty()) {
    Frame& f = stack.back();
    if (f.c == 8) {
      stack.pop_back();
      continue;
    }
    const uint32_t ci = f.c++;
    if (f.node >= nn) return VX_DEVICE_ERROR;
    const uint32_t sl = nodes[size_t(f.node) * 8 + ci];
    const size_t at = size_t(f.id) * 8 + ci;
    if (!(sl & kChild)) {
      node_recs[at] = sl;
      continue;
    }
    const uint32_t child = sl & kIndexMask;
    if (f.depth + 1 == kDepth) {
      if (child >= nl) return VX_DEVICE_ERROR;
      node_recs[at] = next_leaf++ | (sl & (kChild | kFilled));
      leaf_recs.push_back(leaves[size_t(child) * 2]);
      leaf_recs.push_back(leaves[size_t(child) * 2 + 1]);
      continue;
    }
    const uint32_t id = next_node++;
    node_recs[at] = id | (sl & (kChild | kFilled));
    node_recs.resize(size_t(next_node) * 8);
    const uint32_t depth = f.depth + 1;
    stack.push_back(Frame{child, depth, id, 0});  // invalidates f
  }
  const uint64_t total = 3 + node_recs.size() + leaf_recs.size();
  *n_words = total;
  if (!out || cap < total) return VX_OK;
  out[0] = 0x5658534fu;
  out[1] = next_node;
  out[2] = next_leaf;
  std::copy(node_recs.begin(), node_recs.end(), out + 3);
  std::copy(leaf_recs.begin(), leaf_recs.end(), out + 3 + node_recs.size());
  return VX_OK;
}
