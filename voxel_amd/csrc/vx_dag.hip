// Hash DAG edit for gfx950.
//
// Replaces the reference's hash_dag_gpu edit pipeline (detail/edit_pipeline.cuh:31-97)
// with a level-compacted design:
//   records   brushes (or voxels) -> one record per touched 4^3 leaf brick
//             {leaf Morton key, voxel mask, material}, compacted in brush order
//             by the fused scan (replaces pack_leaf_edit_records_kernel,
//             build_leaf_masks/pipeline.cuh:16-36);
//   sort      onesweep sort of (key, record index), 3*D key bits (stable, so
//             records of one leaf stay in brush order: last writer wins);
//   levels    run compaction -> distinct leaves; then, bottom-up, distinct
//             parents per level with child ranges (replaces the per-depth
//             request emission + binary-search parent link,
//             detail/edit/requests.cuh:51-173) — every later kernel runs over
//             exactly one level's entries, not over 9N threads;
//   old refs  one top-down pass per level: each entry reads one parent item
//             (replaces the per-request root walk, detail/edit/snapshots.cuh:14-53);
//   rebuild   leaves then levels D-1..0: build item -> probe -> commit
//             (replaces apply_leaf_masks / process_depth,
//             detail/edit/leaf_update.cuh:52-109, detail/edit/rebuild.cuh:30-89).
// Interning is a true insert-if-absent on a u64-slot table: the probe kernel
// claims an empty slot by CAS with {fingerprint, kTag | item index}; an equal
// item that meets the tagged slot becomes a duplicate of that representative
// (the in-batch dedup the reference does with a temporary map,
// detail/edit/temp_dedup.cuh:12-64); the commit kernel (a fused scan over the
// representatives' widths) copies each representative into the pool at its
// scanned offset and rewrites its slot with the pool offset. No pool space is
// ever allocated for a duplicate, and no thread waits on another.
// Status: first error wins (detail/edit_common.cuh:96-99); the root is
// published only on success (detail/edit/root.cuh:13-25).
#include <cmath>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "vx_dag.hpp"
#include "vx_scan.hpp"
#include "vx_sort.hpp"

using namespace vx;
using namespace vx::dag;

namespace vx {
namespace dag {

// ---- interning --------------------------------------------------------------
struct Table {
  uint64_t* slots;
  uint64_t mask;
  uint32_t* pool;
  uint64_t cap;
  uint32_t* used;
  uint32_t* base_snapshot;
};

__global__ void probe_kernel(ItemSet set, const uint64_t* __restrict__ hashes, const uint32_t* n_ptr, uint64_t* res,
                             Table t, uint32_t* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *t.base_snapshot = *t.used;
  const uint32_t n = *n_ptr;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (res_kind(res[i]) != kResPending) continue;
    const uint32_t* it = set.items + size_t(i) * set.stride;
    const uint32_t w = set.width(it);
    const uint64_t h = hashes[i];
    const uint32_t fp = uint32_t(h >> 32);
    uint64_t idx = h & t.mask;
    const uint64_t limit = t.mask < kMaxProbe ? t.mask + 1 : kMaxProbe;
    uint64_t out = res_make(kResFinal, kEmpty);
    bool done = false;
    for (uint64_t probe = 0; probe < limit && !done; ++probe) {
      uint64_t s = ld_agent(&t.slots[idx]);
      if (s == 0) {
        const uint64_t mine = (uint64_t(fp) << 32) | kTag | i;
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&t.slots[idx]), 0ull,
                                        static_cast<unsigned long long>(mine));
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
          if (words_equal(it, set.items + size_t(j) * set.stride, w)) {
            out = res_make(kResDup, j);
            done = true;
            break;
          }
        } else if (v >= kPoolStart) {
          if (words_equal(it, t.pool + v, w)) {
            out = res_make(kResFinal, v);
            done = true;
            break;
          }
        }
      }
      idx = (idx + 1) & t.mask;
    }
    if (!done) set_status(ctl, kStatusOverflow);
    res[i] = out;
  }
}

struct CommitLoad {
  using Payload = uint64_t;
  const uint64_t* res;
  ItemSet set;
  __device__ uint32_t operator()(uint32_t i, uint64_t& r) const {
    r = res[i];
    return res_kind(r) == kResRep ? set.width(set.items + size_t(i) * set.stride) : 0u;
  }
};

struct CommitStore {
  ItemSet set;
  uint64_t* res;
  Table t;
  uint32_t* ctl;
  const uint32_t* n_ptr;
  __device__ void operator()(uint32_t i, const uint64_t& r, uint32_t w, uint32_t excl, uint32_t incl) const {
    const uint32_t base = *t.base_snapshot;
    if (w) {
      uint64_t* slot = &t.slots[uint32_t(r)];
      const uint64_t fp_hi = ld_agent(slot) & 0xFFFFFFFF00000000ull;
      const uint64_t off = uint64_t(base) + excl;
      if (off + w > t.cap || off >= kTag) {
        set_status(ctl, kStatusOverflow);
        st_agent(slot, fp_hi | kTomb);
        res[i] = res_make(kResFinal, kEmpty);
      } else {
        const uint32_t* it = set.items + size_t(i) * set.stride;
        for (uint32_t k = 0; k < w; ++k) t.pool[off + k] = it[k];
        st_agent(slot, fp_hi | uint32_t(off));
        res[i] = res_make(kResFinal, uint32_t(off));
      }
    }
    if (i == *n_ptr - 1) {
      const uint64_t used = uint64_t(base) + incl;
      *t.used = used > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(used);
    }
  }
};

Table node_table(const Dev& d) {
  return Table{d.node_slots, d.node_mask, d.node_pool, d.node_cap, d.ctl + kCtlNodeUsed, d.ctl + kCtlNodeBase};
}
Table leaf_table(const Dev& d) {
  return Table{d.leaf_slots, d.leaf_mask, d.leaf_pool, d.leaf_cap, d.ctl + kCtlLeafUsed, d.ctl + kCtlLeafBase};
}

vx_status intern(const Dev& d, bool leaves, ItemSet set, const uint64_t* hashes, const uint32_t* n_ptr, uint64_t n_ub,
                 uint64_t* res, void* fused_ws, hipStream_t s) {
  const Table t = leaves ? leaf_table(d) : node_table(d);
  probe_kernel<<<grid_for(n_ub), 256, 0, s>>>(set, hashes, n_ptr, res, t, d.ctl);
  vx_status st = last_launch();
  if (st != VX_OK) return st;
  return scan::fused_dev<scan::kSum>(n_ptr, n_ub, CommitLoad{res, set}, CommitStore{set, res, t, d.ctl, n_ptr},
                                     fused_ws, s);
}

// ---- records ----------------------------------------------------------------
struct BrushJob {
  vx_brush b;
  int32_t bx0, by0, bz0;
  uint32_t nx, ny, nz;
  uint32_t start;
  uint32_t pad;
};

struct Record {
  uint64_t key;
  uint64_t mask;
  uint32_t mat;
};

__device__ __forceinline__ uint64_t brick_mask(const vx_brush& b, int32_t X, int32_t Y, int32_t Z) {
  uint64_t m = 0;
  if (b.kind == 0) {
    const int64_t r2 = int64_t(b.hx) * b.hx;
#pragma unroll
    for (int lz = 0; lz < 4; ++lz) {
      const int64_t dz = int64_t(Z + lz) - b.cz;
#pragma unroll
      for (int ly = 0; ly < 4; ++ly) {
        const int64_t dy = int64_t(Y + ly) - b.cy;
        const int64_t dyz = dy * dy + dz * dz;
#pragma unroll
        for (int lx = 0; lx < 4; ++lx) {
          const int64_t dx = int64_t(X + lx) - b.cx;
          if (dx * dx + dyz <= r2) m |= 1ull << (lx + 4 * ly + 16 * lz);
        }
      }
    }
  } else {
    uint32_t xm = 0, ym = 0, zm = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (abs(X + l - b.cx) <= b.hx) xm |= 1u << l;
      if (abs(Y + l - b.cy) <= b.hy) ym |= 1u << l;
      if (abs(Z + l - b.cz) <= b.hz) zm |= 1u << l;
    }
#pragma unroll
    for (int lz = 0; lz < 4; ++lz)
#pragma unroll
      for (int ly = 0; ly < 4; ++ly)
        if (((zm >> lz) & 1u) && ((ym >> ly) & 1u)) m |= uint64_t(xm) << (4 * ly + 16 * lz);
  }
  return m;
}

struct RecordOut {
  uint64_t* key;
  uint32_t* val;
  uint64_t* mask;
  uint32_t* mat;
  uint32_t* n_rec;
};

struct BrushLoad {
  using Payload = Record;
  const BrushJob* jobs;
  uint32_t n_jobs;
  bool materials;
  __device__ uint32_t operator()(uint32_t i, Record& rec) const {
    uint32_t lo = 0, hi = n_jobs;  // last job with start <= i
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (jobs[mid].start <= i) lo = mid;
      else hi = mid;
    }
    const BrushJob& j = jobs[lo];
    const uint32_t local = i - j.start;
    const uint32_t bx = local % j.nx, by = (local / j.nx) % j.ny, bz = local / (j.nx * j.ny);
    const uint32_t gx = j.bx0 + bx, gy = j.by0 + by, gz = j.bz0 + bz;
    rec.mask = brick_mask(j.b, int32_t(gx * 4), int32_t(gy * 4), int32_t(gz * 4));
    rec.key = morton3(gx, gy, gz);
    rec.mat = materials ? j.b.material : (j.b.material ? 1u : 0u);
    return rec.mask != 0;
  }
};

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
  uint32_t n;
  __device__ void operator()(uint32_t i, const Record& rec, uint32_t v, uint32_t excl, uint32_t incl) const {
    if (v) {
      out.key[excl] = rec.key;
      out.val[excl] = excl;
      out.mask[excl] = rec.mask;
      out.mat[excl] = rec.mat;
    }
    if (i == n - 1) *out.n_rec = incl;
  }
};

// ---- level structure --------------------------------------------------------
struct KeyLoad {
  using Payload = uint64_t;
  const uint64_t* keys;
  uint32_t shift;  // 0: distinct keys; 3: distinct parents
  __device__ uint32_t operator()(uint32_t i, uint64_t& k) const {
    k = keys[i];
    return i == 0 || (keys[i - 1] >> shift) != (k >> shift);
  }
};

struct LevelStore {
  uint64_t* up_keys;
  uint32_t* up_first;
  uint32_t* parent;  // may be null
  uint32_t* up_n;
  const uint32_t* n_ptr;
  uint32_t shift;
  __device__ void operator()(uint32_t i, const uint64_t& k, uint32_t f, uint32_t excl, uint32_t incl) const {
    if (parent) parent[i] = incl - 1;
    if (f) {
      up_keys[excl] = k >> shift;
      up_first[excl] = i;
    }
    if (i == *n_ptr - 1) {
      *up_n = incl;
      up_first[incl] = i + 1;
    }
  }
};

__global__ void root_old_kernel(const uint32_t* ctl, const uint32_t* n0, uint32_t* old0) {
  if (*n0 > 0) old0[0] = ctl[kCtlRoot];
}

__global__ void topdown_kernel(const uint32_t* __restrict__ node_pool, const uint32_t* n_ptr,
                               const uint64_t* __restrict__ keys, const uint32_t* __restrict__ parent,
                               const uint32_t* __restrict__ old_prev, uint32_t* __restrict__ old) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    old[j] = child_of(node_pool, old_prev[parent[j]], uint32_t(keys[j] & 7u));
}

// ---- rebuild ---------------------------------------------------------------
template <uint32_t LW>
__global__ void leaf_build_kernel(Dev d, const uint32_t* n_ptr, const uint32_t* __restrict__ old,
                                  const uint32_t* __restrict__ first, const uint32_t* __restrict__ sval,
                                  const uint64_t* __restrict__ rmask, const uint32_t* __restrict__ rmat,
                                  uint32_t* __restrict__ items, uint64_t* __restrict__ hashes,
                                  uint64_t* __restrict__ res) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t ref = old[j];
    uint32_t w[LW];
    if (ref == kEmpty) {
#pragma unroll
      for (uint32_t k = 0; k < LW; ++k) w[k] = 0;
    } else if (ref == kFull) {
#pragma unroll
      for (uint32_t k = 0; k < LW; ++k) w[k] = 0xffffffffu;
    } else {
#pragma unroll
      for (uint32_t k = 0; k < LW; ++k) w[k] = d.leaf_pool[ref + k];
    }
    const uint32_t r1 = first[j + 1];
    for (uint32_t r = first[j]; r < r1; ++r) {
      const uint32_t idx = sval[r];
      const uint64_t m = rmask[idx];
      const uint32_t mat = rmat[idx];
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
    }
    uint32_t any = 0, all = 0xffffffffu;
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) {
      any |= w[k];
      all &= w[k];
    }
    if (!any) {
      res[j] = res_make(kResFinal, kEmpty);
    } else if (LW == 2 && all == 0xffffffffu) {
      res[j] = res_make(kResFinal, kFull);
    } else {
      uint32_t* it = items + size_t(j) * LW;
#pragma unroll
      for (uint32_t k = 0; k < LW; ++k) it[k] = w[k];
      hashes[j] = hash_words(w, LW);
      res[j] = res_make(kResPending, 0);
    }
  }
}

__global__ void node_build_kernel(Dev d, const uint32_t* n_ptr, const uint32_t* __restrict__ old,
                                  const uint32_t* __restrict__ first, const uint64_t* __restrict__ ckeys,
                                  const uint64_t* __restrict__ cres, uint32_t* __restrict__ items,
                                  uint64_t* __restrict__ hashes, uint64_t* __restrict__ res) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t ch[8];
    children_of(d.node_pool, old[j], ch);
    const uint32_t k1 = first[j + 1];
    for (uint32_t k = first[j]; k < k1; ++k) ch[uint32_t(ckeys[k] & 7u)] = resolve(cres, k);
    uint32_t* it = items + size_t(j) * kNodeStride;
    uint32_t ref = 0;
    const uint32_t w = make_node_item(ch, it, &ref);
    if (!w) {
      res[j] = res_make(kResFinal, ref);
    } else {
      hashes[j] = hash_words(it, w);
      res[j] = res_make(kResPending, 0);
    }
  }
}

__global__ void publish_kernel(uint32_t* ctl, const uint32_t* n0, const uint64_t* res0) {
  if (*n0 == 0 || ctl[kCtlStatus] != kStatusOk) return;
  ctl[kCtlRoot] = resolve(res0, 0);
}

// ---- workspace ---------------------------------------------------------------
struct Levels {
  uint64_t* keys[16];
  uint32_t* first[16];
  uint32_t* parent[16];
  uint32_t* old[16];
  uint64_t* res[16];
  uint64_t cap[16];
};

struct EditWs {
  uint32_t* counts;  // [0] records, [1 + d] entries at level d
  uint64_t *rkey, *rkey_alt, *rmask;
  uint32_t *rval, *rval_alt, *rmat;
  void* sort_ws;
  Levels lv;
  uint32_t* items;
  uint64_t* hashes;
  void* fused_ws;
  BrushJob* jobs;
  size_t bytes;
};

EditWs carve(void* base, uint32_t depth, uint32_t leaf_words, uint64_t r_ub, uint32_t n_jobs, uint64_t candidates) {
  EditWs w{};
  Carver c(base);
  w.counts = c.take<uint32_t>(64);
  const uint64_t R = std::max<uint64_t>(1, r_ub);
  w.rkey = c.take<uint64_t>(R);
  w.rkey_alt = c.take<uint64_t>(R);
  w.rmask = c.take<uint64_t>(R);
  w.rval = c.take<uint32_t>(R);
  w.rval_alt = c.take<uint32_t>(R);
  w.rmat = c.take<uint32_t>(R);
  const uint32_t key_bits = 3 * depth;
  const size_t sws = sort::pairs_u64_dev_workspace_bytes(R, key_bits);
  w.sort_ws = c.take<char>(sws);
  uint64_t max_cap = 1, max_items = 1;
  for (uint32_t d = 0; d <= depth; ++d) {
    uint64_t cap = R;
    if (d < depth && 3 * d < 60) cap = std::min<uint64_t>(R, uint64_t(1) << (3 * d));
    w.lv.cap[d] = cap;
    w.lv.keys[d] = c.take<uint64_t>(cap);
    w.lv.first[d] = c.take<uint32_t>(cap + 1);
    w.lv.parent[d] = c.take<uint32_t>(cap);
    w.lv.old[d] = c.take<uint32_t>(cap);
    w.lv.res[d] = c.take<uint64_t>(cap);
    max_cap = std::max(max_cap, cap);
    max_items = std::max<uint64_t>(max_items, cap * (d == depth ? leaf_words : kNodeStride));
  }
  w.items = c.take<uint32_t>(max_items);
  w.hashes = c.take<uint64_t>(max_cap);
  w.fused_ws = c.take<char>(std::max(scan::fused_dev_workspace_bytes(std::max(max_cap, R)),
                                     scan::fused_workspace_bytes(std::max<uint64_t>(candidates, 1))));
  w.jobs = c.take<BrushJob>(std::max<uint32_t>(1, n_jobs));
  w.bytes = c.size();
  return w;
}

inline void mark(vx_dag* g, uint32_t i, hipStream_t s) {
  if (g->profile) hipEventRecord(g->ev[i], s);
}

// Everything after record generation: sort, levels, old refs, rebuild, publish.
vx_status rebuild(vx_dag* g, EditWs& w, uint64_t r_ub, hipStream_t s) {
  const Dev& d = g->dev;
  const uint32_t D = g->depth;
  uint32_t* cnt = w.counts;
  vx_status st;
  bool in_alt = false;
  mark(g, 1, s);
  if ((st = sort::pairs_u64_dev(w.rkey, w.rval, w.rkey_alt, w.rval_alt, &cnt[0], std::max<uint64_t>(r_ub, 1), 3 * D,
                                w.sort_ws, &in_alt, s)) != VX_OK)
    return st;
  mark(g, 2, s);
  const uint64_t* skey = in_alt ? w.rkey_alt : w.rkey;
  const uint32_t* sval = in_alt ? w.rval_alt : w.rval;
  Levels& L = w.lv;
  // Distinct leaves (level D) and their record ranges.
  if ((st = scan::fused_dev<scan::kSum>(&cnt[0], r_ub, KeyLoad{skey, 0},
                                        LevelStore{L.keys[D], L.first[D], nullptr, &cnt[1 + D], &cnt[0], 0},
                                        w.fused_ws, s)) != VX_OK)
    return st;
  // Distinct ancestors, bottom-up.
  for (uint32_t dd = D; dd-- > 0;) {
    if ((st = scan::fused_dev<scan::kSum>(&cnt[1 + dd + 1], L.cap[dd + 1], KeyLoad{L.keys[dd + 1], 3},
                                          LevelStore{L.keys[dd], L.first[dd], L.parent[dd + 1], &cnt[1 + dd],
                                                     &cnt[1 + dd + 1], 3},
                                          w.fused_ws, s)) != VX_OK)
      return st;
  }
  mark(g, 3, s);
  // Old refs, top-down.
  root_old_kernel<<<1, 1, 0, s>>>(d.ctl, &cnt[1], L.old[0]);
  if ((st = last_launch()) != VX_OK) return st;
  for (uint32_t dd = 1; dd <= D; ++dd) {
    topdown_kernel<<<grid_for(L.cap[dd]), 256, 0, s>>>(d.node_pool, &cnt[1 + dd], L.keys[dd], L.parent[dd],
                                                         L.old[dd - 1], L.old[dd]);
    if ((st = last_launch()) != VX_OK) return st;
  }
  mark(g, 4, s);
  // Leaves.
  if (d.leaf_words == 2)
    leaf_build_kernel<2><<<grid_for(L.cap[D]), 256, 0, s>>>(d, &cnt[1 + D], L.old[D], L.first[D], sval, w.rmask,
                                                              w.rmat, w.items, w.hashes, L.res[D]);
  else
    leaf_build_kernel<16><<<grid_for(L.cap[D]), 256, 0, s>>>(d, &cnt[1 + D], L.old[D], L.first[D], sval, w.rmask,
                                                               w.rmat, w.items, w.hashes, L.res[D]);
  if ((st = last_launch()) != VX_OK) return st;
  if ((st = intern(d, true, ItemSet{w.items, d.leaf_words, d.leaf_words}, w.hashes, &cnt[1 + D], L.cap[D], L.res[D],
                   w.fused_ws, s)) != VX_OK)
    return st;
  mark(g, 5, s);
  // Nodes, bottom-up.
  for (uint32_t dd = D; dd-- > 0;) {
    node_build_kernel<<<grid_for(L.cap[dd]), 256, 0, s>>>(d, &cnt[1 + dd], L.old[dd], L.first[dd], L.keys[dd + 1],
                                                           L.res[dd + 1], w.items, w.hashes, L.res[dd]);
    if ((st = last_launch()) != VX_OK) return st;
    if ((st = intern(d, false, ItemSet{w.items, kNodeStride, 0}, w.hashes, &cnt[1 + dd], L.cap[dd], L.res[dd],
                     w.fused_ws, s)) != VX_OK)
      return st;
  }
  mark(g, 6, s);
  publish_kernel<<<1, 1, 0, s>>>(d.ctl, &cnt[1], L.res[0]);
  st = last_launch();
  mark(g, 7, s);
  return st;
}

vx_status finish(vx_dag* g, const EditWs& w, hipStream_t s) {
  uint32_t status = 0;
  uint32_t counts[64];
  if (hipMemcpyAsync(&status, g->ctl + kCtlStatus, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return VX_DEVICE_ERROR;
  if (g->profile && hipMemcpyAsync(counts, w.counts, sizeof(counts), hipMemcpyDeviceToHost, s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  if (g->profile) {
    for (uint32_t i = 0; i < VX_DAG_STAGES; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, g->ev[i], g->ev[i + 1]) == hipSuccess) g->prof.ms[i] += ms;
    }
    g->prof.edits += 1;
    g->prof.records += counts[0];
    g->prof.leaves += counts[1 + g->depth];
    for (uint32_t d = 0; d < g->depth; ++d) g->prof.nodes += counts[1 + d];
    g->prof.sort_passes = (3 * g->depth + 7) / 8;
  }
  return status == kStatusOk ? VX_OK : VX_OUT_OF_CAPACITY;
}

__global__ void clear_status_kernel(uint32_t* ctl) { ctl[kCtlStatus] = kStatusOk; }

// Brush jobs: brick bounding boxes clipped to the world; empty ones dropped.
uint64_t make_jobs(const vx_brush* b, uint32_t n, uint32_t world, std::vector<BrushJob>* jobs) {
  uint64_t total = 0;
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
    total += uint64_t(cnt[0]) * cnt[1] * cnt[2];
    if (jobs) jobs->push_back(j);
  }
  return total;
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
  out->reserved = 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_voxel_edits_workspace_size(const vx_dag* g, uint32_t count, size_t* bytes) {
  if (!g || !bytes) return VX_INVALID_VALUE;
  *bytes = count ? carve(nullptr, g->depth, g->dev.leaf_words, count, 0, count).bytes : 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_apply_voxel_edits(vx_dag* g, const vx_voxel* edits, uint32_t count, uint32_t op,
                                              uint32_t material, void* ws, size_t ws_bytes, void* stream) {
  if (!g || (!edits && count) || op > 1) return VX_INVALID_VALUE;  // edit.cuh:28-37
  if (count == 0) return VX_OK;
  EditWs w = carve(ws, g->depth, g->dev.leaf_words, count, 0, count);
  if (!ws || ws_bytes < w.bytes) return VX_INVALID_VALUE;
  if (count >= (1u << 30)) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  vx_status st;
  if ((st = fill_u32(w.counts, 64, 0, s)) != VX_OK) return st;
  clear_status_kernel<<<1, 1, 0, s>>>(g->ctl);
  mark(g, 0, s);
  const uint32_t mat = op == 1 ? 0u : (g->materials ? (material & 0xffu) : 1u);
  if (op == 0 && g->materials && mat == 0) return VX_INVALID_VALUE;
  if ((st = scan::fused<scan::kSum>(count, VoxelLoad{edits, 4u << g->depth, mat},
                                    RecordStore{RecordOut{w.rkey, w.rval, w.rmask, w.rmat, &w.counts[0]}, count},
                                    w.fused_ws, s)) != VX_OK)
    return st;
  if ((st = rebuild(g, w, count, s)) != VX_OK) return st;
  return finish(g, w, s);
}

extern "C" vx_status vx_dag_brush_workspace_size(const vx_dag* g, const vx_brush* b, uint32_t n, size_t* bytes) {
  if (!g || !bytes || (!b && n)) return VX_INVALID_VALUE;
  std::vector<BrushJob> jobs;
  const uint64_t cand = make_jobs(b, n, 4u << g->depth, &jobs);
  *bytes = cand ? carve(nullptr, g->depth, g->dev.leaf_words, cand, uint32_t(jobs.size()), cand).bytes : 0;
  return VX_OK;
}

extern "C" vx_status vx_dag_apply_brushes(vx_dag* g, const vx_brush* b, uint32_t n, void* ws, size_t ws_bytes,
                                          void* stream) {
  if (!g || (!b && n)) return VX_INVALID_VALUE;
  std::vector<BrushJob> jobs;
  const uint64_t cand = make_jobs(b, n, 4u << g->depth, &jobs);
  if (cand == 0) return VX_OK;
  if (cand >= (1u << 30)) return VX_INVALID_VALUE;
  EditWs w = carve(ws, g->depth, g->dev.leaf_words, cand, uint32_t(jobs.size()), cand);
  if (!ws || ws_bytes < w.bytes) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  vx_status st;
  if (hipMemcpyAsync(w.jobs, jobs.data(), jobs.size() * sizeof(BrushJob), hipMemcpyHostToDevice, s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if ((st = fill_u32(w.counts, 64, 0, s)) != VX_OK) return st;
  clear_status_kernel<<<1, 1, 0, s>>>(g->ctl);
  mark(g, 0, s);
  if ((st = scan::fused<scan::kSum>(uint32_t(cand), BrushLoad{w.jobs, uint32_t(jobs.size()), g->materials},
                                    RecordStore{RecordOut{w.rkey, w.rval, w.rmask, w.rmat, &w.counts[0]}, uint32_t(cand)},
                                    w.fused_ws, s)) != VX_OK)
    return st;
  if ((st = rebuild(g, w, cand, s)) != VX_OK) return st;
  return finish(g, w, s);
}

extern "C" vx_status vx_dag_get_voxels(const vx_dag* g, const vx_voxel* q, uint32_t n, uint32_t* out, void* stream) {
  if (!g || (n && (!q || !out))) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  get_voxels_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(g->dev, q, n, out);
  return last_launch();
}

// Canonical export: depth-first from the root, children 0..7, ids at first
// visit, memo by (level, ref) — refs are content-unique per table, so the
// export depends only on the voxel/material field.
extern "C" vx_status vx_dag_export_canonical(const vx_dag* g, uint32_t* out, uint64_t cap, uint64_t* n_words) {
  if (!g || !n_words) return VX_INVALID_VALUE;
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  uint32_t ctl[kCtlWords];
  if (hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nu = std::min<uint64_t>(ctl[kCtlNodeUsed], g->dev.node_cap);
  const uint64_t lu = std::min<uint64_t>(ctl[kCtlLeafUsed], g->dev.leaf_cap);
  std::vector<uint32_t> np(nu), lp(lu);
  if (nu && hipMemcpy(np.data(), g->node_pool, nu * 4, hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  if (lu && hipMemcpy(lp.data(), g->leaf_pool, lu * 4, hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const uint32_t D = g->depth, LW = g->dev.leaf_words, root = ctl[kCtlRoot];
  std::unordered_map<uint64_t, uint32_t> node_id, leaf_id;
  std::vector<std::vector<uint32_t>> node_recs;
  std::vector<uint32_t> leaf_refs;
  struct Frame {
    uint32_t level, ref, id, c;
    uint32_t cm, fm;
    std::vector<uint32_t> ids;
  };
  auto leaf_of = [&](uint32_t ref) -> uint32_t {
    auto it = leaf_id.find(ref);
    if (it != leaf_id.end()) return it->second;
    const uint32_t id = uint32_t(leaf_refs.size());
    leaf_id.emplace(ref, id);
    leaf_refs.push_back(ref);
    return id;
  };
  uint32_t root_kind = root == kEmpty ? 0 : (root == kFull ? 1 : 2);
  if (root_kind == 2) {
    std::vector<Frame> stack;
    auto open = [&](uint32_t level, uint32_t ref) -> uint32_t {
      const uint32_t id = uint32_t(node_recs.size());
      node_id.emplace((uint64_t(level) << 32) | ref, id);
      node_recs.emplace_back();
      stack.push_back(Frame{level, ref, id, 0, 0, 0, {}});
      return id;
    };
    open(0, root);
    while (!stack.empty()) {
      Frame& f = stack.back();
      if (f.c == 8) {
        std::vector<uint32_t>& rec = node_recs[f.id];
        rec.push_back(f.level);
        rec.push_back(f.cm | (f.fm << 8));
        rec.insert(rec.end(), f.ids.begin(), f.ids.end());
        stack.pop_back();
        continue;
      }
      const uint32_t c = f.c++;
      if (f.ref >= nu) return VX_DEVICE_ERROR;
      const uint32_t h = np[f.ref], cm = h & 0xffu, fm = (h >> 8) & 0xffu, bit = 1u << c;
      if (!(cm & bit)) {
        if (fm & bit) f.fm |= bit;
        continue;
      }
      const uint32_t child = np[f.ref + 1 + __builtin_popcount(cm & (bit - 1))];
      f.cm |= bit;
      if (f.level + 1 == D) {
        if (child + LW > lu) return VX_DEVICE_ERROR;
        f.ids.push_back(leaf_of(child));
        continue;
      }
      auto it = node_id.find((uint64_t(f.level + 1) << 32) | child);
      if (it != node_id.end()) {
        f.ids.push_back(it->second);
        continue;
      }
      const uint32_t lvl = f.level + 1;
      const uint32_t id = uint32_t(node_recs.size());
      f.ids.push_back(id);
      open(lvl, child);  // invalidates f
    }
  }
  uint64_t total = 6;
  for (auto& r : node_recs) total += r.size();
  total += uint64_t(leaf_refs.size()) * LW;
  *n_words = total;
  if (!out || cap < total) return VX_OK;
  uint64_t k = 0;
  out[k++] = 0x56584443u;
  out[k++] = D;
  out[k++] = LW;
  out[k++] = root_kind;
  out[k++] = uint32_t(node_recs.size());
  out[k++] = uint32_t(leaf_refs.size());
  for (auto& r : node_recs)
    for (uint32_t v : r) out[k++] = v;
  for (uint32_t ref : leaf_refs)
    for (uint32_t i = 0; i < LW; ++i) out[k++] = lp[ref + i];
  return VX_OK;
}

extern "C" vx_status vx_dag_profile_enable(vx_dag* g, int32_t enable) {
  if (!g) return VX_INVALID_VALUE;
  if (enable && !g->profile) {
    for (auto& e : g->ev)
      if (!e && hipEventCreate(&e) != hipSuccess) return VX_DEVICE_ERROR;
  }
  g->profile = enable != 0;
  g->prof = vx_dag_profile{};
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
