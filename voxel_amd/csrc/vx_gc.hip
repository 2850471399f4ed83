// Garbage collection of a scene's interned tables on gfx950.
//
// Edits only append: an item that no longer hangs under the root keeps its
// pool words and its slot (the reference's GPU allocator never frees,
// cuda/hash_table/bump_allocator.cuh:118-121). The reference compacts only its
// CPU Hash DAG (svt/hash_dag.hpp:133-191): live refs per level by
// sort-unique, then leaves and nodes re-interned bottom-up through old->new
// ref maps. The same two sweeps run here, one level per launch group:
//   discovery (top-down): F[0] = {root}; F[l+1] = sort-unique of the child
//     refs of F[l] (a fused scan writes the refs, the onesweep sort orders
//     them, a second fused scan drops repeats);
//   rebuild (bottom-up): the items of F[D] (leaves) and then of F[l], with
//     every child ref replaced by its new ref (binary search in F[l+1]), are
//     interned by the edit's probe / commit / place kernels into fresh slot
//     tables and pools sized to the live words.
// The result is copied over the scene's own buffers at the end, so every
// device pointer into the scene (captured graphs, batchers) stays valid, and
// nothing of the scene changes before that copy: any failure leaves it as it
// was.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "vx_dag.hpp"
#include "vx_scan.hpp"
#include "vx_sort.hpp"

using namespace vx;
using namespace vx::dag;

namespace {

// Child refs of the live nodes of one level, in node order.
struct ExpandLoad {
  using Payload = uint32_t;
  const uint32_t* live;
  const uint32_t* pool;
  __device__ uint32_t operator()(uint32_t j, uint32_t& ref) const {
    ref = live[j];
    return __popc(pool[ref] & 0xffu);
  }
};

struct ExpandStore {
  const uint32_t* pool;
  uint32_t* out;
  uint32_t* n_out;
  __device__ void operator()(uint32_t j, const uint32_t& ref, uint32_t v, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    for (uint32_t k = 0; k < v; ++k) out[excl + k] = pool[ref + 1 + k];
    if (j == n - 1) *n_out = incl;
  }
};

// First occurrence of every value of a sorted array.
struct UniqueLoad {
  using Payload = uint32_t;
  const uint32_t* sorted;
  __device__ uint32_t operator()(uint32_t j, uint32_t& v) const {
    v = sorted[j];
    return j == 0 || sorted[j - 1] != v;
  }
};

struct UniqueStore {
  uint32_t* out;
  uint32_t* n_out;
  __device__ void operator()(uint32_t j, const uint32_t& v, uint32_t flag, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    if (flag) out[excl] = v;
    if (j == n - 1) *n_out = incl;
  }
};

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

template <uint32_t LW>
__global__ void gc_leaf_kernel(const uint32_t* __restrict__ live, const uint32_t* n_ptr,
                               const uint32_t* __restrict__ pool, uint32_t* __restrict__ items,
                               uint64_t* __restrict__ hashes, uint64_t* __restrict__ res) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t* src = pool + live[j];
    uint32_t w[LW];
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) w[k] = src[k];
    uint32_t* it = items + size_t(j) * LW;
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) it[k] = w[k];
    hashes[j] = hash_words(w, LW);
    res[j] = res_make(kResPending, 0);
  }
}

// Index of `ref` in the sorted, repeat-free list[0..n) (present by
// construction: every child of a live node is live).
__device__ __forceinline__ uint32_t index_of(const uint32_t* __restrict__ list, uint32_t n, uint32_t ref) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (list[mid] <= ref) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Node items of one level with child refs mapped to the rebuilt tables. The
// header (child mask | fill mask) is kept as stored.
__global__ void gc_node_kernel(const uint32_t* __restrict__ live, const uint32_t* n_ptr,
                               const uint32_t* __restrict__ pool, const uint32_t* __restrict__ child_live,
                               const uint32_t* child_n_ptr, const uint32_t* __restrict__ child_map,
                               uint32_t* __restrict__ items, uint64_t* __restrict__ hashes,
                               uint64_t* __restrict__ res) {
  const uint32_t n = *n_ptr, cn = *child_n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t ref = live[j];
    const uint32_t h = pool[ref];
    const uint32_t w = 1 + __popc(h & 0xffu);
    uint32_t it[kNodeStride];
    it[0] = h;
    for (uint32_t k = 1; k < w; ++k) it[k] = child_map[index_of(child_live, cn, pool[ref + k])];
    uint32_t* dst = items + size_t(j) * kNodeStride;
    for (uint32_t k = 0; k < w; ++k) dst[k] = it[k];
    hashes[j] = hash_words(it, w);
    res[j] = res_make(kResPending, 0);
  }
}

// New ref of every live item of a level (duplicates followed).
__global__ void gc_map_kernel(const uint32_t* n_ptr, const uint64_t* __restrict__ res, uint32_t* __restrict__ map) {
  const uint32_t n = *n_ptr;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    map[j] = resolve(res, j);
}

__global__ void gc_root_kernel(uint32_t* ctl, const uint32_t* map0) { ctl[kCtlRoot] = map0[0]; }

__global__ void gc_ctl_init_kernel(uint32_t* ctl, uint32_t root) {
  for (uint32_t i = threadIdx.x; i < kCtlWords; i += blockDim.x) ctl[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl[kCtlNodeUsed] = kPoolStart;
    ctl[kCtlLeafUsed] = kPoolStart;
    ctl[kCtlRoot] = root;
  }
}

uint32_t bits_for(uint64_t cap) {
  uint32_t b = 1;
  while (b < 32 && (uint64_t(1) << b) < cap) ++b;
  return b;
}

struct Level {
  DevBuf live;  // sorted, repeat-free refs of the level's live items
  uint64_t n = 0;
  DevBuf map;   // new ref per live item
};

}  // namespace

extern "C" vx_status vx_dag_collect_garbage(vx_dag* g, vx_gc_report* report, void* stream) {
  if (!g) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  uint32_t ctl[kCtlWords];
  if (hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const Dev& od = g->dev;
  const uint32_t D = g->depth, LW = od.leaf_words, root = ctl[kCtlRoot];
  vx_gc_report rep{};
  rep.node_words_before = ctl[kCtlNodeUsed];
  rep.leaf_words_before = ctl[kCtlLeafUsed];
  vx_status st;

  DevBuf cnt_buf;  // [l] live count of level l, [32 + l] raw child count of level l
  if ((st = dev_alloc(cnt_buf, 64 * 4)) != VX_OK) return st;
  uint32_t* cnt = cnt_buf.get<uint32_t>();
  if ((st = fill_u32(cnt, 64, 0, s)) != VX_OK) return st;

  // 1. discovery, top-down. Levels 0..D-1 hold node refs, level D leaf refs.
  std::vector<Level> lv(D + 1);
  uint64_t node_bound = kPoolStart;
  if (root > kFull) {
    if ((st = dev_alloc(lv[0].live, 4)) != VX_OK) return st;
    set_u32_kernel<<<1, 1, 0, s>>>(lv[0].live.get<uint32_t>(), root);
    set_u32_kernel<<<1, 1, 0, s>>>(cnt, 1);
    if ((st = last_launch()) != VX_OK) return st;
    lv[0].n = 1;
    const uint32_t node_bits = bits_for(od.node_cap), leaf_bits = bits_for(od.leaf_cap);
    for (uint32_t l = 0; l < D; ++l) {
      const uint64_t n = lv[l].n;
      if (n == 0) break;  // no live items below either
      DevBuf raw, fws;
      if ((st = dev_alloc(raw, n * 8 * 4)) != VX_OK) return st;
      if ((st = dev_alloc(fws, scan::fused_dev_workspace_bytes(n))) != VX_OK) return st;
      if ((st = scan::fused_dev<scan::kSum>(cnt + l, n, ExpandLoad{lv[l].live.get<uint32_t>(), od.node_pool},
                                            ExpandStore{od.node_pool, raw.get<uint32_t>(), cnt + 32 + l + 1}, fws.p,
                                            s)) != VX_OK)
        return st;
      const uint32_t m = read_u32(cnt + 32 + l + 1, s);
      node_bound += n + m;  // every live node item: header + one ref per child
      if (m == 0) break;    // only EMPTY / FULL children: nothing lives below
      const uint32_t key_bits = l + 1 == D ? leaf_bits : node_bits;
      const size_t sws_bytes = sort::workspace_bytes(m, 4, key_bits, false);
      DevBuf sws;
      if ((st = dev_alloc(sws, sws_bytes)) != VX_OK) return st;
      if ((st = sort::sort_u32(raw.get<uint32_t>(), nullptr, m, key_bits, sws.p, sws_bytes, s)) != VX_OK) return st;
      Level& nx = lv[l + 1];
      if ((st = dev_alloc(nx.live, uint64_t(m) * 4)) != VX_OK) return st;
      DevBuf uws;
      if ((st = dev_alloc(uws, scan::fused_dev_workspace_bytes(m))) != VX_OK) return st;
      if ((st = scan::fused_dev<scan::kSum>(cnt + 32 + l + 1, m, UniqueLoad{raw.get<uint32_t>()},
                                            UniqueStore{nx.live.get<uint32_t>(), cnt + l + 1}, uws.p, s)) != VX_OK)
        return st;
      nx.n = read_u32(cnt + l + 1, s);
    }
  }
  const uint64_t leaf_bound = kPoolStart + lv[D].n * LW;
  for (uint32_t l = 0; l < D; ++l) rep.live_nodes += lv[l].n;
  rep.live_leaves = lv[D].n;

  // 2. fresh tables (same slot counts, so the load factor only drops) and
  // pools for the live words.
  DevBuf nslots, lslots, npool, lpool, nctl;
  const uint64_t n_slots = od.node_mask + 1, l_slots = od.leaf_mask + 1;
  if ((st = dev_alloc(nslots, n_slots * 8)) != VX_OK) return st;
  if ((st = dev_alloc(lslots, l_slots * 8)) != VX_OK) return st;
  if ((st = dev_alloc(npool, node_bound * 4)) != VX_OK) return st;
  if ((st = dev_alloc(lpool, leaf_bound * 4)) != VX_OK) return st;
  if ((st = dev_alloc(nctl, kCtlWords * 4)) != VX_OK) return st;
  if ((st = fill_u32(nslots.get<uint32_t>(), n_slots * 2, 0, s)) != VX_OK) return st;
  if ((st = fill_u32(lslots.get<uint32_t>(), l_slots * 2, 0, s)) != VX_OK) return st;
  if ((st = fill_u32(npool.get<uint32_t>(), kPoolStart, 0, s)) != VX_OK) return st;
  if ((st = fill_u32(lpool.get<uint32_t>(), kPoolStart, 0, s)) != VX_OK) return st;
  gc_ctl_init_kernel<<<1, 64, 0, s>>>(nctl.get<uint32_t>(), root <= kFull ? root : kEmpty);
  if ((st = last_launch()) != VX_OK) return st;
  Dev nd = od;
  nd.node_pool = npool.get<uint32_t>();
  nd.leaf_pool = lpool.get<uint32_t>();
  nd.node_slots = nslots.get<uint64_t>();
  nd.leaf_slots = lslots.get<uint64_t>();
  nd.node_cap = node_bound;
  nd.leaf_cap = leaf_bound;
  nd.ctl = nctl.get<uint32_t>();
  nd.status = nd.ctl + kCtlStatus;
  nd.base = nd.ctl + kCtlNodeBase;

  // 3. rebuild, bottom-up.
  if (root > kFull) {
    uint64_t max_n = 1;
    for (uint32_t l = 0; l <= D; ++l) max_n = std::max(max_n, lv[l].n);
    DevBuf items, hashes, res, fws;
    if ((st = dev_alloc(items, max_n * std::max<uint32_t>(LW, kNodeStride) * 4)) != VX_OK) return st;
    if ((st = dev_alloc(hashes, max_n * 8)) != VX_OK) return st;
    if ((st = dev_alloc(res, max_n * 8)) != VX_OK) return st;
    if ((st = dev_alloc(fws, scan::fused_dev_workspace_bytes(max_n))) != VX_OK) return st;
    for (uint32_t l = D + 1; l-- > 0;) {
      Level& L = lv[l];
      const uint64_t n = L.n;
      if ((st = dev_alloc(L.map, n * 4)) != VX_OK) return st;
      if (n == 0) continue;  // cnt[l] is 0; the level above has no child refs
      if (l == D) {
        if (LW == 2)
          gc_leaf_kernel<2><<<grid_for(n), 256, 0, s>>>(L.live.get<uint32_t>(), cnt + l, od.leaf_pool,
                                                        items.get<uint32_t>(), hashes.get<uint64_t>(),
                                                        res.get<uint64_t>());
        else
          gc_leaf_kernel<16><<<grid_for(n), 256, 0, s>>>(L.live.get<uint32_t>(), cnt + l, od.leaf_pool,
                                                         items.get<uint32_t>(), hashes.get<uint64_t>(),
                                                         res.get<uint64_t>());
        if ((st = last_launch()) != VX_OK) return st;
        if ((st = intern(nd, true, ItemSet{items.get<uint32_t>(), LW, LW}, hashes.get<uint64_t>(), cnt + l, n,
                         res.get<uint64_t>(), fws.p, s)) != VX_OK)
          return st;
      } else {
        const Level& C = lv[l + 1];
        gc_node_kernel<<<grid_for(n), 256, 0, s>>>(L.live.get<uint32_t>(), cnt + l, od.node_pool,
                                                   C.live.get<uint32_t>(), cnt + l + 1, C.map.get<uint32_t>(),
                                                   items.get<uint32_t>(), hashes.get<uint64_t>(),
                                                   res.get<uint64_t>());
        if ((st = last_launch()) != VX_OK) return st;
        if ((st = intern(nd, false, ItemSet{items.get<uint32_t>(), kNodeStride, 0}, hashes.get<uint64_t>(), cnt + l,
                         n, res.get<uint64_t>(), fws.p, s)) != VX_OK)
          return st;
      }
      gc_map_kernel<<<grid_for(n), 256, 0, s>>>(cnt + l, res.get<uint64_t>(), L.map.get<uint32_t>());
      if ((st = last_launch()) != VX_OK) return st;
    }
    gc_root_kernel<<<1, 1, 0, s>>>(nd.ctl, lv[0].map.get<uint32_t>());
    if ((st = last_launch()) != VX_OK) return st;
  }

  uint32_t nctl_h[kCtlWords];
  if (hipMemcpyAsync(nctl_h, nd.ctl, sizeof(nctl_h), hipMemcpyDeviceToHost, s) != hipSuccess) return VX_DEVICE_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nu = nctl_h[kCtlNodeUsed], lu = nctl_h[kCtlLeafUsed];
  // A node item shared by two levels is rebuilt once per level (its child
  // refs map through different tables), so the live words can exceed what
  // the scene held; the scene's capacity still bounds them.
  if (nctl_h[kCtlStatus] != kStatusOk || nu > od.node_cap || lu > od.leaf_cap) return VX_OUT_OF_CAPACITY;

  // 4. copy over the scene's buffers (the only step that changes the scene).
  bool ok = hipMemcpyAsync(g->node_slots, nd.node_slots, n_slots * 8, hipMemcpyDeviceToDevice, s) == hipSuccess &&
            hipMemcpyAsync(g->leaf_slots, nd.leaf_slots, l_slots * 8, hipMemcpyDeviceToDevice, s) == hipSuccess &&
            hipMemcpyAsync(g->node_pool, nd.node_pool, nu * 4, hipMemcpyDeviceToDevice, s) == hipSuccess &&
            hipMemcpyAsync(g->leaf_pool, nd.leaf_pool, lu * 4, hipMemcpyDeviceToDevice, s) == hipSuccess &&
            hipMemcpyAsync(g->ctl, nd.ctl, kCtlWords * 4, hipMemcpyDeviceToDevice, s) == hipSuccess;
  if (!ok || hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  rep.node_words_after = nu;
  rep.leaf_words_after = lu;
  if (report) *report = rep;
  return VX_OK;
}

// ---- growth ----------------------------------------------------------------
// Every live slot of a table (a pool offset, not empty, dead or tagged)
// re-inserted into a larger table at its item's home position: the hash is
// recomputed from the item in the pool (its fingerprint must equal the
// slot's), the slot value itself is unchanged because the pool offsets are.
// Linear probing from the home position keeps every slot reachable from it.
__global__ void rehash_kernel(const uint64_t* __restrict__ old_slots, uint64_t n_old, const uint32_t* __restrict__ pool,
                              uint64_t pool_used, uint32_t leaf_words, uint64_t* new_slots, uint64_t new_mask,
                              uint32_t* status) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n_old; i += stride) {
    const uint64_t sv = old_slots[i];
    const uint32_t v = uint32_t(sv);
    if (v < kPoolStart || (v & kTag)) continue;
    if (v >= pool_used) {  // no item there: the header read below would be out of the used pool
      atomicCAS(status, 0u, 1u);
      continue;
    }
    const uint32_t w = leaf_words ? leaf_words : 1u + uint32_t(__popc(pool[v] & 0xffu));
    if (uint64_t(v) + w > pool_used) {
      atomicCAS(status, 0u, 1u);
      continue;
    }
    const uint64_t h = hash_words(pool + v, w);
    if (uint32_t(h >> 32) != uint32_t(sv >> 32)) {
      atomicCAS(status, 0u, 1u);
      continue;
    }
    uint64_t idx = h & new_mask;
    bool placed = false;
    for (uint64_t p = 0; p <= new_mask && p < kMaxProbe; ++p) {
      const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&new_slots[idx]), 0ull,
                                                static_cast<unsigned long long>(sv));
      if (prev == 0ull) {
        placed = true;
        break;
      }
      idx = (idx + 1) & new_mask;
    }
    if (!placed) atomicCAS(status, 0u, 1u);
  }
}

extern "C" vx_status vx_dag_grow(vx_dag* g, const vx_dag_capacity* cap, void* stream) {
  if (!g || !cap) return VX_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Targets: never below the current sizes; slot counts rounded up to powers of two.
  vx_dag_capacity c = g->cap;
  const uint64_t nw = std::max<uint64_t>(c.node_words, std::min<uint64_t>(cap->node_words, kTag));
  const uint64_t lw = std::max<uint64_t>(c.leaf_words, std::min<uint64_t>(cap->leaf_words, kTag));
  const uint64_t ns = std::max<uint64_t>(c.node_slots, cap->node_slots ? pow2_at_least(cap->node_slots) : 0);
  const uint64_t ls = std::max<uint64_t>(c.leaf_slots, cap->leaf_slots ? pow2_at_least(cap->leaf_slots) : 0);
  if (ns > (1ull << 32) || ls > (1ull << 32)) return VX_INVALID_VALUE;
  const bool grow_np = nw > c.node_words, grow_lp = lw > c.leaf_words;
  const bool grow_ns = ns > c.node_slots, grow_ls = ls > c.leaf_slots;
  if (!grow_np && !grow_lp && !grow_ns && !grow_ls) return VX_OK;
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  uint32_t ctl[kCtlWords];
  if (hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nu = std::min<uint64_t>(ctl[kCtlNodeUsed], c.node_words);
  const uint64_t lu = std::min<uint64_t>(ctl[kCtlLeafUsed], c.leaf_words);
  // New buffers first; the scene is untouched until every step succeeded.
  DevBuf npool, lpool, nslots, lslots, status;
  vx_status st;
  if (grow_np && (st = dev_alloc(npool, nw * 4)) != VX_OK) return st;
  if (grow_lp && (st = dev_alloc(lpool, lw * 4)) != VX_OK) return st;
  if (grow_ns && (st = dev_alloc(nslots, ns * 8)) != VX_OK) return st;
  if (grow_ls && (st = dev_alloc(lslots, ls * 8)) != VX_OK) return st;
  if ((st = dev_alloc(status, 4)) != VX_OK) return st;
  if ((st = fill_u32(status.get<uint32_t>(), 1, 0, s)) != VX_OK) return st;
  if (grow_np && hipMemcpyAsync(npool.p, g->node_pool, nu * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (grow_lp && hipMemcpyAsync(lpool.p, g->leaf_pool, lu * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return VX_DEVICE_ERROR;
  auto rehash = [&](DevBuf& to, uint64_t n_new, const uint64_t* from, uint64_t n_old, const uint32_t* pool,
                    uint64_t used, uint32_t leaf_words) -> vx_status {
    vx_status r = fill_u32(to.get<uint32_t>(), n_new * 2, 0, s);
    if (r != VX_OK) return r;
    const uint32_t blocks = uint32_t(std::min<uint64_t>(std::max<uint64_t>(1, (n_old + 255) / 256), 8192));
    rehash_kernel<<<blocks, 256, 0, s>>>(from, n_old, pool, used, leaf_words, to.get<uint64_t>(), n_new - 1,
                                         status.get<uint32_t>());
    return last_launch();
  };
  if (grow_ns && (st = rehash(nslots, ns, g->node_slots, c.node_slots, g->node_pool, nu, 0)) != VX_OK) return st;
  if (grow_ls && (st = rehash(lslots, ls, g->leaf_slots, c.leaf_slots, g->leaf_pool, lu, g->dev.leaf_words)) != VX_OK)
    return st;
  uint32_t bad = 0;
  if (hipMemcpyAsync(&bad, status.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (bad) return VX_DEVICE_ERROR;  // a slot whose item does not hash to its fingerprint: the table is not intact
  // Swap in the new buffers.
  auto swap_in = [](DevBuf& b, auto*& field) {
    using T = std::remove_reference_t<decltype(*field)>;
    (void)hipFree(field);
    field = b.get<T>();
    b.p = nullptr;
  };
  if (grow_np) swap_in(npool, g->node_pool);
  if (grow_lp) swap_in(lpool, g->leaf_pool);
  if (grow_ns) swap_in(nslots, g->node_slots);
  if (grow_ls) swap_in(lslots, g->leaf_slots);
  g->cap = vx_dag_capacity{nw, lw, ns, ls};
  Dev& d = g->dev;
  d.node_pool = g->node_pool;
  d.leaf_pool = g->leaf_pool;
  d.node_slots = g->node_slots;
  d.leaf_slots = g->leaf_slots;
  d.node_mask = ns - 1;
  d.leaf_mask = ls - 1;
  d.node_cap = nw;
  d.leaf_cap = lw;
  ++g->generation;
  return VX_OK;
}

extern "C" vx_status vx_dag_read_pools(const vx_dag* g, uint32_t* node_out, uint64_t node_cap, uint64_t* node_n,
                                       uint32_t* leaf_out, uint64_t leaf_cap, uint64_t* leaf_n) {
  if (!g || !node_n || !leaf_n) return VX_INVALID_VALUE;
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  uint32_t ctl[kCtlWords];
  if (hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nu = std::min<uint64_t>(ctl[kCtlNodeUsed], g->dev.node_cap);
  const uint64_t lu = std::min<uint64_t>(ctl[kCtlLeafUsed], g->dev.leaf_cap);
  *node_n = nu;
  *leaf_n = lu;
  if (node_out && node_cap >= nu && hipMemcpy(node_out, g->node_pool, nu * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return VX_DEVICE_ERROR;
  if (leaf_out && leaf_cap >= lu && hipMemcpy(leaf_out, g->leaf_pool, lu * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return VX_DEVICE_ERROR;
  return VX_OK;
}
