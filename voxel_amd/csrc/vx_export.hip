// Canonical depth-first export of a scene, computed on gfx950.
//
// The export format (vx.h, vx_dag_export_canonical) numbers the internal
// nodes (level, ref) and the leaves in the order a depth-first pre-order walk
// from the root, children 0..7, first reaches them, each distinct (level, ref)
// once. That order needs no walk: a memoised pre-order walk first reaches a
// node through its smallest root path (paths of one level compared as numbers
// of 3 bits per level), because any later path into a repeated subtree has an
// earlier twin through the subtree's first occurrence. So, level by level
// from the root:
//   expand   a fused scan over the level's nodes (ordered by smallest path)
//            writes every table-backed child ref with the path parent·slot,
//            which leaves the children in ascending path order;
//   group    a stable onesweep sort of (child ref, emitted index) puts each
//            ref's smallest path first in its group; a fused scan over the
//            emitted entries keeps those firsts, still in path order — the
//            next level's node list — and a max-scan over the sorted groups
//            gives every emitted entry the index of its node;
// then one sort of all nodes by (path padded to the deepest node level,
// level) is the walk's first-visit order (an ancestor's padded path equals
// its descendant's only when the descendant extends it by zeros, and the
// lower level then comes first), and leaves are numbered by their own level's
// path order. A last fused scan over the node widths in that order writes the
// records [level, child mask | fill mask << 8, child ids], and one kernel the
// leaf payloads.
#include <vector>

#include "vx_dag.hpp"
#include "vx_scan.hpp"
#include "vx_sort.hpp"

using namespace vx;
using namespace vx::dag;

namespace {

constexpr uint32_t kMagic = 0x56584443u;
constexpr uint32_t kHeaderWords = 6;
constexpr uint32_t kMaxLv = 16;

struct RefHdr {
  uint32_t ref, hdr;
};

// Children of one level's nodes: every item is checked against the used
// words of its pool (an out-of-range ref sets *err and counts as childless).
struct ExpandLoad {
  using Payload = RefHdr;
  const uint32_t* refs;
  const uint32_t* pool;
  uint64_t used;
  uint32_t* err;
  __device__ uint32_t operator()(uint32_t j, RefHdr& p) const {
    p.ref = refs[j];
    p.hdr = 0;
    if (p.ref < kPoolStart || p.ref >= used) {
      *err = 1u;
      return 0;
    }
    const uint32_t h = pool[p.ref];
    const uint32_t w = __popc(h & 0xffu);
    if (uint64_t(p.ref) + 1 + w > used) {
      *err = 1u;
      return 0;
    }
    p.hdr = h;
    return w;
  }
};

struct ExpandStore {
  const uint64_t* paths;
  const uint32_t* pool;
  uint32_t* ekey;
  uint32_t* eref;
  uint64_t* epath;
  uint32_t* child_base;
  uint32_t* n_out;
  __device__ void operator()(uint32_t j, const RefHdr& p, uint32_t v, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    child_base[j] = excl;
    const uint64_t up = paths[j] << 3;
    uint32_t cm = p.hdr & 0xffu;
    for (uint32_t t = 0; t < v; ++t) {
      const uint32_t slot = __builtin_ctz(cm);
      cm &= cm - 1;
      const uint32_t c = pool[p.ref + 1 + t];
      ekey[excl + t] = c;
      eref[excl + t] = c;
      epath[excl + t] = up | slot;
    }
    if (j == n - 1) *n_out = incl;
  }
};

// First entry of each group of the ref-sorted emitted entries.
__global__ void head_kernel(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval, uint32_t m,
                            uint32_t* __restrict__ is_first) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
    is_first[sval[k]] = (k == 0 || skey[k - 1] != skey[k]) ? 1u : 0u;
}

// The firsts, in emitted (= path) order: the next level's nodes.
struct FirstLoad {
  using Payload = scan::NoPayload;
  const uint32_t* is_first;
  __device__ uint32_t operator()(uint32_t e, scan::NoPayload&) const { return is_first[e]; }
};

struct FirstStore {
  const uint32_t* eref;
  const uint64_t* epath;
  uint32_t* refs;
  uint64_t* paths;
  uint32_t* node_of;
  uint32_t* n_out;
  __device__ void operator()(uint32_t e, const scan::NoPayload&, uint32_t v, uint32_t excl, uint32_t incl,
                             uint32_t n) const {
    if (v) {
      refs[excl] = eref[e];
      paths[excl] = epath[e];
      node_of[e] = excl;
    }
    if (e == n - 1) *n_out = incl;
  }
};

// Every other entry of a group takes its first's node index: a max-scan of
// the group-start positions gives each sorted entry its group's start.
struct GroupLoad {
  using Payload = uint32_t;
  const uint32_t* skey;
  __device__ uint32_t operator()(uint32_t k, uint32_t& head) const {
    head = (k == 0 || skey[k - 1] != skey[k]) ? 1u : 0u;
    return head ? k : 0u;
  }
};

struct GroupStore {
  const uint32_t* sval;
  uint32_t* node_of;
  __device__ void operator()(uint32_t k, const uint32_t& head, uint32_t, uint32_t, uint32_t incl, uint32_t) const {
    if (!head) node_of[sval[k]] = node_of[sval[incl]];
  }
};

// Pre-order keys: path padded to the deepest node level, level below it.
__global__ void order_keys_kernel(const uint64_t* __restrict__ paths, uint32_t n, uint32_t pad_bits,
                                  uint32_t level, uint32_t base, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    keys[base + j] = ((paths[j] << pad_bits) << 4) | level;
    vals[base + j] = base + j;
  }
}

__global__ void rank_kernel(const uint32_t* __restrict__ sval, uint32_t n, uint32_t* __restrict__ canon) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) canon[sval[k]] = k;
}

struct Levels {
  const uint32_t* refs[kMaxLv];
  const uint32_t* child_base[kMaxLv];  // levels < D: first emitted index of each node's children
  const uint32_t* node_of[kMaxLv];     // levels >= 1: node index of each emitted entry
  uint32_t off[kMaxLv + 1];            // first global node index of each node level
};

struct LJH {
  uint32_t level, j, hdr;
};

struct RecordLoad {
  using Payload = LJH;
  const uint64_t* skey;
  const uint32_t* sval;
  const uint32_t* pool;
  Levels lv;
  __device__ uint32_t operator()(uint32_t k, LJH& p) const {
    p.level = uint32_t(skey[k] & 15u);
    p.j = sval[k] - lv.off[p.level];
    p.hdr = pool[lv.refs[p.level][p.j]];
    return 2 + __popc(p.hdr & 0xffu);
  }
};

struct RecordStore {
  const uint32_t* canon;
  uint32_t* out;
  uint32_t depth;
  Levels lv;
  __device__ void operator()(uint32_t, const LJH& p, uint32_t v, uint32_t excl, uint32_t, uint32_t) const {
    uint32_t* o = out + kHeaderWords + excl;
    const uint32_t cm = p.hdr & 0xffu, fm = (p.hdr >> 8) & 0xffu;
    o[0] = p.level;
    o[1] = cm | ((fm & ~cm) << 8);
    const uint32_t nl = p.level + 1;
    const uint32_t base = lv.child_base[p.level][p.j];
    const uint32_t* nof = lv.node_of[nl];
    for (uint32_t t = 0; t + 2 < v; ++t) {
      const uint32_t c = nof[base + t];
      o[2 + t] = nl == depth ? c : canon[lv.off[nl] + c];
    }
  }
};

template <uint32_t LW>
__global__ void leaves_kernel(const uint32_t* __restrict__ refs, uint32_t n, const uint32_t* __restrict__ pool,
                              uint64_t used, uint32_t* __restrict__ out, uint32_t* err) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t r = refs[j];
    uint32_t* o = out + size_t(j) * LW;
    if (r < kPoolStart || uint64_t(r) + LW > used) {
      *err = 1u;
      for (uint32_t k = 0; k < LW; ++k) o[k] = 0;
      continue;
    }
#pragma unroll
    for (uint32_t k = 0; k < LW; ++k) o[k] = pool[r + k];
  }
}

__global__ void header_kernel(uint32_t* out, uint32_t d, uint32_t lw, uint32_t root_kind, uint32_t nodes,
                              uint32_t leaves) {
  if (threadIdx.x == 0) {
    out[0] = kMagic;
    out[1] = d;
    out[2] = lw;
    out[3] = root_kind;
    out[4] = nodes;
    out[5] = leaves;
  }
}

__global__ void set_root_kernel(uint32_t* refs, uint64_t* paths, uint32_t root) {
  refs[0] = root;
  paths[0] = 0;
}

uint32_t bits_for(uint64_t v) {
  uint32_t b = 1;
  while (b < 32 && (uint64_t(1) << b) <= v) ++b;
  return b;
}

struct Lv {
  DevBuf refs, paths, child_base, node_of;
  uint64_t n = 0;  // nodes (leaves at level D)
  uint64_t m = 0;  // emitted entries that produced this level (its parents' table-backed children)
};

// out_dev == nullptr: the caller wants only the size (*n_words).
// alloc_out: allocate the output here (host entry point) into *owned.
vx_status export_gpu(const vx_dag* g, uint32_t* out_dev, uint64_t cap, uint64_t* n_words, DevBuf* owned,
                     uint32_t** owned_ptr, hipStream_t s) {
  uint32_t ctl[kCtlWords];
  if (hipMemcpyAsync(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost, s) != hipSuccess) return VX_DEVICE_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return VX_DEVICE_ERROR;
  const uint64_t nu = std::min<uint64_t>(ctl[kCtlNodeUsed], g->dev.node_cap);
  const uint64_t lu = std::min<uint64_t>(ctl[kCtlLeafUsed], g->dev.leaf_cap);
  const uint32_t D = g->depth, LW = g->dev.leaf_words, root = ctl[kCtlRoot];
  if (D == 0 || D >= kMaxLv) return VX_INVALID_VALUE;
  const uint32_t root_kind = root == kEmpty ? 0 : (root == kFull ? 1 : 2);
  vx_status st;
  DevBuf cnt_buf;  // [0] error flag, [1 + l] node count of level l
  if ((st = dev_alloc(cnt_buf, 64 * 4)) != VX_OK) return st;
  uint32_t* cnt = cnt_buf.get<uint32_t>();
  if ((st = fill_u32(cnt, 64, 0, s)) != VX_OK) return st;
  uint32_t* err = cnt;

  std::vector<Lv> lv(D + 1);
  if (root_kind == 2) {
    if ((st = dev_alloc(lv[0].refs, 4)) != VX_OK) return st;
    if ((st = dev_alloc(lv[0].paths, 8)) != VX_OK) return st;
    set_root_kernel<<<1, 1, 0, s>>>(lv[0].refs.get<uint32_t>(), lv[0].paths.get<uint64_t>(), root);
    if ((st = last_launch()) != VX_OK) return st;
    lv[0].n = 1;
    for (uint32_t l = 0; l < D; ++l) {
      Lv& L = lv[l];
      Lv& X = lv[l + 1];
      const uint64_t n = L.n;
      if ((st = dev_alloc(L.child_base, n * 4)) != VX_OK) return st;
      if (n == 0) break;
      DevBuf ekey, eref, epath, fws;
      const uint64_t m_ub = n * 8;
      if ((st = dev_alloc(ekey, m_ub * 4)) != VX_OK) return st;
      if ((st = dev_alloc(eref, m_ub * 4)) != VX_OK) return st;
      if ((st = dev_alloc(epath, m_ub * 8)) != VX_OK) return st;
      if ((st = dev_alloc(fws, scan::fused_workspace_bytes(m_ub))) != VX_OK) return st;
      if ((st = scan::fused<scan::kSum>(uint32_t(n), ExpandLoad{L.refs.get<uint32_t>(), g->node_pool, nu, err},
                                        ExpandStore{L.paths.get<uint64_t>(), g->node_pool, ekey.get<uint32_t>(),
                                                    eref.get<uint32_t>(), epath.get<uint64_t>(),
                                                    L.child_base.get<uint32_t>(), cnt + 32},
                                        fws.p, s)) != VX_OK)
        return st;
      const uint64_t m = read_u32(cnt + 32, s);
      X.m = m;
      if (m == 0) break;
      // Group by ref (stable: each group's smallest path first).
      DevBuf sval, sws;
      if ((st = dev_alloc(sval, m * 4)) != VX_OK) return st;
      if ((st = vx_iota_u32(sval.get<uint32_t>(), uint32_t(m), s)) != VX_OK) return st;
      const uint32_t key_bits = bits_for(l + 1 == D ? lu : nu);
      const size_t sws_bytes = sort::workspace_bytes(uint32_t(m), 4, key_bits, true);
      if ((st = dev_alloc(sws, sws_bytes)) != VX_OK) return st;
      if ((st = sort::sort_u32(ekey.get<uint32_t>(), sval.get<uint32_t>(), uint32_t(m), key_bits, sws.p, sws_bytes,
                               s)) != VX_OK)
        return st;
      DevBuf is_first;
      if ((st = dev_alloc(is_first, m * 4)) != VX_OK) return st;
      head_kernel<<<grid_for(m), 256, 0, s>>>(ekey.get<uint32_t>(), sval.get<uint32_t>(), uint32_t(m),
                                              is_first.get<uint32_t>());
      if ((st = last_launch()) != VX_OK) return st;
      if ((st = dev_alloc(X.refs, m * 4)) != VX_OK) return st;
      if ((st = dev_alloc(X.paths, m * 8)) != VX_OK) return st;
      if ((st = dev_alloc(X.node_of, m * 4)) != VX_OK) return st;
      if ((st = scan::fused<scan::kSum>(uint32_t(m), FirstLoad{is_first.get<uint32_t>()},
                                        FirstStore{eref.get<uint32_t>(), epath.get<uint64_t>(),
                                                   X.refs.get<uint32_t>(), X.paths.get<uint64_t>(),
                                                   X.node_of.get<uint32_t>(), cnt + 1 + l + 1},
                                        fws.p, s)) != VX_OK)
        return st;
      if ((st = scan::fused<scan::kMax>(uint32_t(m), GroupLoad{ekey.get<uint32_t>()},
                                        GroupStore{sval.get<uint32_t>(), X.node_of.get<uint32_t>()}, fws.p, s)) !=
          VX_OK)
        return st;
      X.n = read_u32(cnt + 1 + l + 1, s);
    }
  }
  if (read_u32(err, s)) return VX_DEVICE_ERROR;

  uint64_t nodes = 0, node_words = 0;
  for (uint32_t l = 0; l < D; ++l) {
    nodes += lv[l].n;
    node_words += 2 * lv[l].n + lv[l + 1].m;
  }
  const uint64_t leaves = lv[D].n;
  const uint64_t total = kHeaderWords + node_words + leaves * LW;
  *n_words = total;
  if (nodes >= (1u << 30) || total > 0xFFFFFFFFull) return VX_UNSUPPORTED;
  if (owned) {
    if ((st = dev_alloc(*owned, total * 4)) != VX_OK) return st;
    out_dev = owned->get<uint32_t>();
    *owned_ptr = out_dev;
  } else if (!out_dev || cap < total) {
    return VX_OK;
  }

  header_kernel<<<1, 64, 0, s>>>(out_dev, D, LW, root_kind, uint32_t(nodes), uint32_t(leaves));
  if ((st = last_launch()) != VX_OK) return st;
  if (nodes == 0) return hipStreamSynchronize(s) == hipSuccess ? VX_OK : VX_DEVICE_ERROR;

  // First-visit order of all nodes.
  Levels view{};
  uint32_t base = 0;
  for (uint32_t l = 0; l <= D; ++l) {
    view.refs[l] = lv[l].refs.get<uint32_t>();
    view.child_base[l] = lv[l].child_base.get<uint32_t>();
    view.node_of[l] = lv[l].node_of.get<uint32_t>();
    view.off[l] = base;
    if (l < D) base += uint32_t(lv[l].n);
  }
  DevBuf okey, oval, canon;
  if ((st = dev_alloc(okey, nodes * 8)) != VX_OK) return st;
  if ((st = dev_alloc(oval, nodes * 4)) != VX_OK) return st;
  if ((st = dev_alloc(canon, nodes * 4)) != VX_OK) return st;
  for (uint32_t l = 0; l < D; ++l) {
    if (!lv[l].n) continue;
    order_keys_kernel<<<grid_for(lv[l].n), 256, 0, s>>>(lv[l].paths.get<uint64_t>(), uint32_t(lv[l].n),
                                                         3 * (D - 1 - l), l, view.off[l], okey.get<uint64_t>(),
                                                         oval.get<uint32_t>());
    if ((st = last_launch()) != VX_OK) return st;
  }
  const uint32_t order_bits = 3 * (D - 1) + 4;
  const size_t ows_bytes = sort::workspace_bytes(uint32_t(nodes), 8, order_bits, true);
  DevBuf ows;
  if ((st = dev_alloc(ows, ows_bytes)) != VX_OK) return st;
  if ((st = sort::sort_u64(okey.get<uint64_t>(), oval.get<uint32_t>(), uint32_t(nodes), order_bits, ows.p, ows_bytes,
                           s)) != VX_OK)
    return st;
  rank_kernel<<<grid_for(nodes), 256, 0, s>>>(oval.get<uint32_t>(), uint32_t(nodes), canon.get<uint32_t>());
  if ((st = last_launch()) != VX_OK) return st;

  DevBuf rws;
  if ((st = dev_alloc(rws, scan::fused_workspace_bytes(nodes))) != VX_OK) return st;
  if ((st = scan::fused<scan::kSum>(uint32_t(nodes),
                                    RecordLoad{okey.get<uint64_t>(), oval.get<uint32_t>(), g->node_pool, view},
                                    RecordStore{canon.get<uint32_t>(), out_dev, D, view}, rws.p, s)) != VX_OK)
    return st;
  if (leaves) {
    uint32_t* lo = out_dev + kHeaderWords + node_words;
    if (LW == 2)
      leaves_kernel<2><<<grid_for(leaves), 256, 0, s>>>(lv[D].refs.get<uint32_t>(), uint32_t(leaves), g->leaf_pool,
                                                        lu, lo, err);
    else
      leaves_kernel<16><<<grid_for(leaves), 256, 0, s>>>(lv[D].refs.get<uint32_t>(), uint32_t(leaves), g->leaf_pool,
                                                         lu, lo, err);
    if ((st = last_launch()) != VX_OK) return st;
  }
  if (read_u32(err, s)) return VX_DEVICE_ERROR;
  return VX_OK;
}

}  // namespace

extern "C" vx_status vx_dag_export_canonical_device(const vx_dag* g, uint32_t* d_out, uint64_t cap,
                                                    uint64_t* n_words, void* stream) {
  if (!g || !n_words) return VX_INVALID_VALUE;
  return export_gpu(g, d_out, cap, n_words, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_dag_export_canonical(const vx_dag* g, uint32_t* out, uint64_t cap, uint64_t* n_words) {
  if (!g || !n_words) return VX_INVALID_VALUE;
  if (hipDeviceSynchronize() != hipSuccess) return VX_DEVICE_ERROR;
  hipStream_t s = nullptr;
  if (!out) return export_gpu(g, nullptr, 0, n_words, nullptr, nullptr, s);
  DevBuf buf;
  uint32_t* dev = nullptr;
  vx_status st = export_gpu(g, nullptr, 0, n_words, &buf, &dev, s);
  if (st != VX_OK) return st;
  if (cap < *n_words) return VX_OK;
  if (hipMemcpyAsync(out, dev, *n_words * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return VX_DEVICE_ERROR;
  return hipStreamSynchronize(s) == hipSuccess ? VX_OK : VX_DEVICE_ERROR;
}
