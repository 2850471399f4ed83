// CPU oracle for the pointer-SVO voxel edit — TEST INFRASTRUCTURE ONLY.
//
// Single-threaded restatement of algo::svt::cuda's voxel edit (reference
// algorithm/lib/include/algo/svt/cuda/), used to check the HIP library. Loaded
// only by tests/ and the benchmark's CPU-baseline leg; the product never links
// it. Per batch, in the reference's order (voxel/detail/pipeline.cuh:37-88):
//   1. leaf masks: valid edits -> (leaf key, leaf bit), stable sort by key,
//      one OR-ed mask per key (build_leaf_masks/pipeline.cuh:18-69);
//   2. path allocation: every missing slot on a touched leaf's path gets a
//      new node (all eight slots uniform with the replaced slot's fill) or a
//      new leaf (all ones if that slot was filled), taken from the free list
//      (LIFO) before the bump counter (allocation/batch_kernels.cuh:12-79,
//      materialize_common.cuh:76-88);
//   3. bits set or cleared on every touched leaf (apply_leaf_masks.cuh:15-65);
//   4. collapse, depth 10 (leaves) up to 1: on each touched path the leaf /
//      node at that depth is freed if uniform (leaf all zero or all one; node
//      with eight pointer-free slots of one fill), requests deduplicated in
//      leaf order, freed indices pushed on the free list, the parent slot set
//      to the uniform value (collapse/kernels.cuh:12-92,
//      collapse/pipeline.cuh:19-108, batch_kernels.cuh:81-139).
// The world is 4096^3, ten inner levels of 2x2x2 (config.cuh:16-36).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr uint32_t kDepth = 10;
constexpr uint32_t kWorld = 4096;
constexpr uint32_t kIndexMask = 0x3fffffffu;
constexpr uint32_t kChildBit = 1u << 30;
constexpr uint32_t kFilledBit = 1u << 31;

struct Mask {
  uint32_t key, lo, hi;
};

// make_leaf_key (detail/geometry.cuh:12-25): the ten child indices of the
// path, top level in the highest bits.
uint32_t leaf_key(uint32_t lx, uint32_t ly, uint32_t lz) {
  uint32_t key = 0;
  for (uint32_t d = 0; d < kDepth; ++d) {
    const uint32_t b = kDepth - 1 - d;
    key = (key << 3) | ((lx >> b) & 1u) | (((ly >> b) & 1u) << 1) | (((lz >> b) & 1u) << 2);
  }
  return key;
}

uint32_t child_at(uint32_t key, uint32_t d) { return (key >> ((kDepth - 1 - d) * 3)) & 7u; }

std::vector<Mask> leaf_masks(const uint32_t* xyz, uint32_t n) {
  std::vector<std::pair<uint32_t, uint32_t>> rec;
  rec.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (x >= kWorld || y >= kWorld || z >= kWorld) continue;
    rec.emplace_back(leaf_key(x >> 2, y >> 2, z >> 2), (x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4));
  }
  std::stable_sort(rec.begin(), rec.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<Mask> out;
  for (const auto& r : rec) {
    if (out.empty() || out.back().key != r.first) out.push_back(Mask{r.first, 0, 0});
    if (r.second < 32) out.back().lo |= 1u << r.second;
    else out.back().hi |= 1u << (r.second - 32);
  }
  return out;
}

}  // namespace

struct vxo_svo {
  uint32_t max_nodes, max_leaves;
  std::vector<uint32_t> nodes;   // 8 slots per node
  std::vector<uint32_t> leaves;  // 2 words per leaf
  uint32_t node_count = 1, leaf_count = 0;
  std::vector<uint32_t> free_nodes, free_leaves;  // stacks

  vxo_svo(uint32_t mn, uint32_t ml) : max_nodes(mn), max_leaves(ml), nodes(size_t(mn) * 8), leaves(size_t(ml) * 2) {}

  void reset() {  // detail/reset.cuh:9-21
    std::fill(nodes.begin(), nodes.begin() + 8, 0u);
    node_count = 1;
    leaf_count = 0;
    free_nodes.clear();
    free_leaves.clear();
  }

  uint32_t take_node() {
    if (!free_nodes.empty()) {
      const uint32_t i = free_nodes.back();
      free_nodes.pop_back();
      return i;
    }
    return node_count++;
  }
  uint32_t take_leaf() {
    if (!free_leaves.empty()) {
      const uint32_t i = free_leaves.back();
      free_leaves.pop_back();
      return i;
    }
    return leaf_count++;
  }

  // Storage index at depth `depth` on the path of `key` (depth 10: leaf), or
  // ~0u if a slot on the way holds no pointer.
  uint32_t at_depth(uint32_t key, uint32_t depth, uint32_t* parent, uint32_t* child) const {
    uint32_t node = 0;
    for (uint32_t d = 0; d < depth; ++d) {
      const uint32_t c = child_at(key, d);
      const uint32_t s = nodes[size_t(node) * 8 + c];
      if (!(s & kChildBit)) return ~0u;
      *parent = node;
      *child = c;
      node = s & kIndexMask;
    }
    return node;
  }

  // Number of nodes / leaves a batch would have to take.
  void count_missing(const std::vector<Mask>& masks, uint64_t* n_nodes, uint64_t* n_leaves) const {
    *n_nodes = *n_leaves = 0;
    for (size_t j = 0; j < masks.size(); ++j) {
      uint32_t node = 0, d = 0;
      for (; d < kDepth; ++d) {
        const uint32_t s = nodes[size_t(node) * 8 + child_at(masks[j].key, d)];
        if (!(s & kChildBit)) break;
        node = s & kIndexMask;
      }
      if (d == kDepth) continue;
      // Slots d+1..10 of this path are new; those shared with the previous
      // mask's path were counted there.
      uint32_t common = 0;
      if (j > 0)
        while (common < kDepth && child_at(masks[j].key, common) == child_at(masks[j - 1].key, common)) ++common;
      const uint32_t first = std::max(d + 1, common + 1);
      for (uint32_t e = first; e <= kDepth; ++e) (e == kDepth ? *n_leaves : *n_nodes) += 1;
    }
  }

  int apply(const uint32_t* xyz, uint32_t n, uint32_t op) {
    const std::vector<Mask> masks = leaf_masks(xyz, n);
    if (n == 0) return 0;
    uint64_t need_n = 0, need_l = 0;
    count_missing(masks, &need_n, &need_l);
    const uint64_t new_n = need_n > free_nodes.size() ? need_n - free_nodes.size() : 0;
    const uint64_t new_l = need_l > free_leaves.size() ? need_l - free_leaves.size() : 0;
    if (node_count + new_n > max_nodes || leaf_count + new_l > max_leaves) return 2;
    // 2. paths
    for (const Mask& m : masks) {
      uint32_t node = 0;
      for (uint32_t d = 0; d < kDepth; ++d) {
        uint32_t& s = nodes[size_t(node) * 8 + child_at(m.key, d)];
        if (!(s & kChildBit)) {
          const bool filled = (s & kFilledBit) != 0;
          if (d + 1 < kDepth) {
            const uint32_t nn = take_node();
            for (uint32_t c = 0; c < 8; ++c) nodes[size_t(nn) * 8 + c] = filled ? kFilledBit : 0u;
            s = (nn & kIndexMask) | kChildBit | (filled ? kFilledBit : 0u);
          } else {
            const uint32_t nl = take_leaf();
            leaves[size_t(nl) * 2] = leaves[size_t(nl) * 2 + 1] = filled ? 0xffffffffu : 0u;
            s = (nl & kIndexMask) | kChildBit | (filled ? kFilledBit : 0u);
          }
        }
        node = s & kIndexMask;
      }
    }
    // 3. bits
    for (const Mask& m : masks) {
      uint32_t p = 0, c = 0;
      const uint32_t leaf = at_depth(m.key, kDepth, &p, &c);
      if (leaf == ~0u) continue;
      uint32_t* w = &leaves[size_t(leaf) * 2];
      if (op == 0) {
        w[0] |= m.lo;
        w[1] |= m.hi;
      } else {
        w[0] &= ~m.lo;
        w[1] &= ~m.hi;
      }
    }
    // 4. collapse
    for (uint32_t depth = kDepth; depth >= 1; --depth) {
      std::vector<std::pair<uint32_t, uint32_t>> req;  // (parent, child)
      for (const Mask& m : masks) {
        uint32_t p = 0, c = 0;
        const uint32_t idx = at_depth(m.key, depth, &p, &c);
        if (idx == ~0u) continue;
        bool uniform;
        if (depth == kDepth) {
          const uint32_t lo = leaves[size_t(idx) * 2], hi = leaves[size_t(idx) * 2 + 1];
          uniform = (lo == 0 && hi == 0) || (lo == 0xffffffffu && hi == 0xffffffffu);
        } else {
          bool all_empty = true, all_full = true;
          for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t s = nodes[size_t(idx) * 8 + k];
            if (s & kChildBit) {
              all_empty = all_full = false;
              break;
            }
            if (s & kFilledBit) all_empty = false;
            else all_full = false;
          }
          uniform = all_empty || all_full;
        }
        if (!uniform) continue;
        if (!req.empty() && req.back() == std::make_pair(p, c)) continue;
        req.emplace_back(p, c);
      }
      for (const auto& r : req) {
        uint32_t& s = nodes[size_t(r.first) * 8 + r.second];
        const uint32_t idx = s & kIndexMask;
        bool full;
        if (depth == kDepth) {
          free_leaves.push_back(idx);
          full = leaves[size_t(idx) * 2] == 0xffffffffu && leaves[size_t(idx) * 2 + 1] == 0xffffffffu;
        } else {
          free_nodes.push_back(idx);
          full = true;
          for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t t = nodes[size_t(idx) * 8 + k];
            if ((t & kChildBit) || !(t & kFilledBit)) full = false;
          }
        }
        s = full ? kFilledBit : 0u;
      }
    }
    return 0;
  }

  uint32_t get(uint32_t x, uint32_t y, uint32_t z) const {  // query.cuh:12-39
    if (x >= kWorld || y >= kWorld || z >= kWorld) return 0;
    uint32_t node = 0;
    for (uint32_t d = 0; d < kDepth; ++d) {
      const uint32_t sh = kDepth - 1 - d + 2;
      const uint32_t c = ((x >> sh) & 1u) | (((y >> sh) & 1u) << 1) | (((z >> sh) & 1u) << 2);
      const uint32_t s = nodes[size_t(node) * 8 + c];
      if (!(s & kChildBit)) return (s & kFilledBit) ? 1u : 0u;
      node = s & kIndexMask;
    }
    const uint32_t bit = (x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4);
    return (leaves[size_t(node) * 2 + (bit >> 5)] >> (bit & 31)) & 1u;
  }

  // Canonical export: nodes in pre-order from the root (children 0..7), then
  // leaves in the order the walk reaches them; pointer slots carry the
  // child's canonical id with the slot's two flag bits.
  std::vector<uint32_t> export_words() const {
    std::vector<uint32_t> node_recs, leaf_recs;
    uint32_t next_node = 0, next_leaf = 0;
    struct Frame {
      uint32_t node, depth, id, c;
    };
    std::vector<Frame> st;
    node_recs.resize(8);
    st.push_back(Frame{0, 0, next_node++, 0});
    while (!st.empty()) {
      Frame& f = st.back();
      if (f.c == 8) {
        st.pop_back();
        continue;
      }
      const uint32_t c = f.c++;
      const uint32_t s = nodes[size_t(f.node) * 8 + c];
      const size_t slot = size_t(f.id) * 8 + c;
      if (!(s & kChildBit)) {
        node_recs[slot] = s;
        continue;
      }
      const uint32_t child = s & kIndexMask;
      if (f.depth + 1 == kDepth) {
        node_recs[slot] = next_leaf | (s & (kChildBit | kFilledBit));
        ++next_leaf;
        leaf_recs.push_back(leaves[size_t(child) * 2]);
        leaf_recs.push_back(leaves[size_t(child) * 2 + 1]);
        continue;
      }
      const uint32_t id = next_node++;
      node_recs[slot] = id | (s & (kChildBit | kFilledBit));
      node_recs.resize(size_t(next_node) * 8);
      const uint32_t depth = f.depth + 1;
      st.push_back(Frame{child, depth, id, 0});  // invalidates f
    }
    std::vector<uint32_t> out = {0x5658534fu, next_node, next_leaf};
    out.insert(out.end(), node_recs.begin(), node_recs.end());
    out.insert(out.end(), leaf_recs.begin(), leaf_recs.end());
    return out;
  }
};

extern "C" {

vxo_svo* vxo_svo_create(uint32_t max_nodes, uint32_t max_leaves) {
  if (max_nodes == 0 || max_leaves == 0) return nullptr;
  vxo_svo* s = new vxo_svo(max_nodes, max_leaves);
  s->reset();
  return s;
}

void vxo_svo_destroy(vxo_svo* s) { delete s; }

void vxo_svo_reset(vxo_svo* s) { s->reset(); }

uint32_t vxo_svo_leaf_masks(const uint32_t* xyz, uint32_t n, uint32_t* out) {
  const std::vector<Mask> m = leaf_masks(xyz, n);
  if (out)
    for (size_t i = 0; i < m.size(); ++i) {
      out[3 * i] = m[i].key;
      out[3 * i + 1] = m[i].lo;
      out[3 * i + 2] = m[i].hi;
    }
  return uint32_t(m.size());
}

int vxo_svo_apply(vxo_svo* s, const uint32_t* xyz, uint32_t n, uint32_t op) { return s->apply(xyz, n, op); }

uint32_t vxo_svo_get(const vxo_svo* s, uint32_t x, uint32_t y, uint32_t z) { return s->get(x, y, z); }

void vxo_svo_counters(const vxo_svo* s, uint32_t* out) {
  out[0] = s->node_count;
  out[1] = s->leaf_count;
  out[2] = uint32_t(s->free_nodes.size());
  out[3] = uint32_t(s->free_leaves.size());
}

uint64_t vxo_svo_export(const vxo_svo* s, uint32_t* out, uint64_t cap) {
  const std::vector<uint32_t> w = s->export_words();
  if (out && cap >= w.size()) std::memcpy(out, w.data(), w.size() * 4);
  return w.size();
}

uint32_t vxo_svo_leaf_key(uint32_t lx, uint32_t ly, uint32_t lz) { return leaf_key(lx, ly, lz); }

// make_random_edits (bench/bench_cuda_svt.cu:90-112): xorshift32 (13, 17, 5)
// from 0x9e3779b9, x, y, z drawn in turn, each mod 4096.
void vxo_svo_random_edits(uint32_t count, uint32_t* xyz) {
  uint32_t s = 0x9e3779b9u;
  for (uint64_t i = 0; i < uint64_t(count) * 3; ++i) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    xyz[i] = s % kWorld;
  }
}

}  // extern "C"
