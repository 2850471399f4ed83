// C wrapper around the reference's own CPU Hash DAG (svt/hash_dag.hpp), built
// from the reference sources where they lie, into oracle/_ref/. TEST
// INFRASTRUCTURE ONLY: it pins the oracle's restatement (oracle.cpp) and can
// serve as the "reference" CPU baseline. Nothing here is shipped.
//
// The canonical export walks the reference's private node store, so access
// control is lifted for this translation unit only (after every standard
// header the reference uses has been included with normal access).
#include <algorithm>
#include <array>
#include <bit>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <vector>

#define private public
#include <algo/svt/hash_dag.hpp>
#undef private

using algo::svt::HashDAG;

namespace {

struct Export {
  const HashDAG& d;
  std::unordered_map<uint64_t, uint32_t> node_id, leaf_id;
  std::vector<std::vector<uint32_t>> node_recs;
  std::vector<uint64_t> leaf_recs;

  bool is_full(uint32_t level, uint32_t ptr) const {
    if (level == d.depth_) return ptr == d.full_leaf_ptr_;
    return ptr == d.full_node_ptrs_[level];
  }

  uint32_t leaf(uint32_t ptr) {
    auto it = leaf_id.find(ptr);
    if (it != leaf_id.end()) return it->second;
    const uint32_t id = uint32_t(leaf_recs.size());
    leaf_id.emplace(ptr, id);
    leaf_recs.push_back(d.read_leaf(ptr));
    return id;
  }

  uint32_t node(uint32_t level, uint32_t ptr) {
    const uint64_t key = (uint64_t(level) << 32) | ptr;
    auto it = node_id.find(key);
    if (it != node_id.end()) return it->second;
    const uint32_t id = uint32_t(node_recs.size());
    node_id.emplace(key, id);
    node_recs.emplace_back();
    const auto n = d.read_node(ptr);
    uint32_t cm = 0, fm = 0;
    std::vector<uint32_t> ids;
    for (uint32_t c = 0; c < 8; ++c) {
      if (!(n.child_mask & (1u << c))) continue;
      const uint32_t q = n.child_ptrs[c];
      if (q == 0) continue;
      if (is_full(level + 1, q)) { fm |= 1u << c; continue; }
      cm |= 1u << c;
      ids.push_back(level + 1 == d.depth_ ? leaf(q) : node(level + 1, q));
    }
    auto& rec = node_recs[id];
    rec.push_back(level);
    rec.push_back(cm | (fm << 8));
    rec.insert(rec.end(), ids.begin(), ids.end());
    return id;
  }

  std::vector<uint32_t> run() {
    std::vector<uint32_t> out = {0x56584443u, d.depth_, 2u, 0, 0, 0};
    if (d.root_ == 0) out[3] = 0;
    else if (is_full(0, d.root_)) out[3] = 1;
    else { out[3] = 2; node(0, d.root_); }
    out[4] = uint32_t(node_recs.size());
    out[5] = uint32_t(leaf_recs.size());
    for (auto& r : node_recs) out.insert(out.end(), r.begin(), r.end());
    for (uint64_t l : leaf_recs) { out.push_back(uint32_t(l)); out.push_back(uint32_t(l >> 32)); }
    return out;
  }
};

}  // namespace

extern "C" {

void* vxr_create(uint32_t depth) {
  try { return new HashDAG(depth); } catch (...) { return nullptr; }
}
void vxr_destroy(void* h) { delete static_cast<HashDAG*>(h); }

int vxr_box(void* h, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t x1, uint32_t y1, uint32_t z1, int fill) {
  try {
    const HashDAG::Box b{x0, y0, z0, x1, y1, z1};
    if (fill) static_cast<HashDAG*>(h)->fill_box(b);
    else static_cast<HashDAG*>(h)->clear_box(b);
    return 0;
  } catch (...) { return -1; }
}

// `n` boxes of 7 words each (x0 y0 z0 x1 y1 z1 fill) in order, in one call
// (no per-box call overhead when the reference is timed).
int vxr_boxes(void* h, const uint32_t* ops, uint64_t n) {
  try {
    HashDAG* d = static_cast<HashDAG*>(h);
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t* o = ops + 7 * i;
      const HashDAG::Box b{o[0], o[1], o[2], o[3], o[4], o[5]};
      if (o[6]) d->fill_box(b);
      else d->clear_box(b);
    }
    return 0;
  } catch (...) { return -1; }
}

int vxr_set_voxel(void* h, uint32_t x, uint32_t y, uint32_t z, int value) {
  try { static_cast<HashDAG*>(h)->set_voxel(x, y, z, value != 0); return 0; } catch (...) { return -1; }
}

int vxr_get_voxel(void* h, uint32_t x, uint32_t y, uint32_t z) {
  try { return static_cast<HashDAG*>(h)->get_voxel(x, y, z) ? 1 : 0; } catch (...) { return -1; }
}

uint64_t vxr_node_count(void* h) { return static_cast<HashDAG*>(h)->node_count(); }
uint64_t vxr_leaf_count(void* h) { return static_cast<HashDAG*>(h)->leaf_count(); }
void vxr_collect_garbage(void* h) { static_cast<HashDAG*>(h)->collect_garbage(); }

uint64_t vxr_export(void* h, uint32_t* out, uint64_t cap) {
  Export ex{*static_cast<HashDAG*>(h)};
  const auto v = ex.run();
  if (out && cap >= v.size()) std::copy(v.begin(), v.end(), out);
  return v.size();
}

}  // extern "C"
