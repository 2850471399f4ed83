// CPU oracle for the voxel-editing hot path — TEST INFRASTRUCTURE ONLY.
//
// Single-threaded restatement of the reference's algorithms (file:line cited
// per function) used to check the HIP library. Loaded only by tests/, the
// smoke check and the benchmark's CPU-baseline leg; the product never links it.
//
// Canonical export format (shared contract with the product's exporter):
//   [0] 0x56584443 magic   [1] inner depth D   [2] leaf words (2 or 16)
//   [3] root kind: 0 EMPTY, 1 FULL, 2 table node
//   [4] node count  [5] leaf count
//   node records, canonical id order: level, header, child ids...
//     header = child_mask | fill_mask << 8 (hash_dag_gpu/config.cuh:39-45);
//     ids of children at level < D are node ids, at level D leaf ids.
//   leaf records: leaf words (occupancy lo/hi, or 16 words of u8 materials).
// Ids are assigned at first visit of a depth-first walk from the root that
// visits children 0..7 in order; the memo key is (level, stored ref), which in
// a maximally shared DAG is the subtree content.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
constexpr uint32_t kEmpty = 0;
constexpr uint32_t kFull = 1;

// Stream ids of the workload generators.
constexpr uint64_t kStreamTerrain = 16;

inline uint64_t mix64(uint64_t z) {
  z += kGamma;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t rand_at(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t base = mix64(seed ^ mix64(stream));
  return mix64(base + i * kGamma);
}

inline uint32_t uniform(uint64_t r, uint32_t lo, uint32_t hi) {
  const uint64_t span = uint64_t(hi) - lo + 1;
  return lo + uint32_t(((r >> 32) * span) >> 32);
}

inline uint64_t spread3(uint32_t v) {
  uint64_t out = 0;
  for (uint32_t b = 0; b < 21; ++b) out |= uint64_t((v >> b) & 1u) << (3 * b);
  return out;
}

// ---- terrain ---------------------------------------------------------------
// h(x,z) = base + floor(amp * sum_o 2^-o * n_o(x,z)), n_o = value noise with
// period (period >> o), lattice values uniform in [-1,1) (Q16), smoothstep
// weights, all in integer fixed point so every implementation agrees exactly.
struct TerrainEval {
  vxo_terrain t;
  std::vector<std::vector<int32_t>> lattice;  // per octave, (cells+2)^2
  std::vector<uint32_t> cells;
  uint32_t world;

  TerrainEval(const vxo_terrain& spec, uint32_t world_side)
      : t(spec), world(world_side) {
    for (uint32_t o = 0; o < t.octaves; ++o) {
      const uint32_t period = std::max(1u, t.period >> o);
      const uint32_t n = world / period + 2;
      cells.push_back(n);
      std::vector<int32_t> lat(size_t(n) * n);
      for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < n; ++j) lat[size_t(i) * n + j] = lattice_value(o, i, j);
      lattice.push_back(std::move(lat));
    }
  }

  int32_t lattice_value(uint32_t o, uint32_t i, uint32_t j) const {
    const uint64_t r = rand_at(t.seed, kStreamTerrain + o, (uint64_t(i) << 32) | j);
    return int32_t(int64_t(r >> 47) - 65536);
  }

  static int64_t smooth(int64_t tq) {  // 3t^2 - 2t^3 in Q16
    return (tq * tq * (3 * 65536 - 2 * tq)) >> 32;
  }

  int32_t height(uint32_t x, uint32_t z) const {
    int64_t total = 0;  // Q20
    for (uint32_t o = 0; o < t.octaves; ++o) {
      const uint32_t period = std::max(1u, t.period >> o);
      const uint32_t ix = x / period, iz = z / period;
      const int64_t sx = smooth((int64_t(x % period) << 16) / period);
      const int64_t sz = smooth((int64_t(z % period) << 16) / period);
      const uint32_t n = cells[o];
      const auto& lat = lattice[o];
      const int64_t a = lat[size_t(ix) * n + iz], b = lat[size_t(ix + 1) * n + iz];
      const int64_t c = lat[size_t(ix) * n + iz + 1], d = lat[size_t(ix + 1) * n + iz + 1];
      const int64_t u = a + (((b - a) * sx) >> 16);
      const int64_t v = c + (((d - c) * sx) >> 16);
      const int64_t w = u + (((v - u) * sz) >> 16);
      total += w << (4 - o);
    }
    return int32_t(int64_t(t.base) + ((int64_t(t.amp) * total) >> 20));
  }
};

// ---- shapes ------------------------------------------------------------------
enum class Cover { Outside, Inside, Partial };

struct Shape {
  // Half-open box [lo, hi) or sphere (c, r) over integer voxel positions.
  bool sphere = false;
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int64_t c[3] = {0, 0, 0}, r2 = 0;

  bool contains(int64_t x, int64_t y, int64_t z) const {
    if (sphere) {
      const int64_t dx = x - c[0], dy = y - c[1], dz = z - c[2];
      return dx * dx + dy * dy + dz * dz <= r2;
    }
    return x >= lo[0] && x < hi[0] && y >= lo[1] && y < hi[1] && z >= lo[2] && z < hi[2];
  }

  // Cube of voxels [o, o + e) on each axis (hash_dag.hpp:803-812 for boxes).
  Cover classify(const int64_t o[3], int64_t e) const {
    if (sphere) {
      int64_t near2 = 0, far2 = 0;
      for (int a = 0; a < 3; ++a) {
        const int64_t lo_v = o[a], hi_v = o[a] + e - 1;
        const int64_t nearest = std::clamp(c[a], lo_v, hi_v) - c[a];
        const int64_t farthest = std::max(std::abs(lo_v - c[a]), std::abs(hi_v - c[a]));
        near2 += nearest * nearest;
        far2 += farthest * farthest;
      }
      if (near2 > r2) return Cover::Outside;
      if (far2 <= r2) return Cover::Inside;
      return Cover::Partial;
    }
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      if (!(lo[a] < o[a] + e && hi[a] > o[a])) return Cover::Outside;
      if (!(lo[a] <= o[a] && hi[a] >= o[a] + e)) inside = false;
    }
    return inside ? Cover::Inside : Cover::Partial;
  }
};

Shape shape_of(const vxo_brush& b) {
  Shape s;
  if (b.kind == 0) {
    s.sphere = true;
    s.c[0] = b.cx; s.c[1] = b.cy; s.c[2] = b.cz;
    s.r2 = int64_t(b.hx) * b.hx;
  } else {
    const int64_t cc[3] = {b.cx, b.cy, b.cz}, hh[3] = {b.hx, b.hy, b.hz};
    for (int a = 0; a < 3; ++a) { s.lo[a] = cc[a] - hh[a]; s.hi[a] = cc[a] + hh[a] + 1; }
  }
  return s;
}

// ---- per-level hash-consed stores (hash_dag.hpp:535-656 restated) ------------
// Items are interned by content; equality of refs at one level is equality of
// content, which is what makes the DAG maximally shared.
struct Store {
  std::vector<uint32_t> words;
  std::vector<uint64_t> offs;   // item id -> first word
  std::vector<uint32_t> table;  // open addressing: id + 1, 0 empty
  uint64_t mask = 0;

  uint32_t count() const { return uint32_t(offs.size()); }

  static uint64_t hash(const uint32_t* w, uint32_t n) {
    uint64_t h = 0x84222325cbf29ce4ull ^ n;
    for (uint32_t i = 0; i < n; ++i) h = mix64(h ^ w[i]);
    return h;
  }

  uint32_t len(uint32_t id, bool fixed, uint32_t fixed_len) const {
    if (fixed) return fixed_len;
    return 1 + __builtin_popcount(words[offs[id]] & 0xffu);
  }

  void grow(bool fixed, uint32_t fixed_len) {
    const uint64_t cap = table.empty() ? 1024 : table.size() * 2;
    table.assign(cap, 0);
    mask = cap - 1;
    for (uint32_t id = 0; id < count(); ++id) {
      const uint32_t n = len(id, fixed, fixed_len);
      uint64_t s = hash(&words[offs[id]], n) & mask;
      while (table[s]) s = (s + 1) & mask;
      table[s] = id + 1;
    }
  }

  // find_or_append_entry (hash_dag.hpp:644-656): returns the item id.
  uint32_t intern(const uint32_t* w, uint32_t n, bool fixed) {
    if ((uint64_t(count()) + 1) * 2 > table.size()) grow(fixed, n);
    uint64_t s = hash(w, n) & mask;
    for (;;) {
      const uint32_t e = table[s];
      if (!e) break;
      const uint32_t id = e - 1;
      if (len(id, fixed, n) == n && std::equal(w, w + n, &words[offs[id]])) return id;
      s = (s + 1) & mask;
    }
    const uint32_t id = count();
    offs.push_back(words.size());
    words.insert(words.end(), w, w + n);
    table[s] = id + 1;
    return id;
  }
};

}  // namespace

struct vxo_dag {
  uint32_t depth;       // inner levels D
  bool materials;
  uint32_t leaf_words;  // 2 or 16
  uint32_t world;
  std::vector<Store> nodes;  // levels 0..D-1
  Store leaves;
  uint32_t root = kEmpty;
  std::vector<std::array<uint32_t, 256>> uniform;  // [level][material]
  std::unordered_map<uint64_t, uint32_t> solid;    // (level, y0) -> ref

  vxo_dag(uint32_t d, bool m)
      : depth(d), materials(m), leaf_words(m ? 16 : 2), world(4u << d), nodes(d) {
    uniform.assign(d + 1, {});
    for (auto& u : uniform) u.fill(0xffffffffu);
  }

  // -- leaves --------------------------------------------------------------
  using Leaf = std::array<uint32_t, 16>;

  Leaf read_leaf(uint32_t ref) const {
    Leaf l{};
    if (ref == kEmpty) return l;
    if (ref == kFull) { l[0] = l[1] = 0xffffffffu; return l; }
    const uint32_t* w = &leaves.words[leaves.offs[ref - 2]];
    std::copy(w, w + leaf_words, l.begin());
    return l;
  }

  uint32_t intern_leaf(const Leaf& l) {
    bool any = false, all = true;
    for (uint32_t i = 0; i < leaf_words; ++i) any |= l[i] != 0;
    if (!any) return kEmpty;  // hash_dag.hpp:547-550
    if (!materials) {
      all = l[0] == 0xffffffffu && l[1] == 0xffffffffu;
      if (all) return kFull;  // full leaf is the FULL sentinel
    }
    return leaves.intern(l.data(), leaf_words, true) + 2;
  }

  static uint32_t bit_of(uint32_t lx, uint32_t ly, uint32_t lz) {
    return lx + 4 * ly + 16 * lz;  // hash_dag.hpp:782-790
  }

  uint32_t leaf_get(const Leaf& l, uint32_t bit) const {
    if (!materials) return (bit < 32 ? (l[0] >> bit) : (l[1] >> (bit - 32))) & 1u;
    return (l[bit >> 2] >> ((bit & 3) * 8)) & 0xffu;
  }

  void leaf_set(Leaf& l, uint32_t bit, uint32_t mat) const {
    if (!materials) {
      uint32_t& w = l[bit >> 5];
      const uint32_t m = 1u << (bit & 31);
      w = mat ? (w | m) : (w & ~m);
      return;
    }
    uint32_t& w = l[bit >> 2];
    const uint32_t sh = (bit & 3) * 8;
    w = (w & ~(0xffu << sh)) | ((mat & 0xffu) << sh);
  }

  // -- nodes ---------------------------------------------------------------
  std::array<uint32_t, 8> children_of(uint32_t level, uint32_t ref) const {
    std::array<uint32_t, 8> ch{};
    if (ref == kEmpty) return ch;
    if (ref == kFull) { ch.fill(kFull); return ch; }
    const Store& s = nodes[level];
    const uint32_t* w = &s.words[s.offs[ref - 2]];
    const uint32_t cm = w[0] & 0xffu, fm = (w[0] >> 8) & 0xffu;
    uint32_t k = 1;
    for (uint32_t c = 0; c < 8; ++c) {
      if (cm & (1u << c)) ch[c] = w[k++];
      else ch[c] = (fm & (1u << c)) ? kFull : kEmpty;
    }
    return ch;
  }

  // Canonical compact node (hash_dag_gpu/detail/node.cuh:210-289): uniform
  // nodes collapse to the sentinels before any table access.
  uint32_t intern_node(uint32_t level, const std::array<uint32_t, 8>& ch) {
    uint32_t item[9];
    uint32_t cm = 0, fm = 0, n = 1;
    for (uint32_t c = 0; c < 8; ++c) {
      if (ch[c] == kFull) fm |= 1u << c;
      else if (ch[c] != kEmpty) { cm |= 1u << c; item[n++] = ch[c]; }
    }
    if (cm == 0 && fm == 0) return kEmpty;
    if (cm == 0 && fm == 0xffu) return kFull;
    item[0] = cm | (fm << 8);
    return nodes[level].intern(item, n, false) + 2;
  }

  // Subtree of one material everywhere (the analogue of full_node_ptrs_,
  // hash_dag.hpp:329-341, per material).
  uint32_t uniform_ref(uint32_t level, uint32_t mat) {
    if (mat == 0) return kEmpty;
    if (!materials) return kFull;
    uint32_t& memo = uniform[level][mat];
    if (memo != 0xffffffffu) return memo;
    uint32_t r;
    if (level == depth) {
      Leaf l{};
      for (uint32_t i = 0; i < 16; ++i) l[i] = mat * 0x01010101u;
      r = intern_leaf(l);
    } else {
      std::array<uint32_t, 8> ch;
      ch.fill(uniform_ref(level + 1, mat));
      r = intern_node(level, ch);
    }
    memo = r;
    return r;
  }

  uint32_t extent(uint32_t level) const { return world >> level; }

  // -- shape edit: edit_node / edit_leaf_parent (hash_dag.hpp:369-510) -------
  uint32_t edit(const Shape& s, uint32_t mat, uint32_t level, const int64_t o[3], uint32_t ref) {
    const int64_t e = extent(level);
    const Cover cv = s.classify(o, e);
    if (cv == Cover::Outside) return ref;
    if (cv == Cover::Inside) return uniform_ref(level, mat);
    if (level == depth) {
      Leaf l = read_leaf(ref);
      for (uint32_t lz = 0; lz < 4; ++lz)
        for (uint32_t ly = 0; ly < 4; ++ly)
          for (uint32_t lx = 0; lx < 4; ++lx)
            if (s.contains(o[0] + lx, o[1] + ly, o[2] + lz)) leaf_set(l, bit_of(lx, ly, lz), mat);
      return intern_leaf(l);
    }
    std::array<uint32_t, 8> ch = children_of(level, ref);
    const int64_t h = e / 2;
    bool changed = false;
    for (uint32_t c = 0; c < 8; ++c) {
      const int64_t co[3] = {o[0] + ((c & 1) ? h : 0), o[1] + ((c & 2) ? h : 0), o[2] + ((c & 4) ? h : 0)};
      const uint32_t nc = edit(s, mat, level + 1, co, ch[c]);
      if (nc != ch[c]) { ch[c] = nc; changed = true; }
    }
    if (!changed) return ref;
    return intern_node(level, ch);
  }

  void apply_shape(const Shape& s, uint32_t mat) {
    const int64_t o[3] = {0, 0, 0};
    root = edit(s, mat, 0, o, root);
  }

  // -- terrain -------------------------------------------------------------
  struct Pyramid {
    std::vector<std::vector<uint16_t>> mn, mx;  // level k: (W>>k)^2
  };

  uint32_t solid_ref(uint32_t level, uint32_t y0) {
    if (!materials) return kFull;
    const uint64_t key = (uint64_t(level) << 32) | y0;
    auto it = solid.find(key);
    if (it != solid.end()) return it->second;
    uint32_t r;
    if (level == depth) {
      Leaf l{};
      for (uint32_t ly = 0; ly < 4; ++ly) {
        const uint32_t m = 1 + (y0 + ly) % 255;
        for (uint32_t lz = 0; lz < 4; ++lz)
          for (uint32_t lx = 0; lx < 4; ++lx) leaf_set(l, bit_of(lx, ly, lz), m);
      }
      r = intern_leaf(l);
    } else {
      const uint32_t h = extent(level) / 2;
      std::array<uint32_t, 8> ch;
      for (uint32_t c = 0; c < 8; ++c) ch[c] = solid_ref(level + 1, y0 + ((c & 2) ? h : 0));
      r = intern_node(level, ch);
    }
    solid.emplace(key, r);
    return r;
  }

  uint32_t build(const Pyramid& p, const std::vector<uint16_t>& hmap, uint32_t level,
                 uint32_t x0, uint32_t y0, uint32_t z0) {
    const uint32_t e = extent(level);
    const uint32_t k = __builtin_ctz(e);
    const uint32_t side = world >> k;
    const size_t cell = size_t(z0 >> k) * side + (x0 >> k);
    const uint32_t hmin = p.mn[k][cell], hmax = p.mx[k][cell];
    if (y0 > hmax) return kEmpty;
    if (y0 + e - 1 <= hmin) return solid_ref(level, y0);
    if (level == depth) {
      Leaf l{};
      for (uint32_t lz = 0; lz < 4; ++lz)
        for (uint32_t lx = 0; lx < 4; ++lx) {
          const uint32_t hh = hmap[size_t(z0 + lz) * world + x0 + lx];
          for (uint32_t ly = 0; ly < 4; ++ly) {
            const uint32_t y = y0 + ly;
            if (y <= hh) leaf_set(l, bit_of(lx, ly, lz), materials ? 1 + y % 255 : 1);
          }
        }
      return intern_leaf(l);
    }
    const uint32_t h = e / 2;
    std::array<uint32_t, 8> ch;
    for (uint32_t c = 0; c < 8; ++c)
      ch[c] = build(p, hmap, level + 1, x0 + ((c & 1) ? h : 0), y0 + ((c & 2) ? h : 0), z0 + ((c & 4) ? h : 0));
    return intern_node(level, ch);
  }

  void build_terrain(const vxo_terrain& t) {
    TerrainEval ev(t, world);
    std::vector<uint16_t> hmap(size_t(world) * world);
    for (uint32_t z = 0; z < world; ++z)
      for (uint32_t x = 0; x < world; ++x) {
        const int32_t hv = ev.height(x, z);
        hmap[size_t(z) * world + x] = uint16_t(std::clamp<int32_t>(hv, 0, 65535));
      }
    Pyramid p;
    p.mn.push_back(hmap);
    p.mx.push_back(hmap);
    for (uint32_t k = 1; (world >> k) >= 1 && k <= 31; ++k) {
      const uint32_t side = world >> k, prev = side * 2;
      std::vector<uint16_t> mn(size_t(side) * side), mx(size_t(side) * side);
      for (uint32_t z = 0; z < side; ++z)
        for (uint32_t x = 0; x < side; ++x) {
          const auto& a = p.mn[k - 1];
          const auto& b = p.mx[k - 1];
          const size_t i0 = size_t(2 * z) * prev + 2 * x, i1 = i0 + prev;
          mn[size_t(z) * side + x] = std::min({a[i0], a[i0 + 1], a[i1], a[i1 + 1]});
          mx[size_t(z) * side + x] = std::max({b[i0], b[i0 + 1], b[i1], b[i1 + 1]});
        }
      p.mn.push_back(std::move(mn));
      p.mx.push_back(std::move(mx));
      if (side == 1) break;
    }
    root = build(p, hmap, 0, 0, 0, 0);
  }

  // -- queries -------------------------------------------------------------
  uint32_t get(uint32_t x, uint32_t y, uint32_t z) const {
    if (x >= world || y >= world || z >= world) return 0;
    uint32_t ref = root;
    for (uint32_t level = 0; level < depth; ++level) {
      if (ref == kEmpty) return 0;
      if (ref == kFull) return 1;
      const uint32_t sh = 31 - __builtin_clz(world) - 1 - level;  // hash_dag.hpp:771-780
      const uint32_t c = ((x >> sh) & 1) | (((y >> sh) & 1) << 1) | (((z >> sh) & 1) << 2);
      ref = children_of(level, ref)[c];
    }
    if (ref == kEmpty) return 0;
    if (ref == kFull) return 1;
    return leaf_get(read_leaf(ref), bit_of(x & 3, y & 3, z & 3));
  }
};

namespace {

// Canonical export of any DAG given accessors (shared by the edit DAG and the
// dense brute force).
struct Exporter {
  const vxo_dag& d;
  std::unordered_map<uint64_t, uint32_t> node_id, leaf_id;
  std::vector<std::vector<uint32_t>> node_recs;
  std::vector<vxo_dag::Leaf> leaf_recs;

  explicit Exporter(const vxo_dag& dag) : d(dag) {}

  uint32_t visit_leaf(uint32_t ref) {
    auto it = leaf_id.find(ref);
    if (it != leaf_id.end()) return it->second;
    const uint32_t id = uint32_t(leaf_recs.size());
    leaf_id.emplace(ref, id);
    leaf_recs.push_back(d.read_leaf(ref));
    return id;
  }

  uint32_t visit_node(uint32_t level, uint32_t ref) {
    const uint64_t key = (uint64_t(level) << 32) | ref;
    auto it = node_id.find(key);
    if (it != node_id.end()) return it->second;
    const uint32_t id = uint32_t(node_recs.size());
    node_id.emplace(key, id);
    node_recs.emplace_back();
    const auto ch = d.children_of(level, ref);
    uint32_t cm = 0, fm = 0;
    std::vector<uint32_t> ids;
    for (uint32_t c = 0; c < 8; ++c) {
      if (ch[c] == kEmpty) continue;
      if (ch[c] == kFull) { fm |= 1u << c; continue; }
      cm |= 1u << c;
      ids.push_back(level + 1 == d.depth ? visit_leaf(ch[c]) : visit_node(level + 1, ch[c]));
    }
    std::vector<uint32_t>& rec = node_recs[id];
    rec.push_back(level);
    rec.push_back(cm | (fm << 8));
    rec.insert(rec.end(), ids.begin(), ids.end());
    return id;
  }

  std::vector<uint32_t> run() {
    std::vector<uint32_t> out = {0x56584443u, d.depth, d.leaf_words, 0, 0, 0};
    if (d.root == kEmpty) out[3] = 0;
    else if (d.root == kFull) out[3] = 1;
    else {
      out[3] = 2;
      if (d.depth == 0) visit_leaf(d.root);
      else visit_node(0, d.root);
    }
    out[4] = uint32_t(node_recs.size());
    out[5] = uint32_t(leaf_recs.size());
    for (auto& r : node_recs) out.insert(out.end(), r.begin(), r.end());
    for (auto& l : leaf_recs) out.insert(out.end(), l.begin(), l.begin() + d.leaf_words);
    return out;
  }
};

uint64_t copy_out(const std::vector<uint32_t>& v, uint32_t* out, uint64_t cap) {
  if (out && cap >= v.size()) std::memcpy(out, v.data(), v.size() * 4);
  return v.size();
}

// Bottom-up builder from a dense grid: intern every subtree by content.
uint32_t dense_build(vxo_dag& d, const uint8_t* g, uint32_t level, uint32_t x0, uint32_t y0, uint32_t z0) {
  const uint32_t e = d.extent(level);
  if (level == d.depth) {
    vxo_dag::Leaf l{};
    for (uint32_t lz = 0; lz < 4; ++lz)
      for (uint32_t ly = 0; ly < 4; ++ly)
        for (uint32_t lx = 0; lx < 4; ++lx) {
          const uint8_t m = g[(size_t(z0 + lz) * d.world + (y0 + ly)) * d.world + x0 + lx];
          if (m) d.leaf_set(l, vxo_dag::bit_of(lx, ly, lz), d.materials ? m : 1);
        }
    return d.intern_leaf(l);
  }
  const uint32_t h = e / 2;
  std::array<uint32_t, 8> ch;
  for (uint32_t c = 0; c < 8; ++c)
    ch[c] = dense_build(d, g, level + 1, x0 + ((c & 1) ? h : 0), y0 + ((c & 2) ? h : 0), z0 + ((c & 4) ? h : 0));
  return d.intern_node(level, ch);
}

}  // namespace

extern "C" {

uint64_t vxo_mix64(uint64_t z) { return mix64(z); }
uint64_t vxo_rand(uint64_t seed, uint64_t stream, uint64_t index) { return rand_at(seed, stream, index); }

void vxo_gen_u32(uint64_t seed, uint64_t stream, uint64_t n, uint32_t lo, uint32_t hi, uint32_t* out) {
  const uint64_t base = mix64(seed ^ mix64(stream));
  for (uint64_t i = 0; i < n; ++i) out[i] = uniform(mix64(base + i * kGamma), lo, hi);
}

uint64_t vxo_morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

void vxo_gen_morton(uint64_t seed, uint64_t stream, uint64_t n, uint32_t bits, uint64_t* out) {
  const uint64_t base = mix64(seed ^ mix64(stream));
  const uint64_t m = (uint64_t(1) << bits) - 1;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t r = mix64(base + i * kGamma);
    out[i] = vxo_morton3(uint32_t(r & m), uint32_t((r >> bits) & m), uint32_t((r >> (2 * bits)) & m));
  }
}

void vxo_gen_unique_u64(uint64_t seed, uint64_t stream, uint64_t n, uint64_t* out) {
  // i -> mix64(base + i*gamma) is a bijection on u64, so values are distinct;
  // the reserved key ~0 is replaced by a draw from a disjoint index range.
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = rand_at(seed, stream, i);
    if (k == ~uint64_t(0)) k = rand_at(seed, stream, (uint64_t(1) << 63) + i);
    out[i] = k;
  }
}

void vxo_exclusive_sum_u32(const uint32_t* in, uint64_t n, uint32_t* out) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { const uint32_t v = in[i]; out[i] = acc; acc += v; }
}
void vxo_inclusive_sum_u32(const uint32_t* in, uint64_t n, uint32_t* out) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc += in[i]; out[i] = acc; }
}
void vxo_exclusive_max_u32(const uint32_t* in, uint64_t n, uint32_t* out) {
  uint32_t acc = 0;  // identity of Max over u32 (scan/ops.cuh)
  for (uint64_t i = 0; i < n; ++i) { const uint32_t v = in[i]; out[i] = acc; acc = std::max(acc, v); }
}
void vxo_inclusive_max_u32(const uint32_t* in, uint64_t n, uint32_t* out) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc = std::max(acc, in[i]); out[i] = acc; }
}

// u64 forms (scan<T> with T = uint64_t, cuda/scan/scan.cuh:79-151): sums wrap
// modulo 2^64, max starts from 0 (scan/ops.cuh:22-36).
void vxo_exclusive_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { const uint64_t v = in[i]; out[i] = acc; acc += v; }
}
void vxo_inclusive_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc += in[i]; out[i] = acc; }
}
void vxo_exclusive_max_u64(const uint64_t* in, uint64_t n, uint64_t* out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { const uint64_t v = in[i]; out[i] = acc; acc = std::max(acc, v); }
}
void vxo_inclusive_max_u64(const uint64_t* in, uint64_t n, uint64_t* out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc = std::max(acc, in[i]); out[i] = acc; }
}

// Stable order of n keys of `words` u32 words (words[0] least significant,
// cuda/sort/uint_key.cuh:15-66) by their low key_bits bits:
// std::stable_sort of the indices (sort_by_key semantics, cuda/sort/sort.cuh:118-138).
void vxo_sort_order(const uint32_t* keys, uint32_t words, uint64_t n, uint32_t key_bits, uint32_t* order) {
  std::vector<uint32_t> mask(words, 0u);
  for (uint32_t w = 0; w < words; ++w) {
    const int64_t b = int64_t(key_bits) - 32 * int64_t(w);
    mask[w] = b >= 32 ? 0xffffffffu : (b <= 0 ? 0u : ((1u << b) - 1));
  }
  std::vector<uint32_t> idx(n);
  for (uint64_t i = 0; i < n; ++i) idx[i] = uint32_t(i);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    for (uint32_t w = words; w-- > 0;) {
      const uint32_t x = keys[uint64_t(a) * words + w] & mask[w], y = keys[uint64_t(b) * words + w] & mask[w];
      if (x != y) return x < y;
    }
    return false;
  });
  for (uint64_t i = 0; i < n; ++i) order[i] = idx[i];
}

}  // extern "C"

// std::stable_sort of (key, value) pairs by the low `bits` key bits
// (sort.cuh semantics: stable, ordered by the masked key).
template <class K>
static void sort_pairs(K* keys, uint32_t* vals, uint64_t n, uint32_t bits) {
  const K mask = bits >= sizeof(K) * 8 ? ~K(0) : ((K(1) << bits) - 1);
  struct Pair {
    K key;
    uint32_t val;
  };
  std::vector<Pair> p(n);
  for (uint64_t i = 0; i < n; ++i) p[i] = Pair{keys[i], vals ? vals[i] : 0u};
  std::stable_sort(p.begin(), p.end(), [mask](const Pair& a, const Pair& b) { return (a.key & mask) < (b.key & mask); });
  for (uint64_t i = 0; i < n; ++i) keys[i] = p[i].key;
  if (vals)
    for (uint64_t i = 0; i < n; ++i) vals[i] = p[i].val;
}

extern "C" {

void vxo_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t bits) { sort_pairs(keys, vals, n, bits); }
void vxo_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t bits) { sort_pairs(keys, vals, n, bits); }

void vxo_set_contains_u64(const uint64_t* keys, uint64_t nk, const uint64_t* q, uint64_t nq, uint8_t* out) {
  std::unordered_set<uint64_t> s(keys, keys + nk);
  for (uint64_t i = 0; i < nq; ++i) out[i] = s.count(q[i]) ? 1 : 0;
}

int32_t vxo_terrain_height(const vxo_terrain* t, uint32_t x, uint32_t z) {
  // Lattice for a world just large enough to hold (x, z).
  uint32_t w = std::max(1u, t->period);
  while (w <= std::max(x, z)) w *= 2;
  TerrainEval ev(*t, w);
  return ev.height(x, z);
}

// Heights of n columns (one lattice for all of them).
void vxo_terrain_heights(const vxo_terrain* t, const uint32_t* xs, const uint32_t* zs, uint64_t n, int32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = 0; i < n; ++i) m = std::max(m, std::max(xs[i], zs[i]));
  uint32_t w = std::max(1u, t->period);
  while (w <= m) w *= 2;
  TerrainEval ev(*t, w);
  for (uint64_t i = 0; i < n; ++i) out[i] = ev.height(xs[i], zs[i]);
}

void vxo_gen_brushes(uint64_t seed, uint64_t stream, uint32_t n, uint32_t world, uint32_t mix, vxo_brush* out) {
  const uint64_t base = mix64(seed ^ mix64(stream));
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t r[16];
    for (uint32_t k = 0; k < 16; ++k) r[k] = mix64(base + (uint64_t(i) * 16 + k) * kGamma);
    vxo_brush b{};
    b.cx = int32_t(uniform(r[1], 0, world - 1));
    b.cy = int32_t(uniform(r[2], 0, world - 1));
    b.cz = int32_t(uniform(r[3], 0, world - 1));
    uint32_t add_pct;
    if (mix == 0) {
      b.kind = 0;
      b.hx = int32_t(uniform(r[4], 4, 16));
      add_pct = 50;
    } else {
      b.kind = uniform(r[0], 0, 1);
      if (b.kind == 0) {
        b.hx = int32_t(uniform(r[4], 8, 64));
      } else {
        b.hx = int32_t(uniform(r[4], 4, 48));
        b.hy = int32_t(uniform(r[5], 4, 48));
        b.hz = int32_t(uniform(r[6], 4, 48));
      }
      add_pct = 60;
    }
    const bool add = uniform(r[7], 0, 99) < add_pct;
    b.material = add ? uniform(r[8], 1, 255) : 0;
    out[i] = b;
  }
}

uint64_t vxo_brush_voxels(const vxo_brush* b, uint32_t world) {
  const Shape s = shape_of(*b);
  const int64_t r = b->kind == 0 ? b->hx : 0;
  int64_t lo[3], hi[3];
  const int64_t c[3] = {b->cx, b->cy, b->cz}, h[3] = {b->hx, b->hy, b->hz};
  for (int a = 0; a < 3; ++a) {
    const int64_t ext = b->kind == 0 ? r : h[a];
    lo[a] = std::max<int64_t>(0, c[a] - ext);
    hi[a] = std::min<int64_t>(world - 1, c[a] + ext);
  }
  uint64_t count = 0;
  for (int64_t z = lo[2]; z <= hi[2]; ++z)
    for (int64_t y = lo[1]; y <= hi[1]; ++y)
      for (int64_t x = lo[0]; x <= hi[0]; ++x) count += s.contains(x, y, z);
  return count;
}

vxo_dag* vxo_dag_create(uint32_t inner_depth, int materials) {
  if (inner_depth == 0 || inner_depth > 14) return nullptr;
  return new vxo_dag(inner_depth, materials != 0);
}
void vxo_dag_destroy(vxo_dag* d) { delete d; }

void vxo_dag_build_terrain(vxo_dag* d, const vxo_terrain* t) { d->build_terrain(*t); }

void vxo_dag_apply_brushes(vxo_dag* d, const vxo_brush* b, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t mat = d->materials ? b[i].material : (b[i].material ? 1u : 0u);
    d->apply_shape(shape_of(b[i]), mat);
  }
}

void vxo_dag_apply_voxels(vxo_dag* d, const uint32_t* xyz, uint32_t n, uint32_t material) {
  const uint32_t mat = d->materials ? material : (material ? 1u : 0u);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (x >= d->world || y >= d->world || z >= d->world) continue;  // ignored like the GPU path
    Shape s;
    s.lo[0] = x; s.lo[1] = y; s.lo[2] = z;
    s.hi[0] = x + 1; s.hi[1] = y + 1; s.hi[2] = z + 1;
    d->apply_shape(s, mat);
  }
}

void vxo_dag_box(vxo_dag* d, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t x1, uint32_t y1, uint32_t z1, uint32_t material) {
  if (x0 >= x1 || y0 >= y1 || z0 >= z1) return;  // hash_dag.hpp:345-348
  Shape s;
  s.lo[0] = x0; s.lo[1] = y0; s.lo[2] = z0;
  s.hi[0] = x1; s.hi[1] = y1; s.hi[2] = z1;
  d->apply_shape(s, d->materials ? material : (material ? 1u : 0u));
}

uint32_t vxo_dag_get(const vxo_dag* d, uint32_t x, uint32_t y, uint32_t z) { return d->get(x, y, z); }

// Voxel-by-voxel walk: the next crossing is the axis with the smallest
// (2 k_a + 1) / (2 |d_a|) (compared by cross-multiplication), lowest axis on
// ties; get() at every visited voxel.
void vxo_dag_raycast(const vxo_dag* d, const int32_t* rays, uint32_t n, int32_t* hits) {
  const int64_t world = int64_t(d->world);
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t* q = rays + 8 * size_t(i);
    const int64_t o[3] = {q[0], q[1], q[2]};
    const int64_t dir[3] = {q[3], q[4], q[5]};
    const uint64_t max_steps = uint32_t(q[6]);
    uint64_t ad[3], k[3] = {0, 0, 0};
    int64_t sgn[3];
    for (int a = 0; a < 3; ++a) {
      ad[a] = uint64_t(dir[a] < 0 ? -dir[a] : dir[a]);
      sgn[a] = dir[a] > 0 ? 1 : (dir[a] < 0 ? -1 : 0);
    }
    int32_t* h = hits + 6 * size_t(i);
    for (int c = 0; c < 6; ++c) h[c] = 0;
    for (uint64_t step = 0;; ++step) {
      int64_t v[3];
      bool inside = true;
      for (int a = 0; a < 3; ++a) {
        v[a] = o[a] + sgn[a] * int64_t(k[a]);
        inside = inside && v[a] >= 0 && v[a] < world;
      }
      if (!inside || step > max_steps) break;
      const uint32_t m = d->get(uint32_t(v[0]), uint32_t(v[1]), uint32_t(v[2]));
      if (m) {
        h[0] = int32_t(m);
        h[1] = int32_t(step);
        h[2] = int32_t(v[0]);
        h[3] = int32_t(v[1]);
        h[4] = int32_t(v[2]);
        break;
      }
      int best = -1;
      for (int a = 0; a < 3; ++a) {
        if (!ad[a]) continue;
        if (best < 0 || (2 * k[a] + 1) * ad[best] < (2 * k[best] + 1) * ad[a]) best = a;
      }
      if (best < 0) break;
      ++k[best];
    }
  }
}

uint64_t vxo_dag_export(const vxo_dag* d, uint32_t* out, uint64_t cap) {
  Exporter ex(*d);
  return copy_out(ex.run(), out, cap);
}

uint64_t vxo_dag_stored_nodes(const vxo_dag* d) {
  uint64_t n = 0;
  for (const auto& s : d->nodes) n += s.count();
  return n;
}
uint64_t vxo_dag_stored_leaves(const vxo_dag* d) { return d->leaves.count(); }

uint64_t vxo_dense_export(const uint8_t* grid, uint32_t inner_depth, int materials, uint32_t* out, uint64_t cap) {
  vxo_dag d(inner_depth, materials != 0);
  d.root = dense_build(d, grid, 0, 0, 0, 0);
  Exporter ex(d);
  return copy_out(ex.run(), out, cap);
}

uint64_t vxo_digest(const uint32_t* w, uint64_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (uint64_t i = 0; i < n; ++i) { h ^= w[i]; h *= 0x100000001b3ull; }
  return h;
}

}  // extern "C"
