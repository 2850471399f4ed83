// Drives every function of the CPU oracle once over small and edge-case
// inputs, for the AddressSanitizer / UndefinedBehaviorSanitizer build
// (`make -C oracle sanitize`). TEST INFRASTRUCTURE ONLY: run by
// tests/test_sanitizers.py; a sanitizer finding aborts with a non-zero exit.
// Results are sanity-checked only (the parity of each function is pinned by
// tests/test_oracle.py); the point is that every path runs under the
// sanitizers: generators, scans, sorts, the set, terrain, brushes, both DAG
// kinds (terrain, brushes, voxels, boxes, queries, rays, export), the dense
// export and the pointer SVO.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle.h"

#define REQUIRE(c)                                                      \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      std::exit(3);                                                     \
    }                                                                   \
  } while (0)

static void primitives() {
  for (uint64_t n : {0ull, 1ull, 2ull, 4097ull, 100000ull}) {
    std::vector<uint32_t> a(n), o(n);
    vxo_gen_u32(7, 3, n, 0, 0xFFFFFFFFu, a.data());
    vxo_exclusive_sum_u32(a.data(), n, o.data());
    vxo_inclusive_sum_u32(a.data(), n, o.data());
    vxo_exclusive_max_u32(a.data(), n, o.data());
    vxo_inclusive_max_u32(a.data(), n, o.data());
    std::vector<uint64_t> b(n), p(n);
    for (uint64_t i = 0; i < n; ++i) b[i] = (uint64_t(a[i]) << 32) | a[n - 1 - i];
    vxo_exclusive_sum_u64(b.data(), n, p.data());
    vxo_inclusive_sum_u64(b.data(), n, p.data());
    vxo_exclusive_max_u64(b.data(), n, p.data());
    vxo_inclusive_max_u64(b.data(), n, p.data());
    if (n) REQUIRE(p[n - 1] >= b[0]);

    std::vector<uint32_t> k(a), v(n);
    for (uint64_t i = 0; i < n; ++i) v[i] = uint32_t(i);
    vxo_sort_pairs_u32(k.data(), v.data(), n, 32);
    for (uint64_t i = 1; i < n; ++i) REQUIRE(k[i - 1] <= k[i]);
    k = a;
    vxo_sort_pairs_u32(k.data(), v.data(), n, 5);
    std::vector<uint64_t> m(n);
    vxo_gen_morton(9, 1, n, 14, m.data());
    vxo_sort_pairs_u64(m.data(), v.data(), n, 42);
    for (uint64_t i = 1; i < n; ++i) REQUIRE(m[i - 1] <= m[i]);
    std::vector<uint32_t> words(3 * n), order(n);
    vxo_gen_u32(5, 1, 3 * n, 0, 15, words.data());
    for (uint32_t bits : {96u, 70u, 33u, 4u}) vxo_sort_order(words.data(), 3, n, bits, order.data());

    std::vector<uint64_t> keys(n), q(2 * n);
    vxo_gen_unique_u64(3, 4, n, keys.data());
    for (uint64_t i = 0; i < n; ++i) q[i] = keys[i];
    vxo_gen_unique_u64(3, 5, n, q.data() + n);
    std::vector<uint8_t> hit(2 * n);
    vxo_set_contains_u64(keys.data(), n, q.data(), 2 * n, hit.data());
    for (uint64_t i = 0; i < n; ++i) REQUIRE(hit[i] == 1);
  }
  REQUIRE(vxo_morton3(1, 0, 0) == 1 && vxo_morton3(0, 1, 0) == 2 && vxo_morton3(0, 0, 1) == 4);
  REQUIRE(vxo_mix64(0) != 0 && vxo_rand(1, 2, 3) == vxo_rand(1, 2, 3));
}

static std::vector<uint32_t> export_of(const vxo_dag* d) {
  const uint64_t n = vxo_dag_export(d, nullptr, 0);
  std::vector<uint32_t> out(n);
  REQUIRE(vxo_dag_export(d, out.data(), n) == n);
  return out;
}

static void dags() {
  const vxo_terrain t{11, 96, 24, 32, 4};
  std::vector<uint32_t> xs(2000), zs(2000);
  std::vector<int32_t> hs(2000);
  vxo_gen_u32(2, 9, 2000, 0, 255, xs.data());
  vxo_gen_u32(3, 9, 2000, 0, 255, zs.data());
  vxo_terrain_heights(&t, xs.data(), zs.data(), 2000, hs.data());
  REQUIRE(hs[0] == vxo_terrain_height(&t, xs[0], zs[0]));
  for (int materials = 0; materials < 2; ++materials) {
    REQUIRE(vxo_dag_create(0, materials) == nullptr && vxo_dag_create(15, materials) == nullptr);
    for (uint32_t depth : {1u, 2u, 6u}) {
      const uint32_t world = 4u << depth;
      vxo_dag* d = vxo_dag_create(depth, materials);
      REQUIRE(d);
      vxo_dag_build_terrain(d, &t);
      std::vector<vxo_brush> br(64);
      for (uint32_t mix = 0; mix < 2; ++mix) {
        vxo_gen_brushes(21 + mix, 5, 64, world, mix, br.data());
        for (const vxo_brush& b : br) (void)vxo_brush_voxels(&b, world);
        vxo_dag_apply_brushes(d, br.data(), 64);
      }
      std::vector<uint32_t> xyz(3 * 500);
      vxo_gen_u32(8, 1, xyz.size(), 0, world + 3, xyz.data());  // some outside the world
      vxo_dag_apply_voxels(d, xyz.data(), 500, materials ? 77 : 1);
      vxo_dag_apply_voxels(d, xyz.data(), 250, 0);
      vxo_dag_box(d, 1, 1, 1, world / 2, world / 2 + 1, world - 1, materials ? 5 : 1);
      for (uint32_t i = 0; i < 500; ++i) (void)vxo_dag_get(d, xyz[3 * i] % world, xyz[3 * i + 1] % world, xyz[3 * i + 2] % world);
      std::vector<int32_t> rays(8 * 64), hits(6 * 64);
      for (int i = 0; i < 64; ++i) {
        const int32_t r[8] = {i % 7 - 2, int32_t(world) / 2, i % 5, 1 + i % 3, i % 2 ? -1 : 0, 2, 4 * int32_t(world), 0};
        for (int k = 0; k < 8; ++k) rays[8 * i + k] = r[k];
      }
      vxo_dag_raycast(d, rays.data(), 64, hits.data());
      const std::vector<uint32_t> e = export_of(d);
      REQUIRE(e.size() >= 4);
      (void)vxo_digest(e.data(), e.size());
      (void)vxo_dag_stored_nodes(d);
      (void)vxo_dag_stored_leaves(d);
      vxo_dag_destroy(d);
    }
  }
  // Dense reference export of a 16^3 world (inner depth 2).
  std::vector<uint8_t> grid(16 * 16 * 16);
  for (size_t i = 0; i < grid.size(); ++i) grid[i] = uint8_t((i * 2654435761u >> 7) % 3 == 0 ? 1 + i % 5 : 0);
  for (int materials = 0; materials < 2; ++materials) {
    const uint64_t n = vxo_dense_export(grid.data(), 2, materials, nullptr, 0);
    std::vector<uint32_t> out(n);
    REQUIRE(vxo_dense_export(grid.data(), 2, materials, out.data(), n) == n);
  }
}

static void svo() {
  vxo_svo* s = vxo_svo_create(1u << 20, 1u << 16);
  REQUIRE(s);
  std::vector<uint32_t> xyz(3 * 20000);
  vxo_svo_random_edits(20000, xyz.data());
  std::vector<uint32_t> masks(3 * 20000);
  const uint32_t nm = vxo_svo_leaf_masks(xyz.data(), 20000, masks.data());
  REQUIRE(nm > 0 && nm <= 20000);
  REQUIRE(vxo_svo_apply(s, xyz.data(), 20000, 0) == 0);
  for (int i = 0; i < 1000; ++i) (void)vxo_svo_get(s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  REQUIRE(vxo_svo_apply(s, xyz.data(), 10000, 1) == 0);
  uint32_t c[4];
  vxo_svo_counters(s, c);
  const uint64_t n = vxo_svo_export(s, nullptr, 0);
  std::vector<uint32_t> out(n);
  REQUIRE(vxo_svo_export(s, out.data(), n) == n);
  (void)vxo_svo_leaf_key(1, 2, 3);
  vxo_svo* tiny = vxo_svo_create(4, 2);  // too small: the batch must change nothing
  REQUIRE(vxo_svo_apply(tiny, xyz.data(), 20000, 0) == 2);
  vxo_svo_reset(s);
  vxo_svo_destroy(tiny);
  vxo_svo_destroy(s);
}

int main() {
  primitives();
  dags();
  svo();
  std::printf("oracle sanitizer run: ok\n");
  return 0;
}
