// Drives the C ABI of the host-sanitized build of the library
// (voxel_amd/_build/asan/libvoxel_amd_asan.so: the host half of every
// translation unit compiled with AddressSanitizer and UndefinedBehavior-
// Sanitizer, device code unchanged). Run by tests/test_sanitizers.py. The
// argument-validation paths run everywhere; with a GPU present, small real
// calls of every entry point family run too, with their results checked
// roughly (their parity is the other tests' business). A sanitizer finding
// aborts with a non-zero exit.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <vector>

#include "vx.h"

#define REQUIRE(c)                                                        \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      std::exit(3);                                                       \
    }                                                                     \
  } while (0)
#define OK(call) REQUIRE((call) == VX_OK)

template <class T>
static T* dev(size_t n) {
  void* p = nullptr;
  REQUIRE(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess);
  return static_cast<T*>(p);
}
template <class T>
static T* upload(const std::vector<T>& h) {
  T* d = dev<T>(h.size());
  REQUIRE(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
  return d;
}
template <class T>
static std::vector<T> download(const T* d, size_t n) {
  std::vector<T> h(n);
  REQUIRE(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess);
  return h;
}

static void validation() {
  size_t bytes = 0;
  REQUIRE(vx_sort_workspace_size(100, 3, 8, 1, &bytes) == VX_INVALID_VALUE);
  REQUIRE(vx_sort_workspace_size(100, 4, 40, 1, &bytes) == VX_INVALID_VALUE);
  REQUIRE(vx_sort_pairs_u32(nullptr, nullptr, 5, 32, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_sort_pairs_u64(nullptr, nullptr, 0, 0, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_sort_by_key(nullptr, 4, 5, 32, nullptr, 0, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_scan_workspace_size(100, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_exclusive_sum_u32(nullptr, 5, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_exclusive_sum_u64(nullptr, 5, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_exclusive_sum_u32_noalloc(nullptr, 5, nullptr) == VX_INVALID_VALUE);
  vx_dag* g = nullptr;
  REQUIRE(vx_dag_create(0, 0, nullptr, &g) == VX_INVALID_VALUE);
  REQUIRE(vx_dag_create(14, 0, nullptr, &g) == VX_INVALID_VALUE);
  REQUIRE(vx_dag_create(4, 0, nullptr, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_dag_destroy(nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_dag_apply_brushes(nullptr, nullptr, 0, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_dag_stats_get(nullptr, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_batcher_destroy(nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_slab_set_destroy(nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_accel_set_create(0, 4, 1, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_accel_set_create(17, 4, 1, nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_slab_alloc_destroy(nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_svo_destroy(nullptr) == VX_INVALID_VALUE);
  REQUIRE(vx_table_check_unique(nullptr, 64, nullptr, 16, 2, nullptr, nullptr) == VX_INVALID_VALUE);
  const vx_dag_capacity grow_cap{1 << 20, 1 << 20, 1 << 16, 1 << 16};
  REQUIRE(vx_dag_grow(nullptr, &grow_cap, nullptr) == VX_INVALID_VALUE);
}

static void primitives() {
  const uint32_t n = 100003;
  std::vector<uint32_t> hk(n), hv(n);
  for (uint32_t i = 0; i < n; ++i) {
    hk[i] = (i * 2654435761u) ^ (i >> 3);
    hv[i] = i;
  }
  uint32_t* k = upload(hk);
  uint32_t* v = upload(hv);
  size_t bytes = 0;
  OK(vx_sort_workspace_size(n, 4, 32, 1, &bytes));
  void* ws = dev<char>(bytes);
  OK(vx_sort_pairs_u32(k, v, n, 32, ws, bytes, nullptr));
  std::vector<uint32_t> sk = download(k, n);
  for (uint32_t i = 1; i < n; ++i) REQUIRE(sk[i - 1] <= sk[i]);
  // General sort_by_key: 12-byte keys with two value arrays.
  std::vector<uint32_t> hk3(3 * n);
  for (uint32_t i = 0; i < 3 * n; ++i) hk3[i] = (i * 747796405u) >> 7;
  uint32_t* k3 = upload(hk3);
  uint64_t* v64 = dev<uint64_t>(n);
  vx_value_array arrays[2] = {{v, 4}, {v64, 8}};
  OK(vx_sort_by_key_workspace_size(n, 12, 96, arrays, 2, &bytes));
  void* ws2 = dev<char>(bytes);
  OK(vx_sort_by_key(k3, 12, n, 96, arrays, 2, ws2, bytes, nullptr));
  // Scans u32 and u64.
  std::vector<uint32_t> ones(n, 1);
  uint32_t* a = upload(ones);
  OK(vx_scan_workspace_size(n, &bytes));
  void* ws3 = dev<char>(bytes);
  OK(vx_exclusive_sum_u32(a, n, ws3, bytes, nullptr));
  REQUIRE(download(a, n)[n - 1] == n - 1);
  OK(vx_exclusive_sum_u32_noalloc(a, n, nullptr));
  std::vector<uint64_t> big(n, uint64_t(1) << 40);
  uint64_t* b = upload(big);
  OK(vx_scan_workspace_size_u64(n, &bytes));
  void* ws4 = dev<char>(bytes);
  OK(vx_inclusive_sum_u64(b, n, ws4, bytes, nullptr));
  REQUIRE(download(b, n)[n - 1] == uint64_t(n) << 40);
  for (void* p : {static_cast<void*>(k), static_cast<void*>(v), ws, static_cast<void*>(k3), static_cast<void*>(v64),
                  ws2, static_cast<void*>(a), ws3, static_cast<void*>(b), ws4})
    REQUIRE(hipFree(p) == hipSuccess);
}

static void tables() {
  const uint32_t n = 5000;
  std::vector<uint64_t> hk(n);
  for (uint32_t i = 0; i < n; ++i) hk[i] = 1000 + 7ull * i;
  uint64_t* k = upload(hk);
  uint32_t* st = dev<uint32_t>(n);
  uint8_t* found = dev<uint8_t>(n);
  vx_slab_set* s = nullptr;
  OK(vx_slab_set_create(128, 64, &s));
  OK(vx_slab_set_insert(s, k, n, st, nullptr));
  OK(vx_slab_set_contains(s, k, n, found, st, nullptr));
  for (uint8_t f : download(found, n)) REQUIRE(f == 1);
  OK(vx_slab_set_erase(s, k, n / 2, st, nullptr));
  OK(vx_slab_set_clear(s, nullptr));
  OK(vx_slab_set_destroy(s));
  vx_cuckoo_set* c = nullptr;
  OK(vx_cuckoo_set_create(1 << 14, 64, &c));
  OK(vx_cuckoo_set_insert(c, k, n, st, nullptr));
  OK(vx_cuckoo_set_contains(c, k, n, found, nullptr));
  OK(vx_cuckoo_set_destroy(c));
  std::vector<uint32_t> items(4 * 64);
  for (uint32_t i = 0; i < items.size(); ++i) items[i] = 0x10000 + 131 * (i / 4) + 17 * (i % 4);
  uint32_t* d_items = upload(items);
  uint32_t* ptrs = dev<uint32_t>(64);
  vx_accel_set* as = nullptr;
  OK(vx_accel_set_create(4, 8, 2, &as));
  OK(vx_accel_set_insert_unique_unchecked(as, d_items, 64, ptrs, st, nullptr));
  OK(vx_accel_set_find(as, d_items, 64, ptrs, st, nullptr));
  uint32_t used = 0;
  OK(vx_accel_set_used(as, &used));
  REQUIRE(used == 64);
  OK(vx_accel_set_destroy(as));
  vx_slab_alloc* sa = nullptr;
  OK(vx_slab_alloc_create(1000, &sa));
  uint32_t* ord = dev<uint32_t>(600);
  OK(vx_slab_alloc_allocate(sa, 600, ord, nullptr));
  OK(vx_slab_alloc_deallocate(sa, ord, 300, nullptr));
  OK(vx_slab_alloc_used(sa, &used));
  REQUIRE(used == 300);
  OK(vx_slab_alloc_destroy(sa));
  for (void* p : {static_cast<void*>(k), static_cast<void*>(st), static_cast<void*>(found),
                  static_cast<void*>(d_items), static_cast<void*>(ptrs), static_cast<void*>(ord)})
    REQUIRE(hipFree(p) == hipSuccess);
}

static std::vector<vx_brush> brushes(uint32_t n, int32_t world, uint32_t seed) {
  std::vector<vx_brush> b(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = (i + 1) * 2654435761u ^ seed;
    b[i] = vx_brush{int32_t(r % uint32_t(world)), int32_t((r >> 8) % uint32_t(world)),
                    int32_t((r >> 16) % uint32_t(world)), int32_t(4 + r % 9), int32_t(4 + (r >> 4) % 9),
                    int32_t(4 + (r >> 9) % 9), (r >> 3) & 1u, (r >> 5) % 3 ? 1 + (r >> 11) % 200 : 0};
  }
  return b;
}

static void dags() {
  const uint32_t depth = 6;
  const int32_t world = 4 << depth;
  vx_dag* g[2] = {nullptr, nullptr};
  OK(vx_dag_create(depth, 1, nullptr, &g[0]));
  OK(vx_dag_create(depth - 1, 1, nullptr, &g[1]));
  const vx_terrain t{7, 100, 20, 32, 4};
  OK(vx_dag_build_terrain(g[0], &t, nullptr));
  OK(vx_dag_build_terrain(g[1], &t, nullptr));
  std::vector<vx_brush> br = brushes(64, world, 1);
  size_t bytes = 0;
  OK(vx_dag_brush_workspace_size(g[0], br.data(), 64, &bytes));
  void* ws = dev<char>(bytes);
  OK(vx_dag_apply_brushes(g[0], br.data(), 64, ws, bytes, nullptr));
  REQUIRE(hipFree(ws) == hipSuccess);
  std::vector<vx_brush> bad = br;
  bad[3].kind = 7;
  REQUIRE(vx_dag_brush_workspace_size(g[0], bad.data(), 64, &bytes) == VX_INVALID_VALUE ||
          vx_dag_apply_brushes(g[0], bad.data(), 64, nullptr, 0, nullptr) == VX_INVALID_VALUE);
  std::vector<vx_voxel> vox(1000);
  for (uint32_t i = 0; i < 1000; ++i) vox[i] = vx_voxel{(i * 37) % 300, (i * 11) % 256, (i * 91) % 256};
  vx_voxel* dv = upload(vox);
  OK(vx_dag_voxel_edits_workspace_size(g[0], 1000, &bytes));
  ws = dev<char>(bytes);
  OK(vx_dag_apply_voxel_edits(g[0], dv, 1000, 0, 42, ws, bytes, nullptr));
  OK(vx_dag_apply_voxel_edits(g[0], dv, 500, 1, 0, ws, bytes, nullptr));
  REQUIRE(hipFree(ws) == hipSuccess);
  uint32_t* got = dev<uint32_t>(1000);
  OK(vx_dag_get_voxels(g[0], dv, 1000, got, nullptr));
  std::vector<vx_ray> rays(64);
  for (int i = 0; i < 64; ++i) rays[i] = vx_ray{i, world / 2, i % 5, 1, i % 2 ? -1 : 0, 2, uint32_t(4 * world), 0};
  vx_ray* dr = upload(rays);
  vx_hit* dh = dev<vx_hit>(64);
  OK(vx_dag_raycast(g[0], dr, 64, dh, nullptr));
  uint64_t words = 0;
  OK(vx_dag_export_canonical(g[0], nullptr, 0, &words));
  std::vector<uint32_t> e(words);
  OK(vx_dag_export_canonical(g[0], e.data(), words, &words));
  uint32_t* de = dev<uint32_t>(words);
  OK(vx_dag_export_canonical_device(g[0], de, words, &words, nullptr));
  REQUIRE(download(de, words) == e);
  uint64_t check[6];
  OK(vx_dag_check_unique(g[0], check, nullptr));
  REQUIRE(check[1] == 0 && check[2] == 0 && check[4] == 0 && check[5] == 0);
  // Two scenes in one executor, run and submit / query.
  vx_batcher* b = nullptr;
  OK(vx_batcher_create(g, 2, 1 << 18, 256, 1, &b));
  std::vector<vx_brush> b0 = brushes(20, world, 2), b1 = brushes(20, world / 2, 3);
  const vx_brush* hb[2] = {b0.data(), b1.data()};
  const uint32_t scene_of[2] = {0, 1}, counts[2] = {20, 20};
  uint64_t cand = 0;
  OK(vx_batcher_run(b, scene_of, hb, counts, 2, &cand));
  OK(vx_batcher_submit(b, scene_of, hb, counts, 2, &cand, nullptr));
  int32_t done = 0;
  uint32_t status[2];
  uint64_t used[4];
  OK(vx_batcher_query(b, 1, &done, status, used));
  REQUIRE(done == 1);
  uint32_t key_bits = 0, index_bits = 0;
  OK(vx_batcher_layout(b, &key_bits, &index_bits));
  OK(vx_batcher_destroy(b));
  OK(vx_dag_export_canonical(g[0], nullptr, 0, &words));
  std::vector<uint32_t> before(words);
  OK(vx_dag_export_canonical(g[0], before.data(), words, &words));
  vx_gc_report rep{};
  OK(vx_dag_collect_garbage(g[0], &rep, nullptr));
  std::vector<uint32_t> after(words);
  OK(vx_dag_export_canonical(g[0], after.data(), words, &words));
  REQUIRE(after.size() == words && after == before);
  vx_dag_stats stats{};
  OK(vx_dag_stats_get(g[0], &stats));
  // Growth: pools and tables twice as large, the same canonical DAG.
  const vx_dag_capacity bigger{2 * (stats.node_words_used + 4096), 2 * (stats.leaf_words_used + 4096),
                               2 * stats.node_slots, 2 * stats.leaf_slots};
  REQUIRE(vx_dag_grow(g[0], nullptr, nullptr) == VX_INVALID_VALUE);
  OK(vx_dag_grow(g[0], &bigger, nullptr));
  OK(vx_dag_stats_get(g[0], &stats));
  REQUIRE(stats.node_slots == bigger.node_slots && stats.leaf_slots == bigger.leaf_slots);
  OK(vx_dag_export_canonical(g[0], after.data(), words, &words));
  REQUIRE(after == before);
  uint64_t nn = 0, ln = 0;
  OK(vx_dag_read_pools(g[0], nullptr, 0, &nn, nullptr, 0, &ln));
  for (void* p : {static_cast<void*>(dv), static_cast<void*>(got), static_cast<void*>(dr), static_cast<void*>(dh),
                  static_cast<void*>(de)})
    REQUIRE(hipFree(p) == hipSuccess);
  OK(vx_dag_destroy(g[0]));
  OK(vx_dag_destroy(g[1]));
  // Pointer SVO.
  vx_svo* s = nullptr;
  OK(vx_svo_create(1 << 16, 1 << 14, &s));
  OK(vx_svo_reset(s, nullptr));
  std::vector<vx_voxel> sv(4000);
  for (uint32_t i = 0; i < 4000; ++i) sv[i] = vx_voxel{(i * 977) % 4096, (i * 131) % 4096, (i * 61) % 4096};
  vx_voxel* ds = upload(sv);
  OK(vx_svo_voxel_edits_workspace_size(4000, &bytes));
  ws = dev<char>(bytes);
  OK(vx_svo_apply_voxel_edits(s, ds, 4000, 0, ws, bytes, nullptr));
  OK(vx_svo_apply_voxel_edits(s, ds, 2000, 1, ws, bytes, nullptr));
  REQUIRE(hipFree(ws) == hipSuccess && hipFree(ds) == hipSuccess);
  OK(vx_svo_destroy(s));
}

int main() {
  validation();
  int devices = 0;
  if (hipGetDeviceCount(&devices) == hipSuccess && devices > 0) {
    primitives();
    tables();
    dags();
    std::printf("library host code under sanitizers: ok (validation and device calls)\n");
  } else {
    std::printf("library host code under sanitizers: ok (validation only, no device)\n");
  }
  return 0;
}
