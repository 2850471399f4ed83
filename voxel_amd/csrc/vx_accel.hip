// Content-addressed set / map of multi-word items (the standalone form of
// the interning tables the Hash DAG uses): replaces AccelerationHashSet32<W>
// and AccelerationHashMap32<W> (cuda/hash_table/acceleration_hash/set/host.cuh:15-117,
// map/host.cuh:15-116) and their batch kernels (set/kernels.cuh, map/kernels.cuh).
//
// The reference chains 128-byte slabs of 31 items behind each bucket, with a
// 32-bit "acceleration hash" per slot filtering the full compares, walked by
// a 32-lane warp. Here every lane handles its own item against an
// open-addressing table of 64-bit slots {fingerprint:32 | item index + 1:32}
// kept at most half full (the fingerprint plays the acceleration hash's
// role), over a word pool of items: a lookup reads a short run of adjacent
// slots (one or two cache lines) and then only the items whose fingerprint
// matches. The device functions are in include/vx_accel_device.hpp, so
// callers' kernels use the same code. Capacity is the reference's item
// count, (buckets + overflow slabs) * 31 items; the insert that would exceed
// it reports Overflow (acceleration_hash/common.cuh:21-28 status codes).
#include <new>

#include "../../include/vx_accel_device.hpp"
#include "vx_common.hpp"

namespace {

constexpr uint32_t kItemsPerSlab = 31;

__global__ void accel_clear_kernel(unsigned long long* slots, uint64_t n, uint32_t* used) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) slots[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *used = 0;
}

__global__ void accel_find_kernel(vx_accel_view t, const uint32_t* __restrict__ items, uint32_t n, uint32_t* ptrs,
                                  uint32_t* values, uint32_t* status) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t ptr = 0, value = 0;
    const uint32_t st = vx_accel::find(t, items + uint64_t(i) * t.words, &ptr, &value);
    if (ptrs) ptrs[i] = ptr;
    if (values) values[i] = value;
    if (status) status[i] = st;
  }
}

__global__ void accel_insert_kernel(vx_accel_view t, const uint32_t* __restrict__ items,
                                    const uint32_t* __restrict__ values, uint32_t n, uint32_t* ptrs,
                                    uint32_t* status) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t ptr = 0;
    const uint32_t st =
        vx_accel::insert_unique_unchecked(t, items + uint64_t(i) * t.words, values ? values[i] : 0u, &ptr);
    if (ptrs) ptrs[i] = ptr;
    if (status) status[i] = st;
  }
}

__global__ void accel_read_kernel(vx_accel_view t, const uint32_t* __restrict__ ptrs, uint32_t n, uint32_t* items,
                                  uint32_t* values, uint32_t* ok) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t value = 0;
    uint32_t buf[vx_accel::kMaxWords];
    const bool found = vx_accel::read_item(t, ptrs[i], buf, &value);
    if (items)
      for (uint32_t k = 0; k < t.words; ++k) items[uint64_t(i) * t.words + k] = found ? buf[k] : 0u;
    if (values) values[i] = found ? value : 0u;
    if (ok) ok[i] = found ? 1u : 0u;
  }
}

uint32_t grid_of(uint64_t n) { return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192))); }

}  // namespace

// One owner type for both the set and the map (a map stores a value per item).
struct vx_accel_table {
  vx_accel_view v{};
  uint32_t buckets = 0, overflow_slabs = 0;
  bool map = false;
};
struct vx_accel_set : vx_accel_table {};
struct vx_accel_map : vx_accel_table {};

namespace {

template <class T>
vx_status create(uint32_t words, uint32_t buckets, uint32_t overflow_slabs, bool map, T** out) {
  if (!out || words == 0 || words > vx_accel::kMaxWords || buckets == 0) return VX_INVALID_VALUE;
  const uint64_t slabs = uint64_t(buckets) + overflow_slabs;
  const uint64_t capacity = slabs * kItemsPerSlab;
  if (capacity >= (1ull << 31)) return VX_INVALID_VALUE;
  uint64_t slots = 64;
  while (slots < 2 * capacity) slots <<= 1;
  T* t = new (std::nothrow) T();
  if (!t) return VX_DEVICE_ERROR;
  t->buckets = buckets;
  t->overflow_slabs = overflow_slabs;
  t->map = map;
  t->v.slot_mask = slots - 1;
  t->v.capacity = uint32_t(capacity);
  t->v.words = words;
  bool ok = hipMalloc(&t->v.slots, slots * 8) == hipSuccess &&
            hipMalloc(&t->v.items, capacity * words * 4) == hipSuccess &&
            hipMalloc(&t->v.used, 256) == hipSuccess && (!map || hipMalloc(&t->v.values, capacity * 4) == hipSuccess);
  if (ok) {
    accel_clear_kernel<<<grid_of(slots), 256>>>(t->v.slots, slots, t->v.used);
    ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  }
  if (!ok) {
    if (t->v.slots) (void)hipFree(t->v.slots);
    if (t->v.items) (void)hipFree(t->v.items);
    if (t->v.used) (void)hipFree(t->v.used);
    if (t->v.values) (void)hipFree(t->v.values);
    delete t;
    return VX_OUT_OF_CAPACITY;
  }
  *out = t;
  return VX_OK;
}

vx_status destroy(vx_accel_table* t) {
  if (!t) return VX_INVALID_VALUE;
  (void)hipDeviceSynchronize();
  if (t->v.slots) (void)hipFree(t->v.slots);
  if (t->v.items) (void)hipFree(t->v.items);
  if (t->v.used) (void)hipFree(t->v.used);
  if (t->v.values) (void)hipFree(t->v.values);
  return VX_OK;
}

vx_status clear(vx_accel_table* t, void* stream) {
  if (!t) return VX_INVALID_VALUE;
  accel_clear_kernel<<<grid_of(t->v.slot_mask + 1), 256, 0, static_cast<hipStream_t>(stream)>>>(
      t->v.slots, t->v.slot_mask + 1, t->v.used);
  return vx::last_launch();
}

vx_status find(const vx_accel_table* t, const uint32_t* items, uint32_t n, uint32_t* ptrs, uint32_t* values,
               uint32_t* status, void* stream) {
  if (!t || (n && !items)) return VX_INVALID_VALUE;
  if (n == 0) return VX_OK;
  accel_find_kernel<<<grid_of(n), 256, 0, static_cast<hipStream_t>(stream)>>>(t->v, items, n, ptrs, values, status);
  return vx::last_launch();
}

vx_status insert(vx_accel_table* t, const uint32_t* items, const uint32_t* values, uint32_t n, uint32_t* ptrs,
                 uint32_t* status, void* stream) {
  if (!t || (n && !items) || (t->map && n && !values)) return VX_INVALID_VALUE;
  if (n == 0) return VX_OK;
  accel_insert_kernel<<<grid_of(n), 256, 0, static_cast<hipStream_t>(stream)>>>(t->v, items, values, n, ptrs, status);
  return vx::last_launch();
}

vx_status read(const vx_accel_table* t, const uint32_t* ptrs, uint32_t n, uint32_t* items, uint32_t* values,
               uint32_t* ok, void* stream) {
  if (!t || (n && !ptrs)) return VX_INVALID_VALUE;
  if (n == 0) return VX_OK;
  accel_read_kernel<<<grid_of(n), 256, 0, static_cast<hipStream_t>(stream)>>>(t->v, ptrs, n, items, values, ok);
  return vx::last_launch();
}

vx_status used(const vx_accel_table* t, uint32_t* out) {
  if (!t || !out) return VX_INVALID_VALUE;
  uint32_t u = 0;
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&u, t->v.used, 4, hipMemcpyDeviceToHost) != hipSuccess)
    return VX_DEVICE_ERROR;
  *out = std::min(u, t->v.capacity);
  return VX_OK;
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------
extern "C" vx_status vx_accel_set_create(uint32_t item_words, uint32_t buckets, uint32_t overflow_slabs,
                                         vx_accel_set** out) {
  return create(item_words, buckets, overflow_slabs, false, out);
}
extern "C" vx_status vx_accel_set_destroy(vx_accel_set* t) {
  const vx_status st = destroy(t);
  delete t;
  return st;
}
extern "C" vx_status vx_accel_set_clear(vx_accel_set* t, void* stream) { return clear(t, stream); }
extern "C" vx_status vx_accel_set_find(const vx_accel_set* t, const uint32_t* d_items, uint32_t n, uint32_t* d_ptrs,
                                       uint32_t* d_status, void* stream) {
  return find(t, d_items, n, d_ptrs, nullptr, d_status, stream);
}
extern "C" vx_status vx_accel_set_insert_unique_unchecked(vx_accel_set* t, const uint32_t* d_items, uint32_t n,
                                                          uint32_t* d_ptrs, uint32_t* d_status, void* stream) {
  return insert(t, d_items, nullptr, n, d_ptrs, d_status, stream);
}
extern "C" vx_status vx_accel_set_read(const vx_accel_set* t, const uint32_t* d_ptrs, uint32_t n, uint32_t* d_items,
                                       uint32_t* d_ok, void* stream) {
  return read(t, d_ptrs, n, d_items, nullptr, d_ok, stream);
}
extern "C" vx_status vx_accel_set_view(const vx_accel_set* t, vx_accel_view* out) {
  if (!t || !out) return VX_INVALID_VALUE;
  *out = t->v;
  return VX_OK;
}
extern "C" vx_status vx_accel_set_used(const vx_accel_set* t, uint32_t* items) { return used(t, items); }

extern "C" vx_status vx_accel_map_create(uint32_t item_words, uint32_t buckets, uint32_t overflow_slabs,
                                         vx_accel_map** out) {
  return create(item_words, buckets, overflow_slabs, true, out);
}
extern "C" vx_status vx_accel_map_destroy(vx_accel_map* t) {
  const vx_status st = destroy(t);
  delete t;
  return st;
}
extern "C" vx_status vx_accel_map_clear(vx_accel_map* t, void* stream) { return clear(t, stream); }
extern "C" vx_status vx_accel_map_find(const vx_accel_map* t, const uint32_t* d_items, uint32_t n, uint32_t* d_ptrs,
                                       uint32_t* d_values, uint32_t* d_status, void* stream) {
  return find(t, d_items, n, d_ptrs, d_values, d_status, stream);
}
extern "C" vx_status vx_accel_map_insert_unique_unchecked(vx_accel_map* t, const uint32_t* d_items,
                                                          const uint32_t* d_values, uint32_t n, uint32_t* d_ptrs,
                                                          uint32_t* d_status, void* stream) {
  return insert(t, d_items, d_values, n, d_ptrs, d_status, stream);
}
extern "C" vx_status vx_accel_map_read(const vx_accel_map* t, const uint32_t* d_ptrs, uint32_t n, uint32_t* d_items,
                                       uint32_t* d_values, uint32_t* d_ok, void* stream) {
  return read(t, d_ptrs, n, d_items, d_values, d_ok, stream);
}
extern "C" vx_status vx_accel_map_view(const vx_accel_map* t, vx_accel_view* out) {
  if (!t || !out) return VX_INVALID_VALUE;
  *out = t->v;
  return VX_OK;
}
extern "C" vx_status vx_accel_map_used(const vx_accel_map* t, uint32_t* items) { return used(t, items); }
