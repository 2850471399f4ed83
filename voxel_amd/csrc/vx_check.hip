// Invariant checks of the content-addressed tables, on gfx950.
//
// The edit pipeline's correctness rests on one invariant, which the
// reference also relies on without checking it (insert_unique_unchecked
// requires uniqueness, cuda/hash_table/acceleration_hash/set/common.cuh:51-56;
// the per-depth dedup map guarantees it, svt/hash_dag_gpu/detail/edit/
// temp_dedup.cuh:29-34): no two live slots of a table hold items of equal
// content. Here it is checked on the device, one thread per slot: a live
// slot's item is hashed again, its fingerprint must be the slot's, every slot
// from the hash's home position up to this one must be occupied (otherwise a
// lookup could never reach it), and no earlier slot on that walk may hold an
// equal item — a duplicate always sits on the walk of the later copy, since
// both start at the same home position.
#include "vx_dag.hpp"

using namespace vx;
using namespace vx::dag;

namespace {

// out[0] live slots, out[1] duplicates, out[2] broken slots (tag left over,
// offset outside the used words, fingerprint mismatch, or unreachable).
__global__ void check_unique_kernel(const uint64_t* __restrict__ slots, uint64_t mask, const uint32_t* __restrict__ pool,
                                    uint64_t used, uint32_t item_words, unsigned long long* out) {
  unsigned long long live = 0, dups = 0, bad = 0;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; s <= mask; s += stride) {
    const uint64_t e = slots[s];
    if (e == 0) continue;
    const uint32_t v = uint32_t(e);
    if (v == kTomb) continue;
    if ((v & kTag) || v < kPoolStart || v >= used) {
      ++bad;
      continue;
    }
    const uint32_t w = item_words ? item_words : 1u + __popc(pool[v] & 0xffu);
    if (uint64_t(v) + w > used) {
      ++bad;
      continue;
    }
    const uint64_t h = hash_words(pool + v, w);
    if (uint32_t(h >> 32) != uint32_t(e >> 32)) {
      ++bad;
      continue;
    }
    ++live;
    for (uint64_t i = h & mask; i != s; i = (i + 1) & mask) {
      const uint64_t e2 = slots[i];
      if (e2 == 0) {  // a lookup stops here and never reaches slot s
        ++bad;
        break;
      }
      const uint32_t v2 = uint32_t(e2);
      if (uint32_t(e2 >> 32) != uint32_t(e >> 32) || v2 == kTomb || (v2 & kTag) || v2 < kPoolStart ||
          uint64_t(v2) + w > used)
        continue;
      const uint32_t w2 = item_words ? item_words : 1u + __popc(pool[v2] & 0xffu);
      if (w2 == w && words_equal(pool + v2, pool + v, w)) {
        ++dups;
        break;
      }
    }
  }
  if (live) atomicAdd(&out[0], live);
  if (dups) atomicAdd(&out[1], dups);
  if (bad) atomicAdd(&out[2], bad);
}

vx_status check_table(const uint64_t* slots, uint64_t n_slots, const uint32_t* pool, uint64_t used, uint32_t item_words,
                      uint64_t out[3], hipStream_t s) {
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 3 * sizeof(unsigned long long)) != hipSuccess) return VX_OUT_OF_CAPACITY;
  vx_status st = VX_OK;
  if (hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), s) != hipSuccess) st = VX_DEVICE_ERROR;
  if (st == VX_OK) {
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((n_slots + 255) / 256, 16384)));
    check_unique_kernel<<<grid, 256, 0, s>>>(slots, n_slots - 1, pool, used, item_words, d);
    st = last_launch();
  }
  unsigned long long h[3] = {0, 0, 0};
  if (st == VX_OK && (hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
                      hipStreamSynchronize(s) != hipSuccess))
    st = VX_DEVICE_ERROR;
  (void)hipFree(d);
  for (int i = 0; i < 3; ++i) out[i] = h[i];
  return st;
}

}  // namespace

extern "C" vx_status vx_table_check_unique(const uint64_t* d_slots, uint64_t n_slots, const uint32_t* d_pool,
                                           uint64_t pool_words_used, uint32_t item_words, uint64_t out[3],
                                           void* stream) {
  if (!d_slots || !d_pool || !out || n_slots == 0 || (n_slots & (n_slots - 1)) || item_words > 64 ||
      pool_words_used >= (uint64_t(1) << 31))
    return VX_INVALID_VALUE;
  return check_table(d_slots, n_slots, d_pool, pool_words_used, item_words, out, static_cast<hipStream_t>(stream));
}

extern "C" vx_status vx_dag_check_unique(const vx_dag* g, uint64_t out[6], void* stream) {
  if (!g || !out) return VX_INVALID_VALUE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t ctl[kCtlWords];
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(ctl, g->ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess)
    return VX_DEVICE_ERROR;
  const uint64_t nu = std::min<uint64_t>(ctl[kCtlNodeUsed], g->dev.node_cap);
  const uint64_t lu = std::min<uint64_t>(ctl[kCtlLeafUsed], g->dev.leaf_cap);
  vx_status st = check_table(g->node_slots, g->dev.node_mask + 1, g->node_pool, nu, 0, out, s);
  if (st != VX_OK) return st;
  return check_table(g->leaf_slots, g->dev.leaf_mask + 1, g->leaf_pool, lu, g->dev.leaf_words, out + 3, s);
}
