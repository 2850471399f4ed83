// Caller-side kernels for the GPU tests of the content set / map device API
// (include/vx_accel_device.hpp): the form of use the reference's
// DeviceApiCanBeCalledFromCustomKernel tests exercise
// (test/cuda/test_cuda_acceleration_hash.cu:343-381, 679-729) — a kernel of
// the caller's own, launched with a partial last wave, calls find /
// insert_unique_unchecked / read_item on the table's view.
#include <hip/hip_runtime.h>

#include "vx_accel_device.hpp"

namespace {

__global__ void custom_insert(vx_accel_view t, const uint32_t* items, const uint32_t* values, uint32_t n,
                              uint32_t* ptrs, uint32_t* status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t ptr = 0;
  status[i] = vx_accel::insert_unique_unchecked(t, items + uint64_t(i) * t.words, values ? values[i] : 0u, &ptr);
  ptrs[i] = ptr;
}

__global__ void custom_find(vx_accel_view t, const uint32_t* items, uint32_t n, uint32_t* ptrs, uint32_t* values,
                            uint32_t* status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t ptr = 0, value = 0;
  status[i] = vx_accel::find(t, items + uint64_t(i) * t.words, &ptr, &value);
  ptrs[i] = ptr;
  if (values) values[i] = value;
}

__global__ void custom_read(vx_accel_view t, const uint32_t* ptrs, uint32_t n, uint32_t* items, uint32_t* ok) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t buf[vx_accel::kMaxWords];
  const bool found = vx_accel::read_item(t, ptrs[i], buf, nullptr);
  ok[i] = found ? 1u : 0u;
  for (uint32_t k = 0; k < t.words; ++k) items[uint64_t(i) * t.words + k] = found ? buf[k] : 0u;
}

}  // namespace

// Launches of 96-thread blocks (one full and one partial wave per block).
extern "C" int vxt_custom_insert(const vx_accel_view* t, const uint32_t* items, const uint32_t* values, uint32_t n,
                                 uint32_t* ptrs, uint32_t* status) {
  custom_insert<<<(n + 95) / 96, 96>>>(*t, items, values, n, ptrs, status);
  return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

extern "C" int vxt_custom_find(const vx_accel_view* t, const uint32_t* items, uint32_t n, uint32_t* ptrs,
                               uint32_t* values, uint32_t* status) {
  custom_find<<<(n + 95) / 96, 96>>>(*t, items, n, ptrs, values, status);
  return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

extern "C" int vxt_custom_read(const vx_accel_view* t, const uint32_t* ptrs, uint32_t n, uint32_t* items,
                               uint32_t* ok) {
  custom_read<<<(n + 95) / 96, 96>>>(*t, ptrs, n, items, ok);
  return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
