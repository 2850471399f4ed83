// Device-side generators of the synthetic workloads. Each element is a pure
// function of (seed, stream, index) through a counter-based splitmix64 stream,
// so the host oracle can regenerate any element independently.
#include "vx_common.hpp"

namespace vx {
namespace {

__global__ void gen_u32_kernel(uint64_t base, uint32_t n, uint32_t lo, uint32_t hi, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = uniform_u32(stream_at(base, i), lo, hi);
}

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void gen_morton_kernel(uint64_t base, uint32_t n, uint32_t bits, uint64_t* out) {
  const uint64_t m = (uint64_t(1) << bits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t r = stream_at(base, i);
    out[i] = spread3(uint32_t(r & m)) | (spread3(uint32_t((r >> bits) & m)) << 1) |
             (spread3(uint32_t((r >> (2 * bits)) & m)) << 2);
  }
}

// i -> mix64(base + i*gamma) is a bijection, so draws at distinct indices are
// distinct; the reserved key ~0 is replaced by a draw from index 2^63 + i.
__global__ void gen_unique_kernel(uint64_t base, uint64_t first, uint32_t n, uint64_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t idx = first + i;
    uint64_t k = stream_at(base, idx);
    if (k == ~0ull) k = stream_at(base, (1ull << 63) + idx);
    out[i] = k;
  }
}

__global__ void iota_kernel(uint32_t* out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = i;
}

uint32_t grid_for(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(n, 256), 16384)); }

}  // namespace
}  // namespace vx

extern "C" vx_status vx_gen_u32(uint64_t seed, uint64_t stream_id, uint32_t n, uint32_t lo, uint32_t hi, uint32_t* out,
                                void* stream) {
  if ((!out && n) || lo > hi) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  vx::gen_u32_kernel<<<vx::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(vx::stream_base(seed, stream_id), n,
                                                                                    lo, hi, out);
  return vx::last_launch();
}

extern "C" vx_status vx_gen_morton(uint64_t seed, uint64_t stream_id, uint32_t n, uint32_t coord_bits, uint64_t* out,
                                   void* stream) {
  if ((!out && n) || coord_bits == 0 || coord_bits > 21) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  vx::gen_morton_kernel<<<vx::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(vx::stream_base(seed, stream_id),
                                                                                       n, coord_bits, out);
  return vx::last_launch();
}

extern "C" vx_status vx_gen_unique_u64(uint64_t seed, uint64_t stream_id, uint64_t first, uint32_t n, uint64_t* out,
                                       void* stream) {
  if (!out && n) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  vx::gen_unique_kernel<<<vx::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(vx::stream_base(seed, stream_id),
                                                                                       first, n, out);
  return vx::last_launch();
}

extern "C" vx_status vx_iota_u32(uint32_t* out, uint32_t n, void* stream) {
  if (!out && n) return VX_INVALID_VALUE;
  if (!n) return VX_OK;
  vx::iota_kernel<<<vx::grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(out, n);
  return vx::last_launch();
}

extern "C" const char* vx_version(void) { return "voxel_amd 0.1 gfx950"; }

extern "C" vx_status vx_device_info(int32_t* cu_count, int32_t* lds_bytes_per_block, int64_t* hbm_bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return VX_DEVICE_ERROR;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return VX_DEVICE_ERROR;
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (lds_bytes_per_block) *lds_bytes_per_block = int32_t(p.sharedMemPerBlock);
  if (hbm_bytes) *hbm_bytes = int64_t(p.totalGlobalMem);
  return VX_OK;
}
