// Device-side generators of the synthetic workloads. Each element is a pure
// function of (seed, stream, index) through a counter-based splitmix64 stream,
// so the host oracle can regenerate any element independently.
#include <cmath>

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

// Brush batches of the edit workloads (host side; the batch is small).
// Brush i draws r_k = stream_at(base, 16 i + k): k=0 kind, 1..3 centre,
// 4..6 extents, 7 add-vs-remove, 8 material. Mix 0: spheres r in [4,16], 50%
// add. Mix 1: 50% spheres r in [8,64], 50% boxes with half extents in [4,48],
// 60% add. Added material uniform in [1,255]; removal writes material 0.
extern "C" vx_status vx_gen_brushes(uint64_t seed, uint64_t stream_id, uint32_t n, uint32_t world, uint32_t mix,
                                    vx_brush* out) {
  if ((!out && n) || world == 0 || mix > 1) return VX_INVALID_VALUE;
  const uint64_t base = vx::stream_base(seed, stream_id);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t r[9];
    for (uint32_t k = 0; k < 9; ++k) r[k] = vx::stream_at(base, uint64_t(i) * 16 + k);
    vx_brush b{};
    b.cx = int32_t(vx::uniform_u32(r[1], 0, world - 1));
    b.cy = int32_t(vx::uniform_u32(r[2], 0, world - 1));
    b.cz = int32_t(vx::uniform_u32(r[3], 0, world - 1));
    uint32_t add_pct;
    if (mix == 0) {
      b.kind = 0;
      b.hx = int32_t(vx::uniform_u32(r[4], 4, 16));
      add_pct = 50;
    } else {
      b.kind = vx::uniform_u32(r[0], 0, 1);
      if (b.kind == 0) {
        b.hx = int32_t(vx::uniform_u32(r[4], 8, 64));
      } else {
        b.hx = int32_t(vx::uniform_u32(r[4], 4, 48));
        b.hy = int32_t(vx::uniform_u32(r[5], 4, 48));
        b.hz = int32_t(vx::uniform_u32(r[6], 4, 48));
      }
      add_pct = 60;
    }
    const bool add = vx::uniform_u32(r[7], 0, 99) < add_pct;
    b.material = add ? vx::uniform_u32(r[8], 1, 255) : 0u;
    out[i] = b;
  }
  return VX_OK;
}

// Brush batch sized to a voxel-op target (the batched and serving workloads).
// Batch `index` of (seed, stream_id) draws t = uniform [lo_log2, hi_log2] in
// 1/1024 steps, target = 2^t voxel ops, n = uniform [1, max_brushes] brushes of
// volume target / n each: 50% spheres of radius round(cbrt(3V / 4 pi)), 50%
// cubes of half extent round((cbrt(V) - 1) / 2) (both at least 1), centres
// uniform in the world, 60% add with material uniform [1,255], 40% remove.
// Brush i of the batch uses draws 64 + 8 i .. 64 + 8 i + 5; *scene_out gets
// uniform [0, n_scenes - 1] from draw 2.
extern "C" vx_status vx_gen_sized_batch(uint64_t seed, uint64_t stream_id, uint64_t index, uint32_t lo_log2,
                                        uint32_t hi_log2, uint32_t max_brushes, uint32_t world, uint32_t n_scenes,
                                        vx_brush* out, uint32_t* n_out, uint32_t* scene_out) {
  if (!out || !n_out || !max_brushes || world == 0 || lo_log2 > hi_log2 || hi_log2 > 40) return VX_INVALID_VALUE;
  const uint64_t base = vx::stream_base(seed ^ (index * 0x9E3779B97F4A7C15ull), stream_id);
  const uint32_t tq = vx::uniform_u32(vx::stream_at(base, 0), lo_log2 * 1024, hi_log2 * 1024);
  const double target = std::exp2(double(tq) / 1024.0);
  const uint32_t n = vx::uniform_u32(vx::stream_at(base, 1), 1, max_brushes);
  if (scene_out) *scene_out = n_scenes ? vx::uniform_u32(vx::stream_at(base, 2), 0, n_scenes - 1) : 0;
  const double vol = target / n;
  const int32_t r = std::max<int32_t>(1, int32_t(std::lround(std::cbrt(3.0 * vol / (4.0 * 3.141592653589793)))));
  const int32_t h = std::max<int32_t>(1, int32_t(std::lround((std::cbrt(vol) - 1.0) / 2.0)));
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t d[6];
    for (uint32_t k = 0; k < 6; ++k) d[k] = vx::stream_at(base, 64 + uint64_t(i) * 8 + k);
    vx_brush b{};
    b.kind = vx::uniform_u32(d[0], 0, 1);
    b.cx = int32_t(vx::uniform_u32(d[1], 0, world - 1));
    b.cy = int32_t(vx::uniform_u32(d[2], 0, world - 1));
    b.cz = int32_t(vx::uniform_u32(d[3], 0, world - 1));
    if (b.kind == 0) {
      b.hx = r;
    } else {
      b.hx = b.hy = b.hz = h;
    }
    b.material = vx::uniform_u32(d[4], 0, 99) < 60 ? vx::uniform_u32(d[5], 1, 255) : 0u;
    out[i] = b;
  }
  *n_out = n;
  return VX_OK;
}

// Random voxel edits of the pointer-SVO benchmark (bench/bench_cuda_svt.cu:
// 90-112): one xorshift32 state (13, 17, 5) seeded 0x9e3779b9; voxel i takes
// three consecutive draws, each reduced mod `world`, as x, y, z (host output;
// the stream is sequential by construction).
extern "C" vx_status vx_gen_xorshift_voxels(uint32_t count, uint32_t world, vx_voxel* out) {
  if ((!out && count) || world == 0) return VX_INVALID_VALUE;
  uint32_t s = 0x9e3779b9u;
  auto next = [&s]() {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
  };
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t x = next() % world;
    const uint32_t y = next() % world;
    const uint32_t z = next() % world;
    out[i] = vx_voxel{x, y, z};
  }
  return VX_OK;
}
