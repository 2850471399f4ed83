// sort_by_key for every key and value layout of the reference's API:
// keys of 1..16 u32 words (u32, u64 and UIntKey<W>, words[0] least
// significant, cuda/sort/uint_key.cuh:15-66) and any number of value arrays
// of any trivially copyable element (value_arrays_t, cuda/sort/value_arrays.cuh:12-140).
//
// The onesweep partition moves (key, u32) pairs; every other payload rides on
// a permutation instead of being copied in each radix pass:
//   u32 / u64 keys: the keys sort in place with the record index as the
//     value, then each value array is gathered through the sorted indices;
//   wider keys: LSD over the 32-bit words, least significant first — each
//     word of the current order is gathered and a stable (word, index) sort
//     refines the order — then the keys and the arrays are gathered.
// Gathers read one element per index (random reads of whole elements) and
// write coalesced; each result is copied back in place.
#include "vx_scan.hpp"
#include "vx_sort.hpp"

namespace vx {
namespace sort {
namespace {

constexpr uint32_t kMaxKeyWords = 16;

// dst[i] = src[perm[i]] for elements of k units of type T.
template <class T>
__global__ void gather_kernel(const T* __restrict__ src, const uint32_t* __restrict__ perm, T* __restrict__ dst,
                              uint64_t n_units, uint32_t k) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < n_units; u += stride) {
    if (k == 1) {
      dst[u] = src[perm[u]];
    } else {
      const uint64_t i = u / k, j = u - i * k;
      dst[u] = src[uint64_t(perm[i]) * k + j];
    }
  }
}

// Word w of the key of record perm[i] (keys of W words).
__global__ void key_word_kernel(const uint32_t* __restrict__ keys, uint32_t words, uint32_t w,
                                const uint32_t* __restrict__ perm, uint32_t* __restrict__ out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = keys[uint64_t(perm[i]) * words + w];
}

template <class T>
vx_status gather_as(const void* src, const uint32_t* perm, void* dst, uint32_t n, uint64_t elem, hipStream_t s) {
  const uint32_t k = uint32_t(elem / sizeof(T));
  const uint64_t units = uint64_t(n) * k;
  const uint32_t blocks = uint32_t(std::min<uint64_t>(ceil_div(units, 256), 8192));
  gather_kernel<T><<<blocks, 256, 0, s>>>(static_cast<const T*>(src), perm, static_cast<T*>(dst), units, k);
  return last_launch();
}

// The widest unit that divides the element size and the alignment of both
// buffers.
vx_status gather(const void* src, const uint32_t* perm, void* dst, uint32_t n, uint64_t elem, hipStream_t s) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst) | uintptr_t(elem);
  if ((a & 15) == 0) return gather_as<uint4>(src, perm, dst, n, elem, s);
  if ((a & 7) == 0) return gather_as<uint64_t>(src, perm, dst, n, elem, s);
  if ((a & 3) == 0) return gather_as<uint32_t>(src, perm, dst, n, elem, s);
  if ((a & 1) == 0) return gather_as<uint16_t>(src, perm, dst, n, elem, s);
  return gather_as<uint8_t>(src, perm, dst, n, elem, s);
}

struct Plan {
  uint32_t words;      // key words
  bool native;         // u32 / u64 keys: sorted in place with the index as value
  uint32_t* perm;      // n
  uint32_t* word;      // n (wide keys)
  void* sort_ws;
  size_t sort_bytes;
  void* tmp;           // n * widest gathered element
  size_t bytes;
};

Plan plan(void* ws, uint32_t n, uint32_t key_bytes, uint32_t key_bits, const vx_value_array* arrays,
          uint32_t n_arrays) {
  Plan p{};
  p.words = key_bytes / 4;
  p.native = key_bytes == 4 || key_bytes == 8;
  uint64_t widest = p.native ? 0 : key_bytes;
  for (uint32_t i = 0; i < n_arrays; ++i) widest = std::max<uint64_t>(widest, arrays[i].elem_bytes);
  Carver c(ws);
  p.perm = c.take<uint32_t>(n);
  if (p.native) {
    p.sort_bytes = workspace_bytes(n, key_bytes, key_bits, true);
  } else {
    p.word = c.take<uint32_t>(n);
    p.sort_bytes = workspace_bytes(n, 4, 32, true);
  }
  p.sort_ws = c.take<char>(p.sort_bytes);
  p.tmp = c.take<char>(size_t(n) * widest);
  p.bytes = c.size();
  return p;
}

bool fast_path(uint32_t key_bytes, const vx_value_array* arrays, uint32_t n_arrays) {
  return (key_bytes == 4 || key_bytes == 8) && (n_arrays == 0 || (n_arrays == 1 && arrays[0].elem_bytes == 4));
}

vx_status copy_back(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
  return from_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

}  // namespace

size_t by_key_workspace_bytes(uint32_t n, uint32_t key_bytes, uint32_t key_bits, const vx_value_array* arrays,
                              uint32_t n_arrays) {
  if (n <= 1) return 0;
  if (fast_path(key_bytes, arrays, n_arrays)) return workspace_bytes(n, key_bytes, key_bits, n_arrays == 1);
  return plan(nullptr, n, key_bytes, key_bits, arrays, n_arrays).bytes;
}

vx_status by_key(void* keys, uint32_t key_bytes, uint32_t n, uint32_t key_bits, const vx_value_array* arrays,
                 uint32_t n_arrays, void* ws, size_t ws_bytes, hipStream_t s) {
  if (n <= 1) return VX_OK;
  if (fast_path(key_bytes, arrays, n_arrays)) {
    uint32_t* vals = n_arrays ? static_cast<uint32_t*>(arrays[0].data) : nullptr;
    return key_bytes == 4 ? sort_u32(static_cast<uint32_t*>(keys), vals, n, key_bits, ws, ws_bytes, s)
                          : sort_u64(static_cast<uint64_t*>(keys), vals, n, key_bits, ws, ws_bytes, s);
  }
  const Plan p = plan(ws, n, key_bytes, key_bits, arrays, n_arrays);
  if (!ws || ws_bytes < p.bytes) return VX_INVALID_VALUE;
  vx_status st;
  if ((st = vx_iota_u32(p.perm, n, s)) != VX_OK) return st;
  if (p.native) {
    st = key_bytes == 4 ? sort_u32(static_cast<uint32_t*>(keys), p.perm, n, key_bits, p.sort_ws, p.sort_bytes, s)
                        : sort_u64(static_cast<uint64_t*>(keys), p.perm, n, key_bits, p.sort_ws, p.sort_bytes, s);
    if (st != VX_OK) return st;
  } else {
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 8192);
    for (uint32_t w = 0; 32 * w < key_bits; ++w) {
      key_word_kernel<<<blocks, 256, 0, s>>>(static_cast<const uint32_t*>(keys), p.words, w, p.perm, p.word, n);
      if ((st = last_launch()) != VX_OK) return st;
      const uint32_t bits = std::min(32u, key_bits - 32 * w);
      if ((st = sort_u32(p.word, p.perm, n, bits, p.sort_ws, p.sort_bytes, s)) != VX_OK) return st;
    }
    if ((st = gather(keys, p.perm, p.tmp, n, key_bytes, s)) != VX_OK) return st;
    if ((st = copy_back(keys, p.tmp, uint64_t(n) * key_bytes, s)) != VX_OK) return st;
  }
  for (uint32_t i = 0; i < n_arrays; ++i) {
    if ((st = gather(arrays[i].data, p.perm, p.tmp, n, arrays[i].elem_bytes, s)) != VX_OK) return st;
    if ((st = copy_back(arrays[i].data, p.tmp, uint64_t(n) * arrays[i].elem_bytes, s)) != VX_OK) return st;
  }
  return VX_OK;
}

}  // namespace sort
}  // namespace vx

// ---- C ABI -------------------------------------------------------------------
static vx_status check_by_key_args(const void* keys, uint32_t key_bytes, uint32_t n, uint32_t key_bits,
                                   const vx_value_array* arrays, uint32_t n_arrays) {
  if (key_bytes == 0 || key_bytes % 4 || key_bytes > 4 * vx::sort::kMaxKeyWords) return VX_INVALID_VALUE;
  if (key_bits == 0 || key_bits > 8 * key_bytes || n >= (1u << 30)) return VX_INVALID_VALUE;
  if (n_arrays && !arrays) return VX_INVALID_VALUE;
  for (uint32_t i = 0; i < n_arrays; ++i)
    if (arrays[i].elem_bytes == 0 || (n && !arrays[i].data)) return VX_INVALID_VALUE;
  if (n && !keys) return VX_INVALID_VALUE;
  return VX_OK;
}

extern "C" vx_status vx_sort_by_key_workspace_size(uint32_t n, uint32_t key_bytes, uint32_t key_bits,
                                                   const vx_value_array* arrays, uint32_t n_arrays, size_t* bytes) {
  if (!bytes) return VX_INVALID_VALUE;
  if (key_bytes == 0 || key_bytes % 4 || key_bytes > 4 * vx::sort::kMaxKeyWords || key_bits == 0 ||
      key_bits > 8 * key_bytes || (n_arrays && !arrays))
    return VX_INVALID_VALUE;
  for (uint32_t i = 0; i < n_arrays; ++i)
    if (arrays[i].elem_bytes == 0) return VX_INVALID_VALUE;
  *bytes = vx::sort::by_key_workspace_bytes(n, key_bytes, key_bits, arrays, n_arrays);
  return VX_OK;
}

extern "C" vx_status vx_sort_by_key(void* keys, uint32_t key_bytes, uint32_t n, uint32_t key_bits,
                                    const vx_value_array* arrays, uint32_t n_arrays, void* ws, size_t ws_bytes,
                                    void* stream) {
  const vx_status st = check_by_key_args(keys, key_bytes, n, key_bits, arrays, n_arrays);
  if (st != VX_OK) return st;
  return vx::sort::by_key(keys, key_bytes, n, key_bits, arrays, n_arrays, ws, ws_bytes,
                          static_cast<hipStream_t>(stream));
}
