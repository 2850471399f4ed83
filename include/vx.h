/*
 * vx.h — C ABI of the MI355X (gfx950) voxel-editing library.
 *
 * Plain pointers and sizes only: device pointers are HIP device memory, a
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 * Every entry point returns a vx_status and never throws. All work is
 * stream-ordered on the caller's stream; functions that must report a device
 * outcome (the DAG edits) synchronise that stream before returning, as the
 * reference's edit does (hash_dag_gpu/detail/edit/root.cuh:36-48).
 *
 * Status mapping to the reference's cudaError_t results:
 *   VX_OK              cudaSuccess
 *   VX_INVALID_VALUE   cudaErrorInvalidValue (null pointer, short workspace,
 *                      bad size)              (cuda/sort/sort.cuh:36-92)
 *   VX_OUT_OF_CAPACITY cudaErrorMemoryAllocation from a device-side table or
 *                      pool overflow; the old root stays published
 *                                            (hash_dag_gpu/detail/edit/root.cuh:47)
 *   VX_DEVICE_ERROR    any HIP runtime error
 */
#ifndef VX_H
#define VX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t vx_status;
enum {
  VX_OK = 0,
  VX_INVALID_VALUE = 1,
  VX_OUT_OF_CAPACITY = 2,
  VX_DEVICE_ERROR = 3,
  VX_UNSUPPORTED = 4
};

/* Library / device information. */
const char* vx_version(void);
vx_status vx_device_info(int32_t* cu_count, int32_t* lds_bytes_per_block,
                         int64_t* hbm_bytes);

/* ---------------------------------------------------------------------------
 * Scan. Replaces algo::cuda::scan::{inclusive,exclusive}_{sum,scan} and
 * required_workspace_size (cuda/scan/scan.cuh:79-161). In place, single pass
 * (decoupled look-back). No workspace is needed when n fits one tile
 * (*bytes == 0); otherwise a short or NULL workspace is VX_INVALID_VALUE.
 * ------------------------------------------------------------------------- */
vx_status vx_scan_workspace_size(uint32_t n, size_t* bytes);
vx_status vx_exclusive_sum_u32(uint32_t* data, uint32_t n, void* ws,
                               size_t ws_bytes, void* stream);
vx_status vx_inclusive_sum_u32(uint32_t* data, uint32_t n, void* ws,
                               size_t ws_bytes, void* stream);
vx_status vx_exclusive_max_u32(uint32_t* data, uint32_t n, void* ws,
                               size_t ws_bytes, void* stream);
vx_status vx_inclusive_max_u32(uint32_t* data, uint32_t n, void* ws,
                               size_t ws_bytes, void* stream);

/* Stream compaction: the reference builds it from the fused scan
 * (cuda/scan/fused.cuh:222-269, used at hash_dag_gpu/detail/edit/requests.cuh:51-122).
 * out[k] = in[i] for the k-th i (in order) with flags[i] != 0; *d_count (a
 * device u32) receives the number kept. Same workspace as the scan. */
vx_status vx_compact_u32(const uint32_t* in, const uint32_t* flags, uint32_t n,
                         uint32_t* out, uint32_t* d_count, void* ws,
                         size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Sort. Replaces algo::cuda::sort::{sort_keys, sort_by_key} with one u32
 * value array and required_[sort_by_key_]workspace_size (cuda/sort/sort.cuh:100-167),
 * backed by a onesweep LSD radix sort (cuda/sort/radix/onesweep/host.cuh:12-205).
 * In place, stable, ordered by the low key_bits bits (1..32 for u32 keys,
 * 1..64 for u64 keys). n <= 1 needs no workspace. n must be < 2^30
 * (cuda/sort/radix/onesweep/host.cuh:85-86). vals may be NULL for keys only.
 * ------------------------------------------------------------------------- */
vx_status vx_sort_workspace_size(uint32_t n, uint32_t key_bytes,
                                 uint32_t key_bits, int32_t with_values,
                                 size_t* bytes);
vx_status vx_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n,
                            uint32_t key_bits, void* ws, size_t ws_bytes,
                            void* stream);
vx_status vx_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint32_t n,
                            uint32_t key_bits, void* ws, size_t ws_bytes,
                            void* stream);

/* Sort tuning (the reference selects these by template arguments,
 * cuda/sort/radix.cuh:55-68). radix_bits in {4, 8}, keys_per_thread in
 * {8, 15, 23}, block in {256, 512, 1024}; combinations whose tile does not
 * fit LDS are VX_UNSUPPORTED. Takes effect for later calls of this process.
 * key_bytes selects the table row (4 or 8). */
vx_status vx_sort_set_tuning(uint32_t key_bytes, uint32_t radix_bits,
                             uint32_t keys_per_thread, uint32_t block);
vx_status vx_sort_get_tuning(uint32_t key_bytes, uint32_t* radix_bits,
                             uint32_t* keys_per_thread, uint32_t* block);

/* Device-side generators of the synthetic workloads (counter-based
 * splitmix64 streams, so every element is a pure function of its index). */
vx_status vx_gen_u32(uint64_t seed, uint64_t stream_id, uint32_t n,
                     uint32_t lo, uint32_t hi, uint32_t* out, void* stream);
vx_status vx_gen_morton(uint64_t seed, uint64_t stream_id, uint32_t n,
                        uint32_t coord_bits, uint64_t* out, void* stream);
vx_status vx_gen_unique_u64(uint64_t seed, uint64_t stream_id, uint64_t first,
                            uint32_t n, uint64_t* out, void* stream);
vx_status vx_iota_u32(uint32_t* out, uint32_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
