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
/* The same scans of u64 elements (scan<T> with T = uint64_t, cuda/scan/scan.cuh:79-151;
 * u64 wraparound for sums, identity 0 for max, ops.cuh:22-36); the workspace
 * is sized by vx_scan_workspace_size_u64. */
vx_status vx_scan_workspace_size_u64(uint32_t n, size_t* bytes);
vx_status vx_exclusive_sum_u64(uint64_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream);
vx_status vx_inclusive_sum_u64(uint64_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream);
vx_status vx_exclusive_max_u64(uint64_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream);
vx_status vx_inclusive_max_u64(uint64_t* data, uint32_t n, void* ws, size_t ws_bytes, void* stream);

/* Exclusive sum of u32 that allocates nothing and takes no workspace: for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same result as vx_exclusive_sum_u32; its look-back state lives in a
 * 512 KiB static device array of the library, one per device, held until
 * the process ends, and is cleared by a fill kernel on `stream` at every
 * call. n is at most 2^28 (VX_INVALID_VALUE beyond). Calls that overlap on
 * one device (two streams, two host threads) share that array and race:
 * order them on one stream. */
vx_status vx_exclusive_sum_u32_noalloc(uint32_t* data, uint32_t n, void* stream);

/* Scan tile shape (the reference picks it by template,
 * cuda/scan/block_based.cuh:13-61): workgroup `block` in {256, 512, 1024} and
 * uint4 loads per lane `vectors_per_lane` in {4, 8, 16} (block * 4 * vectors
 * elements per tile; supported pairs: 256x{4,8,16}, 512x{4,8}, 1024x4). */
vx_status vx_scan_set_tuning(uint32_t block, uint32_t vectors_per_lane);
vx_status vx_scan_get_tuning(uint32_t* block, uint32_t* vectors_per_lane);

/* Stream compaction: the reference builds it from the fused scan
 * (cuda/scan/fused.cuh:222-269, used at hash_dag_gpu/detail/edit/requests.cuh:51-122).
 * out[k] = in[i] for the k-th i (in order) with flags[i] != 0; *d_count (a
 * device u32) receives the number kept. Same workspace as the scan. */
vx_status vx_compact_u32(const uint32_t* in, const uint32_t* flags, uint32_t n,
                         uint32_t* out, uint32_t* d_count, void* ws,
                         size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Sort. Replaces algo::cuda::sort::{sort_keys, sort_by_key} for u32 / u64
 * keys with at most one u32 value array, and required_[sort_by_key_]workspace_size
 * (cuda/sort/sort.cuh:100-167; other layouts: vx_sort_by_key below).
 * Backed by a onesweep LSD radix sort (cuda/sort/radix/onesweep/host.cuh:12-205).
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

/* General sort_by_key<Key, Values...> (cuda/sort/sort.cuh:118-167 with
 * value_arrays_t, cuda/sort/value_arrays.cuh:12-140): keys of key_bytes =
 * 4 * W bytes, W = 1..16 u32 words (u32, u64 and UIntKey<W> layouts,
 * words[0] least significant, cuda/sort/uint_key.cuh:15-66), ordered stably
 * by their low key_bits bits; any number of value arrays of any element size
 * (e.g. the u64 voxel masks of build_leaf_masks/pipeline.cuh:94-96 or struct
 * records) are permuted with the keys. In place. n <= 1 needs no workspace;
 * NULL keys or a NULL array with n > 0, key_bits outside 1..8*key_bytes,
 * n >= 2^30 or a short workspace is VX_INVALID_VALUE. u32/u64 keys with at
 * most one 4-byte array take the direct (key, value) path; other layouts
 * sort (key, record index) and gather every array through the order. */
typedef struct vx_value_array {
  void* data;          /* device array of n elements */
  uint64_t elem_bytes; /* bytes per element */
} vx_value_array;
vx_status vx_sort_by_key_workspace_size(uint32_t n, uint32_t key_bytes, uint32_t key_bits,
                                        const vx_value_array* arrays, uint32_t n_arrays, size_t* bytes);
vx_status vx_sort_by_key(void* keys, uint32_t key_bytes, uint32_t n, uint32_t key_bits,
                         const vx_value_array* arrays, uint32_t n_arrays, void* ws, size_t ws_bytes,
                         void* stream);

/* Sort tuning (the reference selects these by template arguments,
 * cuda/sort/radix.cuh:55-68). radix_bits in {4, 8} with keys_per_thread in
 * {8, 15, 23} and block in {256, 512, 1024}, or radix_bits 9 with
 * (keys_per_thread, block) in {(8, 512), (8, 1024), (15, 512), (23, 512), (15, 1024)};
 * combinations whose tile does not fit LDS are VX_UNSUPPORTED. Takes effect
 * for later calls of this process. key_bytes selects the table row: 4 or 8
 * for (key, value) sorts, or'ed with VX_SORT_KEYS_ONLY for sorts without
 * values (the edit pipeline's packed record keys use the 8-byte keys-only
 * row). */
#define VX_SORT_KEYS_ONLY 0x100u
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

/* ---------------------------------------------------------------------------
 * Hash tables. Per-key status codes are the reference's
 * (hash_table/slab_hash/common.cuh:18-25): 0 NotFound, 1 Inserted, 2 Replaced,
 * 3 Erased, 4 Overflow, 5 InvalidKey, 6 Found, 7 AlreadyPresent. Status and
 * result arrays are device memory and may be NULL where marked optional.
 *
 * Slab set of u64 keys: replaces SlabHashSetU32 (slab_hash/set/host.cuh:15-107,
 * device.cuh:32-316) with 64-bit keys and 512-byte slabs (63 slots + link,
 * one u64 per lane of a 64-lane wavefront). Keys ~0 and ~0-1 are reserved
 * (InvalidKey). Overflow slabs come from the slab allocator (vx_slab_alloc
 * below); slabs_used = buckets + overflow slabs in use.
 * ------------------------------------------------------------------------- */
typedef struct vx_slab_set vx_slab_set;
vx_status vx_slab_set_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_set** out);
vx_status vx_slab_set_destroy(vx_slab_set* set);
vx_status vx_slab_set_clear(vx_slab_set* set, void* stream);
vx_status vx_slab_set_insert(vx_slab_set* set, const uint64_t* d_keys, uint32_t n, uint32_t* d_status /*opt*/,
                             void* stream);
vx_status vx_slab_set_contains(const vx_slab_set* set, const uint64_t* d_keys, uint32_t n, uint8_t* d_found /*opt*/,
                               uint32_t* d_status /*opt*/, void* stream);
vx_status vx_slab_set_erase(vx_slab_set* set, const uint64_t* d_keys, uint32_t n, uint32_t* d_status /*opt*/,
                            void* stream);
vx_status vx_slab_set_slabs_used(const vx_slab_set* set, uint32_t* used);

/* Slab set of u32 keys: replaces SlabHashSetU32 (slab_hash/set/host.cuh:15-107)
 * with the same slabs as the u64 set (keys stored zero-extended); keys
 * 0xffffffff and 0xfffffffe are reserved (InvalidKey, slab_hash/common.cuh:10-30). */
typedef struct vx_slab_set_u32 vx_slab_set_u32;
vx_status vx_slab_set_u32_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_set_u32** out);
vx_status vx_slab_set_u32_destroy(vx_slab_set_u32* set);
vx_status vx_slab_set_u32_clear(vx_slab_set_u32* set, void* stream);
vx_status vx_slab_set_u32_insert(vx_slab_set_u32* set, const uint32_t* d_keys, uint32_t n, uint32_t* d_status /*opt*/,
                                 void* stream);
vx_status vx_slab_set_u32_contains(const vx_slab_set_u32* set, const uint32_t* d_keys, uint32_t n,
                                   uint8_t* d_found /*opt*/, uint32_t* d_status /*opt*/, void* stream);
vx_status vx_slab_set_u32_erase(vx_slab_set_u32* set, const uint32_t* d_keys, uint32_t n, uint32_t* d_status /*opt*/,
                                void* stream);
vx_status vx_slab_set_u32_slabs_used(const vx_slab_set_u32* set, uint32_t* used);

/* Reusable slab allocator: replaces SlabAlloc (cuda/hash_table/slab_alloc.cuh:26-251),
 * the default allocator of the slab tables above (whose overflow slabs are
 * its ordinals; clearing a table frees them all). Ordinals 0..capacity-1
 * in a bitmap of one u64 word per 64 slabs; allocate writes one ordinal per
 * request (0xffffffff once every ordinal is taken), deallocate frees them
 * (out-of-range ordinals and double frees are no-ops) for later allocations.
 * used = allocated ordinals (synchronises the device). The view is the
 * layout the device API of include/vx_slab_alloc_device.hpp works on. */
typedef struct vx_slab_alloc_view {
  unsigned long long* bitmap; /* words u64 words, bit set = allocated */
  uint32_t capacity;
  uint32_t words;
} vx_slab_alloc_view;
typedef struct vx_slab_alloc vx_slab_alloc;
vx_status vx_slab_alloc_create(uint32_t capacity, vx_slab_alloc** out);
vx_status vx_slab_alloc_destroy(vx_slab_alloc* alloc);
vx_status vx_slab_alloc_clear(vx_slab_alloc* alloc, void* stream);
vx_status vx_slab_alloc_allocate(vx_slab_alloc* alloc, uint32_t n, uint32_t* d_ordinals, void* stream);
vx_status vx_slab_alloc_deallocate(vx_slab_alloc* alloc, const uint32_t* d_ordinals, uint32_t n, void* stream);
vx_status vx_slab_alloc_used(const vx_slab_alloc* alloc, uint32_t* used);
vx_status vx_slab_alloc_view_get(const vx_slab_alloc* alloc, vx_slab_alloc_view* out);

/* Slab map u32 -> u32: replaces SlabHashMapU32 (slab_hash/map/host.cuh:15-108,
 * device.cuh:49-393); each slot packs key << 32 | value and changes by one
 * 64-bit CAS. Keys 0xffffffff / 0xfffffffe are reserved. find writes
 * 0xffffffff for missing keys. */
typedef struct vx_slab_map vx_slab_map;
vx_status vx_slab_map_create(uint32_t buckets, uint32_t overflow_slabs, vx_slab_map** out);
vx_status vx_slab_map_destroy(vx_slab_map* map);
vx_status vx_slab_map_clear(vx_slab_map* map, void* stream);
vx_status vx_slab_map_insert_or_replace(vx_slab_map* map, const uint32_t* d_keys, const uint32_t* d_values,
                                        uint32_t n, uint32_t* d_status /*opt*/, void* stream);
vx_status vx_slab_map_find(const vx_slab_map* map, const uint32_t* d_keys, uint32_t* d_values, uint32_t n,
                           uint32_t* d_status /*opt*/, void* stream);
vx_status vx_slab_map_erase(vx_slab_map* map, const uint32_t* d_keys, uint32_t n, uint32_t* d_status /*opt*/,
                            void* stream);

/* Cuckoo set of u64 keys: replaces cuckoo_insert / cuckoo_find
 * (cuda/cuckoo_hash.cuh:73-120) — four salted hash functions, eviction chains
 * by 64-bit atomic exchange, the key itself stored in the slot. Keys of one
 * insert call must be unique and not ~0; *d_failed (device u32) counts keys
 * the table could not place within max_evictions (0 = default 512). */
typedef struct vx_cuckoo_set vx_cuckoo_set;
vx_status vx_cuckoo_set_create(uint64_t slots, uint32_t max_evictions, vx_cuckoo_set** out);
vx_status vx_cuckoo_set_destroy(vx_cuckoo_set* set);
vx_status vx_cuckoo_set_clear(vx_cuckoo_set* set, void* stream);
vx_status vx_cuckoo_set_insert(vx_cuckoo_set* set, const uint64_t* d_keys, uint32_t n, uint32_t* d_failed,
                               void* stream);
vx_status vx_cuckoo_set_contains(const vx_cuckoo_set* set, const uint64_t* d_keys, uint32_t n, uint8_t* d_found,
                                 void* stream);

/* Content-addressed set / map of items of item_words u32 words (1..16):
 * replaces AccelerationHashSet32<W> / AccelerationHashMap32<W>
 * (cuda/hash_table/acceleration_hash/set/host.cuh:15-117, map/host.cuh:15-116).
 * Capacity is the reference's, (buckets + overflow_slabs) * 31 items.
 *   find: d_ptrs[i] = a non-zero handle and status 6 (Found) when an equal
 *     item (all words compared) is stored, else 0 and status 0 (NotFound);
 *     a map also returns the item's value (0 when not found).
 *   insert_unique_unchecked: stores every item (the caller guarantees no
 *     equal item is present; equal items inserted anyway are all stored and
 *     find returns one of them, the same for every caller), status 1
 *     (Inserted) with its handle, or 4 (Overflow) with handle 0 once the
 *     capacity is used up.
 *   read: the item (and value) behind each handle; d_ok[i] = 0 for a handle
 *     that names no stored item.
 *   view: the device layout, for the device API of include/vx_accel_device.hpp
 *     (find / insert_unique_unchecked / read_item from the caller's kernels,
 *     the reference's DeviceAccelerationHashSet32 / Map32).
 * d_ptrs / d_status / d_values / d_ok are device arrays and may be NULL. */
typedef struct vx_accel_view {
  unsigned long long* slots; /* slot_mask + 1 slots {fingerprint:32 | item index + 1:32}, 0 = empty */
  uint64_t slot_mask;
  uint32_t* items;           /* capacity * words */
  uint32_t* values;          /* capacity values (map), NULL (set) */
  uint32_t* used;            /* device counter of item indices handed out */
  uint32_t capacity;
  uint32_t words;
} vx_accel_view;
typedef struct vx_accel_set vx_accel_set;
vx_status vx_accel_set_create(uint32_t item_words, uint32_t buckets, uint32_t overflow_slabs, vx_accel_set** out);
vx_status vx_accel_set_destroy(vx_accel_set* set);
vx_status vx_accel_set_clear(vx_accel_set* set, void* stream);
vx_status vx_accel_set_find(const vx_accel_set* set, const uint32_t* d_items, uint32_t n, uint32_t* d_ptrs,
                            uint32_t* d_status, void* stream);
vx_status vx_accel_set_insert_unique_unchecked(vx_accel_set* set, const uint32_t* d_items, uint32_t n,
                                               uint32_t* d_ptrs, uint32_t* d_status, void* stream);
vx_status vx_accel_set_read(const vx_accel_set* set, const uint32_t* d_ptrs, uint32_t n, uint32_t* d_items,
                            uint32_t* d_ok, void* stream);
vx_status vx_accel_set_view(const vx_accel_set* set, vx_accel_view* out);
vx_status vx_accel_set_used(const vx_accel_set* set, uint32_t* items);
typedef struct vx_accel_map vx_accel_map;
vx_status vx_accel_map_create(uint32_t item_words, uint32_t buckets, uint32_t overflow_slabs, vx_accel_map** out);
vx_status vx_accel_map_destroy(vx_accel_map* map);
vx_status vx_accel_map_clear(vx_accel_map* map, void* stream);
vx_status vx_accel_map_find(const vx_accel_map* map, const uint32_t* d_items, uint32_t n, uint32_t* d_ptrs,
                            uint32_t* d_values, uint32_t* d_status, void* stream);
vx_status vx_accel_map_insert_unique_unchecked(vx_accel_map* map, const uint32_t* d_items, const uint32_t* d_values,
                                               uint32_t n, uint32_t* d_ptrs, uint32_t* d_status, void* stream);
vx_status vx_accel_map_read(const vx_accel_map* map, const uint32_t* d_ptrs, uint32_t n, uint32_t* d_items,
                            uint32_t* d_values, uint32_t* d_ok, void* stream);
vx_status vx_accel_map_view(const vx_accel_map* map, vx_accel_view* out);
vx_status vx_accel_map_used(const vx_accel_map* map, uint32_t* items);

/* ---------------------------------------------------------------------------
 * Hash DAG. Replaces algo::svt::hash_dag_gpu (HashDagGpu storage.cuh:16-105,
 * apply/place/destroy_voxel_edits edit.cuh:16-61, get_voxel query.cuh:12-39)
 * with a runtime depth: world side 4 << inner_depth (inner_depth 1..13, i.e.
 * up to 32768^3), 64-bit Morton leaf keys, and an optional per-voxel u8
 * material payload.
 *   materials == 0: occupancy DAG, exactly the reference's canonical form —
 *     refs 0 EMPTY / 1 FULL, node item [child_mask | fill_mask << 8, refs of
 *     table-backed children] (hash_dag_gpu/config.cuh:30-45, detail/node.cuh:210-289),
 *     leaves = u64 occupancy (bit x + 4y + 16z).
 *   materials != 0: leaves are 64 u8 materials (0 = empty) in 16 u32 words;
 *     FULL is never used (a solid region is a shared subtree of materials).
 * Nodes and leaves are interned by content in two device hash tables (u64
 * slots: 32-bit fingerprint | pool offset) over two word pools; every
 * distinct subtree is stored once. A failed edit (pool or table full) returns
 * VX_OUT_OF_CAPACITY and leaves the old root published.
 * ------------------------------------------------------------------------- */
typedef struct vx_dag vx_dag;

typedef struct vx_dag_capacity {
  uint64_t node_words; /* node pool, u32 words */
  uint64_t leaf_words; /* leaf pool, u32 words */
  uint64_t node_slots; /* node table slots (rounded up to a power of two) */
  uint64_t leaf_slots; /* leaf table slots (rounded up to a power of two) */
} vx_dag_capacity;

typedef struct vx_voxel { uint32_t x, y, z; } vx_voxel; /* = hash_dag_gpu::VoxelEdit */

/* Brush: sphere (kind 0, radius hx) or box (kind 1, half extents hx, hy, hz)
 * around an integer centre; covers voxel (x,y,z) iff dx^2+dy^2+dz^2 <= r^2 or
 * |dx|<=hx && |dy|<=hy && |dz|<=hz. material 0 removes, 1..255 adds (the
 * occupancy DAG treats any non-zero material as "place"). Brushes of one call
 * apply in array order: per voxel, the last brush wins. A kind above 1 or a
 * material above 255 makes the call VX_INVALID_VALUE. */
typedef struct vx_brush {
  int32_t cx, cy, cz;
  int32_t hx, hy, hz;
  uint32_t kind;
  uint32_t material;
} vx_brush;

/* Heightfield: solid iff y <= base + floor(amp * sum_o 2^-o * n_o(x, z)),
 * n_o = value noise of period (period >> o) over lattice values uniform in
 * [-1, 1) from the seeded stream; material 1 + (y mod 255). */
typedef struct vx_terrain {
  uint64_t seed;
  uint32_t base, amp, period, octaves;
} vx_terrain;

typedef struct vx_dag_stats {
  uint64_t node_words_used, leaf_words_used;
  uint64_t node_slots, leaf_slots;
  uint32_t root;
  uint32_t inner_depth;
  uint32_t leaf_words; /* 2 (occupancy) or 16 (materials) */
  uint32_t last_status; /* first-error word of the scene's last direct edit call: 0 ok, 1 overflow */
} vx_dag_stats;

vx_status vx_dag_create(uint32_t inner_depth, int32_t materials, const vx_dag_capacity* cap, vx_dag** out);
vx_status vx_dag_destroy(vx_dag* dag);
/* Clear every table and pool, root = EMPTY (storage.cuh:62-73). */
vx_status vx_dag_reset(vx_dag* dag, void* stream);
vx_status vx_dag_stats_get(const vx_dag* dag, vx_dag_stats* out);
/* Replace the scene by a heightfield (reset + hierarchical build). */
vx_status vx_dag_build_terrain(vx_dag* dag, const vx_terrain* terrain, void* stream);

/* Drop-in for apply_voxel_edits_workspace_size / apply_voxel_edits
 * (edit.cuh:16-41): op 0 = Place, 1 = Destroy; material is the placed
 * material for a materials DAG. Out-of-world voxels are ignored; count 0 is a
 * no-op; NULL edits with count > 0, count >= 2^29 or a short workspace is
 * VX_INVALID_VALUE. */
vx_status vx_dag_voxel_edits_workspace_size(const vx_dag* dag, uint32_t count, size_t* bytes);
vx_status vx_dag_apply_voxel_edits(vx_dag* dag, const vx_voxel* d_edits, uint32_t count, uint32_t op,
                                   uint32_t material, void* ws, size_t ws_bytes, void* stream);

/* Brush edit batch; h_brushes is HOST memory (it is small and the host needs
 * it to size the work). A batch (and a batch set, and a batcher's capacity)
 * spans fewer than 2^29 candidate leaf bricks (VX_INVALID_VALUE otherwise). */
vx_status vx_dag_brush_workspace_size(const vx_dag* dag, const vx_brush* h_brushes, uint32_t n, size_t* bytes);
vx_status vx_dag_apply_brushes(vx_dag* dag, const vx_brush* h_brushes, uint32_t n, void* ws, size_t ws_bytes,
                               void* stream);

/* Batched edit over several resident scenes in one pipeline (one segmented
 * sort / scan over all batches; no counterpart in the reference, which edits
 * one DAG per call). Batch i (h_brushes[i][0..counts[i]), host memory) applies
 * to dags[i]; batches of one scene apply in array order (the result equals
 * applying them one by one), different scenes are independent. All scenes must
 * be of one kind (occupancy or materials). A scene whose tables overflow keeps
 * its old root (VX_OUT_OF_CAPACITY is returned); the others publish. */
vx_status vx_dag_batch_workspace_size(vx_dag* const* dags, const vx_brush* const* h_brushes, const uint32_t* counts,
                                      uint32_t n_batches, size_t* bytes);
vx_status vx_dag_apply_batches(vx_dag* const* dags, const vx_brush* const* h_brushes, const uint32_t* counts,
                               uint32_t n_batches, void* ws, size_t ws_bytes, void* stream);

/* Persistent batched executor over a fixed set of resident scenes: owns a
 * workspace sized for max_candidates brick candidates / max_jobs brushes and
 * its own stream. With use_graph != 0 the whole device pipeline is captured
 * into a hipGraph on the first run and replayed afterwards (every count it
 * works on is device-resident). run: batch i applies h_brushes[i] (host) to
 * scenes[scene_of_batch[i]], in array order per scene; *candidates receives
 * the batch set's candidate count (VX_INVALID_VALUE if above the capacity).
 * No counterpart in the reference. */
typedef struct vx_batcher vx_batcher;
vx_status vx_batcher_create(vx_dag* const* scenes, uint32_t n_scenes, uint64_t max_candidates, uint32_t max_jobs,
                            int32_t use_graph, vx_batcher** out);
vx_status vx_batcher_destroy(vx_batcher* b);
vx_status vx_batcher_info(const vx_batcher* b, uint64_t* workspace_bytes, int32_t* graph_ready);
vx_status vx_batcher_run(vx_batcher* b, const uint32_t* scene_of_batch, const vx_brush* const* h_brushes,
                         const uint32_t* counts, uint32_t n_batches, uint64_t* candidates);
/* Asynchronous form of vx_batcher_run for continuous batching: enqueues the
 * batch set on the executor's stream (brush upload, graph launch, read-back
 * of every scene's pool usage and status) and returns at once. One
 * submission per executor is in flight; vx_batcher_query completes it before
 * the next (VX_INVALID_VALUE otherwise). With `after` != NULL the device work
 * waits for `after`'s last submission (on the device: the upload still
 * overlaps it); submissions of two executors over disjoint scene sets may run
 * concurrently (each executor keeps its own per-scene status words), and
 * submissions that share a scene must be ordered with `after`. */
vx_status vx_batcher_submit(vx_batcher* b, const uint32_t* scene_of_batch, const vx_brush* const* h_brushes,
                            const uint32_t* counts, uint32_t n_batches, uint64_t* candidates,
                            const vx_batcher* after);
/* Completion of the last submission: *done = 0 while the device works on it
 * (wait != 0 blocks until it is done), else 1 and the outcome is returned:
 * VX_OUT_OF_CAPACITY if a scene overflowed (that scene keeps its old root;
 * scene_status[i] != 0 names it), VX_OK otherwise. scene_status (n_scenes
 * u32) and scene_used (2 * n_scenes u64: node words, leaf words used after
 * the submission) may be NULL. */
vx_status vx_batcher_query(vx_batcher* b, int32_t wait, int32_t* done, uint32_t* scene_status,
                           uint64_t* scene_used);
/* Record layout of the executor's sort: *key_bits = scene bits + 3 bits per
 * level of the deepest scene; *index_bits > 0 when (key, record index) pack
 * into one u64 sorted by its upper bits, 0 when records sort as (u64 key,
 * u32 index) pairs (key_bits + index bits above 64). */
vx_status vx_batcher_layout(const vx_batcher* b, uint32_t* key_bits, uint32_t* index_bits);

/* get_voxel for n device coordinates: d_out[i] = material (1 = solid in an
 * occupancy DAG), 0 if empty or out of the world (query.cuh:12-39). */
vx_status vx_dag_get_voxels(const vx_dag* dag, const vx_voxel* d_xyz, uint32_t n, uint32_t* d_out, void* stream);

/* Ray traversal (the consumer side of the DAG, next to query.cuh:12-39; the
 * reference has no GPU ray caster). A ray starts at the centre of voxel
 * (ox, oy, oz) with the integer direction (dx, dy, dz) and visits voxels in
 * the order of its boundary crossings: crossing j of axis a happens at
 * t = (2j + 1) / (2|d_a|); simultaneous crossings are taken x before y
 * before z. It visits the origin voxel and then at most max_steps further
 * voxels, and stops at the first non-empty one (material, or 1 in an
 * occupancy DAG) or when it leaves the world (a miss, as is an origin outside
 * the world or a zero direction from an empty origin). Exact integer
 * arithmetic throughout; empty subtrees are crossed in one step. */
typedef struct vx_ray {
  int32_t ox, oy, oz;
  int32_t dx, dy, dz;
  uint32_t max_steps;
  uint32_t pad;
} vx_ray;
typedef struct vx_hit {
  uint32_t material; /* 0: miss */
  uint32_t steps;    /* voxels crossed before the hit voxel */
  int32_t x, y, z;   /* hit voxel */
  uint32_t pad;
} vx_hit;
vx_status vx_dag_raycast(const vx_dag* dag, const vx_ray* d_rays, uint32_t n, vx_hit* d_hits, void* stream);

/* Canonical depth-first export (host buffer of u32 words):
 *   [magic 0x56584443, inner_depth, leaf_words, root kind (0 EMPTY, 1 FULL,
 *    2 node), node count, leaf count], then per node in first-visit order
 *   [level, child_mask | fill_mask << 8, child ids...], then leaf payloads.
 * Child ids of level inner_depth-1 nodes are leaf ids. *n_words receives the
 * size; words are written only if cap suffices. Computed on the GPU
 * (first-visit order from every node's smallest root path); synchronises
 * the device and allocates temporaries. The reference checks its GPU DAG
 * through get_voxel only (test/cuda/test_cuda_hash_dag_gpu.cu:181-535); this
 * export is the bit-exact comparison form of the canonical DAG. */
vx_status vx_dag_export_canonical(const vx_dag* dag, uint32_t* out, uint64_t cap, uint64_t* n_words);
/* The same words into device memory d_out, ordered on `stream` (which it
 * synchronises); with d_out NULL or cap short only *n_words is computed. */
vx_status vx_dag_export_canonical_device(const vx_dag* dag, uint32_t* d_out, uint64_t cap, uint64_t* n_words,
                                         void* stream);

/* Garbage collection of the interned tables: keeps exactly the items
 * reachable from the root, each once, and rebuilds both slot tables over them
 * (the GPU counterpart of the CPU HashDAG::collect_garbage,
 * svt/hash_dag.hpp:133-191; the reference's GPU tables only grow,
 * cuda/hash_table/bump_allocator.cuh:118-121). The voxel field, the export
 * and every later edit are unchanged; pool offsets (refs) change. Synchronous
 * on `stream`; no other call may use the scene meanwhile. Allocates
 * temporaries (two slot tables and the live pool words); if anything fails
 * the scene is left as it was. `report` may be NULL. */
typedef struct vx_gc_report {
  uint64_t node_words_before, node_words_after;
  uint64_t leaf_words_before, leaf_words_after;
  uint64_t live_nodes;  /* live (level, node item) pairs */
  uint64_t live_leaves; /* live leaf items */
} vx_gc_report;
vx_status vx_dag_collect_garbage(vx_dag* dag, vx_gc_report* report, void* stream);

/* Growth of a scene's pools and slot tables to at least `cap` (a field at or
 * below the current size keeps it; slot counts round up to powers of two).
 * The used pool words move to the larger pools unchanged, so every ref and
 * slot value stays valid; a table that grows is rebuilt by re-inserting each
 * live slot at its item's home position in the larger table (the hash
 * recomputed from the item). Counterpart of the reference's growing GPU
 * tables, whose bump allocator hands out further slabs as buckets fill
 * (cuda/hash_table/bump_allocator.cuh:118-121, slab_hash/set/host.cuh:107);
 * an open-addressing table grows by rebuilding instead. Synchronous on
 * `stream`; no other call may use the scene meanwhile (executors that hold
 * the scene take its new buffers at their next submission and capture their
 * graph again). Allocation failure: VX_OUT_OF_CAPACITY, a slot whose item does
 * not hash to its fingerprint: VX_DEVICE_ERROR; either way the scene is left
 * as it was. */
vx_status vx_dag_grow(vx_dag* dag, const vx_dag_capacity* cap, void* stream);

/* Host copy of the used words [0, used) of both pools (inspection and
 * tests; no counterpart in the reference). *node_n / *leaf_n receive the
 * sizes; a pool's words are written only if its cap suffices. */
vx_status vx_dag_read_pools(const vx_dag* dag, uint32_t* node_out, uint64_t node_cap, uint64_t* node_n,
                            uint32_t* leaf_out, uint64_t leaf_cap, uint64_t* leaf_n);

/* Device-side check of the tables' invariant (no counterpart in the
 * reference, which relies on it without checking: insert_unique_unchecked
 * requires uniqueness, cuda/hash_table/acceleration_hash/set/common.cuh:51-56,
 * guaranteed per depth by svt/hash_dag_gpu/detail/edit/temp_dedup.cuh:29-34):
 * no two live slots hold items of equal content. Synchronous. out[0..2] for
 * the node table, out[3..5] for the leaf table: live slots, duplicates, and
 * broken slots (a tag left over, an offset outside the used words, a
 * fingerprint that is not the item's, or a slot a lookup cannot reach). */
vx_status vx_dag_check_unique(const vx_dag* dag, uint64_t out[6], void* stream);

/* The same check over a caller's table: n_slots (a power of two) u64 slots
 * {fingerprint:32 | pool offset:32} (0 empty, 1 dead) over a u32 pool whose
 * words [0, pool_words_used) are in use; items of item_words words, or DAG
 * node items (1 + popcount of the header's low byte) when item_words is 0.
 * out[3]: live slots, duplicates, broken slots. */
vx_status vx_table_check_unique(const uint64_t* d_slots, uint64_t n_slots, const uint32_t* d_pool,
                                uint64_t pool_words_used, uint32_t item_words, uint64_t out[3], void* stream);

/* ---- pointer SVO voxel edits (algo::svt::cuda) -------------------------------
 * The mutable sparse voxel octree of svt/cuda: 4096^3 world, ten inner
 * levels of 2x2x2 (config.cuh:16-36); nodes of eight u32 slots {child
 * index:30 | pointer bit 30 | filled bit 31}, leaves of two u32 masks (bit
 * x + 4y + 16z of the 4^3 block), LIFO free-index stacks. */
typedef struct vx_svo vx_svo;
/* GpuSvoCounters (svt/cuda/device_view.cuh:24-30). */
typedef struct vx_svo_counters {
  uint32_t node_count, leaf_count, free_node_count, free_leaf_count;
} vx_svo_counters;
/* LeafMask (svt/cuda/voxel/types.cuh:20-25). */
typedef struct vx_leaf_mask {
  uint32_t leaf_key, voxel_data_low, voxel_data_high;
} vx_leaf_mask;
/* DeviceGpuSvo (svt/cuda/device_view.cuh:34-42): device pointers for kernels
 * of the caller's own. */
typedef struct vx_svo_view {
  uint32_t* nodes; /* 8 slots per node */
  uint32_t* leaves; /* 2 words per leaf */
  vx_svo_counters* counters;
  uint32_t* free_node_indices;
  uint32_t* free_leaf_indices;
  uint32_t max_node_count, max_leaf_count;
} vx_svo_view;
/* GpuSvo<MaxNodeCount, MaxLeafCount> (svt/cuda/storage.cuh:12-122): capacities
 * are runtime values here. The tree starts reset (as after vx_svo_reset),
 * where the reference's constructor leaves it unusable until reset_svo. */
vx_status vx_svo_create(uint32_t max_node_count, uint32_t max_leaf_count, vx_svo** out);
vx_status vx_svo_destroy(vx_svo* svo);
/* reset_svo (svt/cuda/svo.cuh:14-22, detail/reset.cuh:9-21): empty root,
 * counters {1, 0, 0, 0}. */
vx_status vx_svo_reset(vx_svo* svo, void* stream);
/* GpuSvo::view() (svt/cuda/storage.cuh:40-49). */
vx_status vx_svo_view_get(const vx_svo* svo, vx_svo_view* out);
/* Host copy of the counters (synchronises the stream). */
vx_status vx_svo_counters_get(const vx_svo* svo, vx_svo_counters* out, void* stream);
/* voxel_edits_to_leaf_masks[_workspace_size] (svt/cuda/voxel/edit.cuh:14-41):
 * one mask per touched leaf, sorted by leaf key; out-of-world edits dropped;
 * `leaf_masks` holds `count` entries, *leaf_count (device) receives the
 * number written. */
vx_status vx_voxel_edits_to_leaf_masks_workspace_size(uint32_t count, size_t* bytes);
vx_status vx_voxel_edits_to_leaf_masks(const vx_voxel* d_edits, uint32_t count, vx_leaf_mask* d_leaf_masks,
                                       uint32_t* d_leaf_count, void* workspace, size_t workspace_bytes, void* stream);
/* apply_voxel_edits_workspace_size / place_voxel_edits / destroy_voxel_edits
 * (svt/cuda/voxel/edit.cuh:43-92): op 0 Place, 1 Destroy. Paths to touched
 * leaves are materialised (new slots inherit the replaced slot's fill), the
 * leaf bits set or cleared, and uniform leaves / nodes on the touched paths
 * collapsed into filled / empty slots, deepest first, their indices pushed on
 * the free stacks. Synchronises the stream once (the reference's default
 * HostLeafCountDispatch does too). If the batch needs more nodes or leaves
 * than the capacity holds it returns VX_OUT_OF_CAPACITY and changes nothing
 * (the reference skips the allocations that do not fit). */
vx_status vx_svo_voxel_edits_workspace_size(uint32_t count, size_t* bytes);
vx_status vx_svo_apply_voxel_edits(vx_svo* svo, const vx_voxel* d_edits, uint32_t count, uint32_t op,
                                   void* workspace, size_t workspace_bytes, void* stream);
/* get_voxel (svt/cuda/query.cuh:12-39), batched: out[i] = 1 if set. */
vx_status vx_svo_get_voxels(const vx_svo* svo, const vx_voxel* d_xyz, uint32_t n, uint32_t* d_out, void* stream);
/* Canonical export (host): [0x5658534F, node count, leaf count], 8 slot words
 * per node in pre-order from the root (a pointer slot holds the child's
 * canonical id with the slot's own flag bits), then 2 words per leaf in the
 * order the walk reaches them. *n_words receives the size. */
vx_status vx_svo_export_canonical(const vx_svo* svo, uint32_t* out, uint64_t cap, uint64_t* n_words);

/* Seeded brush batch (host output). mix 0: spheres r in [4,16], 50% add;
 * mix 1: 50% spheres r in [8,64] / 50% boxes half extents in [4,48], 60% add;
 * added material uniform in [1,255]. Brush i uses draws 16i..16i+8 of the
 * (seed, stream_id) splitmix64 stream. */
vx_status vx_gen_brushes(uint64_t seed, uint64_t stream_id, uint32_t n, uint32_t world, uint32_t mix,
                         vx_brush* out);

/* Brush batch sized to a voxel-op target (batched and serving workloads):
 * target = 2^t, t uniform in [lo_log2, hi_log2]; n uniform in [1, max_brushes]
 * brushes of volume target/n (50% spheres, 50% cubes), 60% add / 40% remove;
 * *scene_out = uniform scene in [0, n_scenes). `out` holds max_brushes. */
vx_status vx_gen_sized_batch(uint64_t seed, uint64_t stream_id, uint64_t index, uint32_t lo_log2, uint32_t hi_log2,
                             uint32_t max_brushes, uint32_t world, uint32_t n_scenes, vx_brush* out, uint32_t* n_out,
                             uint32_t* scene_out);

/* Random voxel edits of the pointer-SVO benchmark (make_random_edits,
 * bench/bench_cuda_svt.cu:90-112): xorshift32 (13, 17, 5) from 0x9e3779b9,
 * three draws per voxel, each mod `world` (host output). */
vx_status vx_gen_xorshift_voxels(uint32_t count, uint32_t world, vx_voxel* out);

/* Exact number of voxel operations of a brush batch: sum over brushes of the
 * voxels each covers inside the world (overlaps count once per brush). */
vx_status vx_brush_voxel_ops(const vx_brush* h_brushes, uint32_t n, uint32_t world, uint64_t* ops);

/* Per-stage edit timing (hipEvents recorded on the edit's stream) and the
 * sizes each stage worked on, accumulated over edits while enabled.
 * Stages: 0 records, 1 sort, 2 level structure, 3 old-ref walk,
 * 4 leaves (build + probe + commit), 5 nodes (all levels), 6 publish.
 * Kernels: launches and summed duration of each kernel kind of the edit
 * pipeline (a hipEvent pair around every launch of a direct edit call;
 * graph replays are not timed). */
#define VX_DAG_STAGES 7
enum {
  VX_K_RECORDS = 0,
  VX_K_SORT_HISTOGRAM,
  VX_K_SORT_SCAN,
  VX_K_SORT_PARTITION,
  VX_K_SORT_COPY,
  VX_K_RUNS,
  VX_K_UP,
  VX_K_TOPDOWN,
  VX_K_LEAF_BUILD,
  VX_K_LEAF_PROBE,
  VX_K_LEAF_COMMIT,
  VX_K_LEAF_PLACE,
  VX_K_NODE_BUILD,
  VX_K_NODE_PROBE,
  VX_K_NODE_COMMIT,
  VX_K_NODE_PLACE,
  VX_K_PUBLISH,
  VX_K_FILL,
  VX_K_OTHER,
  VX_K_TAIL_UP,      /* single-workgroup kernels for the small top heights */
  VX_K_TAIL_TOPDOWN, /* not launched any more: every height's old refs are VX_K_TOPDOWN (kept so the
                        * kinds after it keep their values) */
  VX_K_TAIL_NODES,
  VX_DAG_KERNELS
};
typedef struct vx_dag_profile {
  double ms[VX_DAG_STAGES];
  uint64_t edits;        /* edit calls measured */
  uint64_t records;      /* leaf-brick records (sum) */
  uint64_t leaves;       /* distinct leaves touched (sum) */
  uint64_t nodes;        /* distinct inner nodes touched (sum) */
  uint64_t sort_passes;  /* radix passes per edit */
  uint64_t sort_record_bytes; /* bytes one record occupies in the sort: 12 (key, index) or 8 (packed) */
  uint64_t kernel_launches[VX_DAG_KERNELS];
  double kernel_ms[VX_DAG_KERNELS];
  uint64_t pending_leaves; /* leaves not found among the interned items by the build (sum): the probe's work */
  uint64_t height_entries[16]; /* entries per height (sum; [0] leaves, [h] inner nodes at height h) */
} vx_dag_profile;
/* enable: 0 off, VX_PROFILE_ON per-stage and per-kernel events; with
 * VX_PROFILE_NO_STAGES added, per-kernel events only (each stage event costs
 * the stream a few microseconds). Clears the accumulated profile and the
 * kernel period. */
#define VX_PROFILE_ON 1
#define VX_PROFILE_NO_STAGES 2
vx_status vx_dag_profile_enable(vx_dag* dag, int32_t enable);
vx_status vx_dag_profile_read(const vx_dag* dag, vx_dag_profile* out);
/* Which kernel kinds get per-launch events while profiled (bit k = VX_K_k;
 * default all): timing only the kernel under study keeps the event overhead
 * out of the other launches. */
vx_status vx_dag_profile_kernel_mask(vx_dag* dag, uint64_t mask);
/* period > 1: of the launches of a timed kind within one call, only the one at
 * index (call number mod period) gets events; over `period` calls each index
 * is timed once (one event pair per call for a kind launched `period` times
 * per call). */
vx_status vx_dag_profile_kernel_period(vx_dag* dag, uint32_t period);

#ifdef __cplusplus
}
#endif
#endif
