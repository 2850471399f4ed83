/*
 * CPU oracle for the voxel-editing hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, on the CPU and single-threaded, what the reference
 * computes on the path (yamaatsu1112/voxel, algorithm/lib/include/algo/...):
 *   - exclusive / inclusive prefix sum        (cuda/scan/scan.cuh:122-151)
 *   - stable LSD sort of (key, value) pairs   (cuda/sort/sort.cuh:100-167)
 *   - the Hash DAG box edit with per-level hash consing
 *                                             (svt/hash_dag.hpp:343-533)
 *   - the GPU DAG's canonical node form (FULL/EMPTY sentinels + fill mask)
 *                                             (svt/hash_dag_gpu/detail/node.cuh:210-289)
 * plus the seeded generators that define the synthetic workloads.
 *
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg
 * may load this library, and only as the checker. The product library never
 * links or calls it.
 *
 * Parity status of every function is listed next to its declaration.
 */
#ifndef VOXEL_ORACLE_H
#define VOXEL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators (definitions of the synthetic inputs) ------------------ */
/* pinned: known-answer values of splitmix64 (published test vector). */
uint64_t vxo_mix64(uint64_t z);
/* Counter-based stream: r(seed, stream, i). */
uint64_t vxo_rand(uint64_t seed, uint64_t stream, uint64_t index);
/* Uniform integer in [lo, hi] from r(seed, stream, i) by multiply-shift. */
void vxo_gen_u32(uint64_t seed, uint64_t stream, uint64_t n, uint32_t lo,
                 uint32_t hi, uint32_t* out);
/* 3*coord_bits-bit Morton codes of uniform coord_bits-bit (x, y, z). */
void vxo_gen_morton(uint64_t seed, uint64_t stream, uint64_t n,
                    uint32_t coord_bits, uint64_t* out);
/* n distinct u64 keys (never ~0), in generation order. */
void vxo_gen_unique_u64(uint64_t seed, uint64_t stream, uint64_t n,
                        uint64_t* out);
uint64_t vxo_morton3(uint32_t x, uint32_t y, uint32_t z);

/* ---- scan (pinned: closed form + reference fixed vector) ---------------- */
void vxo_exclusive_sum_u32(const uint32_t* in, uint64_t n, uint32_t* out);
void vxo_inclusive_sum_u32(const uint32_t* in, uint64_t n, uint32_t* out);
void vxo_exclusive_max_u32(const uint32_t* in, uint64_t n, uint32_t* out);
void vxo_inclusive_max_u32(const uint32_t* in, uint64_t n, uint32_t* out);

/* ---- stable sort (pinned: std::stable_sort on the masked key) ----------- */
/* Sorts in place by the low key_bits bits; vals may be NULL. */
void vxo_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint64_t n,
                        uint32_t key_bits);
void vxo_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint64_t n,
                        uint32_t key_bits);
/* u64 scans (wrapping sums, maxima). */
void vxo_exclusive_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out);
void vxo_inclusive_sum_u64(const uint64_t* in, uint64_t n, uint64_t* out);
void vxo_exclusive_max_u64(const uint64_t* in, uint64_t n, uint64_t* out);
void vxo_inclusive_max_u64(const uint64_t* in, uint64_t n, uint64_t* out);
/* Stable order of n keys of `words` u32 words (word 0 least significant),
 * compared on their low key_bits bits: order[i] = index of the i-th key. */
void vxo_sort_order(const uint32_t* keys, uint32_t words, uint64_t n, uint32_t key_bits, uint32_t* order);

/* ---- hash set truth (pinned by construction) ---------------------------- */
/* out[i] = 1 if queries[i] is one of keys[0..nk). */
void vxo_set_contains_u64(const uint64_t* keys, uint64_t nk,
                          const uint64_t* queries, uint64_t nq, uint8_t* out);

/* ---- terrain + brushes (definitions of the DAG workloads) -------------- */
/* Heightfield: solid iff y <= base + floor(amp * sum_o 2^-o * n_o(x, z)),
 * n_o = value noise of period (period >> o). The 512^3 case is base 256,
 * amp 64, period 64; a world k times larger scales all three by k. */
typedef struct vxo_terrain {
  uint64_t seed;
  uint32_t base;
  uint32_t amp;
  uint32_t period;
  uint32_t octaves; /* 4 in every configuration */
} vxo_terrain;

typedef struct vxo_brush {
  int32_t cx, cy, cz; /* centre voxel */
  int32_t hx, hy, hz; /* sphere: radius in hx; box: half extents */
  uint32_t kind;      /* 0 sphere, 1 box */
  uint32_t material;  /* 0 remove, 1..255 add with this material */
} vxo_brush;

int32_t vxo_terrain_height(const vxo_terrain* t, uint32_t x, uint32_t z);
void vxo_terrain_heights(const vxo_terrain* t, const uint32_t* xs, const uint32_t* zs, uint64_t n, int32_t* out);
/* Brush mixes: 0 = 512^3 reference case (spheres r[4,16], 50/50 add);
 * 1 = large mix (50% spheres r[8,64], 50% boxes h[4,48], 60/40 add). */
void vxo_gen_brushes(uint64_t seed, uint64_t stream, uint32_t n,
                     uint32_t world, uint32_t mix, vxo_brush* out);
/* Voxels a brush covers inside the world (inclusive of overlap). */
uint64_t vxo_brush_voxels(const vxo_brush* b, uint32_t world);

/* ---- CPU Hash DAG (restatement of svt/hash_dag.hpp, extended) ----------
 * inner_depth D: world side 4 << D. materials = 0: occupancy-only DAG with
 * FULL sentinels (pinned against the compiled reference HashDAG, Appendix
 * semantics of hash_dag.hpp). materials = 1: 4^3 leaves carry one u8
 * material per voxel (0 = empty), no FULL sentinel — parity unpinned by the
 * reference (it has no materials); pinned by a dense brute-force builder. */
typedef struct vxo_dag vxo_dag;
vxo_dag* vxo_dag_create(uint32_t inner_depth, int materials);
void vxo_dag_destroy(vxo_dag* d);
void vxo_dag_build_terrain(vxo_dag* d, const vxo_terrain* t);
void vxo_dag_apply_brushes(vxo_dag* d, const vxo_brush* b, uint32_t n);
/* Voxel edits (reference place/destroy semantics; material used if
 * materials != 0).  xyz is n triples; out-of-world triples are ignored. */
void vxo_dag_apply_voxels(vxo_dag* d, const uint32_t* xyz, uint32_t n,
                          uint32_t material);
/* Half-open box edit (hash_dag.hpp:43-49). */
void vxo_dag_box(vxo_dag* d, uint32_t x0, uint32_t y0, uint32_t z0,
                 uint32_t x1, uint32_t y1, uint32_t z1, uint32_t material);
/* 0 if empty, else material (1 in occupancy mode). */
uint32_t vxo_dag_get(const vxo_dag* d, uint32_t x, uint32_t y, uint32_t z);
/* Ray walk, one voxel per crossing (definition in include/vx.h, vx_ray):
 * crossing j of axis a at t = (2j+1)/(2|d_a|), ties x before y before z.
 * rays: n x {ox, oy, oz, dx, dy, dz, max_steps, pad} (int32 / uint32);
 * hits: n x {material, steps, x, y, z, pad}. */
void vxo_dag_raycast(const vxo_dag* d, const int32_t* rays, uint32_t n, int32_t* hits);
/* Canonical depth-first export (format in the .cpp header comment). Returns
 * the number of u32 words; writes them if out != NULL and cap suffices. */
uint64_t vxo_dag_export(const vxo_dag* d, uint32_t* out, uint64_t cap);
uint64_t vxo_dag_stored_nodes(const vxo_dag* d);
uint64_t vxo_dag_stored_leaves(const vxo_dag* d);

/* Dense brute force (tiny worlds): builds the canonical export directly from
 * a dense material grid of side 4 << D (x fastest, then y, then z). */
uint64_t vxo_dense_export(const uint8_t* grid, uint32_t inner_depth,
                          int materials, uint32_t* out, uint64_t cap);

/* ---- CPU pointer SVO (restatement of svt/cuda's voxel edit) -------------
 * 4096^3 world, ten inner levels, 8-slot nodes (child index | bit 30 pointer
 * | bit 31 filled), 4^3 leaves as two u32 masks, LIFO free lists. Pinned by
 * the reference's own SVO tests (the .cuh files under test/cuda/svt, re-expressed on the
 * CPU) and by invariants: get() equals a set model of the field, live
 * counters equal the reachable counts, no uniform node on a touched path. */
typedef struct vxo_svo vxo_svo;
vxo_svo* vxo_svo_create(uint32_t max_nodes, uint32_t max_leaves);
void vxo_svo_destroy(vxo_svo* s);
void vxo_svo_reset(vxo_svo* s);
/* Leaf masks of a voxel batch, sorted by leaf key: out = 3 words per mask
 * {leaf_key, voxel_data_low, voxel_data_high}; returns the mask count. */
uint32_t vxo_svo_leaf_masks(const uint32_t* xyz, uint32_t n, uint32_t* out);
/* op 0 place, 1 destroy; returns 0, or 2 (out of capacity, nothing changed). */
int vxo_svo_apply(vxo_svo* s, const uint32_t* xyz, uint32_t n, uint32_t op);
uint32_t vxo_svo_get(const vxo_svo* s, uint32_t x, uint32_t y, uint32_t z);
/* {node_count, leaf_count, free_node_count, free_leaf_count} */
void vxo_svo_counters(const vxo_svo* s, uint32_t* out);
/* Canonical export: [0x5658534F, nodes, leaves], 8 slot words per node in
 * pre-order (pointer slots: canonical child id | the slot's flag bits), then
 * 2 words per leaf in the order reached. */
uint64_t vxo_svo_export(const vxo_svo* s, uint32_t* out, uint64_t cap);
uint32_t vxo_svo_leaf_key(uint32_t lx, uint32_t ly, uint32_t lz);
/* The reference benchmark's random voxel batch (xorshift32 from 0x9e3779b9). */
void vxo_svo_random_edits(uint32_t count, uint32_t* xyz);

/* FNV-1a 64 over u32 words (digest of a canonical export). */
uint64_t vxo_digest(const uint32_t* words, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
