// Ray traversal over the hash DAG: integer DDA with empty-subtree skipping.
//
// The voxel order is fixed by exact integer arithmetic (include/vx.h, vx_ray):
// crossing j of axis a at t = (2j + 1) / (2|d_a|), ties x < y < z. The state
// is the number of crossings k_a taken on each axis. Each iteration descends
// the DAG to the current voxel, restarting below the deepest node it shares
// with the previous voxel (the path is kept in LDS); an EMPTY ref at level l
// means the whole
// aligned cube of side world >> l is empty, and the ray leaves it in one
// step: the first crossing that exits the cube is found per axis, and the
// crossings of the other axes before it are counted in closed form — the
// same state a voxel-by-voxel walk would reach.
#include "vx_dag.hpp"

using namespace vx;
using namespace vx::dag;

namespace {

struct Ray {
  int64_t o[3];
  int32_t sgn[3];
  uint64_t ad[3];
  uint64_t k[3];
};

// Whether crossing (a, ja) of an axis with |d_a| = da comes before crossing
// (b, jb) of an axis with |d_b| = db.
__device__ __forceinline__ bool before(int a, uint64_t ja, uint64_t da, int b, uint64_t jb, uint64_t db) {
  const uint64_t lhs = (2 * ja + 1) * db, rhs = (2 * jb + 1) * da;
  return lhs < rhs || (lhs == rhs && a < b);
}

// Leave the aligned cube [lo, lo + side) that holds the current voxel: returns
// false if the ray cannot move (zero direction). The winning axis travels as
// values (its index, crossing and |d|), so every array index is a constant and
// the ray stays in registers.
__device__ __forceinline__ bool exit_cube(Ray& r, const int64_t (&lo)[3], int64_t side) {
  int best = -1;
  uint64_t jbest = 0, Y = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!r.ad[a]) continue;
    const int64_t j = r.sgn[a] > 0 ? lo[a] + side - 1 - r.o[a] : r.o[a] - lo[a];
    if (best < 0 || before(a, uint64_t(j), r.ad[a], best, jbest, Y)) {
      best = a;
      jbest = uint64_t(j);
      Y = r.ad[a];
    }
  }
  if (best < 0) return false;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    if (b == best) {
      r.k[b] = jbest + 1;
      continue;
    }
    if (!r.ad[b]) continue;
    const uint64_t X = (2 * jbest + 1) * r.ad[b];
    // crossings (b, j) before (best, jbest): (2j+1)Y < X, or == X when b < best
    uint64_t cnt;
    if (b < best) cnt = X >= Y ? (X - Y) / (2 * Y) + 1 : 0;
    else cnt = X > Y ? (X - Y - 1) / (2 * Y) + 1 : 0;
    if (cnt > r.k[b]) r.k[b] = cnt;
  }
  return true;
}

// Refs of the current descent path, one column per thread (level-major, so
// the threads of a wave hit distinct banks): path[l] is the node at level l
// that holds the last voxel visited.
constexpr uint32_t kRayBlock = 256;
constexpr uint32_t kMaxLevels = 16;

__global__ __launch_bounds__(kRayBlock) void raycast_kernel(Dev d, const vx_ray* __restrict__ rays, uint32_t n,
                                                            vx_hit* __restrict__ hits) {
  __shared__ uint32_t s_path[kMaxLevels][kRayBlock];
  uint32_t* path = &s_path[0][threadIdx.x];  // path[l * kRayBlock]
  const int64_t world = int64_t(4) << d.depth;
  const uint32_t log_world = d.depth + 2;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const vx_ray q = rays[i];
    Ray r;
    const int32_t dir[3] = {q.dx, q.dy, q.dz};
    r.o[0] = q.ox;
    r.o[1] = q.oy;
    r.o[2] = q.oz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r.sgn[a] = dir[a] > 0 ? 1 : (dir[a] < 0 ? -1 : 0);
      r.ad[a] = uint64_t(dir[a] < 0 ? -int64_t(dir[a]) : int64_t(dir[a]));
      r.k[a] = 0;
    }
    vx_hit h{0u, 0u, 0, 0, 0, 0u};
    path[0] = d.ctl[kCtlRoot];
    uint32_t plen = 0;                 // path[0..plen] hold the last descent
    uint32_t px = 0, py = 0, pz = 0;   // the last voxel visited
    for (;;) {
      int64_t v[3];
      bool inside = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        v[a] = r.o[a] + int64_t(r.sgn[a]) * int64_t(r.k[a]);
        inside = inside && v[a] >= 0 && v[a] < world;
      }
      const uint64_t steps = r.k[0] + r.k[1] + r.k[2];
      if (!inside || steps > q.max_steps) break;
      const uint32_t x = uint32_t(v[0]), y = uint32_t(v[1]), z = uint32_t(v[2]);
      // The new voxel shares the nodes of levels below the first differing
      // coordinate bit with the last one: the descent restarts there, not at
      // the root (the same nodes a descent from the root would read).
      uint32_t lvl = 0;
      const uint32_t diff = (x ^ px) | (y ^ py) | (z ^ pz);
      if (plen > 0 && diff != 0) {
        const uint32_t common = log_world - 1 - (31u - uint32_t(__builtin_clz(diff)));
        lvl = common < plen ? common : plen;
      }
      uint32_t ref = path[lvl * kRayBlock];
      for (; lvl < d.depth && ref > kFull; ++lvl) {
        const uint32_t sh = log_world - 1 - lvl;
        ref = child_of(d.node_pool, ref, ((x >> sh) & 1u) | (((y >> sh) & 1u) << 1) | (((z >> sh) & 1u) << 2));
        path[(lvl + 1) * kRayBlock] = ref;
      }
      plen = lvl;
      px = x;
      py = y;
      pz = z;
      uint32_t m = 0;
      int64_t side = 1;
      if (ref == kFull) {
        m = 1;
      } else if (ref == kEmpty) {
        side = world >> lvl;
      } else {
        const uint32_t bit = (x & 3u) + 4u * (y & 3u) + 16u * (z & 3u);
        m = d.leaf_words == 2 ? (d.leaf_pool[ref + (bit >> 5)] >> (bit & 31)) & 1u
                              : (d.leaf_pool[ref + (bit >> 2)] >> ((bit & 3) * 8)) & 0xffu;
      }
      if (m) {
        h = vx_hit{m, uint32_t(steps), int32_t(x), int32_t(y), int32_t(z), 0u};
        break;
      }
      const int64_t lo[3] = {v[0] & ~(side - 1), v[1] & ~(side - 1), v[2] & ~(side - 1)};
      if (!exit_cube(r, lo, side)) break;
    }
    hits[i] = h;
  }
}

}  // namespace

extern "C" vx_status vx_dag_raycast(const vx_dag* g, const vx_ray* rays, uint32_t n, vx_hit* hits, void* stream) {
  if (!g || ((!rays || !hits) && n)) return VX_INVALID_VALUE;
  if (n == 0) return VX_OK;
  if (g->depth + 1 > kMaxLevels) return VX_INVALID_VALUE;
  raycast_kernel<<<grid_for(n), kRayBlock, 0, static_cast<hipStream_t>(stream)>>>(g->dev, rays, n, hits);
  return last_launch();
}
