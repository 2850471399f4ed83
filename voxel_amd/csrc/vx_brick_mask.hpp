// Voxel masks of one 4x4x4 brick under a brush: bit lx + 4 ly + 16 lz (the
// reference's leaf bit order, detail/node.cuh:90-94) is set iff the voxel lies
// inside. Sphere: dx^2 + dy^2 + dz^2 <= r^2, box: |dx| <= hx, |dy| <= hy,
// |dz| <= hz, with d the voxel's offset from the brush centre (the brush
// semantics of algorithm/brush.cuh). Host and device: the host build is what
// the CPU tests check against a voxel-by-voxel evaluation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "vx.h"

namespace vx::dag {

// Sphere voxels of one brick, row by row: voxel (lx, ly, lz) is inside iff
// dx^2 <= r^2 - dy^2 - dz^2. Every |d| <= r + 3 for a brick inside the
// brush's bounding box, so for r < 2^14 all sums stay below 2^30 in int32.
__host__ __device__ __forceinline__ uint64_t sphere_mask32(int32_t dx0, int32_t dy0, int32_t dz0, int32_t r2) {
  int32_t qx[4], qy[4], qz[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    qx[l] = (dx0 + l) * (dx0 + l);
    qy[l] = (dy0 + l) * (dy0 + l);
    qz[l] = (dz0 + l) * (dz0 + l);
  }
  // Outside iff rem - dx^2 < 0: collect sign bits (plain ALU work, no
  // compare-to-mask round trips), then invert.
  uint32_t out[2] = {0u, 0u};
#pragma unroll
  for (int lz = 0; lz < 4; ++lz)
#pragma unroll
    for (int ly = 0; ly < 4; ++ly) {
      const int32_t rem = r2 - qy[ly] - qz[lz];
#pragma unroll
      for (int lx = 0; lx < 4; ++lx)
        out[lz >> 1] |= (uint32_t(rem - qx[lx]) >> 31) << (lx + 4 * ly + 16 * (lz & 1));
    }
  return ~((uint64_t(out[1]) << 32) | out[0]);
}

// The same test for radii up to kSphere16MaxR, two voxels per 32-bit
// operation. Every |d| <= r + 3 = 103, so r^2 - dx^2 - dy^2 - dz^2 lies in
// [-21827, 10000]. With 32768 added, r^2 - dy^2 - dz^2 of the voxels at
// lz = 2P (low half) and lz = 2P + 1 (high half) sit in the two 16-bit halves
// of one word; subtracting dx^2 from both halves with one 32-bit subtraction
// never borrows across them (each half stays in [10941, 42768]), and bit 15
// of a half is set iff its voxel is inside. Those two bits lie 16 apart, as
// the two voxels' mask bits do, so a shift and an and-or place both: three
// operations per voxel pair.
constexpr int32_t kSphere16MaxR = 100;

// The dy, dz part of sphere_mask16, which every brick of one brick row (same
// y and z) shares: t[4 P + ly] holds r^2 - dy^2 - dz^2 + bias of voxel row
// (ly, lz = 2P) in its low half and (ly, lz = 2P + 1) in its high half.
__host__ __device__ __forceinline__ void sphere_row16(int32_t dy0, int32_t dz0, int32_t r2, uint32_t (&t)[8]) {
  constexpr uint32_t kBias = 32768u, kPair = 0x10001u;
#pragma unroll
  for (int P = 0; P < 2; ++P) {
    const int32_t z0 = dz0 + 2 * P, z1 = z0 + 1;
    const uint32_t zz = (uint32_t(r2 - z1 * z1 + int32_t(kBias)) << 16) | uint32_t(r2 - z0 * z0 + int32_t(kBias));
#pragma unroll
    for (int ly = 0; ly < 4; ++ly) t[4 * P + ly] = zz - uint32_t((dy0 + ly) * (dy0 + ly)) * kPair;
  }
}

// The brick at x offset dx0 of a row with terms t (sphere_row16).
__host__ __device__ __forceinline__ uint64_t sphere_mask16_row(int32_t dx0, const uint32_t (&t)[8]) {
  constexpr uint32_t kPair = 0x10001u;
  uint32_t qx[4];  // dx^2 in both halves
#pragma unroll
  for (int l = 0; l < 4; ++l) qx[l] = uint32_t((dx0 + l) * (dx0 + l)) * kPair;
  uint32_t out[2] = {0u, 0u};
#pragma unroll
  for (int P = 0; P < 2; ++P)
#pragma unroll
    for (int ly = 0; ly < 4; ++ly)
#pragma unroll
      for (int lx = 0; lx < 4; ++lx) {
        const uint32_t pos = uint32_t(lx + 4 * ly);
        out[P] |= ((t[4 * P + ly] - qx[lx]) >> (15u - pos)) & (kPair << pos);
      }
  return (uint64_t(out[1]) << 32) | out[0];
}

__host__ __device__ __forceinline__ uint64_t sphere_mask16(int32_t dx0, int32_t dy0, int32_t dz0, int32_t r2) {
  uint32_t t[8];
  sphere_row16(dy0, dz0, r2, t);
  return sphere_mask16_row(dx0, t);
}

// The same test in 64-bit arithmetic, for radii of 2^14 and more (out of
// line: it is rare and would otherwise cost every caller its registers).
__host__ __device__ __noinline__ inline uint64_t sphere_mask64(int64_t dx0, int64_t dy0, int64_t dz0, int64_t r2) {
  uint64_t m = 0;
  for (int lz = 0; lz < 4; ++lz)
    for (int ly = 0; ly < 4; ++ly) {
      const int64_t rem = r2 - (dy0 + ly) * (dy0 + ly) - (dz0 + lz) * (dz0 + lz);
      for (int lx = 0; lx < 4; ++lx)
        if ((dx0 + lx) * (dx0 + lx) <= rem) m |= 1ull << (lx + 4 * ly + 16 * lz);
    }
  return m;
}

// Box voxels of one brick in 64-bit arithmetic, for half extents of 2^30 and
// more (out of line, like sphere_mask64).
// (Arguments by value: a reference would force the caller's brush copy into
// scratch memory.)
__host__ __device__ __noinline__ inline uint64_t box_mask64(int64_t dx, int64_t dy, int64_t dz, int64_t hx, int64_t hy,
                                                            int64_t hz) {
  uint32_t xm = 0, ym = 0, zm = 0;
  for (int l = 0; l < 4; ++l) {
    xm |= uint32_t(llabs(dx + l) <= hx) << l;
    ym |= uint32_t(llabs(dy + l) <= hy) << l;
    zm |= uint32_t(llabs(dz + l) <= hz) << l;
  }
  uint64_t m = 0;
  for (int lz = 0; lz < 4; ++lz)
    for (int ly = 0; ly < 4; ++ly)
      if (((zm >> lz) & 1u) && ((ym >> ly) & 1u)) m |= uint64_t(xm) << (4 * ly + 16 * lz);
  return m;
}

// The offsets l in 0..3 with |d + l| <= h, as 4 bits: the interval
// [max(0, -h - d), min(3, h - d)] (h < 2^30 and |d| <= h + 3, so no sum
// overflows).
__host__ __device__ __forceinline__ uint32_t span4(int32_t d, int32_t h) {
  const int32_t lo = -h - d > 0 ? -h - d : 0, hi = h - d < 3 ? h - d : 3;
  const uint32_t ul = uint32_t(lo & 3), uh = uint32_t(hi & 3);
  return lo <= hi ? (2u << uh) - (1u << ul) : 0u;
}

// Box voxels of one brick for half extents below 2^30: one 16-bit z layer
// (the x bits in every nibble whose y bit is set), copied into the layers
// whose z bit is set by one multiplication per 32-bit half.
__host__ __device__ __forceinline__ uint64_t box_mask32(int32_t dx, int32_t dy, int32_t dz, int32_t hx, int32_t hy,
                                                       int32_t hz) {
  const uint32_t xm = span4(dx, hx), ym = span4(dy, hy), zm = span4(dz, hz);
  const uint32_t ynib = ((ym & 1u) | ((ym & 2u) << 3) | ((ym & 4u) << 6) | ((ym & 8u) << 9)) * 0xfu;
  const uint32_t layer = (xm * 0x1111u) & ynib;
  const uint32_t lo = layer * ((zm & 1u) | ((zm & 2u) << 15));
  const uint32_t hi = layer * (((zm >> 2) & 1u) | ((zm & 8u) << 13));
  return (uint64_t(hi) << 32) | lo;
}

// Voxel mask of brick (X, Y, Z) (its minimum corner) under brush b. The
// brick lies inside the brush's clipped bounding box, so every offset from
// the centre is at most the half extent + 3: 32-bit arithmetic is exact below
// the stated extents.
__host__ __device__ __forceinline__ uint64_t brick_mask(const vx_brush& b, int32_t X, int32_t Y, int32_t Z) {
  if (b.kind == 0) {
    if (b.hx <= kSphere16MaxR) return sphere_mask16(X - b.cx, Y - b.cy, Z - b.cz, b.hx * b.hx);
    if (b.hx < (1 << 14)) return sphere_mask32(X - b.cx, Y - b.cy, Z - b.cz, b.hx * b.hx);
    return sphere_mask64(int64_t(X) - b.cx, int64_t(Y) - b.cy, int64_t(Z) - b.cz, int64_t(b.hx) * b.hx);
  }
  if (b.hx >= (1 << 30) || b.hy >= (1 << 30) || b.hz >= (1 << 30))
    return box_mask64(int64_t(X) - b.cx, int64_t(Y) - b.cy, int64_t(Z) - b.cz, b.hx, b.hy, b.hz);
  return box_mask32(X - b.cx, Y - b.cy, Z - b.cz, b.hx, b.hy, b.hz);
}

}  // namespace vx::dag
