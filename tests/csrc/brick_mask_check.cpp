// Brick voxel masks (voxel_amd/csrc/vx_brick_mask.hpp, host build) against a
// voxel-by-voxel evaluation of the brush definitions in 64-bit arithmetic:
// sphere dx^2 + dy^2 + dz^2 <= r^2, box |dx| <= hx && |dy| <= hy && |dz| <= hz.
// Exhaustive over every brick offset for small extents, seeded samples over
// the whole range each mask is used for (a brick meeting the brush's bounding
// box: every voxel offset within the half extent + 3). Prints "brick masks: ok".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>

#include "vx_brick_mask.hpp"

using namespace vx::dag;

static uint64_t ref_sphere(int64_t dx0, int64_t dy0, int64_t dz0, int64_t r) {
  uint64_t m = 0;
  for (int lz = 0; lz < 4; ++lz)
    for (int ly = 0; ly < 4; ++ly)
      for (int lx = 0; lx < 4; ++lx) {
        const int64_t x = dx0 + lx, y = dy0 + ly, z = dz0 + lz;
        if (x * x + y * y + z * z <= r * r) m |= 1ull << (lx + 4 * ly + 16 * lz);
      }
  return m;
}

static uint64_t ref_box(int64_t dx0, int64_t dy0, int64_t dz0, int64_t hx, int64_t hy, int64_t hz) {
  uint64_t m = 0;
  for (int lz = 0; lz < 4; ++lz)
    for (int ly = 0; ly < 4; ++ly)
      for (int lx = 0; lx < 4; ++lx)
        if (llabs(dx0 + lx) <= hx && llabs(dy0 + ly) <= hy && llabs(dz0 + lz) <= hz)
          m |= 1ull << (lx + 4 * ly + 16 * lz);
  return m;
}

static long failures = 0, checks = 0;

static void expect(uint64_t got, uint64_t want, const char* what, int64_t a, int64_t b, int64_t c, int64_t d) {
  ++checks;
  if (got != want && failures++ < 10)
    std::printf("MISMATCH %s (%lld %lld %lld | %lld): got %016llx want %016llx\n", what, (long long)a, (long long)b,
                (long long)c, (long long)d, (unsigned long long)got, (unsigned long long)want);
}

// A brush at the origin and a brick at offset (dx0, dy0, dz0): brick_mask takes
// the brick corner, so the centre is put at minus the offset.
static uint64_t via_brick(uint32_t kind, int32_t dx0, int32_t dy0, int32_t dz0, int32_t hx, int32_t hy, int32_t hz) {
  const vx_brush b{-dx0, -dy0, -dz0, hx, hy, hz, kind, 1u};
  return brick_mask(b, 0, 0, 0);
}

int main() {
  std::mt19937_64 rng(20240611);
  auto uni = [&rng](int64_t lo, int64_t hi) { return lo + int64_t(rng() % uint64_t(hi - lo + 1)); };

  // Spheres, exhaustive for r <= 14 (both the 16-bit and 32-bit forms), then
  // samples up to the 16-bit form's limit and through the 32-bit range.
  for (int32_t r = 0; r <= 14; ++r)
    for (int32_t dz = -r - 3; dz <= r; ++dz)
      for (int32_t dy = -r - 3; dy <= r; ++dy)
        for (int32_t dx = -r - 3; dx <= r; ++dx) {
          const uint64_t want = ref_sphere(dx, dy, dz, r);
          expect(sphere_mask16(dx, dy, dz, r * r), want, "sphere16", dx, dy, dz, r);
          expect(sphere_mask32(dx, dy, dz, r * r), want, "sphere32", dx, dy, dz, r);
          expect(via_brick(0, dx, dy, dz, r, 0, 0), want, "brick sphere", dx, dy, dz, r);
        }
  for (int i = 0; i < 3000000; ++i) {
    const int32_t r = int32_t(uni(0, kSphere16MaxR));
    // Offsets near the surface are where a lost borrow or sign would show.
    const bool shell = i & 1;
    auto pick = [&]() { return int32_t(uni(-r - 3, r)); };
    int32_t dx = pick(), dy = pick(), dz = pick();
    if (shell) {
      // A voxel of the brick near the sphere surface along a random direction.
      const double t = double(uni(0, 1 << 20)) / double(1 << 20), u = double(uni(0, 1 << 20)) / double(1 << 20);
      const double th = 6.283185307179586 * t, ph = 3.141592653589793 * u;
      dx = int32_t(r * std::cos(th) * std::sin(ph)) - int32_t(uni(0, 3));
      dy = int32_t(r * std::sin(th) * std::sin(ph)) - int32_t(uni(0, 3));
      dz = int32_t(r * std::cos(ph)) - int32_t(uni(0, 3));
      if (dx < -r - 3) dx = -r - 3;
      if (dy < -r - 3) dy = -r - 3;
      if (dz < -r - 3) dz = -r - 3;
      if (dx > r) dx = r;
      if (dy > r) dy = r;
      if (dz > r) dz = r;
    }
    const uint64_t want = ref_sphere(dx, dy, dz, r);
    expect(sphere_mask16(dx, dy, dz, r * r), want, "sphere16", dx, dy, dz, r);
    expect(via_brick(0, dx, dy, dz, (i & 2) ? r : -r, 0, 0), want, "brick sphere", dx, dy, dz, r);
  }
  for (int i = 0; i < 1000000; ++i) {
    const int32_t r = int32_t(uni(kSphere16MaxR + 1, (1 << 14) - 1));
    const int32_t dx = int32_t(uni(-r - 3, r)), dy = int32_t(uni(-r - 3, r)), dz = int32_t(uni(-r - 3, r));
    expect(via_brick(0, dx, dy, dz, r, 0, 0), ref_sphere(dx, dy, dz, r), "brick sphere32", dx, dy, dz, r);
  }

  // Boxes: exhaustive for half extents <= 6 per axis, then samples up to
  // just below 2^30 (the 32-bit form's limit) and beyond (the 64-bit form).
  for (int32_t hz = 0; hz <= 6; hz += 3)
    for (int32_t hy = 0; hy <= 6; ++hy)
      for (int32_t hx = 0; hx <= 6; ++hx)
        for (int32_t dz = -hz - 3; dz <= hz; ++dz)
          for (int32_t dy = -hy - 3; dy <= hy; ++dy)
            for (int32_t dx = -hx - 3; dx <= hx; ++dx) {
              const uint64_t want = ref_box(dx, dy, dz, hx, hy, hz);
              expect(box_mask32(dx, dy, dz, hx, hy, hz), want, "box32", dx, dy, dz, hx);
              expect(via_brick(1, dx, dy, dz, hx, hy, hz), want, "brick box", dx, dy, dz, hx);
            }
  for (int i = 0; i < 3000000; ++i) {
    const int32_t top = (i % 3 == 0) ? 64 : (1 << 30) - 1;
    const int32_t hx = int32_t(uni(0, top)), hy = int32_t(uni(0, top)), hz = int32_t(uni(0, top));
    // Offsets near a face (where the interval ends) half of the time.
    auto pick = [&](int32_t h) {
      if (i & 1) return int32_t(uni(-h - 3, h));
      return (i & 2) ? int32_t(h - uni(0, 3)) : int32_t(-h - uni(0, 3));
    };
    const int32_t dx = pick(hx), dy = pick(hy), dz = pick(hz);
    const uint64_t want = ref_box(dx, dy, dz, hx, hy, hz);
    expect(box_mask32(dx, dy, dz, hx, hy, hz), want, "box32", dx, dy, dz, hx);
    expect(via_brick(1, dx, dy, dz, hx, hy, hz), want, "brick box", dx, dy, dz, hx);
  }
  for (int i = 0; i < 200000; ++i) {
    const int32_t hx = int32_t(uni(1 << 30, 0x7ffffff0)), hy = int32_t(uni(0, 100)), hz = int32_t(uni(0, 100));
    const int32_t dx = int32_t(uni(-(1 << 30), 1 << 30)), dy = int32_t(uni(-hy - 3, hy)), dz = int32_t(uni(-hz - 3, hz));
    expect(via_brick(1, dx, dy, dz, hx, hy, hz), ref_box(dx, dy, dz, hx, hy, hz), "brick box64", dx, dy, dz, hx);
  }
  if (failures) {
    std::printf("brick masks: %ld of %ld checks failed\n", failures, checks);
    return 1;
  }
  std::printf("brick masks: ok (%ld checks)\n", checks);
  return 0;
}
