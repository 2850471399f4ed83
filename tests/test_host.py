"""Host-side logic of the C-ABI library (no GPU needed): workload generators
and voxel-op counting agree with the oracle's definitions."""
import os

import numpy as np

from oracle import oracle as O


def test_brush_generator_matches_oracle():
    from voxel_amd.dag import gen_brushes
    for seed, n, world, mix in ((1, 64, 512, 0), (3, 4096, 16384, 1), (17, 100, 1 << 15, 1)):
        assert np.array_equal(gen_brushes(seed, O.STREAM_BRUSH, n, world, mix),
                              O.gen_brushes(seed, O.STREAM_BRUSH, n, world, mix))


def test_voxel_op_count_matches_brute_force():
    from voxel_amd.dag import brush_voxel_ops
    br = O.gen_brushes(3, O.STREAM_BRUSH, 300, 1024, 1)
    assert brush_voxel_ops(br, 1024) == O.brush_voxels(br, 1024)
    br = O.gen_brushes(1, O.STREAM_BRUSH, 64, 512, 0)
    assert brush_voxel_ops(br, 512) == O.brush_voxels(br, 512)


def test_voxel_op_count_known_answers():
    from voxel_amd.dag import BRUSH_DTYPE, brush_voxel_ops
    b = np.zeros(3, BRUSH_DTYPE)
    b[0] = (10, 10, 10, 1, 0, 0, 0, 5)      # sphere r=1: 7 voxels
    b[1] = (0, 0, 0, 2, 3, 4, 1, 0)         # box clipped at the corner: 3*4*5
    b[2] = (100, 100, 100, 2, 0, 0, 0, 1)   # r=2: 33 voxels
    assert brush_voxel_ops(b, 64) == 7 + 60 + 0
    assert brush_voxel_ops(b, 128) == 7 + 60 + 33


def test_sized_batches_match_workload_definition():
    """Batching / serving requests (vx_gen_sized_batch): 1..max brushes, 50%
    spheres / 50% cubes, 60/40 add/remove, and a voxel-op count whose log2 is
    uniform over [lo, hi] (2^8..2^20 for batching) up to the rounding of the
    brush sizes."""
    from voxel_amd.dag import brush_voxel_ops, gen_sized_batch
    n, lo, hi, max_brushes, world = 1500, 8, 20, 16, 1 << 15
    logs, counts, kinds, adds, scenes = [], [], [], [], []
    for i in range(n):
        br, s = gen_sized_batch(3, 6, i, lo, hi, max_brushes, world, 8)
        counts.append(len(br))
        kinds += br["kind"].tolist()
        adds += (br["material"] > 0).tolist()
        scenes.append(s)
        logs.append(np.log2(brush_voxel_ops(br, world)))
    counts, kinds, adds, logs = map(np.asarray, (counts, kinds, adds, logs))
    assert counts.min() == 1 and counts.max() == max_brushes
    assert sorted(set(scenes)) == list(range(8))
    m = len(kinds)
    assert abs(kinds.mean() - 0.5) < 5 * np.sqrt(0.25 / m)
    assert abs(adds.mean() - 0.6) < 5 * np.sqrt(0.24 / m)
    q = np.quantile(logs, [0.1, 0.5, 0.9])
    want = [lo + 0.1 * (hi - lo), (lo + hi) / 2, lo + 0.9 * (hi - lo)]
    assert all(abs(a - b) < 1.0 for a, b in zip(q, want)), q
    assert logs.min() > lo - 2 and logs.max() < hi + 1


def test_full_size_sort_check_rejects_planted_mistakes():
    """The bench's size-independent check of a stable (key, index) sort, run
    on CPU tensors: it accepts numpy's stable argsort and rejects an output
    with two tied values swapped (unstable), a key that does not match its
    index, keys out of order, and a repeated value."""
    import sys

    import torch
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    rng = np.random.default_rng(5)
    src = rng.integers(0, 1 << 32, 5000, dtype=np.uint64).astype(np.uint32)
    src[::7] = src[3]  # ties
    order = np.argsort(src, kind="stable").astype(np.int32)
    tk = torch.from_numpy(src.view(np.int32).copy())
    good_k, good_v = tk[torch.from_numpy(order).long()].clone(), torch.from_numpy(order.copy())
    assert bench.sorted_pairs_ok(torch, tk, good_k, good_v, unsigned32=True)["ok"]
    tied = np.nonzero(src[order][1:] == src[order][:-1])[0][0]
    v = good_v.clone()
    v[tied], v[tied + 1] = good_v[tied + 1], good_v[tied]
    assert not bench.sorted_pairs_ok(torch, tk, good_k, v, unsigned32=True)["ok"]
    k = good_k.clone()
    k[10] ^= 1
    assert not bench.sorted_pairs_ok(torch, tk, k, good_v, unsigned32=True)["ok"]
    k, v = good_k.clone(), good_v.clone()
    k[[0, -1]], v[[0, -1]] = good_k[[-1, 0]], good_v[[-1, 0]]
    assert not bench.sorted_pairs_ok(torch, tk, k, v, unsigned32=True)["ok"]
    v = good_v.clone()
    v[1] = v[0]
    assert not bench.sorted_pairs_ok(torch, tk, good_k, v, unsigned32=True)["ok"]


def test_brick_masks_match_voxel_by_voxel_evaluation():
    """The brick voxel masks of the record kernel (vx_brick_mask.hpp, built for
    the host) against the brush definitions evaluated voxel by voxel in 64-bit
    arithmetic: exhaustive for small extents, seeded samples over each form's
    whole range (tests/csrc/brick_mask_check.cpp)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "_build", "brick_mask_check")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    subprocess.check_call(["/opt/rocm/llvm/bin/clang++", "-O2", "-std=c++20", "-D__HIP_PLATFORM_AMD__",
                           "-I/opt/rocm/include", "-I" + os.path.join(root, "include"),
                           "-I" + os.path.join(root, "voxel_amd", "csrc"),
                           os.path.join(root, "tests", "csrc", "brick_mask_check.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "brick masks: ok" in r.stdout, r.stdout[-3000:]
