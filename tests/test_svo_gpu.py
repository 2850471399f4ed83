"""GPU pointer-SVO voxel edits (svt/cuda API) through the C ABI against the
CPU oracle: the reference's own SVO tests (test/cuda/svt/*.cuh, in
svo_cases.py), exact counters {node_count, leaf_count, free_node_count,
free_leaf_count} and canonical exports on random mixed batches, leaf masks on
large batches, and the reference benchmark's two workloads (a radius-128
sphere and 2^24 random voxels, bench/bench_cuda_svt.cu:46-112) at full size."""
import numpy as np
import pytest
import torch

import svo_cases as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class _Gpu:
    """The svo_cases interface over voxel_amd.svo.GpuSvo."""

    def __init__(self, max_nodes, max_leaves):
        from voxel_amd.svo import GpuSvo
        self.s = GpuSvo(max_nodes, max_leaves)

    def place(self, xyz):
        self.s.place_voxel_edits(xyz)

    def destroy(self, xyz):
        self.s.destroy_voxel_edits(xyz)

    def apply(self, xyz, op):
        self.s.apply_voxel_edits(xyz, op)

    def get_many(self, xyz):
        return self.s.get_voxels(xyz).cpu().numpy().astype(np.uint32)

    def counters(self):
        return self.s.counters()

    def export(self):
        return self.s.export()


def _gpu_masks(xyz):
    from voxel_amd.svo import voxel_edits_to_leaf_masks
    return voxel_edits_to_leaf_masks(xyz).cpu().numpy().view(np.uint32)


def test_leaf_masks_match_reference_tests(gpu):
    S.check_leaf_masks(_gpu_masks)


@pytest.mark.parametrize("check", S.ALL_TREE_CHECKS, ids=lambda f: f.__name__)
def test_reference_tree_checks(gpu, check):
    check(_Gpu)


@pytest.mark.parametrize("n,span,seed", [(1, 4096, 0), (63, 64, 1), (64, 64, 2), (65, 8, 3), (4097, 4096, 4),
                                         (100000, 32, 5), (300000, 4100, 6)])
def test_leaf_masks_match_oracle(gpu, n, span, seed):
    """Sizes around the 64-record wave chunks and the 8192-record tiles, dense
    duplicates (leaves spanning many chunks) and out-of-world edits."""
    rng = np.random.default_rng(seed)
    xyz = rng.integers(0, span, (n, 3)).astype(np.uint32)
    assert np.array_equal(_gpu_masks(xyz), O.svo_leaf_masks(xyz))


def _same(g, o):
    assert g.counters() == o.counters()
    assert np.array_equal(g.export(), o.export())


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_mixed_batches_match_oracle(gpu, seed):
    """Place / destroy batches over nested regions (dense corners, a filled
    block carved and refilled, scattered far voxels): counters and canonical
    export equal the oracle's after every batch."""
    rng = np.random.default_rng(seed)
    g, o = _Gpu(1 << 18, 1 << 18), O.Svo(1 << 18, 1 << 18)
    for step in range(14):
        box = int(rng.choice([8, 16, 64, 256]))
        n = int(rng.integers(1, 20000))
        base = rng.integers(0, 4096 - box, 3) if step % 3 == 0 else np.zeros(3, np.int64)
        xyz = (base + rng.integers(0, box, (n, 3))).astype(np.uint32)
        if step % 5 == 4:
            xyz = np.concatenate([xyz, S.box_edits_array(16)])
        if step % 4 == 3:
            xyz[:7] = rng.integers(0, 4104, (7, 3))
        op = int(rng.integers(0, 2))
        g.apply(xyz, op)
        assert o.apply(xyz, op) == 0
        _same(g, o)
    q = rng.integers(0, 300, (20000, 3)).astype(np.uint32)
    assert np.array_equal(g.get_many(q), o.get_many(q))


def test_fill_carve_refill_matches_oracle(gpu):
    g, o = _Gpu(1 << 12, 1 << 12), O.Svo(1 << 12, 1 << 12)
    for s in (g, o):
        s.place(S.box_edits_array(32))
    _same(g, o)
    for xyz, op in (([(5, 6, 7)], 1), ([(5, 6, 7)], 0), (S.box_edits_array(8), 1), ([(31, 31, 31), (0, 0, 0)], 1),
                    (S.box_edits_array(32), 0)):
        g.apply(xyz, op)
        o.apply(xyz, op)
        _same(g, o)


def test_capacity_overflow_changes_nothing(gpu):
    from voxel_amd import OutOfCapacity
    g = _Gpu(16, 4)
    g.place([(0, 0, 0)])
    e, c = g.export(), g.counters()
    with pytest.raises(OutOfCapacity):
        g.place([(4 * i, 0, 0) for i in range(1, 8)])
    assert np.array_equal(g.export(), e) and g.counters() == c


def test_invalid_arguments(gpu):
    from voxel_amd import InvalidValue
    from voxel_amd.svo import GpuSvo
    s = GpuSvo(256, 64)
    with pytest.raises(InvalidValue):
        s.apply_voxel_edits([(0, 0, 0)], 2)
    with pytest.raises(InvalidValue):
        s.apply_voxel_edits([(0, 0, 0)] * 100, 0, workspace=torch.empty(64, dtype=torch.uint8, device="cuda"))
    s.place_voxel_edits(np.zeros((0, 3), np.uint32))
    assert s.counters() == (1, 0, 0, 0)


def test_fresh_svo_is_reset_without_reset_call(gpu):
    """A tree straight from vx_svo_create is already reset (the reference's
    constructor only allocates): counters {1, 0, 0, 0}, and edits on it
    equal to the oracle's."""
    from voxel_amd.svo import GpuSvo
    g = GpuSvo(1 << 16, 1 << 14, reset=False)
    assert g.counters() == (1, 0, 0, 0)
    rng = np.random.default_rng(12)
    xyz = rng.integers(0, 4096, (4000, 3)).astype(np.uint32)
    ref = O.Svo(1 << 16, 1 << 14)
    g.place_voxel_edits(xyz)
    ref.apply(xyz, 0)
    g.destroy_voxel_edits(xyz[:2000])
    ref.apply(xyz[:2000], 1)
    assert g.counters() == ref.counters() and np.array_equal(g.export(), ref.export())


def test_reference_sphere_workload(gpu):
    """bench_cuda_svt.cu sphere case: r = 128 at the world centre, 8.78M voxels
    / 144k leaves, placed on a reset SVO of the benchmark's capacity
    (GpuSvo<2^26, 2^24>); then the half-overlap sphere destroyed."""
    sph = S.sphere_edits(2048, 2048, 2048, 128)
    assert len(sph) == 8782785
    g, o = _Gpu(1 << 26, 1 << 24), O.Svo(1 << 26, 1 << 24)
    g.place(sph)
    o.place(sph)
    _same(g, o)
    assert len(_gpu_masks(sph)) == 144416
    half = S.sphere_edits(2048 + 64, 2048, 2048, 128)
    g.destroy(half)
    o.destroy(half)
    _same(g, o)


def test_reference_random_workload(gpu):
    """bench_cuda_svt.cu random case: 2^24 xorshift voxels, 15.80M leaves;
    counters, leaf masks and every edited voxel equal the oracle's."""
    r = S.random_edits(1 << 24)
    g, o = _Gpu(1 << 26, 1 << 24), O.Svo(1 << 26, 1 << 24)
    m = _gpu_masks(r)
    assert len(m) == 15799303
    assert np.array_equal(m, O.svo_leaf_masks(r))
    g.place(r)
    o.place(r)
    assert g.counters() == o.counters()
    assert g.get_many(r).all()
    q = np.random.default_rng(0).integers(0, 4096, (1 << 16, 3)).astype(np.uint32)
    assert np.array_equal(g.get_many(q), o.get_many(q))
    d = r[: 1 << 22]
    g.destroy(d)
    o.destroy(d)
    assert g.counters() == o.counters()
    assert not g.get_many(d).any()
