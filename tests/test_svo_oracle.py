"""The CPU pointer-SVO oracle pinned by the reference's own SVO tests
(test/cuda/svt/*.cuh, re-expressed in svo_cases.py) and by invariants that a
plausible slip (a lost fill bit, a skipped collapse level, a wrong free-list
order) breaks: the field equals a set model, the live counters equal the
reachable counts, and no uniform node or leaf survives on the tree."""
import numpy as np
import pytest

import svo_cases as S
from oracle import oracle as O


def _make(max_nodes, max_leaves):
    return O.Svo(max_nodes, max_leaves)


def test_leaf_masks_match_reference_tests():
    S.check_leaf_masks(O.svo_leaf_masks)


@pytest.mark.parametrize("check", S.ALL_TREE_CHECKS, ids=lambda f: f.__name__)
def test_reference_tree_checks(check):
    check(_make)


def test_leaf_key_matches_reference_definition():
    rng = np.random.default_rng(0)
    for lx, ly, lz in rng.integers(0, 1024, (200, 3)).tolist():
        assert O.svo_leaf_key(lx, ly, lz) == S.leaf_key(lx, ly, lz)


def test_random_edit_generator_first_values():
    # xorshift32 (13, 17, 5) from 0x9e3779b9 restated in Python.
    s, want = 0x9E3779B9, []
    for _ in range(30):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        want.append(s % 4096)
    assert O.svo_random_edits(10).ravel().tolist() == want


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_sequences_match_set_model(seed):
    """Mixed place / destroy batches in a 64^3 corner plus far voxels; after
    each batch the field equals a set, the export is canonical and the
    counters are the reachable ones."""
    rng = np.random.default_rng(seed)
    s = O.Svo(1 << 16, 1 << 16)
    model = set()
    for step in range(12):
        n = int(rng.integers(1, 3000))
        box = int(rng.choice([8, 16, 64]))
        xyz = rng.integers(0, box, (n, 3)).astype(np.uint32)
        if step % 4 == 3:  # a few far and out-of-world voxels
            xyz[:5] = rng.integers(0, 4096 + 8, (5, 3))
        op = int(rng.integers(0, 2))
        assert s.apply(xyz, op) == 0
        for x, y, z in xyz.tolist():
            if max(x, y, z) >= 4096:
                continue
            (model.discard if op else model.add)((x, y, z))
        q = np.concatenate([xyz, rng.integers(0, 64, (500, 3)).astype(np.uint32)])
        got = s.get_many(q)
        want = [1 if (x, y, z) in model else 0 for x, y, z in q.tolist()]
        assert got.tolist() == want, step
        S.assert_canonical(s.export(), s.counters())


def test_filled_regions_inherit_fill_and_recollapse():
    """Place a full 16^3 block (collapses to a filled slot), carve one voxel
    (paths materialise with the fill bit), refill it (collapses back)."""
    s = O.Svo(256, 64)
    s.place(S.box_edits(16, None))
    full = s.export()
    assert S.live(s.counters()) == (MAX_D := S.MAX_DEPTH - 2, 0)
    s.destroy([(5, 6, 7)])
    assert S.live(s.counters()) == (S.MAX_DEPTH, 1)
    nodes, leaves = S.parse_export(s.export())
    slots = S.path_slots(nodes, 5, 6, 7)
    # slots below the old filled cut carry the inherited fill bit
    assert all(v & S.FILLED_BIT for v in slots[MAX_D:])
    bit =(5 & 3) | ((6 & 3) << 2) | ((7 & 3) << 4)
    word = [0xFFFFFFFF, 0xFFFFFFFF]
    word[bit >> 5] &= ~(1 << (bit & 31)) & 0xFFFFFFFF
    assert leaves.tolist() == [word]
    assert s.get(5, 6, 7) == 0 and s.get(4, 6, 7) == 1
    s.place([(5, 6, 7)])
    assert np.array_equal(s.export(), full)
    assert S.live(s.counters()) == (S.MAX_DEPTH - 2, 0)


def test_capacity_overflow_changes_nothing():
    s = O.Svo(16, 4)
    s.place([(0, 0, 0)])
    e, c = s.export(), s.counters()
    assert s.place([(4 * i, 0, 0) for i in range(1, 8)]) == 2
    assert np.array_equal(s.export(), e) and s.counters() == c


def test_product_workload_generators():
    """The library's generators of the reference benchmark workloads (used by
    the benchmark) equal the oracle's and the reference's stated sizes."""
    from voxel_amd import svo
    assert np.array_equal(svo.random_edits(1 << 16), O.svo_random_edits(1 << 16))
    sph = svo.sphere_edits(2048, 2048, 2048, 128)
    assert len(sph) == 8782785
    assert np.array_equal(sph, S.sphere_edits(2048, 2048, 2048, 128))
