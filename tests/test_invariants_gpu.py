"""Table invariants and repeat-run determinism on the GPU.

The device-side duplicate-content check (no two live slots of a table hold
equal items; every live slot reachable from its hash's home position, with the
item's fingerprint) must flag planted duplicates and broken slots in hand-built
tables, and must find none after every kind of edit: voxel edits, brush
batches, heightfield scenes, many scenes in one pipeline, served requests and
garbage collection. Repeated runs and different sort tunings give the same
scenes."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _chunk(h, c):
    """One 64-bit chunk (two item words) folded into the item hash."""
    h = ((h ^ c) * 0x9E3779B97F4A7C15) & M64
    return h ^ (h >> 32)


def _final(h):
    """MurmurHash3's fmix64 (published constants)."""
    h ^= h >> 33
    h = (h * 0xFF51AFD7ED558CCD) & M64
    h ^= h >> 33
    h = (h * 0xC4CEB9FE1A85EC53) & M64
    return h ^ (h >> 33)


def _hash(words):
    """The tables' item hash: the words two at a time as 64-bit chunks (low
    word first), a last odd word alone, then the final avalanche."""
    w = [int(x) for x in words]
    h = 0x243F6A8885A308D3 ^ len(w)
    k = 0
    while k + 1 < len(w):
        h = _chunk(h, w[k] | (w[k + 1] << 32))
        k += 2
    if k < len(w):
        h = _chunk(h, w[k])
    return _final(h)


def _build(items, n_slots=64):
    """A table laid out as the library lays out its own: items appended to a
    pool from offset 16, slot {fingerprint | offset} at the first free slot
    from the item's home position. Returns (slots, pool, slot of each item)."""
    pool = [0] * 16
    slots = [0] * n_slots
    where = []
    for it in items:
        off = len(pool)
        pool += [int(x) for x in it]
        h = _hash(it)
        i = h & (n_slots - 1)
        while slots[i]:
            i = (i + 1) & (n_slots - 1)
        slots[i] = ((h >> 32) << 32) | off
        where.append(i)
    return np.array(slots, np.uint64), np.array(pool, np.uint32), where


def _check(slots, pool, item_words, gpu):
    from voxel_amd._lib import check, lib
    import voxel_amd.dag  # noqa: F401  (registers the signatures)
    ds = torch.from_numpy(slots.view(np.int64).copy()).to(gpu)
    dp = torch.from_numpy(pool.view(np.int32).copy()).to(gpu)
    out = (C.c_uint64 * 3)()
    check(lib().vx_table_check_unique(C.c_void_p(ds.data_ptr()), len(slots), C.c_void_p(dp.data_ptr()), len(pool),
                                      item_words, out, None), "vx_table_check_unique")
    return list(out)


def _items(n, words, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 1 << 32, words, dtype=np.uint64).astype(np.uint32) for _ in range(n)]


def test_clean_table_has_no_findings(gpu):
    slots, pool, _ = _build(_items(20, 2, 1))
    assert _check(slots, pool, 2, gpu) == [20, 0, 0]


def test_planted_duplicates_are_found(gpu):
    items = _items(12, 16, 2)
    slots, pool, _ = _build(items + [items[3], items[7], items[3]])
    live, dups, bad = _check(slots, pool, 16, gpu)
    assert (live, dups, bad) == (15, 3, 0)


def test_dead_slots_are_skipped(gpu):
    items = _items(10, 2, 3)
    slots, pool, where = _build(items + [items[4]])
    slots[where[4]] = (int(slots[where[4]]) & ~0xFFFFFFFF) | 1  # the first copy dies (kTomb)
    assert _check(slots, pool, 2, gpu) == [10, 0, 0]


def test_broken_slots_are_found(gpu):
    items = _items(10, 2, 4)
    slots, pool, where = _build(items)
    s = slots.copy()
    s[where[0]] ^= np.uint64(1 << 40)  # fingerprint not the item's
    assert _check(s, pool, 2, gpu)[2] == 1
    s = slots.copy()
    s[where[1]] |= np.uint64(0x80000000)  # a tag left over
    assert _check(s, pool, 2, gpu)[2] == 1
    s = slots.copy()
    s[where[2]] = np.uint64((int(slots[where[2]]) & ~0xFFFFFFFF) | 0x7000)  # offset beyond the used words
    assert _check(s, pool, 2, gpu)[2] == 1
    # An item moved past an empty slot on its walk: no lookup reaches it.
    s = slots.copy()
    i = where[5]
    j = (i + 2) % len(s)
    while s[j] or s[(j - 1) % len(s)]:
        j = (j + 1) % len(s)
    s[j], s[i] = s[i], np.uint64(0)
    assert _check(s, pool, 2, gpu)[2] >= 1


def test_node_items_use_header_widths(gpu):
    """item_words 0: DAG node items [child mask | fill mask << 8, one ref per
    set child-mask bit], 1..9 words."""
    nodes = [np.array([0b1011 | (0b100 << 8), 40, 41, 42], np.uint32), np.array([0x01, 77], np.uint32),
             np.array([0xFF] + list(range(100, 108)), np.uint32), np.array([0 | (0xF0 << 8)], np.uint32)]
    slots, pool, _ = _build(nodes + [nodes[2], nodes[3]])
    assert _check(slots, pool, 0, gpu) == [6, 2, 0]


def _clean(d):
    r = d.check_unique()
    for kind in ("nodes", "leaves"):
        assert r[kind]["duplicates"] == 0 and r[kind]["broken"] == 0, (kind, r)
    return r


def _dag(depth, materials, **kw):
    from voxel_amd.dag import HashDagGpu
    return HashDagGpu(depth, materials, **kw)


@pytest.mark.parametrize("materials", [False, True])
def test_no_duplicates_after_voxel_and_brush_edits(gpu, materials):
    from voxel_amd.dag import terrain_for_world
    d = _dag(7, materials, node_words=1 << 23, leaf_words=1 << 24)
    d.build_terrain(terrain_for_world(512, 5))
    r = _clean(d)
    assert r["nodes"]["live"] > 0 and r["leaves"]["live"] > 0
    for k in range(3):
        d.apply_brushes(O.gen_brushes(40 + k, O.STREAM_BRUSH, 200, 512, 1))
        _clean(d)
    rng = np.random.default_rng(6)
    xyz = torch.from_numpy(rng.integers(0, 512, (1 << 14, 3)).astype(np.uint32).view(np.int32)).to(gpu)
    d.place_voxel_edits(xyz, material=9 if materials else 1)
    _clean(d)
    d.destroy_voxel_edits(xyz[: 1 << 13])
    _clean(d)


def test_no_duplicates_across_scenes_in_one_pipeline_and_after_gc(gpu):
    from voxel_amd.dag import apply_batches, terrain_for_world
    depths = [5, 7, 6]
    dags = [_dag(dd, True, node_words=1 << 22, leaf_words=1 << 23) for dd in depths]
    for k, (dd, d) in enumerate(zip(depths, dags)):
        d.build_terrain(terrain_for_world(4 << dd, 20 + k))
    order = [0, 1, 2, 1, 0, 2, 1]
    batches = [O.gen_brushes(300 + i, O.STREAM_BRUSH, 40, 4 << depths[s], 1) for i, s in enumerate(order)]
    apply_batches([dags[s] for s in order], batches)
    for d in dags:
        _clean(d)
    before = dags[1].export()
    dags[1].collect_garbage()
    _clean(dags[1])
    assert np.array_equal(dags[1].export(), before)


def test_no_duplicates_after_serving(gpu):
    from voxel_amd.dag import terrain_for_world
    from voxel_amd.serving import Server, make_requests, poisson_arrivals, serve
    depths = [5, 6, 6]
    dags = [_dag(dd, True) for dd in depths]
    for k, (dd, d) in enumerate(zip(depths, dags)):
        d.build_terrain(terrain_for_world(4 << dd, 30 + k))
    reqs = make_requests(36, len(depths), [4 << dd for dd in depths], seed=31, lo_log2=6, hi_log2=12, max_brushes=8)
    srv = Server(dags, max_candidates=20_000, max_jobs=256)
    res = serve(srv, reqs, poisson_arrivals(len(reqs), 4000.0, 2))
    srv.close()
    assert res["completed"] == len(reqs)
    for d in dags:
        _clean(d)


def _scene_after_batches(materials):
    from voxel_amd.dag import terrain_for_world
    d = _dag(7, materials, node_words=1 << 23, leaf_words=1 << 24)
    d.build_terrain(terrain_for_world(512, 8))
    for k in range(3):
        d.apply_brushes(O.gen_brushes(60 + k, O.STREAM_BRUSH, 300, 512, 1))
    return d


@pytest.mark.parametrize("materials", [False, True])
def test_repeated_runs_give_the_same_scene(gpu, materials):
    """Same scene and batches three times on fresh scenes: the same canonical
    DAG and point queries, and the invariant each time. (The raw pools are not
    reproduced: which of several equal items of a batch becomes the stored
    one is decided by a CAS race, so pool offsets, and with them accidental
    sharing of node items between levels, differ from run to run.)"""
    runs = [_scene_after_batches(materials) for _ in range(3)]
    rng = np.random.default_rng(9)
    q = torch.from_numpy(rng.integers(0, 512, (1 << 16, 3)).astype(np.uint32).view(np.int32)).to(gpu)
    e0, g0 = runs[0].export(), runs[0].get_voxels(q).cpu()
    for d in runs:
        _clean(d)
        assert np.array_equal(d.export(), e0)
        assert torch.equal(d.get_voxels(q).cpu(), g0)


def test_sort_tunings_give_the_same_scene(gpu):
    from voxel_amd.primitives import SORT_KEYS_ONLY, get_sort_tuning, set_sort_tuning
    row = 8 | SORT_KEYS_ONLY
    saved = get_sort_tuning(row)
    exports = []
    try:
        for tn in ((9, 15, 512), (8, 15, 1024), (4, 23, 512)):
            set_sort_tuning(row, *tn)
            d = _scene_after_batches(True)
            _clean(d)
            exports.append(d.export())
    finally:
        set_sort_tuning(row, *saved)
    for e in exports[1:]:
        assert np.array_equal(e, exports[0])
