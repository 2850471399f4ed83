"""The CPU oracle pinned against published vectors, the reference's own
fixed vectors, closed forms, the compiled reference HashDAG (when built) and
a dense brute force."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_splitmix64_published_vector():
    # splitmix64 seeded with 0: first outputs (Vigna's reference implementation).
    G = 0x9E3779B97F4A7C15
    got = [O.lib().vxo_mix64((k * G) & (2**64 - 1)) for k in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_scan_reference_fixed_vector():
    a = [3, 1, 4, 1, 5, 9, 2, 6, 5]
    assert O.inclusive_sum(a).tolist() == [3, 4, 8, 9, 14, 23, 25, 31, 36]
    assert O.exclusive_sum(a).tolist() == [0, 3, 4, 8, 9, 14, 23, 25, 31]
    assert O.inclusive_max(a).tolist() == [3, 3, 4, 4, 5, 9, 9, 9, 9]


def test_scan_matches_numpy_cumsum():
    a = O.gen_u32(123, O.STREAM_SCAN, 100_000, 0, 0xFFFFFFFF)
    assert np.array_equal(O.inclusive_sum(a), np.cumsum(a, dtype=np.uint64).astype(np.uint32))
    assert np.array_equal(O.inclusive_max(a), np.maximum.accumulate(a))


def test_sort_is_stable_lexicographic():
    k = O.gen_u32(7, O.STREAM_KEYS, 50_000, 0, 255)
    v = np.arange(k.size, dtype=np.uint32)
    sk, sv = O.sort_pairs(k, v)
    order = np.lexsort((v, k))
    assert np.array_equal(sk, k[order]) and np.array_equal(sv, v[order])


def test_sort_lower_bits_reference_vectors():
    k, _ = O.sort_pairs(np.array([0x30, 0x21, 0x12, 0x03, 0x34, 0x25, 0x16, 0x07], np.uint32), None, 4)
    assert k.tolist() == [0x30, 0x21, 0x12, 0x03, 0x34, 0x25, 0x16, 0x07]
    k, _ = O.sort_pairs(np.array([0x1ff, 0x281, 0x142, 0x003, 0x3c0, 0x224, 0x165, 0x026], np.uint32), None, 8)
    assert k.tolist() == [0x003, 0x224, 0x026, 0x142, 0x165, 0x281, 0x3c0, 0x1ff]


def test_morton_closed_form():
    assert O.lib().vxo_morton3(1, 0, 0) == 1 and O.lib().vxo_morton3(0, 1, 0) == 2
    assert O.lib().vxo_morton3(0, 0, 1) == 4 and O.lib().vxo_morton3(2, 0, 0) == 8
    assert O.lib().vxo_morton3(16383, 16383, 16383) == (1 << 42) - 1


def test_unique_keys_are_unique_and_never_reserved():
    k = O.gen_unique_u64(2, O.STREAM_HASH, 200_000)
    assert np.unique(k).size == k.size and not (k == np.uint64(2**64 - 1)).any()


def _dense_after(world, terrain, brushes, materials):
    h = np.array([[O.terrain_height(terrain, x, z) for x in range(world)] for z in range(world)])
    y = np.arange(world)[None, :, None]
    g = np.where(y <= h[:, None, :], (1 + (y % 255)) if materials else 1, 0).astype(np.uint8)
    g = np.broadcast_to(g, (world,) * 3).copy()
    zz, yy, xx = np.meshgrid(*(np.arange(world),) * 3, indexing="ij")
    for b in brushes:
        if b["kind"] == 0:
            m = (xx - b["cx"]) ** 2 + (yy - b["cy"]) ** 2 + (zz - b["cz"]) ** 2 <= int(b["hx"]) ** 2
        else:
            m = (abs(xx - b["cx"]) <= b["hx"]) & (abs(yy - b["cy"]) <= b["hy"]) & (abs(zz - b["cz"]) <= b["hz"])
        g[m] = b["material"] if materials else (1 if b["material"] else 0)
    return g


@pytest.mark.parametrize("materials", [False, True])
def test_dag_terrain_and_brushes_match_dense_brute_force(materials):
    D = 4
    W = 4 << D
    t = O.terrain_for_world(W, 1)
    br = O.gen_brushes(7, O.STREAM_BRUSH, 40, W, 1)
    for f in ("hx", "hy", "hz"):
        br[f] = np.maximum(1, br[f] // 8)
    d = O.Dag(D, materials)
    d.build_terrain(t)
    d.apply_brushes(br)
    g = _dense_after(W, t, br, materials)
    assert np.array_equal(d.export(), O.dense_export(g, D, materials))
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, W, (300, 3)):
        assert d.get(int(x), int(y), int(z)) == g[z, y, x]


def test_dag_canonical_form_reference_known_answers():
    # test/svt/test_svt_hash_dag.cpp: fill/clear of the whole volume uses the
    # reserved FULL/EMPTY states; identical content gives identical export.
    d = O.Dag(3, False)
    d.box((0, 0, 0), (32, 32, 32), 1)
    e = d.export()
    assert e[3] == 1 and e[4] == 0 and e[5] == 0
    d.box((0, 0, 0), (32, 32, 32), 0)
    assert d.export()[3] == 0
    a, b = O.Dag(3, False), O.Dag(3, False)
    a.box((1, 2, 3), (9, 9, 9), 1)
    for x in range(1, 9):
        for y in range(2, 9):
            for z in range(3, 9):
                b.apply_voxels([x, y, z], 1)
    assert np.array_equal(a.export(), b.export())


def test_dag_matches_reference_golden_vectors():
    """Exports of the reference's own HashDAG on fixed edit scripts, recorded
    by tests/golden/make_golden.py from the compiled reference."""
    with open(os.path.join(GOLDEN, "ref_hash_dag.json")) as f:
        cases = json.load(f)
    assert cases
    for case in cases:
        d = O.Dag(case["depth"], False)
        for op in case["ops"]:
            d.box(tuple(op[0:3]), tuple(op[3:6]), op[6])
        e = d.export()
        assert int(e[4]) == case["nodes"] and int(e[5]) == case["leaves"], case["name"]
        assert O.digest(e) == int(case["digest"], 16), case["name"]
        if "export" in case:
            assert e.tolist() == case["export"], case["name"]


@pytest.mark.skipif(not O.ref_available(), reason="reference not built here")
def test_dag_matches_compiled_reference_random_boxes():
    rng = np.random.default_rng(0)
    for D in (3, 5):
        W = 4 << D
        r, o = O.RefDag(D), O.Dag(D, False)
        for _ in range(300):
            lo = rng.integers(0, W, 3)
            hi = np.minimum(lo + rng.integers(1, W // 2, 3), W)
            fill = bool(rng.integers(0, 2))
            r.box(tuple(map(int, lo)), tuple(map(int, hi)), fill)
            o.box(tuple(map(int, lo)), tuple(map(int, hi)), int(fill))
        assert np.array_equal(r.export(), o.export())
        # counts: canonical = stored - retained full chain (hash_dag.hpp:329-341)
        r.collect_garbage()
        e = o.export()
        nodes, leaves = r.counts()
        assert nodes - D == e[4] or e[3] != 2
        assert leaves - 1 == e[5] or e[3] != 2


def test_config0_occupancy_matches_reference_golden():
    """512^3 config-0 terrain + 64 spheres: the oracle's hierarchical terrain
    builder and sphere edit equal the reference HashDAG fed column and row boxes."""
    with open(os.path.join(GOLDEN, "ref_config0.json")) as f:
        g = json.load(f)
    d = O.Dag(g["depth"], False)
    world = 4 << g["depth"]
    d.build_terrain(O.terrain_for_world(world, g["seed"]))
    e = d.export()
    assert (int(e[4]), int(e[5]), "%016x" % O.digest(e)) == (g["terrain"]["nodes"], g["terrain"]["leaves"],
                                                             g["terrain"]["digest"])
    d.apply_brushes(O.gen_brushes(g["seed"], O.STREAM_BRUSH, g["brushes"], world, g["mix"]))
    e = d.export()
    assert (int(e[4]), int(e[5]), "%016x" % O.digest(e)) == (g["edited"]["nodes"], g["edited"]["leaves"],
                                                             g["edited"]["digest"])


@pytest.mark.skipif(not O.ref_available(), reason="reference not built here")
def test_reference_box_scripts_reproduce_config0_golden():
    """The column / row box scripts the benchmark feeds the compiled reference
    HashDAG (one native call per script) reproduce the committed golden."""
    with open(os.path.join(GOLDEN, "ref_config0.json")) as f:
        g = json.load(f)
    world = 4 << g["depth"]
    r = O.RefDag(g["depth"])
    r.boxes(O.terrain_column_boxes(O.terrain_for_world(world, g["seed"]), world))
    e = r.export()
    assert "%016x" % O.digest(e) == g["terrain"]["digest"]
    r.boxes(O.sphere_row_boxes(O.gen_brushes(g["seed"], O.STREAM_BRUSH, g["brushes"], world, g["mix"]), world))
    e = r.export()
    assert (int(e[4]), int(e[5]), "%016x" % O.digest(e)) == (g["edited"]["nodes"], g["edited"]["leaves"],
                                                             g["edited"]["digest"])


def test_oracle_raycast_pins():
    """The oracle's ray walk against what the geometry fixes: a straight
    segment between two voxel centres crosses exactly |dx|+|dy|+|dz| voxel
    boundaries, so a ray aimed at the only solid voxel hits it after that many
    steps; an axis-aligned ray hits the first solid voxel of its column
    (dense scan); rays that leave the world, start outside it, or cannot move
    miss."""
    import numpy as np
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    d = O.Dag(4, True)
    world = 64
    tgt = (37, 21, 50)
    d.box(tgt, tuple(c + 1 for c in tgt), 77)
    rays = []
    for _ in range(200):
        o = rng.integers(0, world, 3)
        dv = np.array(tgt) - o
        if not dv.any():
            continue
        rays.append([*o, *dv, 10_000, 0])
    hits = d.raycast(np.array(rays))
    for r, h in zip(rays, hits):
        assert h[0] == 77 and tuple(h[2:5]) == tgt
        assert h[1] == abs(r[3]) + abs(r[4]) + abs(r[5])
    # max_steps one short of the target: miss
    short = np.array(rays)
    short[:, 6] = np.abs(short[:, 3:6]).sum(1) - 1
    assert (d.raycast(short)[:, 0] == 0).all()
    # axis-aligned rays in a terrain scene vs a dense column scan
    t = O.Dag(4, True)
    t.build_terrain(O.terrain_for_world(world, 9))
    grid = np.array([[[t.get(x, y, z) for x in range(world)] for y in range(world)] for z in range(world)])
    for _ in range(100):
        o = rng.integers(0, world, 3)
        axis = int(rng.integers(0, 3))
        sign = int(rng.choice([-1, 1]))
        dv = [0, 0, 0]
        dv[axis] = sign * int(rng.integers(1, 5))
        h = t.raycast(np.array([[*o, *dv, 10_000, 0]]))[0]
        cells = []
        p = list(o)
        while 0 <= p[axis] < world:
            cells.append(tuple(p))
            p[axis] += sign
        solid = [c for c in cells if grid[c[2], c[1], c[0]]]
        if solid:
            c = solid[0]
            assert h[0] == grid[c[2], c[1], c[0]] and tuple(h[2:5]) == c and h[1] == abs(c[axis] - o[axis])
        else:
            assert h[0] == 0
    out = t.raycast(np.array([[-1, 5, 5, 1, 0, 0, 100, 0], [3, 3, 3, 0, 0, 0, 100, 0]]))
    assert out[0, 0] == 0 and (out[1, 0] == 0) == (t.get(3, 3, 3) == 0)
