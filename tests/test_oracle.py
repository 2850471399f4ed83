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


# ---- pins of the scans' max forms, the u64 scans and the general sort order ----
def test_exclusive_max_closed_form():
    a = O.gen_u32(321, O.STREAM_SCAN, 100_000, 0, 0xFFFFFFFF)
    want = np.concatenate([[0], np.maximum.accumulate(a)[:-1]]).astype(np.uint32)
    assert np.array_equal(O.exclusive_max(a), want)
    assert O.exclusive_max([3, 1, 4, 1, 5, 9, 2, 6, 5]).tolist() == [0, 3, 3, 4, 4, 5, 9, 9, 9]


def test_u64_scans_closed_forms():
    lo = O.gen_u32(11, O.STREAM_SCAN, 50_000, 0, 0xFFFFFFFF).astype(np.uint64)
    hi = O.gen_u32(12, O.STREAM_SCAN, 50_000, 0, 0xFFFFFFFF).astype(np.uint64)
    a = (hi << np.uint64(32)) | lo
    inc = np.cumsum(a, dtype=np.uint64)  # numpy integer sums wrap modulo 2^64
    assert np.array_equal(O.inclusive_sum_u64(a), inc)
    assert np.array_equal(O.exclusive_sum_u64(a), np.concatenate([np.zeros(1, np.uint64), inc[:-1]]))
    mx = np.maximum.accumulate(a)
    assert np.array_equal(O.inclusive_max_u64(a), mx)
    assert np.array_equal(O.exclusive_max_u64(a), np.concatenate([np.zeros(1, np.uint64), mx[:-1]]))
    # test_cuda_scan.cu:716-813: odd numbers sum to squares; (37 i) mod 1021 over 2049 elements.
    odd = np.arange(1, 34, 2, dtype=np.uint64)
    assert O.inclusive_sum_u64(odd).tolist() == [(k + 1) ** 2 for k in range(17)]
    b = (np.arange(2049, dtype=np.uint64) * 37) % 1021
    assert O.exclusive_max_u64(b)[0] == 0 and O.inclusive_max_u64(b)[-1] == 1020


def test_sort_order_matches_lexsort():
    n, words = 20_000, 3
    k = O.gen_u32(5, O.STREAM_KEYS, n * words, 0, 15).reshape(n, words)  # many ties
    for bits in (96, 70, 33, 4):
        masked = [k[:, w] & np.uint32((1 << max(0, min(32, bits - 32 * w))) - 1) for w in range(words)]
        want = np.lexsort([np.arange(n)] + masked)  # last key primary, index breaks ties
        assert np.array_equal(O.sort_order(k, bits), want), bits


def test_sort_order_reference_fixed_word_keys():
    keys = np.array([[0, 0, 1], [3, 0, 0], [0, 2, 0], [0xFFFFFFFF] * 3, [0, 0, 0], [1, 0, 1]], np.uint32)
    assert O.sort_order(keys, 96).tolist() == [4, 1, 2, 0, 5, 3]


# ---- the workload generators against the workload definitions -----------------
def _lattice(seed, i, j):
    """Lattice value of octave 0 at lattice point (i, j), in 1/65536 units."""
    r = O.lib().vxo_rand(seed, 16, (i << 32) | j)
    return (r >> 47) - 65536


def test_terrain_at_lattice_points_and_midpoints():
    """One octave: at lattice points the height is base + floor(amp * v) with
    v the lattice value in [-1, 1); halfway between two lattice points along x
    the smoothstep weight is exactly 1/2."""
    seed, base, amp, period = 9, 512, 128, 64
    t = O.Terrain(seed, base, amp, period, 1)
    ii, jj = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    h = O.terrain_heights(t, (ii * period).ravel(), (jj * period).ravel())
    want = [base + ((amp * _lattice(seed, int(i), int(j))) >> 16) for i, j in zip(ii.ravel(), jj.ravel())]
    assert h.tolist() == want
    mid = O.terrain_heights(t, (ii * period + period // 2).ravel(), (jj * period).ravel())
    want = []
    for i, j in zip(ii.ravel(), jj.ravel()):
        a, b = _lattice(seed, int(i), int(j)), _lattice(seed, int(i) + 1, int(j))
        want.append(base + ((amp * (a + ((b - a) >> 1))) >> 16))
    assert mid.tolist() == want


def test_terrain_monotone_between_lattice_points():
    """One octave, z on the lattice: the smoothstep is monotone, so heights
    between two lattice points stay between (and move towards) their values."""
    seed, period = 4, 64
    t = O.Terrain(seed, 256, 64, period, 1)
    for i in range(6):
        xs = np.arange(i * period, (i + 1) * period + 1)
        h = O.terrain_heights(t, xs, np.full(xs.size, 3 * period))
        d = np.diff(h.astype(np.int64))
        assert (d >= 0).all() or (d <= 0).all()


def test_terrain_config0_bounds_and_scale():
    """Config-0 heightfield (base 256, amplitude 64, periods 64 / 2^o, weights
    0.5^o, o < 4): every height within base +- 1.875 * amp (+1 for the floor),
    the range well used, and amplitude 0 gives the flat base."""
    rng = np.random.default_rng(1)
    xs, zs = rng.integers(0, 512, 200_000), rng.integers(0, 512, 200_000)
    h = O.terrain_heights(O.terrain_for_world(512, 1), xs, zs)
    assert h.min() >= 256 - 120 - 1 and h.max() <= 256 + 120
    assert h.max() - h.min() > 64
    flat = O.terrain_heights(O.Terrain(1, 256, 0, 64, 4), xs[:1000], zs[:1000])
    assert (flat == 256).all()
    # x32 scale (config 2): the same field stretched, heights x32 (to the floor).
    big = O.terrain_for_world(16384, 3)
    assert (big.base, big.amp, big.period, big.octaves) == (8192, 2048, 2048, 4)


def _binom_ok(k, n, p, sigmas=5.0):
    return abs(k - n * p) <= sigmas * np.sqrt(n * p * (1 - p))


def test_brush_generator_matches_workload_definition():
    """mix 0 (config 0): spheres, r in [4, 16], 50/50 add/remove; mix 1
    (configs 2-4): 50% spheres r in [8, 64], 50% boxes with half extents in
    [4, 48], 60/40 add/remove; added material in [1, 255]; centres uniform in
    the world — ranges hit at both ends, splits within 5 sigma."""
    n, world = 100_000, 16384
    b0 = O.gen_brushes(1, O.STREAM_BRUSH, n, world, 0)
    assert (b0["kind"] == 0).all()
    assert b0["hx"].min() == 4 and b0["hx"].max() == 16
    assert _binom_ok(int((b0["material"] > 0).sum()), n, 0.5)
    b1 = O.gen_brushes(3, O.STREAM_BRUSH, n, world, 1)
    sph, box = b1[b1["kind"] == 0], b1[b1["kind"] == 1]
    assert len(sph) + len(box) == n and _binom_ok(len(sph), n, 0.5)
    assert sph["hx"].min() == 8 and sph["hx"].max() == 64
    for ax in ("hx", "hy", "hz"):
        assert box[ax].min() == 4 and box[ax].max() == 48
    assert _binom_ok(int((b1["material"] > 0).sum()), n, 0.6)
    for b in (b0, b1):
        add = b[b["material"] > 0]["material"]
        assert add.min() == 1 and add.max() == 255
        for ax in ("cx", "cy", "cz"):
            assert b[ax].min() >= 0 and b[ax].max() <= world - 1
            counts = np.bincount(b[ax] * 8 // world, minlength=8)
            assert all(_binom_ok(int(c), n, 1 / 8) for c in counts)
    # Radii uniform: each value of [8, 64] about equally often.
    rc = np.bincount(sph["hx"], minlength=65)[8:65]
    assert all(_binom_ok(int(c), len(sph), 1 / 57) for c in rc)
