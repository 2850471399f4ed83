"""The reference's pointer-SVO known-answer tests (test/cuda/svt/*.cuh),
written once against a small interface so that they pin the CPU oracle and
check the GPU library alike. `make(max_nodes, max_leaves)` returns a reset SVO
with place(xyz), destroy(xyz), get_many(xyz), counters() -> (node_count,
leaf_count, free_node_count, free_leaf_count), export(); `leaf_masks(xyz)`
returns (m, 3) {leaf_key, low, high} rows."""
import numpy as np

MAX_DEPTH = 10
WORLD = 4096
NODE_BIT = 1 << 30
FILLED_BIT = 1 << 31


def leaf_key(lx, ly, lz):
    key = 0
    for d in range(MAX_DEPTH):
        b = MAX_DEPTH - 1 - d
        key = (key << 3) | ((lx >> b) & 1) | (((ly >> b) & 1) << 1) | (((lz >> b) & 1) << 2)
    return key


def expected_leaf_key(x, y, z):
    return leaf_key(x >> 2, y >> 2, z >> 2)


def live(c):
    return c[0] - c[2], c[1] - c[3]


def box_edits(size, omitted):
    return [(x, y, z) for z in range(size) for y in range(size) for x in range(size) if (x, y, z) != omitted]


def box_edits_array(size):
    """Every voxel of [0, size)^3, z slowest."""
    z, y, x = np.mgrid[0:size, 0:size, 0:size]
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.uint32)


def parse_export(e):
    """(nodes (n, 8), leaves (m, 2)) of a canonical export."""
    assert e[0] == 0x5658534F
    n, m = int(e[1]), int(e[2])
    nodes = np.asarray(e[3:3 + 8 * n], np.uint32).reshape(n, 8)
    leaves = np.asarray(e[3 + 8 * n:3 + 8 * n + 2 * m], np.uint32).reshape(m, 2)
    return nodes, leaves


def path_slots(nodes, x, y, z):
    """Slot words along the path of voxel (x, y, z) in an export (node ids)."""
    out, node = [], 0
    for d in range(MAX_DEPTH):
        sh = MAX_DEPTH - 1 - d + 2
        c = ((x >> sh) & 1) | (((y >> sh) & 1) << 1) | (((z >> sh) & 1) << 2)
        s = int(nodes[node][c])
        out.append(s)
        if not s & NODE_BIT:
            break
        node = s & 0x3FFFFFFF
    return out


# ---- leaf_masks_tests.cuh ------------------------------------------------------
def check_leaf_masks(leaf_masks):
    m = leaf_masks([(0, 0, 0), (1, 0, 0), (0, 0, 2), (4, 0, 0), (1023, 1023, 1023)])
    assert m.tolist() == [[expected_leaf_key(0, 0, 0), 0b11, 1], [expected_leaf_key(4, 0, 0), 1, 0],
                          [expected_leaf_key(1023, 1023, 1023), 0, 1 << 31]]
    m = leaf_masks([(0, 0, 0), (WORLD, 0, 0)])
    assert m.tolist() == [[expected_leaf_key(0, 0, 0), 1, 0]]
    m = leaf_masks([(8, 0, 0), (0, 0, 4), (0, 4, 0), (0, 0, 0)])
    assert expected_leaf_key(0, 0, 0) < expected_leaf_key(0, 4, 0) < expected_leaf_key(0, 0, 4) < \
        expected_leaf_key(8, 0, 0)
    assert m[:, 0].tolist() == [expected_leaf_key(0, 0, 0), expected_leaf_key(0, 4, 0), expected_leaf_key(0, 0, 4),
                                expected_leaf_key(8, 0, 0)]
    assert len(leaf_masks([])) == 0


# ---- structure_tests.cuh -------------------------------------------------------
def check_structure(make):
    s = make(256, 64)
    s.place([(0, 0, 0), (1, 0, 0), (0, 0, 2)])
    assert live(s.counters()) == (MAX_DEPTH, 1)
    nodes, leaves = parse_export(s.export())
    slots = path_slots(nodes, 0, 0, 0)
    assert len(slots) == MAX_DEPTH and all(v & NODE_BIT and not v & FILLED_BIT for v in slots)
    assert leaves[slots[-1] & 0x3FFFFFFF].tolist() == [0b11, 1]

    s = make(256, 64)
    s.place([(0, 0, 0), (2048, 0, 0)])
    assert live(s.counters()) == (1 + 2 * (MAX_DEPTH - 1), 2)
    nodes, leaves = parse_export(s.export())
    root = nodes[0]
    assert [bool(v & NODE_BIT) for v in root] == [True, True] + [False] * 6
    a, b = path_slots(nodes, 0, 0, 0), path_slots(nodes, 2048, 0, 0)
    assert len(a) == len(b) == MAX_DEPTH
    assert a[-1] != b[-1]
    assert leaves[a[-1] & 0x3FFFFFFF].tolist() == [1, 0] and leaves[b[-1] & 0x3FFFFFFF].tolist() == [1, 0]


# ---- place_tests.cuh -----------------------------------------------------------
def check_place_queries(make):
    s = make(256, 64)
    s.place([(0, 0, 0), (1, 0, 0), (0, 0, 2), (4, 0, 0), (512, 512, 512), (1023, 1023, 1023)])
    q = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 0, 2), (0, 0, 3), (4, 0, 0), (512, 512, 512), (1023, 1023, 1023),
         (1022, 1023, 1023), (WORLD, 0, 0)]
    assert s.get_many(q).tolist() == [1, 1, 0, 1, 0, 1, 1, 1, 0, 0]


# ---- edit_invariants_tests.cuh -------------------------------------------------
def check_invariants(make):
    s = make(256, 64)
    s.place([(0, 0, 0)])
    c1 = s.counters()
    s.place([(0, 0, 0)] * 3)
    assert s.counters() == c1
    assert s.get_many([(0, 0, 0), (1, 0, 0)]).tolist() == [1, 0]

    s = make(256, 64)
    s.place([(0, 0, 0), (1, 0, 0), (0, 0, 2)])
    s.destroy([(1, 0, 0)])
    assert live(s.counters())[1] == 1
    assert s.get_many([(0, 0, 0), (1, 0, 0), (0, 0, 2)]).tolist() == [1, 0, 1]
    assert parse_export(s.export())[1].tolist() == [[1, 1]]

    s = make(256, 64)
    s.place([(3, 3, 1), (0, 0, 2)])
    assert s.get_many([(3, 3, 1), (0, 0, 2), (2, 3, 1)]).tolist() == [1, 1, 0]
    assert parse_export(s.export())[1].tolist() == [[1 << 31, 1]]
    s.destroy([(3, 3, 1)])
    assert s.get_many([(3, 3, 1), (0, 0, 2)]).tolist() == [0, 1]
    assert parse_export(s.export())[1].tolist() == [[0, 1]]

    s = make(256, 64)
    c0 = s.counters()
    s.place([(WORLD, 0, 0), (0, WORLD, 0), (0, 0, WORLD)])
    assert s.counters() == c0
    s.place([(0, 0, 0), (WORLD, 0, 0)])
    assert live(s.counters()) == (MAX_DEPTH, 1)
    assert s.get_many([(0, 0, 0), (WORLD, 0, 0)]).tolist() == [1, 0]


# ---- collapse_tests.cuh --------------------------------------------------------
def check_collapse(make):
    for size, nodes_after in ((4, MAX_DEPTH), (8, MAX_DEPTH - 1), (16, MAX_DEPTH - 2)):
        s = make(256, 64)
        final = (size - 1,) * 3
        s.place(box_edits(size, final))
        assert live(s.counters())[1] == 1
        if size > 4:
            assert live(s.counters())[0] == MAX_DEPTH
        assert s.get_many([(0, 0, 0), final]).tolist() == [1, 0]
        s.place([final])
        assert live(s.counters()) == (nodes_after, 0)
        assert s.get_many([(0, 0, 0), final, (size, 0, 0)]).tolist() == [1, 1, 0]


# ---- destroy_tests.cuh ---------------------------------------------------------
def check_destroy(make):
    s = make(256, 64)
    s.place(box_edits(4, None))
    s.destroy([(0, 0, 0)])
    assert live(s.counters())[1] == 1
    assert s.get_many([(0, 0, 0), (1, 0, 0), (3, 3, 3)]).tolist() == [0, 1, 1]

    s = make(256, 64)
    s.place([(0, 0, 0), (1, 0, 0), (0, 0, 2), (4, 0, 0), (512, 512, 512), (1023, 1023, 1023)])
    s.destroy([(512, 512, 512), (1023, 1023, 1023)])
    c = s.counters()
    assert c[3] > 0 and c[2] > 0
    assert s.get_many([(0, 0, 0), (512, 512, 512), (1023, 1023, 1023)]).tolist() == [1, 0, 0]
    s.place([(1023, 1023, 1023)])
    c2 = s.counters()
    assert c2[3] < c[3] and c2[2] < c[2]
    assert s.get_many([(1023, 1023, 1023)]).tolist() == [1]


ALL_TREE_CHECKS = (check_structure, check_place_queries, check_invariants, check_collapse, check_destroy)


def assert_canonical(e, counters):
    """Every stored node except the root has a pointer or mixed fill, no leaf
    is all zero or all one, and the live counters equal the reachable ones."""
    nodes, leaves = parse_export(e)
    assert live(counters) == (len(nodes), len(leaves))
    for i in range(1, len(nodes)):
        row = nodes[i]
        ptr = (row & NODE_BIT) != 0
        if not ptr.any():
            f = (row & FILLED_BIT) != 0
            assert f.any() and not f.all(), i
    for lo, hi in leaves.tolist():
        assert (lo, hi) not in ((0, 0), (0xFFFFFFFF, 0xFFFFFFFF))


def sphere_edits(cx, cy, cz, r):
    """make_voxel_sphere_edits (bench/bench_cuda_svt.cu:56-80): z, y, x order."""
    rng = np.arange(-r, r + 1)
    dz, dy, dx = np.meshgrid(rng, rng, rng, indexing="ij")
    inside = dx * dx + dy * dy + dz * dz <= r * r
    return np.stack([(cx + dx)[inside], (cy + dy)[inside], (cz + dz)[inside]], 1).astype(np.uint32)


def random_edits(count):
    """make_random_edits (bench/bench_cuda_svt.cu:90-112): xorshift32 from
    0x9e3779b9, three draws per voxel, each mod 4096 (the oracle's C loop)."""
    from oracle import oracle as O
    return O.svo_random_edits(count)
