"""GPU Hash DAG edits vs the CPU oracle (canonical export, bit-exact) and the
reference's own known-answer tests (test/cuda/test_cuda_hash_dag_gpu.cu)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dag(depth, materials, **kw):
    from voxel_amd.dag import HashDagGpu
    return HashDagGpu(depth, materials, **kw)


def _dev_xyz(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).reshape(-1, 3).view(np.int32)).to(dev)


def spread_edits(count):
    """make_spread_edits (test_cuda_hash_dag_gpu.cu:142-156): every 19th edit is out of range."""
    e = []
    for i in range(count):
        if i % 19 == 0:
            e.append((1024 + i, 0, 0))
        else:
            e.append(((i * 37) & 1023, (i * 91 + 3) & 1023, (i * 53 + 7) & 1023))
    return np.array(e, np.uint32)


def test_single_voxel_place_destroy(gpu):
    d = _dag(8, False)
    d.place_voxel_edits(_dev_xyz([[0, 0, 0], [1, 1, 1], [1023, 1023, 1023]], gpu))
    q = _dev_xyz([[0, 0, 0], [1, 1, 1], [1, 0, 0], [2, 0, 0], [1023, 1023, 1023], [1022, 1023, 1023], [1024, 0, 0]], gpu)
    assert d.get_voxels(q).cpu().tolist() == [1, 1, 0, 0, 1, 0, 0]
    d.destroy_voxel_edits(_dev_xyz([[0, 0, 0]], gpu))
    assert d.get_voxels(q).cpu().tolist() == [0, 1, 0, 0, 1, 0, 0]


def test_eight_partial_siblings_make_a_nine_word_parent(gpu):
    d = _dag(8, False)
    pts = [(x * 4, y * 4, z * 4) for z in range(2) for y in range(2) for x in range(2)]
    d.place_voxel_edits(_dev_xyz(pts, gpu))
    e = d.export()
    ref = O.Dag(8, False)
    ref.apply_voxels(np.array(pts, np.uint32), 1)
    assert np.array_equal(e, ref.export())
    # deepest node (level 7) has 8 table-backed children: header + 8 refs
    nodes = e[4]
    k, widths = 6, []
    for _ in range(nodes):
        lvl, hdr = e[k], e[k + 1]
        cm = hdr & 0xFF
        widths.append((lvl, 1 + bin(cm).count("1")))
        k += 2 + bin(cm).count("1")
    assert (7, 9) in widths


@pytest.mark.parametrize("count", [31, 32, 33, 255, 256, 257, 4099])
def test_boundary_batch_sizes_match_truth(gpu, count):
    d = _dag(8, False)
    edits = spread_edits(count)
    d.place_voxel_edits(_dev_xyz(edits, gpu))
    valid = edits[(edits < 1024).all(axis=1)]
    truth = {tuple(v) for v in valid.tolist()}
    extra = np.array([((i * 13 + 5) & 1023, (i * 17 + 11) & 1023, (i * 29 + 23) & 1023) for i in range(32)] +
                     [(1024, 0, 0)], np.uint32)
    q = np.concatenate([edits, extra])
    got = d.get_voxels(_dev_xyz(q, gpu)).cpu().numpy()
    exp = np.array([1 if tuple(v) in truth else 0 for v in q.tolist()])
    assert np.array_equal(got, exp)
    ref = O.Dag(8, False)
    ref.apply_voxels(edits, 1)
    assert np.array_equal(d.export(), ref.export())


def test_invalid_arguments_keep_existing_dag(gpu):
    from voxel_amd import InvalidValue
    from voxel_amd._lib import lib
    import ctypes as C
    d = _dag(8, False)
    d.place_voxel_edits(_dev_xyz([[0, 0, 0]], gpu))
    e = _dev_xyz([[2, 0, 0]], gpu)
    need = d.apply_voxel_edits_workspace_size(1)
    ws = torch.empty(need, dtype=torch.uint8, device=gpu)
    assert lib().vx_dag_apply_voxel_edits(d.h, None, 1, 0, 1, C.c_void_p(ws.data_ptr()), need, None) == 1
    assert lib().vx_dag_apply_voxel_edits(d.h, C.c_void_p(e.data_ptr()), 1, 0, 1, None, need, None) == 1
    assert lib().vx_dag_apply_voxel_edits(d.h, C.c_void_p(e.data_ptr()), 1, 0, 1, C.c_void_p(ws.data_ptr()),
                                          need - 1, None) == 1
    with pytest.raises(InvalidValue):
        d.apply_voxel_edits(e, 7)
    assert d.get_voxels(_dev_xyz([[0, 0, 0], [2, 0, 0]], gpu)).cpu().tolist() == [1, 0]


def test_overflow_does_not_publish_partial_root(gpu):
    """HashDagGpu<1,0> analogue: tiny pools must fail with the old root kept."""
    from voxel_amd import OutOfCapacity
    d = _dag(8, False, node_words=64, leaf_words=64, node_slots=64, leaf_slots=64)
    with pytest.raises(OutOfCapacity):
        d.place_voxel_edits(_dev_xyz(spread_edits(512), gpu))
    q = _dev_xyz([[37, 94, 60], [74, 185, 113], [111, 276, 166], [148, 367, 219]], gpu)
    assert d.get_voxels(q).cpu().tolist() == [0, 0, 0, 0]
    assert d.export()[3] == 0


@pytest.mark.parametrize("materials", [False, True])
@pytest.mark.parametrize("depth,n,mix,seed", [(4, 30, 0, 1), (6, 64, 0, 2), (7, 64, 0, 1), (8, 200, 1, 5)])
def test_brushes_on_empty_scene_match_oracle(gpu, materials, depth, n, mix, seed):
    d = _dag(depth, materials)
    br = O.gen_brushes(seed, O.STREAM_BRUSH, n, 4 << depth, mix)
    d.apply_brushes(br)
    ref = O.Dag(depth, materials)
    ref.apply_brushes(br)
    assert np.array_equal(d.export(), ref.export())


@pytest.mark.parametrize("materials", [False, True])
def test_sphere_radii_across_arithmetic_paths_match_oracle(gpu, materials):
    """Sphere masks switch from packed 16-bit to 32-bit arithmetic above radius
    100: radii 95..106 and 126 at unaligned centres (some clipped by the world
    boundary), mixed add / remove, on a 512^3 world."""
    depth = 7
    rng = np.random.default_rng(7)
    radii = list(range(95, 107)) + [126, 1, 2, 3]
    br = np.zeros(len(radii), O.BRUSH_DTYPE)
    for i, r in enumerate(radii):
        br[i]["cx"], br[i]["cy"], br[i]["cz"] = rng.integers(-20, 532, 3)
        br[i]["hx"] = r
        br[i]["kind"] = 0
        br[i]["material"] = 0 if i % 4 == 3 else int(rng.integers(1, 256))
    d = _dag(depth, materials)
    d.apply_brushes(br)
    ref = O.Dag(depth, materials)
    ref.apply_brushes(br)
    assert np.array_equal(d.export(), ref.export())


@pytest.mark.parametrize("materials", [False, True])
def test_huge_sphere_caps_match_oracle(gpu, materials):
    """Spheres of radius 2^14 .. 2^20 whose centres lie far outside a 512^3
    world, so only a curved cap crosses it: the brick rows' touched runs come
    from 64-bit squares and an integer square root near 2^40, and the masks
    from the 64-bit sphere test."""
    depth = 7
    br = np.zeros(4, O.BRUSH_DTYPE)
    caps = [(256, 250, -16300, (1 << 14) + 37, 9), (-20000, 300, 200, 20090, 0),
            (700, 100, 256, (1 << 20) - 3, 31), (200, (1 << 20) + 500, 300, (1 << 20) + 433, 0)]
    for i, (cx, cy, cz, r, m) in enumerate(caps):
        br[i]["cx"], br[i]["cy"], br[i]["cz"], br[i]["hx"], br[i]["kind"], br[i]["material"] = cx, cy, cz, r, 0, m or 200
    br[1]["material"] = 0  # a removal
    br[3]["material"] = 0
    from voxel_amd.dag import terrain_for_world
    d = _dag(depth, materials, node_words=1 << 24, leaf_words=1 << 25)
    d.build_terrain(terrain_for_world(4 << depth, 2))
    d.apply_brushes(br)
    ref = O.Dag(depth, materials)
    ref.build_terrain(O.terrain_for_world(4 << depth, 2))
    ref.apply_brushes(br)
    assert np.array_equal(d.export(), ref.export())


@pytest.mark.parametrize("materials", [False, True])
def test_successive_batches_match_oracle(gpu, materials):
    depth = 7
    d = _dag(depth, materials)
    ref = O.Dag(depth, materials)
    for seed in range(4):
        br = O.gen_brushes(10 + seed, O.STREAM_BRUSH, 40, 4 << depth, 0)
        d.apply_brushes(br)
        ref.apply_brushes(br)
        assert np.array_equal(d.export(), ref.export()), seed
    vox = O.gen_u32(3, O.STREAM_KEYS, 3000 * 3, 0, (4 << depth) + 20)  # some out of range
    d.apply_voxel_edits(_dev_xyz(vox, gpu), 0, material=77)
    ref.apply_voxels(vox, 77)
    assert np.array_equal(d.export(), ref.export())
    d.apply_voxel_edits(_dev_xyz(vox[: 1500 * 3], gpu), 1)
    ref.apply_voxels(vox[: 1500 * 3], 0)
    assert np.array_equal(d.export(), ref.export())


def test_get_voxels_matches_oracle(gpu):
    depth = 6
    d = _dag(depth, True)
    ref = O.Dag(depth, True)
    br = O.gen_brushes(4, O.STREAM_BRUSH, 60, 4 << depth, 0)
    d.apply_brushes(br)
    ref.apply_brushes(br)
    q = O.gen_u32(9, O.STREAM_KEYS, 5000 * 3, 0, (4 << depth) - 1).reshape(-1, 3)
    got = d.get_voxels(_dev_xyz(q, gpu)).cpu().numpy()
    exp = np.array([ref.get(*map(int, v)) for v in q])
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("materials", [False, True])
@pytest.mark.parametrize("depth,seed", [(3, 1), (5, 2), (7, 1), (8, 10)])
def test_terrain_matches_oracle(gpu, materials, depth, seed):
    from voxel_amd.dag import terrain_for_world
    world = 4 << depth
    d = _dag(depth, materials, node_words=1 << 24, leaf_words=1 << 25, node_slots=1 << 21, leaf_slots=1 << 21)
    d.build_terrain(terrain_for_world(world, seed))
    ref = O.Dag(depth, materials)
    ref.build_terrain(O.terrain_for_world(world, seed))
    assert np.array_equal(d.export(), ref.export())


@pytest.mark.parametrize("materials", [False, True])
def test_config0_terrain_plus_spheres_matches_oracle(gpu, materials):
    """512^3 heightfield (seed 1) + one batch of 64 spheres r[4,16] 50/50."""
    from voxel_amd.dag import terrain_for_world
    depth, world = 7, 512
    d = _dag(depth, materials, node_words=1 << 24, leaf_words=1 << 25, node_slots=1 << 21, leaf_slots=1 << 21)
    d.build_terrain(terrain_for_world(world, 1))
    br = O.gen_brushes(1, O.STREAM_BRUSH, 64, world, 0)
    d.apply_brushes(br)
    ref = O.Dag(depth, materials)
    ref.build_terrain(O.terrain_for_world(world, 1))
    ref.apply_brushes(br)
    assert np.array_equal(d.export(), ref.export())


@pytest.mark.parametrize("materials", [False, True])
def test_batched_scenes_equal_sequential_and_oracle(gpu, materials):
    """Batches over scenes of different depths in one pipeline == each batch
    applied in order (and == the CPU oracle)."""
    from voxel_amd.dag import apply_batches, terrain_for_world
    depths = [5, 7, 6]
    dags = [_dag(dd, materials, node_words=1 << 22, leaf_words=1 << 23) for dd in depths]
    refs = []
    for k, (dd, d) in enumerate(zip(depths, dags)):
        d.build_terrain(terrain_for_world(4 << dd, 10 + k))
        r = O.Dag(dd, materials)
        r.build_terrain(O.terrain_for_world(4 << dd, 10 + k))
        refs.append(r)
    order = [0, 1, 0, 2, 1, 1, 0, 2]
    batches, targets = [], []
    for i, s in enumerate(order):
        br = O.gen_brushes(100 + i, O.STREAM_BRUSH, 5 + 3 * i, 4 << depths[s], 0)
        batches.append(br)
        targets.append(dags[s])
        refs[s].apply_brushes(br)
    apply_batches(targets, batches)
    for s in range(3):
        assert np.array_equal(dags[s].export(), refs[s].export()), s


def test_batch_of_one_scene_matches_plain(gpu):
    from voxel_amd.dag import apply_batches
    a, b = _dag(6, True), _dag(6, True)
    bs = [O.gen_brushes(7 + i, O.STREAM_BRUSH, 20, 256, 0) for i in range(3)]
    apply_batches([a, a, a], bs)
    for br in bs:
        b.apply_brushes(br)
    assert np.array_equal(a.export(), b.export())


@pytest.mark.parametrize("use_graph", [False, True])
def test_batcher_replays_match_sequential(gpu, use_graph):
    """The persistent executor (graph captured once, replayed for new batch
    sets) gives the same scenes as plain per-batch calls."""
    from voxel_amd.dag import Batcher, batch_candidates, terrain_for_world
    depths = [4, 6, 5]
    a = [_dag(d, True, node_words=1 << 22, leaf_words=1 << 23) for d in depths]
    b = [_dag(d, True, node_words=1 << 22, leaf_words=1 << 23) for d in depths]
    for k, d in enumerate(depths):
        a[k].build_terrain(terrain_for_world(4 << d, 30 + k))
        b[k].build_terrain(terrain_for_world(4 << d, 30 + k))
    rounds = []
    for r in range(3):
        rounds.append([(s, O.gen_brushes(200 + 10 * r + i, O.STREAM_BRUSH, 4 + i, 4 << depths[s], 0))
                       for i, s in enumerate([0, 1, 2, 1, 0])])
    cap = max(batch_candidates(a, rb) for rb in rounds)
    ex = Batcher(a, cap, 64, use_graph=use_graph)
    for rb in rounds:
        ex.run(rb)
        for s, br in rb:
            b[s].apply_brushes(br)
        for k in range(3):
            assert np.array_equal(a[k].export(), b[k].export()), k
    assert ex.info()[1] == use_graph


def _brush_truth(brushes, xyz, base):
    """Material at each sample after the brushes are applied in order (last
    writer wins; material 0 erases), evaluated straight from the shapes:
    spheres |p - c|^2 <= r^2, boxes |p - c| <= h per axis."""
    out = base.copy()
    order = np.argsort(xyz[:, 0], kind="stable")
    xs = xyz[order, 0]
    p = xyz.astype(np.int64)
    for b in brushes:
        kind, hx = int(b["kind"]), int(b["hx"])
        lo = np.searchsorted(xs, int(b["cx"]) - hx, "left")
        hi = np.searchsorted(xs, int(b["cx"]) + hx, "right")
        idx = order[lo:hi]
        d = p[idx] - np.array([b["cx"], b["cy"], b["cz"]], np.int64)
        if kind == 0:
            inside = (d * d).sum(1) <= hx * hx
        else:
            inside = (np.abs(d) <= np.array([b["hx"], b["hy"], b["hz"]], np.int64)).all(1)
        out[idx[inside]] = int(b["material"])
    return out


def test_full_size_edit_sampled_truth(gpu):
    """16384^3 materials heightfield (inner depth 12, seed 3) + one batch of
    4096 brushes (50% spheres r[8,64], 50% boxes h[4,48], 60/40 add/remove):
    2^18 sampled voxels — half uniform over the world, half inside brush
    bounding boxes — hold exactly the heightfield material overwritten by the
    brushes in order."""
    from voxel_amd.dag import STREAM_BRUSH, gen_brushes, terrain_for_world
    depth = 12
    world = 4 << depth
    g = _dag(depth, True, node_words=1 << 28, leaf_words=1 << 29, node_slots=1 << 26, leaf_slots=1 << 26)
    t = terrain_for_world(world, 3)
    g.build_terrain(t)
    brushes = gen_brushes(3, STREAM_BRUSH, 4096, world, 1)
    g.apply_brushes(brushes)
    rng = np.random.default_rng(12)
    n = 1 << 17
    uni = rng.integers(0, world, (n, 3))
    pick = brushes[rng.integers(0, len(brushes), n)]
    ext = np.where(pick["kind"] == 0, pick["hx"], 0)[:, None] + np.where(
        pick["kind"] == 0, 0, 1)[:, None] * np.stack([pick["hx"], pick["hy"], pick["hz"]], 1)
    ctr = np.stack([pick["cx"], pick["cy"], pick["cz"]], 1).astype(np.int64)
    near = np.clip(ctr + rng.integers(-1 << 20, 1 << 20, (n, 3)) % (2 * ext + 1) - ext, 0, world - 1)
    xyz = np.concatenate([uni, near]).astype(np.int64)
    ot = O.terrain_for_world(world, 3)
    heights = {}
    base = np.zeros(len(xyz), np.int64)
    for i, (x, y, z) in enumerate(xyz):
        h = heights.get((x, z))
        if h is None:
            h = heights[(x, z)] = O.terrain_height(ot, int(x), int(z))
        if y <= h:
            base[i] = 1 + y % 255
    want = _brush_truth(brushes, xyz, base)
    got = g.get_voxels(_dev_xyz(xyz, gpu)).cpu().numpy()
    assert np.array_equal(got.astype(np.int64), want)


def test_kernel_profile_counts_every_launch(gpu):
    """Per-kernel timing while profiled: one partition launch per radix pass,
    one records launch per edit, and the kernel times add up to (at most) the
    stage times they fall in."""
    from voxel_amd.dag import KERNELS, STREAM_BRUSH, gen_brushes, terrain_for_world
    d = _dag(7, True, node_words=1 << 24, leaf_words=1 << 25)
    d.build_terrain(terrain_for_world(512, 1))
    d.profile(True)
    for k in range(3):
        d.apply_brushes(gen_brushes(1, STREAM_BRUSH + (k << 8), 64, 512, 0))
    p = d.profile_read()
    launches = dict(zip(KERNELS, list(p.kernel_launches)))
    ms = dict(zip(KERNELS, list(p.kernel_ms)))
    assert p.edits == 3
    assert launches["records"] == 3
    assert launches["sort_partition"] == 3 * p.sort_passes
    assert launches["leaf_probe"] == 3
    # Node heights run as one launch per height below the tail, then one
    # single-workgroup launch for the rest.
    assert launches["node_probe"] % 3 == 0 and launches["node_probe"] <= 3 * 7
    assert launches["tail_nodes"] == (3 if launches["node_probe"] < 3 * 7 else 0)
    assert launches["tail_up"] == launches["tail_nodes"]
    # Old refs: one root-walk launch for heights >= 2, one parent step for height 1.
    assert launches["tail_topdown"] == 0 and launches["topdown"] == 3 * 2
    assert sum(ms.values()) <= sum(p.ms) * 1.05 + 0.05
    assert all(v > 0 for k, v in ms.items() if launches[k])


def test_kernel_sampling_without_stage_events(gpu):
    """Profiling without stage events records no stage times, and a kernel
    period of P times one launch per edit of a kind launched P times per
    edit, rotating over the launches (P edits time each launch once)."""
    from voxel_amd.dag import KERNELS, STREAM_BRUSH, gen_brushes, terrain_for_world
    d = _dag(7, True, node_words=1 << 24, leaf_words=1 << 25)
    d.build_terrain(terrain_for_world(512, 1))
    d.profile(True)
    d.apply_brushes(gen_brushes(1, STREAM_BRUSH, 64, 512, 0))
    passes = d.profile_read().sort_passes
    assert passes > 1
    d.profile(True, stages=False)
    d.profile_kernels(["sort_partition"])
    d.profile_period(passes)
    for k in range(passes):
        d.apply_brushes(gen_brushes(1, STREAM_BRUSH + ((k + 1) << 8), 64, 512, 0))
    p = d.profile_read()
    i = KERNELS.index("sort_partition")
    assert p.edits == passes
    assert list(p.ms) == [0.0] * len(p.ms)
    assert p.kernel_launches[i] == passes
    assert sum(p.kernel_launches) == passes and p.kernel_ms[i] > 0


def test_pair_sort_layout_matches_packed(gpu, monkeypatch):
    """The (key, index) pair sort, used when a packed key would not fit 64
    bits, gives the same scene as the packed-key sort."""
    from voxel_amd.dag import STREAM_BRUSH, gen_brushes, terrain_for_world
    br = [gen_brushes(41, STREAM_BRUSH + (k << 8), 96, 1024, 1) for k in range(2)]
    out = []
    for force in ("0", "1"):
        monkeypatch.setenv("VOXEL_AMD_SORT_PAIRS", force)
        d = _dag(8, True, node_words=1 << 23, leaf_words=1 << 24)
        d.build_terrain(terrain_for_world(1024, 5))
        for b in br:
            d.apply_brushes(b)
        out.append(d.export())
    assert np.array_equal(out[0], out[1])


def test_batcher_keeps_its_sort_tuning(gpu):
    """A sort tuning change after the executor is created does not change the
    layout of the workspace it carved: runs without a graph still sort into
    the space sized at creation and give the sequential result."""
    from voxel_amd.dag import Batcher, batch_candidates, terrain_for_world
    from voxel_amd.primitives import SORT_KEYS_ONLY, get_sort_tuning, set_sort_tuning
    depths = [5, 6]
    a = [_dag(d, True, node_words=1 << 22, leaf_words=1 << 23) for d in depths]
    b = [_dag(d, True, node_words=1 << 22, leaf_words=1 << 23) for d in depths]
    for k, d in enumerate(depths):
        a[k].build_terrain(terrain_for_world(4 << d, 50 + k))
        b[k].build_terrain(terrain_for_world(4 << d, 50 + k))
    rb = [(i % 2, O.gen_brushes(300 + i, O.STREAM_BRUSH, 30, 4 << depths[i % 2], 1)) for i in range(4)]
    ex = Batcher(a, batch_candidates(a, rb), 256, use_graph=False)
    rows = [8 | SORT_KEYS_ONLY, 8]
    saved = [get_sort_tuning(r) for r in rows]
    try:
        for r in rows:  # 8-bit digits in 2048-key tiles: twice the look-back rows of the default
            set_sort_tuning(r, 8, 8, 256)
        ex.run(rb)
    finally:
        for r, t in zip(rows, saved):
            set_sort_tuning(r, *t)
    for s, br in rb:
        b[s].apply_brushes(br)
    for k in range(2):
        assert np.array_equal(a[k].export(), b[k].export()), k


def _dense_cube(count):
    """make_dense_edits (bench/bench_cuda_hash_dag_gpu.cu:156-166)."""
    side = max(1, int(np.ceil(np.cbrt(count))))
    while side ** 3 < count:
        side += 1
    i = np.arange(count, dtype=np.uint64)
    return np.stack([i % side, (i // side) % side, i // (side * side)], 1).astype(np.uint32)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("workload", ["dense", "random"])
def test_reference_benchmark_workloads_match_oracle(gpu, workload):
    """The reference Hash DAG benchmark's inputs at 2^20 voxels on the 1024^3
    world (bench/bench_cuda_hash_dag_gpu.cu:156-190): place on an empty DAG
    (canonical export equal to the oracle's), get_voxel of the placed voxels
    (all set), destroy half (equal to the oracle again), destroy all (empty)."""
    from voxel_amd.svo import random_edits
    n = 1 << 20
    xyz = _dense_cube(n) if workload == "dense" else random_edits(n, 1024)
    d = _dag(8, False, node_words=1 << 24, leaf_words=1 << 24, node_slots=1 << 22, leaf_slots=1 << 22)
    dx = _dev_xyz(xyz, gpu)
    d.place_voxel_edits(dx)
    ref = O.Dag(8, False)
    ref.apply_voxels(xyz, 1)
    assert np.array_equal(d.export(), ref.export())
    assert int(d.get_voxels(dx).sum().item()) == n
    half = xyz[: n // 2]
    d.destroy_voxel_edits(_dev_xyz(half, gpu))
    ref.apply_voxels(half, 0)
    assert np.array_equal(d.export(), ref.export())
    d.destroy_voxel_edits(dx)
    assert d.export()[3] == 0


def test_corner_miss_spheres_leave_scene_unchanged(gpu):
    """Spheres whose clipped bounding boxes hold brick rows but no voxel of the
    sphere (centre (-100,-100,-100), r 150: the box reaches the world corner,
    the sphere does not) give no records; the scene is unchanged."""
    from voxel_amd.dag import terrain_for_world
    br = np.zeros(3, O.BRUSH_DTYPE)
    br[0] = (-100, -100, -100, 150, 0, 0, 0, 7)
    br[1] = (-100, -100, -100, 150, 0, 0, 0, 0)
    br[2] = (1100, 1100, -100, 130, 0, 0, 0, 9)  # nearest world voxel at distance^2 21858 > 130^2
    d = _dag(8, True, node_words=1 << 23, leaf_words=1 << 24)
    d.build_terrain(terrain_for_world(1024, 6))
    before = d.export()
    d.apply_brushes(br)
    assert np.array_equal(d.export(), before)
    ref = O.Dag(8, True)
    ref.build_terrain(O.terrain_for_world(1024, 6))
    ref.apply_brushes(br)
    assert np.array_equal(before, ref.export())


def test_invalid_brush_kind_or_material_is_rejected(gpu):
    from voxel_amd import InvalidValue
    d = _dag(5, True)
    for field, value in (("kind", 2), ("material", 256)):
        br = np.zeros(1, O.BRUSH_DTYPE)
        br[0] = (10, 10, 10, 3, 3, 3, 1, 5)
        br[0][field] = value
        with pytest.raises(InvalidValue):
            d.apply_brushes(br)
    assert d.export()[3] == 0
