"""Growth of a scene's pools and slot tables (vx_dag_grow): the voxel field,
the canonical export and every ref survive; the grown tables hold each live
item once, reachable from its home slot; later edits (direct, through
graph-captured executors that held the old buffers, and through the serving
loop, which grows a full scene instead of rejecting its requests) match the
CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pair(depth, materials, seed, **cap):
    """A small-capacity scene and the oracle, both on the same terrain."""
    from voxel_amd.dag import HashDagGpu, terrain_for_world
    d = HashDagGpu(depth, materials, **cap)
    d.build_terrain(terrain_for_world(4 << depth, seed))
    ref = O.Dag(depth, materials)
    ref.build_terrain(O.terrain_for_world(4 << depth, seed))
    return d, ref


@pytest.mark.parametrize("materials", [False, True])
def test_grow_keeps_scene_and_later_edits_match_oracle(gpu, materials):
    depth = 7
    world = 4 << depth
    d, ref = _pair(depth, materials, 5, node_words=1 << 20, leaf_words=1 << 21, node_slots=1 << 16,
                   leaf_slots=1 << 17)
    for i in range(2):
        br = O.gen_brushes(90 + i, O.STREAM_BRUSH, 40, world, 1)
        d.apply_brushes(br)
        ref.apply_brushes(br)
    before = d.export()
    assert np.array_equal(before, ref.export())
    q = torch.from_numpy(O.gen_u32(4, O.STREAM_KEYS, 30000 * 3, 0, world - 1).reshape(-1, 3).view(np.int32).copy()
                         ).to(gpu)  # (n, 3) voxel coordinates
    v0 = d.get_voxels(q).cpu()
    st0, u0 = d.stats(), d.check_unique()
    pools0 = d.read_pools()
    d.grow(node_words=1 << 22, leaf_words=1 << 23, node_slots=1 << 18, leaf_slots=1 << 19)
    st1, u1 = d.stats(), d.check_unique()
    assert (st1.node_slots, st1.leaf_slots) == (1 << 18, 1 << 19)
    assert (st1.node_words_used, st1.leaf_words_used, st1.root) == (st0.node_words_used, st0.leaf_words_used,
                                                                    st0.root)
    pools1 = d.read_pools()
    assert np.array_equal(pools0[0], pools1[0]) and np.array_equal(pools0[1], pools1[1])
    for t in ("nodes", "leaves"):  # every live slot moved, none twice, each reachable from its home slot
        assert u1[t]["live"] == u0[t]["live"] and u1[t]["duplicates"] == 0 and u1[t]["broken"] == 0, t
    assert np.array_equal(d.export(), before)
    assert torch.equal(d.get_voxels(q).cpu(), v0)
    # Later edits intern against the grown tables (lookups must find the old items).
    for i in range(3):
        br = O.gen_brushes(95 + i, O.STREAM_BRUSH, 60, world, i % 2)
        d.apply_brushes(br)
        ref.apply_brushes(br)
        assert np.array_equal(d.export(), ref.export()), i
    u2 = d.check_unique()
    assert u2["nodes"]["duplicates"] == 0 and u2["leaves"]["duplicates"] == 0


def test_grow_pools_only_slots_only_and_no_op(gpu):
    depth = 6
    world = 4 << depth
    d, ref = _pair(depth, True, 6, node_words=1 << 19, leaf_words=1 << 20, node_slots=1 << 15,
                   leaf_slots=1 << 16)
    br = O.gen_brushes(101, O.STREAM_BRUSH, 30, world, 1)
    d.apply_brushes(br)
    ref.apply_brushes(br)
    e = d.export()
    st = d.stats()
    d.grow(node_words=1 << 18, leaf_words=1 << 19)  # below the current sizes: nothing changes
    assert (d.stats().node_slots, d.stats().leaf_slots) == (st.node_slots, st.leaf_slots)
    assert (d._cap.node_words, d._cap.leaf_words) == (1 << 19, 1 << 20)
    d.grow(node_words=1 << 21)  # one pool only
    assert d._cap.node_words == 1 << 21 and d.stats().node_slots == st.node_slots
    d.grow(leaf_slots=(1 << 16) + 1)  # one table only, rounded up to a power of two
    assert d.stats().leaf_slots == 1 << 17 and d.stats().node_slots == st.node_slots
    assert np.array_equal(d.export(), e)
    br = O.gen_brushes(102, O.STREAM_BRUSH, 30, world, 0)
    d.apply_brushes(br)
    ref.apply_brushes(br)
    assert np.array_equal(d.export(), ref.export())


def test_grow_rejects_bad_arguments(gpu):
    import ctypes as C

    from voxel_amd import InvalidValue
    from voxel_amd._lib import check, lib
    from voxel_amd.dag import Capacity, HashDagGpu
    with pytest.raises(InvalidValue):
        check(lib().vx_dag_grow(None, C.byref(Capacity(1 << 20, 1 << 20, 1 << 16, 1 << 16)), None), "grow")
    d = HashDagGpu(5, True)
    with pytest.raises(InvalidValue):
        check(lib().vx_dag_grow(d.h, None, None), "grow")
    with pytest.raises(InvalidValue):
        d.grow(node_slots=(1 << 32) + 1)
    d.close()


@pytest.mark.parametrize("depths", [[6], [5, 6]])
def test_executors_take_grown_buffers(gpu, depths):
    """Executors captured their pipeline (and, with one scene, its
    descriptor) before the growth: their next run re-captures and edits the
    grown scene."""
    from voxel_amd.dag import Batcher, HashDagGpu, batch_candidates, terrain_for_world
    a = [HashDagGpu(x, True, node_words=1 << 20, leaf_words=1 << 21, node_slots=1 << 16, leaf_slots=1 << 17)
         for x in depths]
    refs = []
    for k, x in enumerate(depths):
        a[k].build_terrain(terrain_for_world(4 << x, 70 + k))
        r = O.Dag(x, True)
        r.build_terrain(O.terrain_for_world(4 << x, 70 + k))
        refs.append(r)
    rounds = [[(s, O.gen_brushes(400 + 10 * r + i, O.STREAM_BRUSH, 6, 4 << depths[s], 0))
               for i, s in enumerate([k % len(depths) for k in range(4)])] for r in range(4)]
    cap = max(batch_candidates(a, rb) for rb in rounds)
    ex = Batcher(a, cap, 64, use_graph=True)
    for r, rb in enumerate(rounds):
        ex.run(rb)
        for s, br in rb:
            refs[s].apply_brushes(br)
        for k in range(len(depths)):
            assert np.array_equal(a[k].export(), refs[k].export()), (r, k)
        g = a[r % len(depths)]
        g.grow(node_words=2 * g._cap.node_words, leaf_words=2 * g._cap.leaf_words,
               node_slots=2 * g.stats().node_slots, leaf_slots=2 * g.stats().leaf_slots)
    assert ex.info()[1]  # the last run replayed a captured graph
    ex.close()


def test_serving_grows_full_scenes_instead_of_rejecting(gpu):
    """Scenes far too small for their requests: with growth and without
    collection, every request completes and every scene equals the oracle's
    replay in arrival order; without growth or collection, requests that do
    not fit are rejected."""
    from voxel_amd.dag import HashDagGpu, terrain_for_world
    from voxel_amd.serving import Server, make_requests, poisson_arrivals, serve
    depths = [5, 6, 5]

    def pow2(v):
        p = 1
        while p < v:
            p <<= 1
        return p

    def scenes():
        # Each scene just holds its terrain: a few thousand words to spare.
        out = []
        for k, x in enumerate(depths):
            big = HashDagGpu(x, True, node_words=1 << 20, leaf_words=1 << 21, node_slots=1 << 16, leaf_slots=1 << 16)
            big.build_terrain(terrain_for_world(4 << x, 80 + k))
            st = big.stats()
            big.close()
            g = HashDagGpu(x, True, node_words=st.node_words_used + 2048, leaf_words=st.leaf_words_used + 4096,
                           node_slots=pow2(st.node_words_used), leaf_slots=pow2(st.leaf_words_used // 8))
            g.build_terrain(terrain_for_world(4 << x, 80 + k))
            out.append(g)
        return out

    reqs = make_requests(40, len(depths), [4 << x for x in depths], seed=22, lo_log2=8, hi_log2=13, max_brushes=8)
    grown = scenes()
    srv = Server(grown, max_candidates=1 << 16, max_jobs=256, collect_on_full=False)
    res = serve(srv, reqs, poisson_arrivals(len(reqs), 3000.0, 4))
    srv.close()
    assert res["rejected"] == 0 and res["completed"] == len(reqs)
    assert res["growths"] >= 1
    for s, x in enumerate(depths):
        ref = O.Dag(x, True)
        ref.build_terrain(O.terrain_for_world(4 << x, 80 + s))
        for r in reqs:
            if r.scene == s:
                ref.apply_brushes(r.brushes)
        assert np.array_equal(grown[s].export(), ref.export()), s
    fixed = scenes()
    srv = Server(fixed, max_candidates=1 << 16, max_jobs=256, collect_on_full=False, grow_on_full=False)
    for r in reqs:
        r.status, r.finished = "pending", -1.0
    res = serve(srv, reqs, poisson_arrivals(len(reqs), 3000.0, 4))
    srv.close()
    assert res["rejected"] > 0 and res["growths"] == 0
