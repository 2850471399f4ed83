"""Continuous-batching server: final DAGs equal the CPU oracle replaying the
requests in arrival order; a scene's result is the same when its requests
are served alone and when they are mixed with other scenes' requests."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _scenes(depths, materials=True, seeds=None):
    from voxel_amd.dag import HashDagGpu, terrain_for_world
    out = []
    for k, d in enumerate(depths):
        g = HashDagGpu(d, materials, node_words=1 << 22, leaf_words=1 << 23)
        g.build_terrain(terrain_for_world(4 << d, (seeds or [20 + i for i in range(len(depths))])[k]))
        out.append(g)
    return out


@pytest.mark.parametrize("executor", ["direct", "graph"])
def test_served_scenes_equal_oracle_replay(gpu, executor):
    from voxel_amd.serving import Server, make_requests, poisson_arrivals, serve
    depths = [5, 6, 4, 6]
    scenes = _scenes(depths)
    reqs = make_requests(48, len(depths), [4 << d for d in depths], seed=20, lo_log2=6, hi_log2=12, max_brushes=8)
    srv = Server(scenes) if executor == "direct" else Server(scenes, max_candidates=20_000, max_jobs=256)
    res = serve(srv, reqs, poisson_arrivals(len(reqs), 4000.0, 1))
    assert res["completed"] == len(reqs) and res["rejected"] == 0
    assert res["ticks"] < len(reqs)  # requests were batched
    for s, d in enumerate(depths):
        ref = O.Dag(d, True)
        ref.build_terrain(O.terrain_for_world(4 << d, 20 + s))
        for r in reqs:  # arrival order
            if r.scene == s:
                ref.apply_brushes(r.brushes)
        assert np.array_equal(scenes[s].export(), ref.export()), s


def test_result_same_alone_and_in_mix(gpu):
    from voxel_amd.serving import Server, make_requests, poisson_arrivals, serve
    depths = [5, 5, 5]
    reqs = make_requests(36, len(depths), [4 << d for d in depths], seed=21, lo_log2=6, hi_log2=11, max_brushes=6)
    mixed = _scenes(depths)
    serve(Server(mixed), reqs, poisson_arrivals(len(reqs), 5000.0, 2))
    for s in range(len(depths)):
        alone = _scenes([depths[s]], seeds=[20 + s])
        mine = [r for r in reqs if r.scene == s]
        for r in mine:
            r.scene = 0
        serve(Server(alone), mine, poisson_arrivals(len(mine), 5000.0, 3))
        for r in mine:
            r.scene = s
        assert np.array_equal(alone[0].export(), mixed[s].export()), s


def test_admission_by_free_pool_space_keeps_scene_order(gpu):
    """A request whose worst-case growth does not fit its scene's free pool
    words is deferred, later requests of that scene wait behind it, and other
    scenes' requests are still admitted."""
    from voxel_amd.dag import HashDagGpu
    from voxel_amd.serving import Request, Server
    tiny = HashDagGpu(5, True, node_words=4096, leaf_words=4096, node_slots=1024, leaf_slots=1024)
    roomy = HashDagGpu(5, True)
    big = O.gen_brushes(1, O.STREAM_BRUSH, 4, 128, 0)
    small = np.zeros(1, big.dtype)
    small[0] = (3, 3, 3, 1, 0, 0, 0, 9)
    srv = Server([tiny, roomy])
    admitted = srv.admit([Request(0, 0, big), Request(1, 0, small), Request(2, 1, small)])
    assert [r.rid for r in admitted] == [2]
    assert [r.rid for r in srv.admit([Request(3, 0, small)])] == [3]


def _replay(depth, seed, reqs, scene):
    ref = O.Dag(depth, True)
    ref.build_terrain(O.terrain_for_world(4 << depth, seed))
    for r in reqs:
        if r.scene == scene and r.status == "done":
            ref.apply_brushes(r.brushes)
    return ref.export()


def test_next_tick_is_submitted_while_one_runs(gpu):
    """Two executors: a tick is enqueued while the previous one is still in
    flight (overlapped), ticks over disjoint scenes run without a device
    wait (concurrent), and the final scenes still equal the oracle replay."""
    from voxel_amd.serving import Server, make_requests, serve
    depths = [6, 6, 5, 6]
    scenes = _scenes(depths)
    reqs = make_requests(120, len(depths), [4 << d for d in depths], seed=23, lo_log2=8, hi_log2=13, max_brushes=8)
    srv = Server(scenes, max_candidates=40_000, max_jobs=512)
    res = serve(srv, reqs, [0.0] * len(reqs), max_tick=3)  # saturated: always something to submit
    assert res["completed"] == len(reqs) and res["rejected"] == 0
    assert res["overlapped_ticks"] > 0 and res["concurrent_ticks"] > 0
    srv.close()
    for s, d in enumerate(depths):
        assert np.array_equal(scenes[s].export(), _replay(d, 20 + s, reqs, s)), s


def test_overflowing_scene_is_rejected_others_publish(gpu):
    """A scene whose slot table overflows inside a tick keeps its old root and
    its requests of that tick are rejected; the other scene of the tick
    publishes; the loop keeps serving."""
    from voxel_amd.dag import HashDagGpu, terrain_for_world
    from voxel_amd.serving import Request, Server, serve
    small = HashDagGpu(5, True, node_words=1 << 20, leaf_words=1 << 20, node_slots=64, leaf_slots=64)
    roomy = HashDagGpu(5, True, node_words=1 << 22, leaf_words=1 << 23)
    roomy.build_terrain(terrain_for_world(128, 21))
    before = small.export()  # empty: 64 slots hold no terrain
    br = O.gen_brushes(5, O.STREAM_BRUSH, 20, 128, 1)
    reqs = [Request(0, 0, br), Request(1, 1, br)]
    srv = Server([small, roomy], max_candidates=1 << 18, max_jobs=256)
    res = serve(srv, reqs, [0.0, 0.0])
    srv.close()
    assert [r.status for r in reqs] == ["rejected", "done"] and res["completed"] == 1
    assert np.array_equal(small.export(), before)
    ref = O.Dag(5, True)
    ref.build_terrain(O.terrain_for_world(128, 21))
    ref.apply_brushes(br)
    assert np.array_equal(roomy.export(), ref.export())


def test_full_scene_is_collected_and_served(gpu):
    """Requests that repaint one region fill a small scene with garbage;
    admission then finds no room, the server collects the scene's garbage and
    carries on; the result equals the oracle replay of every request."""
    from voxel_amd.dag import BRUSH_DTYPE, HashDagGpu, terrain_for_world
    from voxel_amd.serving import Request, Server, serve
    g = HashDagGpu(5, True, node_words=40_000, leaf_words=60_000, node_slots=1 << 14, leaf_slots=1 << 14)
    g.build_terrain(terrain_for_world(128, 20))
    reqs = []
    for i in range(40):
        b = np.zeros(1, BRUSH_DTYPE)
        b[0] = (64, 64, 64, 12, 0, 0, 0, 1 + (i % 250))
        reqs.append(Request(i, 0, b))
    srv = Server([g], max_candidates=1 << 16, max_jobs=64)
    res = serve(srv, reqs, [0.0] * len(reqs), max_tick=1)
    srv.close()
    assert res["completed"] == len(reqs) and res["collections"] > 0
    assert np.array_equal(g.export(), _replay(5, 20, reqs, 0))
