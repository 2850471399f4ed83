"""Records the CPU oracle's canonical DAG digests for the plain-edit, batching
and serving workloads at their full sizes (oracle/oracle.cpp replaying every
batch in order; the digest is FNV-1a over the canonical export words):

  plain: the 16384^3 materials heightfield of inner depth 12 (seed 3) after
    each of the benchmark's first two batches of 4096 brushes (brush stream
    seed 3, 50% spheres r[8,64] / 50% boxes h[4,48], 60/40 add/remove);
  batching: 8 materials heightfield scenes of inner depth 8, 9, 10, 11, 12, 9,
    10, 11 (worlds 1024^3..16384^3, heightfield seeds 10..17); 256 warm-up
    batches (request stream seed 103) then 256 batches (seed 3), each 2^8..2^20
    voxel ops of 1..16 brushes (50% spheres / 50% cubes, 60/40 add/remove);
  serving: 16 materials heightfield scenes of inner depth 10, 11, 12, 13 x 4
    (worlds 4096^3..32768^3, seeds 20..35); 1000 calibration requests (seed
    1020) then 2000 requests (seed 20), 1..64 brushes of 2^6..2^18 voxel ops
    each, applied per scene in request order.

The inputs come from the library's seeded host generators (the request
streams the benchmark and the GPU tests replay). Scenes are independent, so
they run in parallel worker processes; the largest (inner depth 13) takes a
few minutes of one core. Run:

    make -C oracle && python tests/golden/make_workload_golden.py [plain] [batching] [serving]

(no names: all three; named sections replace theirs in the existing file).
"""
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workloads.json")

BATCH_DEPTHS = [8, 9, 10, 11, 12, 9, 10, 11]
BATCH_SEEDS = (103, 3)      # warm-up stream, then the timed stream
BATCH_COUNT = 256
SERVE_DEPTHS = [10, 11, 12, 13] * 4
SERVE_SEEDS = ((1020, 1000), (20, 2000))  # (request stream seed, count): calibration, then Poisson run


def batching_requests():
    from voxel_amd.serving import make_requests
    worlds = [4 << d for d in BATCH_DEPTHS]
    return [make_requests(BATCH_COUNT, len(BATCH_DEPTHS), worlds, s, 8, 20, 16) for s in BATCH_SEEDS]


def serving_requests():
    from voxel_amd.serving import make_requests
    worlds = [4 << d for d in SERVE_DEPTHS]
    return [make_requests(n, len(SERVE_DEPTHS), worlds, s, 6, 18, 64) for s, n in SERVE_SEEDS]


def replay(depth, seed, batches):
    from oracle import oracle as O
    t0 = time.time()
    d = O.Dag(depth, True)
    d.build_terrain(O.terrain_for_world(4 << depth, seed))
    for br in batches:
        d.apply_brushes(br)
    e = d.export()
    out = {"depth": depth, "terrain_seed": seed, "batches": len(batches), "words": int(e.size),
           "nodes": int(e[4]), "leaves": int(e[5]), "digest": "%016x" % O.digest(e)}
    del e
    print(f"depth {depth} seed {seed}: {out['nodes']} nodes {out['leaves']} leaves "
          f"({time.time() - t0:.0f} s)", flush=True)
    return out


PLAIN_DEPTH, PLAIN_SEED, PLAIN_BRUSHES, PLAIN_BATCHES = 12, 3, 4096, 2


def plain_batches():
    """The benchmark's batches k = 0, 1 (bench.py: brush stream STREAM_BRUSH + (k << 8), config-2 mix)."""
    from voxel_amd import dag as V
    world = 4 << PLAIN_DEPTH
    return [V.gen_brushes(PLAIN_SEED, V.STREAM_BRUSH + (k << 8), PLAIN_BRUSHES, world, 1) for k in range(PLAIN_BATCHES)]


def plain_digests():
    """Digest after every batch (one oracle scene, batches applied in order)."""
    from oracle import oracle as O
    t0 = time.time()
    d = O.Dag(PLAIN_DEPTH, True)
    d.build_terrain(O.terrain_for_world(4 << PLAIN_DEPTH, PLAIN_SEED))
    out = []
    for k, br in enumerate(plain_batches()):
        d.apply_brushes(br)
        e = d.export()
        out.append({"after_batch": k, "words": int(e.size), "nodes": int(e[4]), "leaves": int(e[5]),
                    "digest": "%016x" % O.digest(e)})
        del e
        print(f"plain after batch {k}: {out[-1]['nodes']} nodes {out[-1]['leaves']} leaves "
              f"({time.time() - t0:.0f} s)", flush=True)
    return out


def scene_jobs(depths, seed0, streams):
    jobs = []
    for s, depth in enumerate(depths):
        batches = [r.brushes for reqs in streams for r in reqs if r.scene == s]
        jobs.append((depth, seed0 + s, batches))
    return jobs


def main():
    workers = int(os.environ.get("GOLDEN_WORKERS", "4"))
    want = set(sys.argv[1:]) or {"plain", "batching", "serving"}
    doc = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            doc = json.load(f)
    doc["about"] = ("CPU oracle canonical digests (FNV-1a over the export words) for the plain-edit, batching and "
                    "serving workloads; generated by tests/golden/make_workload_golden.py")
    if "plain" in want:
        doc["plain"] = {"depth": PLAIN_DEPTH, "terrain_seed": PLAIN_SEED, "brush_seed": PLAIN_SEED,
                        "brushes_per_batch": PLAIN_BRUSHES, "mix": 1, "digests": plain_digests()}
    jobs = {}
    if "batching" in want:
        jobs["batching"] = scene_jobs(BATCH_DEPTHS, 10, batching_requests())
    if "serving" in want:
        jobs["serving"] = scene_jobs(SERVE_DEPTHS, 20, serving_requests())
    # Largest scenes first, so the pool's tail is short.
    order = sorted(((k, i) for k in jobs for i in range(len(jobs[k]))), key=lambda ki: -jobs[ki[0]][ki[1]][0])
    res = {k: [None] * len(v) for k, v in jobs.items()}
    if order:
        with cf.ProcessPoolExecutor(max_workers=workers) as ex:
            futs = {ex.submit(replay, *jobs[k][i]): (k, i) for k, i in order}
            for f in cf.as_completed(futs):
                k, i = futs[f]
                res[k][i] = f.result()
    if "batching" in res:
        doc["batching"] = {"depths": BATCH_DEPTHS, "terrain_seeds": list(range(10, 18)),
                           "request_seeds": list(BATCH_SEEDS), "batches_per_stream": BATCH_COUNT, "log2_ops": [8, 20],
                           "max_brushes": 16, "scenes": res["batching"]}
    if "serving" in res:
        doc["serving"] = {"depths": SERVE_DEPTHS, "terrain_seeds": list(range(20, 36)),
                          "request_streams": [list(x) for x in SERVE_SEEDS], "log2_ops": [6, 18], "max_brushes": 64,
                          "scenes": res["serving"]}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
