"""Continuous-batching edit server over resident scenes.

Requests (scene, brush batch) arrive over time; every tick the server admits
the pending requests whose worst-case table growth fits the free pool space of
their scene (admission by free HBM reserved for node/leaf growth), batches
all admitted requests into ONE multi-scene pipeline call
(`vx_dag_apply_batches`), and completes them together. Requests that arrive
while a tick runs on the GPU wait for the next tick. Per scene the requests
apply in arrival order, so every scene's final DAG equals replaying its
requests one by one, and a request's effect is the same alone or in the mix.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import dag as V


@dataclasses.dataclass
class Request:
    rid: int
    scene: int
    brushes: np.ndarray
    arrival: float = 0.0
    voxel_ops: int = 0
    done: float = -1.0
    status: str = "pending"


def growth_bound(d: V.HashDagGpu, brushes: np.ndarray) -> tuple[int, int]:
    """Worst-case (node words, leaf words) one request can add to its scene:
    every touched brick a new leaf, every touched ancestor a new 9-word item."""
    world = d.world
    bricks = 0
    for b in brushes:
        ext = (b["hx"], b["hx"], b["hx"]) if b["kind"] == 0 else (b["hx"], b["hy"], b["hz"])
        n = 1
        for c, e in zip((b["cx"], b["cy"], b["cz"]), ext):
            lo, hi = max(0, int(c) - int(e)), min(world - 1, int(c) + int(e))
            n *= max(0, (hi >> 2) - (lo >> 2) + 1)
        bricks += n
    leaf_words = bricks * (16 if d.materials else 2)
    nodes = sum(min(bricks, 8 ** lv) for lv in range(d.inner_depth))
    return nodes * 9, leaf_words


class Server:
    """Tick executor. With `max_candidates` set, ticks run on a persistent
    Batcher (pipeline captured once into a hipGraph, replayed every tick) and
    admission also keeps each tick within its candidate capacity."""

    def __init__(self, scenes: list[V.HashDagGpu], max_candidates: int | None = None, max_jobs: int = 8192,
                 use_graph: bool = True):
        self.scenes = scenes
        self.caps = [(d._cap.node_words, d._cap.leaf_words) for d in scenes]
        self.ticks = 0
        self.tick_sizes = []
        self.max_candidates = max_candidates
        self.max_jobs = max_jobs
        self.ex = V.Batcher(scenes, max_candidates, max_jobs, use_graph) if max_candidates else None

    def admit(self, pending: list[Request]) -> list[Request]:
        """Requests whose worst-case growth fits their scene's free pool words
        (cumulative within the tick) and, with a Batcher, whose candidates fit
        its capacity; the rest stay pending, in order."""
        used = {}
        out = []
        cand = jobs = 0
        blocked = set()
        for r in pending:
            if r.scene in blocked:
                continue  # keep per-scene arrival order: nothing overtakes a deferred request
            d = self.scenes[r.scene]
            if r.scene not in used:
                st = d.stats()
                used[r.scene] = [st.node_words_used, st.leaf_words_used]
            nw, lw = growth_bound(d, r.brushes)
            cap_n, cap_l = self.caps[r.scene]
            u = used[r.scene]
            c = V.batch_candidates(self.scenes, [(r.scene, r.brushes)]) if self.ex else 0
            fits_pool = u[0] + nw <= cap_n and u[1] + lw <= cap_l
            fits_exec = not self.ex or (cand + c <= self.max_candidates and jobs + len(r.brushes) <= self.max_jobs)
            if fits_pool and fits_exec:
                u[0] += nw
                u[1] += lw
                cand += c
                jobs += len(r.brushes)
                out.append(r)
            else:
                blocked.add(r.scene)
        return out

    def tick(self, reqs: list[Request], workspace=None) -> None:
        """One pipeline call for every request of the tick, in arrival order."""
        if not reqs:
            return
        if self.ex:
            self.ex.run([(r.scene, r.brushes) for r in reqs])
        else:
            V.apply_batches([self.scenes[r.scene] for r in reqs], [r.brushes for r in reqs], workspace=workspace)
        self.ticks += 1
        self.tick_sizes.append(len(reqs))


def make_requests(n: int, n_scenes: int, worlds: list[int], seed: int, lo_log2: int = 6, hi_log2: int = 18,
                  max_brushes: int = 64) -> list[Request]:
    """Seeded requests: uniform scene, 1..max_brushes brushes sized to a
    log-uniform [2^lo, 2^hi] voxel-op target."""
    out = []
    for i in range(n):
        # The scene draw picks the world size used to place the brushes.
        _, s = V.gen_sized_batch(seed, V.STREAM_BATCH + 1, i, lo_log2, hi_log2, 1, 64, n_scenes)
        br, _ = V.gen_sized_batch(seed, V.STREAM_BATCH + 2, i, lo_log2, hi_log2, max_brushes, worlds[s], 1)
        out.append(Request(i, s, br, voxel_ops=V.brush_voxel_ops(br, worlds[s])))
    return out


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def serve(server: Server, reqs: list[Request], arrivals: np.ndarray, max_tick: int = 1 << 30) -> dict:
    """Real-time loop: arrivals are offsets in seconds from the start."""
    import torch
    for r, a in zip(reqs, arrivals):
        r.arrival = float(a)
    pending: list[Request] = []
    nxt = 0
    t0 = time.perf_counter()
    rejected = 0
    while nxt < len(reqs) or pending:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            pending.append(reqs[nxt])
            nxt += 1
        if not pending:
            wait = reqs[nxt].arrival - now
            if wait > 0:
                time.sleep(min(wait, 0.001))
            continue
        batch = server.admit(pending[:max_tick])
        if not batch:
            # Nothing fits even alone with the GPU idle: reject the oldest.
            pending[0].status = "rejected"
            pending[0].done = time.perf_counter() - t0
            pending.pop(0)
            rejected += 1
            continue
        server.tick(batch)
        torch.cuda.synchronize()
        done = time.perf_counter() - t0
        ids = {r.rid for r in batch}
        for r in batch:
            r.done = done
            r.status = "done"
        pending = [r for r in pending if r.rid not in ids]
    makespan = time.perf_counter() - t0
    lat = np.array([(r.done - r.arrival) * 1e3 for r in reqs if r.status == "done"])
    ops = sum(r.voxel_ops for r in reqs if r.status == "done")
    return {
        "requests": len(reqs),
        "completed": int(lat.size),
        "rejected": rejected,
        "p50_ms": float(np.percentile(lat, 50)) if lat.size else math.nan,
        "p99_ms": float(np.percentile(lat, 99)) if lat.size else math.nan,
        "mean_ms": float(lat.mean()) if lat.size else math.nan,
        "edits_per_s": ops / makespan,
        "makespan_s": makespan,
        "ticks": server.ticks,
        "mean_tick_requests": float(np.mean(server.tick_sizes)) if server.tick_sizes else 0.0,
    }
