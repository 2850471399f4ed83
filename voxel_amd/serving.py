"""Continuous-batching edit server over resident scenes.

Requests (scene, brush batch) arrive over time. The server keeps up to two
ticks in flight on two persistent executors (vx_batcher, each with its own
workspace, pinned staging buffer, stream and captured hipGraph):

- admission: a pending request is admitted when its worst-case growth
  (every touched brick a new leaf, every touched ancestor a new 9-word node)
  fits its scene's pool words, counted as the usage read back by the scene's
  last completed tick plus the bounds of its requests still in flight — the
  free HBM reserved for table and pool growth — and when the tick stays
  within the executor's candidate and brush capacity; a deferred request
  blocks the later requests of its scene (per-scene arrival order), other
  scenes proceed;
- submission (vx_batcher_submit) returns at once: the host prepares and
  uploads the next tick while the previous one runs. A tick whose scenes are
  disjoint from the tick in flight runs concurrently with it on the device;
  a tick that shares a scene waits for it on the device (vx_batcher_submit's
  `after`), never on the host;
- completion is polled (vx_batcher_query): a tick's requests complete when
  the host sees its event, each scene's pool usage comes back with it, and a
  scene that overflowed keeps its old root and its requests of that tick
  are marked rejected;
- a request that does not fit its scene even with nothing in flight triggers
  a garbage collection of that scene (vx_dag_collect_garbage) after the
  executors drain; if it still does not fit, the scene's pools (and slot
  tables, in proportion) grow to twice their size or to what the request
  needs (vx_dag_grow), provided the device's free HBM holds the growth plus
  a reserve; if it still does not fit it is rejected.

Per scene the requests apply in arrival order, so every scene's final DAG
equals replaying its requests one by one, and a request's effect is the same
alone or in the mix.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import dag as V
from ._lib import check


@dataclasses.dataclass
class Request:
    rid: int
    scene: int
    brushes: np.ndarray
    arrival: float = 0.0
    voxel_ops: int = 0
    done: float = -1.0      # seconds from the start of serve()
    status: str = "pending"
    finished: float = -1.0  # host clock (time.perf_counter) at completion
    bound: tuple = (0, 0)   # reserved (node words, leaf words) while admitted


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


def brick_candidates(d: V.HashDagGpu, brushes: np.ndarray) -> int:
    """Candidate leaf bricks of a request: the bricks of each brush's box
    clipped to the world (what an executor's capacity counts)."""
    world = d.world
    total = 0
    for b in brushes:
        ext = (b["hx"], b["hx"], b["hx"]) if b["kind"] == 0 else (b["hx"], b["hy"], b["hz"])
        n = 1
        for c, e in zip((b["cx"], b["cy"], b["cz"]), ext):
            lo, hi = max(0, int(c) - int(e)), min(world - 1, int(c) + int(e))
            n *= max(0, (hi >> 2) - (lo >> 2) + 1)
        total += n
    return total


@dataclasses.dataclass
class Tick:
    ex: int
    reqs: list
    scenes: set
    bounds: dict  # scene -> [node words, leaf words] reserved by the tick


class Server:
    """Continuous-batching executor. With `max_candidates` set, ticks run on
    `executors` persistent Batchers (pipeline captured once into a hipGraph
    per executor, replayed every tick), submitted asynchronously; without it,
    each tick is one synchronous multi-scene call (vx_dag_apply_batches)."""

    def __init__(self, scenes: list[V.HashDagGpu], max_candidates: int | None = None, max_jobs: int = 8192,
                 use_graph: bool = True, executors: int = 2, collect_on_full: bool = True,
                 grow_on_full: bool = True, hbm_reserve: int = 1 << 30):
        if not 1 <= executors <= 2:
            raise ValueError("one or two executors")
        self.scenes = scenes
        self.caps = [(d._cap.node_words, d._cap.leaf_words) for d in scenes]
        self.used = []
        for d in scenes:
            st = d.stats()
            self.used.append([st.node_words_used, st.leaf_words_used])
        self.ticks = 0
        self.tick_sizes = []
        self.overlapped = 0  # ticks submitted while another tick was still in flight
        self.concurrent = 0  # ... of them, with no scene in common (no device wait)
        self.collections = 0
        self.growths = 0
        self.grow_on_full = grow_on_full
        self.hbm_reserve = hbm_reserve  # device bytes a growth must leave free
        self.max_candidates = max_candidates
        self.max_jobs = max_jobs
        self.collect_on_full = collect_on_full
        self.exs = [V.Batcher(scenes, max_candidates, max_jobs, use_graph) for _ in range(executors)] \
            if max_candidates else []
        self.ex = self.exs[0] if self.exs else None
        self.inflight: list[Tick] = []

    def close(self):
        self.drain()
        for e in self.exs:
            e.close()
        self.exs, self.ex = [], None

    # -- admission --------------------------------------------------------------
    def _reserved(self, s: int) -> list[int]:
        r = [0, 0]
        for t in self.inflight:
            b = t.bounds.get(s)
            if b:
                r[0] += b[0]
                r[1] += b[1]
        return r

    def admit(self, pending: list[Request]) -> list[Request]:
        """Requests whose worst-case growth fits their scene's free pool words
        (usage known from the last completed tick + everything reserved by
        ticks in flight + this tick) and, with executors, whose candidates fit
        the executor's capacity; the rest stay pending, in order."""
        extra = {}
        out = []
        cand = jobs = 0
        blocked = set()
        for r in pending:
            if r.scene in blocked:
                continue  # keep per-scene arrival order: nothing overtakes a deferred request
            d = self.scenes[r.scene]
            e = extra.setdefault(r.scene, [0, 0])
            res = self._reserved(r.scene)
            nw, lw = growth_bound(d, r.brushes)
            cap_n, cap_l = self.caps[r.scene]
            u = self.used[r.scene]
            c = brick_candidates(d, r.brushes) if self.exs else 0
            fits_pool = u[0] + res[0] + e[0] + nw <= cap_n and u[1] + res[1] + e[1] + lw <= cap_l
            fits_exec = not self.exs or (cand + c <= self.max_candidates and jobs + len(r.brushes) <= self.max_jobs)
            if fits_pool and fits_exec:
                e[0] += nw
                e[1] += lw
                cand += c
                jobs += len(r.brushes)
                r.bound = (nw, lw)
                out.append(r)
            else:
                blocked.add(r.scene)
        return out

    # -- ticks ------------------------------------------------------------------
    def can_submit(self) -> bool:
        return not self.exs or len(self.inflight) < len(self.exs)

    def busy(self) -> bool:
        return bool(self.inflight)

    def submit(self, reqs: list[Request]) -> None:
        """Start one tick for `reqs` (arrival order kept per scene)."""
        if not reqs:
            return
        self.ticks += 1
        self.tick_sizes.append(len(reqs))
        bounds = {}
        for r in reqs:
            b = bounds.setdefault(r.scene, [0, 0])
            b[0] += r.bound[0]
            b[1] += r.bound[1]
        scenes = set(bounds)
        if not self.exs:
            self._tick_direct(reqs)
            return
        busy = {t.ex for t in self.inflight}
        ex = next(i for i in range(len(self.exs)) if i not in busy)
        after = None
        for t in self.inflight:
            if t.scenes & scenes:
                after = self.exs[t.ex]
        if self.inflight:
            self.overlapped += 1
            if after is None:
                self.concurrent += 1
        self.exs[ex].submit([(r.scene, r.brushes) for r in reqs], after=after)
        self.inflight.append(Tick(ex, reqs, scenes, bounds))

    def _tick_direct(self, reqs: list[Request]) -> None:
        from ._lib import OutOfCapacity
        failed = set()
        try:
            V.apply_batches([self.scenes[r.scene] for r in reqs], [r.brushes for r in reqs])
        except OutOfCapacity:
            failed = {s for s in {r.scene for r in reqs} if self.scenes[s].stats().last_status != 0}
        for s in {r.scene for r in reqs}:
            st = self.scenes[s].stats()
            self.used[s] = [st.node_words_used, st.leaf_words_used]
        now = time.perf_counter()
        for r in reqs:
            r.status = "rejected" if r.scene in failed else "done"
            r.finished = now

    def poll(self, wait: bool = False) -> list[Tick]:
        """Complete the ticks the device has finished (with wait, at least the
        oldest one). A scene that overflowed keeps its old root; its requests
        of that tick are marked rejected."""
        done = []
        for k, t in enumerate(list(self.inflight)):
            ok, status, st, used = self.exs[t.ex].query(wait=wait and k == 0)
            if not ok:
                continue
            if status not in (0, 2):  # VX_OK, VX_OUT_OF_CAPACITY (per scene, below)
                check(status, "vx_batcher_query")
            now = time.perf_counter()
            for s in t.scenes:
                self.used[s] = [int(used[s][0]), int(used[s][1])]
            for r in t.reqs:
                r.status = "rejected" if st[r.scene] != 0 else "done"
                r.finished = now
            self.inflight.remove(t)
            done.append(t)
        return done

    def drain(self) -> None:
        while self.inflight:
            self.poll(wait=True)

    def tick(self, reqs: list[Request]) -> None:
        """One tick, waited for (the synchronous form)."""
        for r in reqs:
            r.bound = growth_bound(self.scenes[r.scene], r.brushes)
        self.submit(reqs)
        self.drain()

    def make_room(self, r: Request) -> bool:
        """After the executors drain: garbage-collect r's scene, then, if r
        still does not fit, grow the scene within the free HBM; report whether
        r fits it now."""
        if not (self.collect_on_full or self.grow_on_full):
            return False
        self.drain()
        d = self.scenes[r.scene]
        if self.collect_on_full:
            d.collect_garbage()
            self.collections += 1
            st = d.stats()
            self.used[r.scene] = [st.node_words_used, st.leaf_words_used]
            if self.admit([r]):
                return True
        return self.grow_on_full and self._grow_for(r) and bool(self.admit([r]))

    def _grow_for(self, r: Request) -> bool:
        """Grow r's scene so that r's worst-case growth fits: each short pool
        to twice its size or to what r needs, the slot tables in the same
        proportion; only if the device's free HBM keeps `hbm_reserve` after."""
        import torch
        d = self.scenes[r.scene]
        nw, lw = growth_bound(d, r.brushes)
        u = self.used[r.scene]
        cap_n, cap_l = self.caps[r.scene]
        word_cap = 1 << 31
        want_n = cap_n if u[0] + nw <= cap_n else min(word_cap, max(2 * cap_n, u[0] + nw))
        want_l = cap_l if u[1] + lw <= cap_l else min(word_cap, max(2 * cap_l, u[1] + lw))
        if (want_n, want_l) == (cap_n, cap_l) or u[0] + nw > want_n or u[1] + lw > want_l:
            return False
        st = d.stats()
        slots_n = pow2_at_least(-(-st.node_slots * want_n // cap_n))
        slots_l = pow2_at_least(-(-st.leaf_slots * want_l // cap_l))
        need = 4 * (want_n - cap_n + want_l - cap_l) + 8 * (slots_n - st.node_slots + slots_l - st.leaf_slots)
        free, _ = torch.cuda.mem_get_info()
        if free < need + self.hbm_reserve:
            return False
        d.grow(want_n, want_l, slots_n, slots_l)
        self.caps[r.scene] = (d._cap.node_words, d._cap.leaf_words)
        self.growths += 1
        return True


def pow2_at_least(v: int) -> int:
    p = 1
    while p < v:
        p <<= 1
    return p


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


def serve(server: Server, reqs: list[Request], arrivals, max_tick: int = 1 << 30) -> dict:
    """Real-time loop: arrivals are offsets in seconds from the start. Ticks
    are submitted as soon as an executor is free and requests are pending;
    completions are polled between arrivals."""
    for r, a in zip(reqs, arrivals):
        r.arrival = float(a)
    pending: list[Request] = []
    nxt = 0
    t0 = time.perf_counter()
    ticks0 = server.ticks
    while nxt < len(reqs) or pending or server.busy():
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            pending.append(reqs[nxt])
            nxt += 1
        if server.busy():
            server.poll(wait=False)
        if pending and server.can_submit():
            batch = server.admit(pending[:max_tick])
            if batch:
                server.submit(batch)
                ids = {r.rid for r in batch}
                pending = [r for r in pending if r.rid not in ids]
                continue
            if not server.busy():
                # Nothing fits with the GPU idle: collect the oldest request's
                # scene, else reject the request.
                if server.make_room(pending[0]):
                    continue
                pending[0].status = "rejected"
                pending[0].finished = time.perf_counter()
                pending.pop(0)
                continue
        if server.busy():
            # Only completions can change anything before the next arrival.
            if nxt >= len(reqs) or reqs[nxt].arrival > time.perf_counter() - t0:
                time.sleep(0)
            continue
        if nxt < len(reqs):
            wait = reqs[nxt].arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.001))
    makespan = time.perf_counter() - t0
    for r in reqs:
        r.done = r.finished - t0
    lat = np.array([(r.done - r.arrival) * 1e3 for r in reqs if r.status == "done"])
    ops = sum(r.voxel_ops for r in reqs if r.status == "done")
    ticks = server.ticks - ticks0
    return {
        "requests": len(reqs),
        "completed": int(lat.size),
        "rejected": sum(1 for r in reqs if r.status == "rejected"),
        "p50_ms": float(np.percentile(lat, 50)) if lat.size else math.nan,
        "p99_ms": float(np.percentile(lat, 99)) if lat.size else math.nan,
        "mean_ms": float(lat.mean()) if lat.size else math.nan,
        "edits_per_s": ops / makespan,
        "makespan_s": makespan,
        "ticks": ticks,
        "mean_tick_requests": float(np.mean(server.tick_sizes[ticks0:])) if ticks else 0.0,
        "overlapped_ticks": server.overlapped,
        "concurrent_ticks": server.concurrent,
        "collections": server.collections,
        "growths": server.growths,
    }
