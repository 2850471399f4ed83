"""Host mirror of the reference's ``algo::svt::hash_dag_gpu`` API over the C ABI.

``HashDagGpu`` owns the device tables and pools (storage.cuh:16-105);
``place_voxel_edits`` / ``destroy_voxel_edits`` / ``apply_voxel_edits`` and
``apply_voxel_edits_workspace_size`` follow edit.cuh:16-61; ``get_voxel`` is a
batched form of query.cuh:12-39. Brush batches, terrain builds and the
canonical export extend the API for the synthetic workloads.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib, register
from .primitives import _ptr, _stream

PLACE = 0
DESTROY = 1


class Capacity(C.Structure):
    _fields_ = [("node_words", C.c_uint64), ("leaf_words", C.c_uint64), ("node_slots", C.c_uint64),
                ("leaf_slots", C.c_uint64)]


class Terrain(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("base", C.c_uint32), ("amp", C.c_uint32), ("period", C.c_uint32),
                ("octaves", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("node_words_used", C.c_uint64), ("leaf_words_used", C.c_uint64), ("node_slots", C.c_uint64),
                ("leaf_slots", C.c_uint64), ("root", C.c_uint32), ("inner_depth", C.c_uint32),
                ("leaf_words", C.c_uint32), ("last_status", C.c_uint32)]


class GcReport(C.Structure):
    _fields_ = [("node_words_before", C.c_uint64), ("node_words_after", C.c_uint64),
                ("leaf_words_before", C.c_uint64), ("leaf_words_after", C.c_uint64), ("live_nodes", C.c_uint64),
                ("live_leaves", C.c_uint64)]


BRUSH_DTYPE = np.dtype([("cx", "<i4"), ("cy", "<i4"), ("cz", "<i4"), ("hx", "<i4"), ("hy", "<i4"), ("hz", "<i4"),
                        ("kind", "<u4"), ("material", "<u4")])

_P = C.c_void_p
_u32, _u64, _i32, _sz = C.c_uint32, C.c_uint64, C.c_int32, C.c_size_t
register({
    "vx_dag_create": (_i32, [_u32, _i32, _P, C.POINTER(_P)]),
    "vx_dag_destroy": (_i32, [_P]),
    "vx_dag_reset": (_i32, [_P, _P]),
    "vx_dag_stats_get": (_i32, [_P, _P]),
    "vx_dag_build_terrain": (_i32, [_P, _P, _P]),
    "vx_dag_voxel_edits_workspace_size": (_i32, [_P, _u32, C.POINTER(_sz)]),
    "vx_dag_apply_voxel_edits": (_i32, [_P, _P, _u32, _u32, _u32, _P, _sz, _P]),
    "vx_dag_brush_workspace_size": (_i32, [_P, _P, _u32, C.POINTER(_sz)]),
    "vx_dag_apply_brushes": (_i32, [_P, _P, _u32, _P, _sz, _P]),
    "vx_dag_get_voxels": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_dag_raycast": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_dag_export_canonical": (_i32, [_P, _P, _u64, C.POINTER(_u64)]),
    "vx_dag_export_canonical_device": (_i32, [_P, _P, _u64, C.POINTER(_u64), _P]),
    "vx_dag_collect_garbage": (_i32, [_P, _P, _P]),
    "vx_dag_grow": (_i32, [_P, _P, _P]),
    "vx_dag_read_pools": (_i32, [_P, _P, _u64, C.POINTER(_u64), _P, _u64, C.POINTER(_u64)]),
    "vx_dag_check_unique": (_i32, [_P, C.POINTER(_u64), _P]),
    "vx_table_check_unique": (_i32, [_P, _u64, _P, _u64, _u32, C.POINTER(_u64), _P]),
    "vx_brush_voxel_ops": (_i32, [_P, _u32, _u32, C.POINTER(_u64)]),
    "vx_gen_brushes": (_i32, [_u64, _u64, _u32, _u32, _u32, _P]),
    "vx_gen_sized_batch": (_i32, [_u64, _u64, _u64, _u32, _u32, _u32, _u32, _u32, _P, C.POINTER(_u32),
                                  C.POINTER(_u32)]),
    "vx_dag_batch_workspace_size": (_i32, [_P, _P, _P, _u32, C.POINTER(_sz)]),
    "vx_dag_apply_batches": (_i32, [_P, _P, _P, _u32, _P, _sz, _P]),
    "vx_batcher_create": (_i32, [_P, _u32, _u64, _u32, _i32, C.POINTER(_P)]),
    "vx_batcher_destroy": (_i32, [_P]),
    "vx_batcher_info": (_i32, [_P, C.POINTER(_u64), C.POINTER(_i32)]),
    "vx_batcher_run": (_i32, [_P, _P, _P, _P, _u32, C.POINTER(_u64)]),
    "vx_batcher_layout": (_i32, [_P, C.POINTER(_u32), C.POINTER(_u32)]),
    "vx_batcher_submit": (_i32, [_P, _P, _P, _P, _u32, C.POINTER(_u64), _P]),
    "vx_batcher_query": (_i32, [_P, _i32, C.POINTER(_i32), _P, _P]),
    "vx_dag_profile_enable": (_i32, [_P, _i32]),
    "vx_dag_profile_read": (_i32, [_P, _P]),
    "vx_dag_profile_kernel_mask": (_i32, [_P, _u64]),
    "vx_dag_profile_kernel_period": (_i32, [_P, _u32]),
})

STAGES = ("records", "sort", "levels", "old_refs", "leaves", "nodes", "publish")


KERNELS = ("records", "sort_histogram", "sort_scan", "sort_partition", "sort_copy", "runs", "up", "topdown",
           "leaf_build", "leaf_probe", "leaf_commit", "leaf_place", "node_build", "node_probe", "node_commit",
           "node_place", "publish", "fill", "other", "tail_up", "tail_topdown", "tail_nodes")


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * 7), ("edits", C.c_uint64), ("records", C.c_uint64), ("leaves", C.c_uint64),
                ("nodes", C.c_uint64), ("sort_passes", C.c_uint64), ("sort_record_bytes", C.c_uint64),
                ("kernel_launches", C.c_uint64 * len(KERNELS)), ("kernel_ms", C.c_double * len(KERNELS)),
                ("pending_leaves", C.c_uint64), ("height_entries", C.c_uint64 * 16)]


STREAM_BRUSH = 5


def gen_brushes(seed: int, stream_id: int, n: int, world: int, mix: int) -> np.ndarray:
    """Seeded brush batch (see vx_gen_brushes in include/vx.h)."""
    out = np.zeros(n, BRUSH_DTYPE)
    check(lib().vx_gen_brushes(seed, stream_id, n, world, mix, out.ctypes.data_as(_P)), "vx_gen_brushes")
    return out


STREAM_BATCH = 6


def gen_sized_batch(seed: int, stream_id: int, index: int, lo_log2: int, hi_log2: int, max_brushes: int,
                    world: int, n_scenes: int = 1):
    """(brushes, scene) of a batch sized to a log-uniform voxel-op target."""
    out = np.zeros(max_brushes, BRUSH_DTYPE)
    n, s = C.c_uint32(), C.c_uint32()
    check(lib().vx_gen_sized_batch(seed, stream_id, index, lo_log2, hi_log2, max_brushes, world, n_scenes,
                                   out.ctypes.data_as(_P), C.byref(n), C.byref(s)), "vx_gen_sized_batch")
    return out[: n.value].copy(), s.value


def brush_voxel_ops(brushes: np.ndarray, world: int) -> int:
    """Exact voxel operations of a brush batch (per-brush voxels inside the world)."""
    b = np.ascontiguousarray(brushes, BRUSH_DTYPE)
    n = C.c_uint64()
    check(lib().vx_brush_voxel_ops(b.ctypes.data_as(_P), len(b), world, C.byref(n)), "vx_brush_voxel_ops")
    return n.value


# Upper estimates of a heightfield scene's interned nodes and leaves per inner
# depth (measured on the 512^3-formula terrain scaled to the world: nodes grow
# ~4x per level, leaves ~2-3x).
_NODES_EST = {8: 8e4, 9: 3e5, 10: 1.2e6, 11: 4.5e6, 12: 1.6e7, 13: 6.4e7}
_LEAVES_EST = {8: 1.6e5, 9: 5e5, 10: 1.5e6, 11: 3.2e6, 12: 5e6, 13: 8e6}


def scene_capacity(inner_depth: int, materials: bool = True, headroom: float = 1.5,
                   edit_leaves: float = 1e6) -> dict:
    """Pool / table sizes for a heightfield scene of this depth plus room for
    `edit_leaves` new leaves (and their ancestors) from edits."""
    d = min(max(inner_depth, 8), 13)
    shrink = 4.0 ** (min(inner_depth, 8) - 8)
    nodes = _NODES_EST[d] * shrink * headroom + 2 * edit_leaves
    leaves = _LEAVES_EST[d] * shrink * headroom + edit_leaves
    lw = 16 if materials else 2
    node_words = int(min(2**31 - 64, max(1 << 20, 7.5 * nodes)))
    leaf_words = int(min(2**31 - 64, max(1 << 20, lw * leaves)))
    node_slots = 1 << max(16, int(2.2 * nodes).bit_length())
    leaf_slots = 1 << max(16, int(2.2 * leaves).bit_length())
    return dict(node_words=node_words, leaf_words=leaf_words, node_slots=node_slots, leaf_slots=leaf_slots)


def terrain_for_world(world: int, seed: int, octaves: int = 4) -> Terrain:
    """The 512^3 heightfield (base 256, amplitude 64, period 64) scaled to `world`."""
    return Terrain(seed, world // 2, world // 8, world // 8, octaves)


class HashDagGpu:
    """Device Hash DAG of side 4 << inner_depth (occupancy, or per-voxel materials)."""

    def __init__(self, inner_depth: int, materials: bool = False, node_words: int = 1 << 22,
                 leaf_words: int = 1 << 22, node_slots: int = 1 << 20, leaf_slots: int = 1 << 20):
        self.h = C.c_void_p()
        cap = Capacity(node_words, leaf_words, node_slots, leaf_slots)
        check(lib().vx_dag_create(inner_depth, int(materials), C.byref(cap), C.byref(self.h)), "vx_dag_create")
        self._cap = cap
        self.inner_depth = inner_depth
        self.world = 4 << inner_depth
        self.materials = materials

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vx_dag_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None):
        check(lib().vx_dag_reset(self.h, _stream(stream)), "vx_dag_reset")

    def stats(self) -> Stats:
        s = Stats()
        check(lib().vx_dag_stats_get(self.h, C.byref(s)), "vx_dag_stats_get")
        return s

    def build_terrain(self, terrain: Terrain, stream=None):
        check(lib().vx_dag_build_terrain(self.h, C.byref(terrain), _stream(stream)), "vx_dag_build_terrain")

    # -- reference editing API (edit.cuh) -----------------------------------
    def apply_voxel_edits_workspace_size(self, count: int) -> int:
        n = C.c_size_t()
        check(lib().vx_dag_voxel_edits_workspace_size(self.h, count, C.byref(n)), "workspace size")
        return n.value

    def apply_voxel_edits(self, edits: torch.Tensor | None, op: int, material: int = 1, workspace=None,
                          stream=None, count: int | None = None):
        """edits: int32 CUDA tensor [n, 3] of x, y, z (device memory)."""
        n = (edits.shape[0] if edits is not None else 0) if count is None else count
        need = self.apply_voxel_edits_workspace_size(n)
        if workspace is None and need:
            workspace = torch.empty(need, dtype=torch.uint8, device="cuda")
        wsb = 0 if workspace is None else workspace.numel()
        check(lib().vx_dag_apply_voxel_edits(self.h, _ptr(edits), n, op, material, _ptr(workspace), wsb,
                                             _stream(stream)), "vx_dag_apply_voxel_edits")

    def place_voxel_edits(self, edits, material: int = 1, **kw):
        self.apply_voxel_edits(edits, PLACE, material, **kw)

    def destroy_voxel_edits(self, edits, **kw):
        self.apply_voxel_edits(edits, DESTROY, 0, **kw)

    # -- brushes -------------------------------------------------------------
    def brush_workspace_size(self, brushes: np.ndarray) -> int:
        b = np.ascontiguousarray(brushes, BRUSH_DTYPE)
        n = C.c_size_t()
        check(lib().vx_dag_brush_workspace_size(self.h, b.ctypes.data_as(_P), len(b), C.byref(n)), "brush ws")
        return n.value

    def apply_brushes(self, brushes: np.ndarray, workspace=None, stream=None):
        b = np.ascontiguousarray(brushes, BRUSH_DTYPE)
        if workspace is None:  # a caller's workspace is checked against the need by the library
            need = self.brush_workspace_size(b)
            if need:
                workspace = torch.empty(need, dtype=torch.uint8, device="cuda")
        wsb = 0 if workspace is None else workspace.numel()
        check(lib().vx_dag_apply_brushes(self.h, b.ctypes.data_as(_P), len(b), _ptr(workspace), wsb,
                                         _stream(stream)), "vx_dag_apply_brushes")

    # -- queries ------------------------------------------------------------------
    def get_voxels(self, xyz: torch.Tensor, stream=None) -> torch.Tensor:
        n = xyz.shape[0]
        out = torch.empty(n, dtype=torch.int32, device=xyz.device)
        check(lib().vx_dag_get_voxels(self.h, _ptr(xyz), n, _ptr(out), _stream(stream)), "vx_dag_get_voxels")
        return out

    def raycast(self, rays: torch.Tensor, stream=None) -> torch.Tensor:
        """rays: (n, 8) int32 device tensor {ox, oy, oz, dx, dy, dz, max_steps,
        0} (vx_ray); returns (n, 6) int32 {material, steps, x, y, z, 0}."""
        r = rays.contiguous()
        n = r.shape[0]
        out = torch.empty((n, 6), dtype=torch.int32, device=r.device)
        check(lib().vx_dag_raycast(self.h, _ptr(r), n, _ptr(out), _stream(stream)), "vx_dag_raycast")
        return out

    def profile(self, enable: bool = True, stages: bool = True):
        """Per-stage and per-kernel events (stages=False: per-kernel only)."""
        mode = (1 | (0 if stages else 2)) if enable else 0
        check(lib().vx_dag_profile_enable(self.h, mode), "vx_dag_profile_enable")

    def profile_period(self, period: int):
        """Time one launch per call of each timed kind, rotating over `period` positions."""
        check(lib().vx_dag_profile_kernel_period(self.h, period), "vx_dag_profile_kernel_period")

    def profile_kernels(self, kinds=None):
        """Time only these kernel kinds (names from KERNELS; None = all)."""
        mask = (1 << 64) - 1 if kinds is None else sum(1 << KERNELS.index(k) for k in kinds)
        check(lib().vx_dag_profile_kernel_mask(self.h, mask), "vx_dag_profile_kernel_mask")

    def profile_read(self) -> Profile:
        p = Profile()
        check(lib().vx_dag_profile_read(self.h, C.byref(p)), "vx_dag_profile_read")
        return p

    # -- maintenance ---------------------------------------------------------------
    def collect_garbage(self, stream=None) -> GcReport:
        """Keep only the items reachable from the root (HashDAG::collect_garbage,
        svt/hash_dag.hpp:133-191); the voxel field is unchanged."""
        r = GcReport()
        check(lib().vx_dag_collect_garbage(self.h, C.byref(r), _stream(stream)), "vx_dag_collect_garbage")
        return r

    def grow(self, node_words: int = 0, leaf_words: int = 0, node_slots: int = 0, leaf_slots: int = 0,
             stream=None) -> None:
        """Grow the pools and slot tables to at least these sizes (0, or a size
        at or below the current one, keeps it); every ref stays valid
        (vx_dag_grow)."""
        cap = Capacity(node_words, leaf_words, node_slots, leaf_slots)
        check(lib().vx_dag_grow(self.h, C.byref(cap), _stream(stream)), "vx_dag_grow")
        st = self.stats()
        word_cap = 1 << 31  # pool offsets stay below the tag bit
        self._cap = Capacity(max(self._cap.node_words, min(node_words, word_cap)),
                             max(self._cap.leaf_words, min(leaf_words, word_cap)), st.node_slots, st.leaf_slots)

    def read_pools(self) -> tuple[np.ndarray, np.ndarray]:
        """Host copies of the used words of the node and leaf pools."""
        nn, ln = C.c_uint64(), C.c_uint64()
        check(lib().vx_dag_read_pools(self.h, None, 0, C.byref(nn), None, 0, C.byref(ln)), "pool sizes")
        npool, lpool = np.empty(nn.value, np.uint32), np.empty(ln.value, np.uint32)
        check(lib().vx_dag_read_pools(self.h, npool.ctypes.data_as(_P), nn.value, C.byref(nn),
                                      lpool.ctypes.data_as(_P), ln.value, C.byref(ln)), "vx_dag_read_pools")
        return npool, lpool

    def check_unique(self, stream=None) -> dict:
        """Device-side invariant check of both interning tables: live slots,
        duplicate items and broken slots per table."""
        out = (_u64 * 6)()
        check(lib().vx_dag_check_unique(self.h, out, _stream(stream)), "vx_dag_check_unique")
        return {"nodes": {"live": out[0], "duplicates": out[1], "broken": out[2]},
                "leaves": {"live": out[3], "duplicates": out[4], "broken": out[5]}}

    def export(self) -> np.ndarray:
        """Canonical export (host array of u32 words)."""
        n = C.c_uint64()
        check(lib().vx_dag_export_canonical(self.h, None, 0, C.byref(n)), "export size")
        out = np.empty(n.value, np.uint32)
        check(lib().vx_dag_export_canonical(self.h, out.ctypes.data_as(_P), n.value, C.byref(n)), "export")
        return out

    def export_device(self, stream=None) -> torch.Tensor:
        """Canonical export left in device memory (int32 tensor of the u32 words)."""
        n = C.c_uint64()
        s = _stream(stream)
        check(lib().vx_dag_export_canonical_device(self.h, None, 0, C.byref(n), s), "export size")
        out = torch.empty(n.value, dtype=torch.int32, device="cuda")
        check(lib().vx_dag_export_canonical_device(self.h, _ptr(out), n.value, C.byref(n), s), "export")
        return out


class _BatchArgs:
    """Keeps the ctypes arrays of one batched call alive."""

    def __init__(self, dags, batches):
        if len(dags) != len(batches) or not dags:
            raise ValueError("one brush array per batch")
        self.arrays = [np.ascontiguousarray(b, BRUSH_DTYPE) for b in batches]
        n = len(dags)
        self.dags = (_P * n)(*[d.h.value for d in dags])
        self.ptrs = (_P * n)(*[a.ctypes.data for a in self.arrays])
        self.counts = (_u32 * n)(*[len(a) for a in self.arrays])
        self.n = n


def batch_workspace_size(dags, batches) -> int:
    a = _BatchArgs(dags, batches)
    n = C.c_size_t()
    check(lib().vx_dag_batch_workspace_size(a.dags, a.ptrs, a.counts, a.n, C.byref(n)), "batch workspace")
    return n.value


def apply_batches(dags, batches, workspace=None, stream=None):
    """Apply batches[i] to dags[i], all in one pipeline (same result as in order)."""
    a = _BatchArgs(dags, batches)
    need = C.c_size_t()
    check(lib().vx_dag_batch_workspace_size(a.dags, a.ptrs, a.counts, a.n, C.byref(need)), "batch workspace")
    if workspace is None and need.value:
        workspace = torch.empty(need.value, dtype=torch.uint8, device="cuda")
    wsb = 0 if workspace is None else workspace.numel()
    check(lib().vx_dag_apply_batches(a.dags, a.ptrs, a.counts, a.n, _ptr(workspace), wsb, _stream(stream)),
          "vx_dag_apply_batches")


def batch_candidates(scenes, batches) -> int:
    """Brick candidates of (scene index, brushes) batches (sizes a Batcher)."""
    total = 0
    for s, br in batches:
        world = scenes[s].world
        for b in br:
            ext = (b["hx"],) * 3 if b["kind"] == 0 else (b["hx"], b["hy"], b["hz"])
            n = 1
            for c, e in zip((b["cx"], b["cy"], b["cz"]), ext):
                lo, hi = max(0, int(c) - int(e)), min(world - 1, int(c) + int(e))
                n *= max(0, (hi >> 2) - (lo >> 2) + 1)
            total += n
    return total


class Batcher:
    """Persistent batched executor (vx_batcher): one pipeline per batch set,
    captured once into a hipGraph and replayed."""

    def __init__(self, scenes, max_candidates: int, max_jobs: int, use_graph: bool = True):
        self.scenes = list(scenes)
        arr = (_P * len(self.scenes))(*[d.h.value for d in self.scenes])
        self.h = C.c_void_p()
        check(lib().vx_batcher_create(arr, len(self.scenes), max_candidates, max_jobs, int(use_graph),
                                      C.byref(self.h)), "vx_batcher_create")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().vx_batcher_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        b, g = C.c_uint64(), C.c_int32()
        check(lib().vx_batcher_info(self.h, C.byref(b), C.byref(g)), "vx_batcher_info")
        return b.value, bool(g.value)

    def layout(self) -> tuple[int, int]:
        """(key bits, index bits) of the record sort; index bits 0 = (key, index) pairs."""
        k, i = C.c_uint32(), C.c_uint32()
        check(lib().vx_batcher_layout(self.h, C.byref(k), C.byref(i)), "vx_batcher_layout")
        return k.value, i.value

    @staticmethod
    def _args(batches):
        arrays = [np.ascontiguousarray(br, BRUSH_DTYPE) for _, br in batches]
        n = len(arrays)
        scene_of = (_u32 * max(1, n))(*[s for s, _ in batches])
        ptrs = (_P * max(1, n))(*[a.ctypes.data for a in arrays])
        counts = (_u32 * max(1, n))(*[len(a) for a in arrays])
        return arrays, scene_of, ptrs, counts, n

    def run(self, batches) -> int:
        """batches: list of (scene index, brush array). Returns candidates."""
        arrays, scene_of, ptrs, counts, n = self._args(batches)
        cand = C.c_uint64()
        check(lib().vx_batcher_run(self.h, scene_of, ptrs, counts, n, C.byref(cand)), "vx_batcher_run")
        return cand.value

    def submit(self, batches, after: "Batcher | None" = None) -> int:
        """Enqueue without waiting (vx_batcher_submit); `after` orders the
        device work behind another executor's last submission."""
        arrays, scene_of, ptrs, counts, n = self._args(batches)
        cand = C.c_uint64()
        check(lib().vx_batcher_submit(self.h, scene_of, ptrs, counts, n, C.byref(cand),
                                      after.h if after is not None else None), "vx_batcher_submit")
        return cand.value

    def query(self, wait: bool = False):
        """(done, status, per-scene status, per-scene (node, leaf) words used)
        of the last submission; status is the vx_status of the completed one."""
        n = len(self.scenes)
        done = C.c_int32()
        st = np.zeros(n, np.uint32)
        used = np.zeros(2 * n, np.uint64)
        r = lib().vx_batcher_query(self.h, int(wait), C.byref(done), st.ctypes.data_as(_P), used.ctypes.data_as(_P))
        if not done.value:
            return False, 0, None, None
        return True, r, st, used.reshape(n, 2)

