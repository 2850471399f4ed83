"""Host mirror of the reference's ``algo::svt::cuda`` pointer-SVO API over the C ABI.

``GpuSvo`` owns the device buffers (storage.cuh:12-122) and is reset by
``reset`` (svo.cuh:14-22); ``place_voxel_edits`` / ``destroy_voxel_edits`` /
``apply_voxel_edits_workspace_size`` follow voxel/edit.cuh:43-92,
``voxel_edits_to_leaf_masks`` edit.cuh:14-41 and ``get_voxels`` is a batched
query.cuh:12-39.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib, register
from .primitives import _ptr, _stream

PLACE = 0
DESTROY = 1
MAX_DEPTH = 10
WORLD = 4096
DEFAULT_MAX_NODES = 131072   # kDefaultMaxNodeCount (config.cuh:32)
DEFAULT_MAX_LEAVES = 524288  # kDefaultMaxLeafCount (config.cuh:33)


class Counters(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("leaf_count", C.c_uint32), ("free_node_count", C.c_uint32),
                ("free_leaf_count", C.c_uint32)]


class View(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("leaves", C.c_void_p), ("counters", C.c_void_p),
                ("free_node_indices", C.c_void_p), ("free_leaf_indices", C.c_void_p),
                ("max_node_count", C.c_uint32), ("max_leaf_count", C.c_uint32)]


_P = C.c_void_p
_u32, _u64, _i32, _sz = C.c_uint32, C.c_uint64, C.c_int32, C.c_size_t
register({
    "vx_svo_create": (_i32, [_u32, _u32, C.POINTER(_P)]),
    "vx_svo_destroy": (_i32, [_P]),
    "vx_svo_reset": (_i32, [_P, _P]),
    "vx_svo_view_get": (_i32, [_P, _P]),
    "vx_svo_counters_get": (_i32, [_P, _P, _P]),
    "vx_voxel_edits_to_leaf_masks_workspace_size": (_i32, [_u32, C.POINTER(_sz)]),
    "vx_voxel_edits_to_leaf_masks": (_i32, [_P, _u32, _P, _P, _P, _sz, _P]),
    "vx_svo_voxel_edits_workspace_size": (_i32, [_u32, C.POINTER(_sz)]),
    "vx_svo_apply_voxel_edits": (_i32, [_P, _P, _u32, _u32, _P, _sz, _P]),
    "vx_svo_get_voxels": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_svo_export_canonical": (_i32, [_P, _P, _u64, C.POINTER(_u64)]),
    "vx_gen_xorshift_voxels": (_i32, [_u32, _u32, _P]),
})


def _edits(xyz, device="cuda") -> torch.Tensor:
    """(n, 3) voxel coordinates as a contiguous int32 device tensor (u32 bits)."""
    if isinstance(xyz, torch.Tensor):
        return xyz.to(device=device, dtype=torch.int32).contiguous()
    a = np.ascontiguousarray(np.asarray(xyz, np.int64).reshape(-1, 3) & 0xFFFFFFFF, np.uint32)
    return torch.from_numpy(a.view(np.int32)).to(device)


def sphere_edits(cx: int, cy: int, cz: int, r: int) -> np.ndarray:
    """(n, 3) u32 voxels of the solid sphere |v - c|^2 <= r^2 in z, y, x order
    (make_voxel_sphere_edits, bench/bench_cuda_svt.cu:56-80)."""
    d = np.arange(-r, r + 1, dtype=np.int64)
    dz, dy, dx = np.meshgrid(d, d, d, indexing="ij")
    inside = dx * dx + dy * dy + dz * dz <= r * r
    return np.stack([(cx + dx)[inside], (cy + dy)[inside], (cz + dz)[inside]], 1).astype(np.uint32)


def random_edits(count: int, world: int = WORLD) -> np.ndarray:
    """(count, 3) u32 voxels of the reference SVO benchmark's random workload
    (make_random_edits, bench/bench_cuda_svt.cu:90-112)."""
    out = np.empty((count, 3), np.uint32)
    check(lib().vx_gen_xorshift_voxels(count, world, out.ctypes.data_as(_P)), "vx_gen_xorshift_voxels")
    return out


def voxel_edits_to_leaf_masks_workspace_size(count: int) -> int:
    n = C.c_size_t()
    check(lib().vx_voxel_edits_to_leaf_masks_workspace_size(count, C.byref(n)), "leaf mask workspace size")
    return n.value


def voxel_edits_to_leaf_masks(edits, workspace=None, stream=None) -> torch.Tensor:
    """(m, 3) int32 device tensor {leaf_key, voxel_data_low, voxel_data_high}
    (u32 bits), one row per touched leaf, sorted by leaf key."""
    e = _edits(edits)
    n = e.shape[0]
    out = torch.empty((max(n, 1), 3), dtype=torch.int32, device=e.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=e.device)
    need = voxel_edits_to_leaf_masks_workspace_size(n)
    if workspace is None and n:
        workspace = torch.empty(need, dtype=torch.uint8, device=e.device)
    wsb = 0 if workspace is None else workspace.numel()
    check(lib().vx_voxel_edits_to_leaf_masks(_ptr(e), n, _ptr(out), _ptr(cnt), _ptr(workspace), wsb,
                                             _stream(stream)), "vx_voxel_edits_to_leaf_masks")
    return out[: int(cnt.item())]


class GpuSvo:
    """Device pointer SVO of 4096^3 voxels (GpuSvo<MaxNodeCount, MaxLeafCount>)."""

    def __init__(self, max_node_count: int = DEFAULT_MAX_NODES, max_leaf_count: int = DEFAULT_MAX_LEAVES,
                 reset: bool = True):
        self.h = C.c_void_p()
        check(lib().vx_svo_create(max_node_count, max_leaf_count, C.byref(self.h)), "vx_svo_create")
        self.max_node_count = max_node_count
        self.max_leaf_count = max_leaf_count
        if reset:
            self.reset()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vx_svo_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None):
        check(lib().vx_svo_reset(self.h, _stream(stream)), "vx_svo_reset")

    def view(self) -> View:
        v = View()
        check(lib().vx_svo_view_get(self.h, C.byref(v)), "vx_svo_view_get")
        return v

    def counters(self, stream=None) -> tuple[int, int, int, int]:
        c = Counters()
        check(lib().vx_svo_counters_get(self.h, C.byref(c), _stream(stream)), "vx_svo_counters_get")
        return (c.node_count, c.leaf_count, c.free_node_count, c.free_leaf_count)

    @staticmethod
    def apply_voxel_edits_workspace_size(count: int) -> int:
        n = C.c_size_t()
        check(lib().vx_svo_voxel_edits_workspace_size(count, C.byref(n)), "workspace size")
        return n.value

    def apply_voxel_edits(self, edits, op: int, workspace=None, stream=None):
        e = _edits(edits)
        n = e.shape[0]
        if workspace is None:
            workspace = torch.empty(max(self.apply_voxel_edits_workspace_size(n), 1), dtype=torch.uint8,
                                    device=e.device)
        check(lib().vx_svo_apply_voxel_edits(self.h, _ptr(e), n, op, _ptr(workspace), workspace.numel(),
                                             _stream(stream)), "vx_svo_apply_voxel_edits")

    def place_voxel_edits(self, edits, **kw):
        self.apply_voxel_edits(edits, PLACE, **kw)

    def destroy_voxel_edits(self, edits, **kw):
        self.apply_voxel_edits(edits, DESTROY, **kw)

    def get_voxels(self, xyz, stream=None) -> torch.Tensor:
        q = _edits(xyz)
        out = torch.empty(q.shape[0], dtype=torch.int32, device=q.device)
        check(lib().vx_svo_get_voxels(self.h, _ptr(q), q.shape[0], _ptr(out), _stream(stream)), "vx_svo_get_voxels")
        return out

    def export(self) -> np.ndarray:
        n = C.c_uint64()
        check(lib().vx_svo_export_canonical(self.h, None, 0, C.byref(n)), "export size")
        out = np.empty(n.value, np.uint32)
        check(lib().vx_svo_export_canonical(self.h, out.ctypes.data_as(_P), n.value, C.byref(n)), "export")
        return out
