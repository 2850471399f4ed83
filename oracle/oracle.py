"""ctypes view of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py``, as the checker. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_hash_dag.so")

# Stream ids of the workload generators (shared with the product).
STREAM_KEYS = 1
STREAM_VALUES = 2
STREAM_SCAN = 3
STREAM_HASH = 4
STREAM_BRUSH = 5
STREAM_BATCH = 6
STREAM_ARRIVAL = 7


class Terrain(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("base", C.c_uint32), ("amp", C.c_uint32),
                ("period", C.c_uint32), ("octaves", C.c_uint32)]


def terrain_for_world(world: int, seed: int, octaves: int = 4) -> "Terrain":
    """The 512^3 heightfield (base 256, amplitude 64, period 64) scaled to `world`."""
    return Terrain(seed, world // 2, world // 8, world // 8, octaves)


class Brush(C.Structure):
    _fields_ = [("cx", C.c_int32), ("cy", C.c_int32), ("cz", C.c_int32),
                ("hx", C.c_int32), ("hy", C.c_int32), ("hz", C.c_int32),
                ("kind", C.c_uint32), ("material", C.c_uint32)]


BRUSH_DTYPE = np.dtype([("cx", "<i4"), ("cy", "<i4"), ("cz", "<i4"), ("hx", "<i4"),
                        ("hy", "<i4"), ("hz", "<i4"), ("kind", "<u4"), ("material", "<u4")])

_lib = None
_ref = None


def build() -> None:
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int32
        sig = {
            "vxo_mix64": (u64, [u64]),
            "vxo_rand": (u64, [u64, u64, u64]),
            "vxo_gen_u32": (None, [u64, u64, u64, u32, u32, P]),
            "vxo_gen_morton": (None, [u64, u64, u64, u32, P]),
            "vxo_gen_unique_u64": (None, [u64, u64, u64, P]),
            "vxo_morton3": (u64, [u32, u32, u32]),
            "vxo_exclusive_sum_u32": (None, [P, u64, P]),
            "vxo_inclusive_sum_u32": (None, [P, u64, P]),
            "vxo_exclusive_max_u32": (None, [P, u64, P]),
            "vxo_inclusive_max_u32": (None, [P, u64, P]),
            "vxo_exclusive_sum_u64": (None, [P, u64, P]),
            "vxo_inclusive_sum_u64": (None, [P, u64, P]),
            "vxo_exclusive_max_u64": (None, [P, u64, P]),
            "vxo_inclusive_max_u64": (None, [P, u64, P]),
            "vxo_sort_order": (None, [P, u32, u64, u32, P]),
            "vxo_sort_pairs_u32": (None, [P, P, u64, u32]),
            "vxo_sort_pairs_u64": (None, [P, P, u64, u32]),
            "vxo_set_contains_u64": (None, [P, u64, P, u64, P]),
            "vxo_terrain_height": (i32, [P, u32, u32]),
            "vxo_terrain_heights": (None, [P, P, P, u64, P]),
            "vxo_gen_brushes": (None, [u64, u64, u32, u32, u32, P]),
            "vxo_brush_voxels": (u64, [P, u32]),
            "vxo_dag_create": (P, [u32, C.c_int]),
            "vxo_dag_destroy": (None, [P]),
            "vxo_dag_build_terrain": (None, [P, P]),
            "vxo_dag_apply_brushes": (None, [P, P, u32]),
            "vxo_dag_apply_voxels": (None, [P, P, u32, u32]),
            "vxo_dag_box": (None, [P, u32, u32, u32, u32, u32, u32, u32]),
            "vxo_dag_get": (u32, [P, u32, u32, u32]),
            "vxo_dag_raycast": (None, [P, P, u32, P]),
            "vxo_dag_export": (u64, [P, P, u64]),
            "vxo_dag_stored_nodes": (u64, [P]),
            "vxo_dag_stored_leaves": (u64, [P]),
            "vxo_dense_export": (u64, [P, u32, C.c_int, P, u64]),
            "vxo_digest": (u64, [P, u64]),
            "vxo_svo_create": (P, [u32, u32]),
            "vxo_svo_destroy": (None, [P]),
            "vxo_svo_reset": (None, [P]),
            "vxo_svo_leaf_masks": (u32, [P, u32, P]),
            "vxo_svo_apply": (C.c_int, [P, P, u32, u32]),
            "vxo_svo_get": (u32, [P, u32, u32, u32]),
            "vxo_svo_counters": (None, [P, P]),
            "vxo_svo_export": (u64, [P, P, u64]),
            "vxo_svo_leaf_key": (u32, [u32, u32, u32]),
            "vxo_svo_random_edits": (None, [u32, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own CPU HashDAG, compiled from its sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        P = C.c_void_p
        u32, u64 = C.c_uint32, C.c_uint64
        sig = {
            "vxr_create": (P, [u32]),
            "vxr_destroy": (None, [P]),
            "vxr_box": (C.c_int, [P, u32, u32, u32, u32, u32, u32, C.c_int]),
            "vxr_boxes": (C.c_int, [P, P, u64]),
            "vxr_set_voxel": (C.c_int, [P, u32, u32, u32, C.c_int]),
            "vxr_get_voxel": (C.c_int, [P, u32, u32, u32]),
            "vxr_node_count": (u64, [P]),
            "vxr_leaf_count": (u64, [P]),
            "vxr_collect_garbage": (None, [P]),
            "vxr_export": (u64, [P, P, u64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- generators ---------------------------------------------------------------
def gen_u32(seed, stream, n, lo, hi):
    out = np.empty(n, np.uint32)
    lib().vxo_gen_u32(seed, stream, n, lo, hi, _ptr(out))
    return out


def gen_morton(seed, stream, n, bits):
    out = np.empty(n, np.uint64)
    lib().vxo_gen_morton(seed, stream, n, bits, _ptr(out))
    return out


def gen_unique_u64(seed, stream, n):
    out = np.empty(n, np.uint64)
    lib().vxo_gen_unique_u64(seed, stream, n, _ptr(out))
    return out


def gen_brushes(seed, stream, n, world, mix):
    out = np.zeros(n, BRUSH_DTYPE)
    lib().vxo_gen_brushes(seed, stream, n, world, mix, _ptr(out))
    return out


def brush_voxels(brushes: np.ndarray, world: int) -> int:
    b = np.ascontiguousarray(brushes, BRUSH_DTYPE)
    return sum(int(lib().vxo_brush_voxels(C.c_void_p(b.ctypes.data + i * 32), world)) for i in range(len(b)))


def terrain_height(t: Terrain, x, z):
    return lib().vxo_terrain_height(C.byref(t), x, z)


def terrain_heights(t: Terrain, xs, zs) -> np.ndarray:
    xs = np.ascontiguousarray(xs, np.uint32)
    zs = np.ascontiguousarray(zs, np.uint32)
    out = np.empty(xs.size, np.int32)
    lib().vxo_terrain_heights(C.byref(t), _ptr(xs), _ptr(zs), xs.size, _ptr(out))
    return out


# ---- primitives ---------------------------------------------------------------
def exclusive_sum(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().vxo_exclusive_sum_u32(_ptr(a), a.size, _ptr(out))
    return out


def inclusive_sum(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().vxo_inclusive_sum_u32(_ptr(a), a.size, _ptr(out))
    return out


def exclusive_max(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().vxo_exclusive_max_u32(_ptr(a), a.size, _ptr(out))
    return out


def inclusive_max(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib().vxo_inclusive_max_u32(_ptr(a), a.size, _ptr(out))
    return out


def _scan64(name, a):
    a = np.ascontiguousarray(a, np.uint64)
    out = np.empty_like(a)
    getattr(lib(), name)(_ptr(a), a.size, _ptr(out))
    return out


def exclusive_sum_u64(a):
    return _scan64("vxo_exclusive_sum_u64", a)


def inclusive_sum_u64(a):
    return _scan64("vxo_inclusive_sum_u64", a)


def exclusive_max_u64(a):
    return _scan64("vxo_exclusive_max_u64", a)


def inclusive_max_u64(a):
    return _scan64("vxo_inclusive_max_u64", a)


def sort_order(key_words: np.ndarray, key_bits: int) -> np.ndarray:
    """Stable order of keys given as an (n, W) u32 array (column 0 least
    significant) by their low key_bits bits."""
    k = np.ascontiguousarray(key_words, np.uint32)
    if k.ndim == 1:
        k = k.reshape(-1, 1)
    order = np.empty(k.shape[0], np.uint32)
    lib().vxo_sort_order(_ptr(k), k.shape[1], k.shape[0], key_bits, _ptr(order))
    return order


def sort_by_key(key_words: np.ndarray, arrays, key_bits: int):
    """sort_by_key: keys (n,) or (n, W) u32 words and any arrays (first axis
    n), permuted by the stable order; keys keep their shape."""
    order = sort_order(key_words, key_bits)
    k = np.ascontiguousarray(key_words, np.uint32)
    return k[order], [np.asarray(a)[order] for a in arrays]


def sort_pairs(keys, vals=None, key_bits=None):
    keys = np.array(keys, copy=True)
    if keys.dtype == np.uint32:
        f, kb = lib().vxo_sort_pairs_u32, 32
    elif keys.dtype == np.uint64:
        f, kb = lib().vxo_sort_pairs_u64, 64
    else:
        raise TypeError(keys.dtype)
    v = None if vals is None else np.array(vals, np.uint32, copy=True)
    f(_ptr(keys), None if v is None else _ptr(v), keys.size, kb if key_bits is None else key_bits)
    return keys, v


def set_contains(keys, queries):
    keys = np.ascontiguousarray(keys, np.uint64)
    queries = np.ascontiguousarray(queries, np.uint64)
    out = np.empty(queries.size, np.uint8)
    lib().vxo_set_contains_u64(_ptr(keys), keys.size, _ptr(queries), queries.size, _ptr(out))
    return out


def digest(words: np.ndarray) -> int:
    w = np.ascontiguousarray(words, np.uint32)
    return int(lib().vxo_digest(_ptr(w), w.size))


# ---- DAG ----------------------------------------------------------------------
class Dag:
    """CPU Hash DAG (restatement of svt/hash_dag.hpp, extended with materials)."""

    def __init__(self, inner_depth: int, materials: bool):
        self.h = lib().vxo_dag_create(inner_depth, int(materials))
        if not self.h:
            raise ValueError("bad depth")
        self.depth = inner_depth
        self.world = 4 << inner_depth

    def __del__(self):
        if getattr(self, "h", None):
            lib().vxo_dag_destroy(self.h)
            self.h = None

    def build_terrain(self, t: Terrain):
        lib().vxo_dag_build_terrain(self.h, C.byref(t))

    def apply_brushes(self, brushes):
        b = np.ascontiguousarray(brushes, BRUSH_DTYPE)
        lib().vxo_dag_apply_brushes(self.h, _ptr(b), len(b))

    def apply_voxels(self, xyz, material):
        a = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
        lib().vxo_dag_apply_voxels(self.h, _ptr(a), len(a), material)

    def box(self, lo, hi, material):
        lib().vxo_dag_box(self.h, *lo, *hi, material)

    def get(self, x, y, z) -> int:
        return int(lib().vxo_dag_get(self.h, x, y, z))

    def raycast(self, rays: np.ndarray) -> np.ndarray:
        """rays: (n, 8) int32 {ox, oy, oz, dx, dy, dz, max_steps, 0}; returns
        (n, 6) int32 {material, steps, x, y, z, 0}."""
        r = np.ascontiguousarray(rays, np.int32).reshape(-1, 8)
        out = np.zeros((len(r), 6), np.int32)
        lib().vxo_dag_raycast(self.h, _ptr(r), len(r), _ptr(out))
        return out

    def export(self) -> np.ndarray:
        n = lib().vxo_dag_export(self.h, None, 0)
        out = np.empty(n, np.uint32)
        lib().vxo_dag_export(self.h, _ptr(out), n)
        return out

    def stored(self):
        return int(lib().vxo_dag_stored_nodes(self.h)), int(lib().vxo_dag_stored_leaves(self.h))


def dense_export(grid: np.ndarray, inner_depth: int, materials: bool) -> np.ndarray:
    """Canonical export of a dense grid[z, y, x] of u8 materials (brute force)."""
    g = np.ascontiguousarray(grid, np.uint8)
    n = lib().vxo_dense_export(_ptr(g), inner_depth, int(materials), None, 0)
    out = np.empty(n, np.uint32)
    lib().vxo_dense_export(_ptr(g), inner_depth, int(materials), _ptr(out), n)
    return out


def terrain_column_boxes(terrain, world: int) -> np.ndarray:
    """The heightfield as one solid column box per (x, z): (x, 0, z) to
    (x + 1, min(world, h + 1), z + 1), z-major (the reference has no terrain
    builder; this is how its HashDAG is fed the config-0 scene)."""
    ops = np.empty((world * world, 7), np.uint32)
    k = 0
    for z in range(world):
        for x in range(world):
            h = terrain_height(terrain, x, z)
            ops[k] = (x, 0, z, x + 1, min(world, h + 1), z + 1, 1)
            k += 1
    return ops


def sphere_row_boxes(brushes, world: int) -> np.ndarray:
    """Sphere brushes as x-row boxes in brush order (fill = add), clipped to
    the world: for each (y, z) row, x in [cx - s, cx + s] with
    s = isqrt(r^2 - dy^2 - dz^2)."""
    import math
    ops = []
    for b in np.asarray(brushes):
        r, cx, cy, cz = int(b["hx"]), int(b["cx"]), int(b["cy"]), int(b["cz"])
        fill = 1 if int(b["material"]) else 0
        for dz in range(-r, r + 1):
            z = cz + dz
            if not 0 <= z < world:
                continue
            for dy in range(-r, r + 1):
                y = cy + dy
                rem = r * r - dy * dy - dz * dz
                if rem < 0 or not 0 <= y < world:
                    continue
                s = math.isqrt(rem)
                x0, x1 = max(0, cx - s), min(world, cx + s + 1)
                if x0 < x1:
                    ops.append((x0, y, z, x1, y + 1, z + 1, fill))
    return np.array(ops, np.uint32).reshape(-1, 7)


class RefDag:
    """The reference's own HashDAG (occupancy only), from oracle/_ref."""

    def __init__(self, inner_depth: int):
        self.h = ref().vxr_create(inner_depth)
        if not self.h:
            raise ValueError("bad depth")
        self.world = 4 << inner_depth

    def __del__(self):
        if getattr(self, "h", None):
            ref().vxr_destroy(self.h)
            self.h = None

    def box(self, lo, hi, fill: bool):
        rc = ref().vxr_box(self.h, *lo, *hi, int(fill))
        if rc:
            raise RuntimeError("reference box edit failed")

    def boxes(self, ops: np.ndarray):
        """(n, 7) u32 half-open boxes (x0 y0 z0 x1 y1 z1 fill), applied in order."""
        o = np.ascontiguousarray(ops, np.uint32).reshape(-1, 7)
        if ref().vxr_boxes(self.h, _ptr(o), o.shape[0]):
            raise RuntimeError("reference box edit failed")

    def set_voxel(self, x, y, z, value: bool):
        rc = ref().vxr_set_voxel(self.h, x, y, z, int(value))
        if rc:
            raise RuntimeError("reference set_voxel failed")

    def get(self, x, y, z) -> int:
        return int(ref().vxr_get_voxel(self.h, x, y, z))

    def counts(self):
        return int(ref().vxr_node_count(self.h)), int(ref().vxr_leaf_count(self.h))

    def collect_garbage(self):
        ref().vxr_collect_garbage(self.h)

    def export(self) -> np.ndarray:
        n = ref().vxr_export(self.h, None, 0)
        out = np.empty(n, np.uint32)
        ref().vxr_export(self.h, _ptr(out), n)
        return out


# ---- pointer SVO ----------------------------------------------------------------
SVO_WORLD = 4096
SVO_MAGIC = 0x5658534F


def svo_leaf_key(lx: int, ly: int, lz: int) -> int:
    return int(lib().vxo_svo_leaf_key(lx, ly, lz))


def svo_random_edits(count: int) -> np.ndarray:
    out = np.empty((count, 3), np.uint32)
    lib().vxo_svo_random_edits(count, _ptr(out))
    return out


def svo_leaf_masks(xyz) -> np.ndarray:
    """(n, 3) u32 voxels -> (m, 3) u32 {leaf_key, low, high}, sorted by key."""
    a = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
    out = np.empty((max(len(a), 1), 3), np.uint32)
    m = lib().vxo_svo_leaf_masks(_ptr(a), len(a), _ptr(out))
    return out[:m].copy()


class Svo:
    """CPU pointer SVO (restatement of algo::svt::cuda's voxel edit)."""

    def __init__(self, max_nodes: int = 1 << 20, max_leaves: int = 1 << 20):
        self.h = lib().vxo_svo_create(max_nodes, max_leaves)
        if not self.h:
            raise ValueError("bad capacity")

    def __del__(self):
        if getattr(self, "h", None):
            lib().vxo_svo_destroy(self.h)
            self.h = None

    def reset(self):
        lib().vxo_svo_reset(self.h)

    def apply(self, xyz, op: int) -> int:
        a = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
        return int(lib().vxo_svo_apply(self.h, _ptr(a), len(a), op))

    def place(self, xyz) -> int:
        return self.apply(xyz, 0)

    def destroy(self, xyz) -> int:
        return self.apply(xyz, 1)

    def get(self, x: int, y: int, z: int) -> int:
        return int(lib().vxo_svo_get(self.h, x, y, z))

    def get_many(self, xyz) -> np.ndarray:
        a = np.asarray(xyz, np.int64).reshape(-1, 3)
        return np.array([self.get(int(x) & 0xFFFFFFFF, int(y) & 0xFFFFFFFF, int(z) & 0xFFFFFFFF)
                         for x, y, z in a], np.uint32)

    def counters(self) -> tuple[int, int, int, int]:
        out = np.zeros(4, np.uint32)
        lib().vxo_svo_counters(self.h, _ptr(out))
        return tuple(int(v) for v in out)

    def export(self) -> np.ndarray:
        n = lib().vxo_svo_export(self.h, None, 0)
        out = np.empty(n, np.uint32)
        lib().vxo_svo_export(self.h, _ptr(out), n)
        return out
