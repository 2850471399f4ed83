"""ctypes binding of the gfx950 C ABI (include/vx.h).

The shared library must have been built (``python voxel_amd/build.py``);
loading fails loudly otherwise — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoxel_amd.so")
HEADER = os.path.join(HERE, "..", "include", "vx.h")

VX_OK = 0
VX_INVALID_VALUE = 1
VX_OUT_OF_CAPACITY = 2
VX_DEVICE_ERROR = 3
VX_UNSUPPORTED = 4

SORT_KEYS_ONLY = 0x100  # VX_SORT_KEYS_ONLY: the keys-only row of the sort tuning table

STATUS_NAMES = {0: "VX_OK", 1: "VX_INVALID_VALUE", 2: "VX_OUT_OF_CAPACITY", 3: "VX_DEVICE_ERROR",
                4: "VX_UNSUPPORTED"}


class VxError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")
        self.status = status


class InvalidValue(VxError, ValueError):
    pass


class OutOfCapacity(VxError, MemoryError):
    pass


def check(status: int, what: str) -> None:
    if status == VX_OK:
        return
    if status == VX_INVALID_VALUE:
        raise InvalidValue(status, what)
    if status == VX_OUT_OF_CAPACITY:
        raise OutOfCapacity(status, what)
    raise VxError(status, what)


_P = C.c_void_p
_u32, _u64, _i32, _i64, _sz = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_size_t
_pu32, _pu64, _psz = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)

SIGNATURES = {
    "vx_version": (C.c_char_p, []),
    "vx_device_info": (_i32, [C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i64)]),
    "vx_scan_workspace_size": (_i32, [_u32, _psz]),
    "vx_exclusive_sum_u32": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_inclusive_sum_u32": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_exclusive_max_u32": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_inclusive_max_u32": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_compact_u32": (_i32, [_P, _P, _u32, _P, _P, _P, _sz, _P]),
    "vx_scan_set_tuning": (_i32, [_u32, _u32]),
    "vx_scan_get_tuning": (_i32, [_pu32, _pu32]),
    "vx_sort_workspace_size": (_i32, [_u32, _u32, _u32, _i32, _psz]),
    "vx_sort_pairs_u32": (_i32, [_P, _P, _u32, _u32, _P, _sz, _P]),
    "vx_sort_pairs_u64": (_i32, [_P, _P, _u32, _u32, _P, _sz, _P]),
    "vx_sort_set_tuning": (_i32, [_u32, _u32, _u32, _u32]),
    "vx_sort_get_tuning": (_i32, [_u32, _pu32, _pu32, _pu32]),
    "vx_gen_u32": (_i32, [_u64, _u64, _u32, _u32, _u32, _P, _P]),
    "vx_gen_morton": (_i32, [_u64, _u64, _u32, _u32, _P, _P]),
    "vx_gen_unique_u64": (_i32, [_u64, _u64, _u64, _u32, _P, _P]),
    "vx_iota_u32": (_i32, [_P, _u32, _P]),
    "vx_exclusive_sum_u32_noalloc": (_i32, [_P, _u32, _P]),
    "vx_scan_workspace_size_u64": (_i32, [_u32, _psz]),
    "vx_exclusive_sum_u64": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_inclusive_sum_u64": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_exclusive_max_u64": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_inclusive_max_u64": (_i32, [_P, _u32, _P, _sz, _P]),
    "vx_sort_by_key_workspace_size": (_i32, [_u32, _u32, _u32, _P, _u32, _psz]),
    "vx_sort_by_key": (_i32, [_P, _u32, _u32, _u32, _P, _u32, _P, _sz, _P]),
}

_lib = None


def header_functions() -> list[str]:
    """Names of every entry point declared in include/vx.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"voxel_amd native library missing: {LIB_PATH} (run voxel_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        _apply_tuning(L)
    return _lib


def _apply_tuning(L) -> None:
    """Apply the autotuner's table (voxel_amd/tuning.json) if one was written."""
    import json
    path = os.path.join(HERE, "tuning.json")
    if not os.path.exists(path):
        return
    with open(path) as f:
        t = json.load(f)
    for key, kb in (("sort_u32", 4), ("sort_u64", 8), ("sort_u64_keys", 8 | SORT_KEYS_ONLY)):
        if key in t:
            L.vx_sort_set_tuning(kb, *t[key])
    if "scan" in t:
        L.vx_scan_set_tuning(*t["scan"])


def register(extra: dict) -> None:
    """Attach signatures of further entry points (used by submodules)."""
    SIGNATURES.update(extra)
    if _lib is not None:
        for name, (res, args) in extra.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
