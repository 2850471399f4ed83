"""Host mirror of the reference's slab / cuckoo hash tables over the C ABI.

``SlabHashSetU64`` and ``SlabHashMapU32`` follow SlabHashSetU32 / SlabHashMapU32
(slab_hash/set/host.cuh:15-107, slab_hash/map/host.cuh:15-108): batch
insert / contains (find) / erase with per-key status codes. ``CuckooSetU64``
follows cuckoo_hash.cuh:73-120.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, lib, register
from .primitives import _ptr, _stream

NOT_FOUND, INSERTED, REPLACED, ERASED, OVERFLOW, INVALID_KEY, FOUND, ALREADY_PRESENT = range(8)

_P = C.c_void_p
_u32, _u64, _i32 = C.c_uint32, C.c_uint64, C.c_int32
register({
    "vx_slab_set_create": (_i32, [_u32, _u32, C.POINTER(_P)]),
    "vx_slab_set_destroy": (_i32, [_P]),
    "vx_slab_set_clear": (_i32, [_P, _P]),
    "vx_slab_set_insert": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_slab_set_contains": (_i32, [_P, _P, _u32, _P, _P, _P]),
    "vx_slab_set_erase": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_slab_set_slabs_used": (_i32, [_P, C.POINTER(_u32)]),
    "vx_slab_map_create": (_i32, [_u32, _u32, C.POINTER(_P)]),
    "vx_slab_map_destroy": (_i32, [_P]),
    "vx_slab_map_clear": (_i32, [_P, _P]),
    "vx_slab_map_insert_or_replace": (_i32, [_P, _P, _P, _u32, _P, _P]),
    "vx_slab_map_find": (_i32, [_P, _P, _P, _u32, _P, _P]),
    "vx_slab_map_erase": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_cuckoo_set_create": (_i32, [_u64, _u32, C.POINTER(_P)]),
    "vx_cuckoo_set_destroy": (_i32, [_P]),
    "vx_cuckoo_set_clear": (_i32, [_P, _P]),
    "vx_cuckoo_set_insert": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_cuckoo_set_contains": (_i32, [_P, _P, _u32, _P, _P]),
})


def slab_buckets_for(n_keys: int, load: float = 0.8, slots_per_slab: int = 63) -> int:
    """Bucket count giving `load` average occupancy of the bucket slabs."""
    return max(1, int(-(-n_keys // int(load * slots_per_slab))))


class _Owned:
    _destroy = ""

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            getattr(lib(), self._destroy)(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SlabHashSetU64(_Owned):
    _destroy = "vx_slab_set_destroy"

    def __init__(self, bucket_count: int, slab_capacity: int):
        self.h = C.c_void_p()
        check(lib().vx_slab_set_create(bucket_count, slab_capacity, C.byref(self.h)), "vx_slab_set_create")

    def clear(self, stream=None):
        check(lib().vx_slab_set_clear(self.h, _stream(stream)), "vx_slab_set_clear")

    def insert_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None):
        check(lib().vx_slab_set_insert(self.h, _ptr(keys), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_set_insert")

    def contains_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        found = torch.empty(keys.numel(), dtype=torch.uint8, device=keys.device)
        check(lib().vx_slab_set_contains(self.h, _ptr(keys), keys.numel(), _ptr(found), _ptr(statuses),
                                         _stream(stream)), "vx_slab_set_contains")
        return found

    def erase_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None):
        check(lib().vx_slab_set_erase(self.h, _ptr(keys), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_set_erase")

    def slabs_used(self) -> int:
        v = C.c_uint32()
        check(lib().vx_slab_set_slabs_used(self.h, C.byref(v)), "vx_slab_set_slabs_used")
        return v.value


class SlabHashMapU32(_Owned):
    _destroy = "vx_slab_map_destroy"

    def __init__(self, bucket_count: int, slab_capacity: int):
        self.h = C.c_void_p()
        check(lib().vx_slab_map_create(bucket_count, slab_capacity, C.byref(self.h)), "vx_slab_map_create")

    def clear(self, stream=None):
        check(lib().vx_slab_map_clear(self.h, _stream(stream)), "vx_slab_map_clear")

    def insert_or_replace_batch(self, keys, values, statuses=None, stream=None):
        check(lib().vx_slab_map_insert_or_replace(self.h, _ptr(keys), _ptr(values), keys.numel(), _ptr(statuses),
                                                  _stream(stream)), "vx_slab_map_insert_or_replace")

    def find_batch(self, keys, statuses=None, stream=None) -> torch.Tensor:
        vals = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
        check(lib().vx_slab_map_find(self.h, _ptr(keys), _ptr(vals), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_map_find")
        return vals

    def erase_batch(self, keys, statuses=None, stream=None):
        check(lib().vx_slab_map_erase(self.h, _ptr(keys), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_map_erase")


class CuckooSetU64(_Owned):
    _destroy = "vx_cuckoo_set_destroy"

    def __init__(self, slots: int, max_evictions: int = 0):
        self.h = C.c_void_p()
        check(lib().vx_cuckoo_set_create(slots, max_evictions, C.byref(self.h)), "vx_cuckoo_set_create")

    def clear(self, stream=None):
        check(lib().vx_cuckoo_set_clear(self.h, _stream(stream)), "vx_cuckoo_set_clear")

    def insert_batch(self, keys: torch.Tensor, stream=None) -> torch.Tensor:
        failed = torch.zeros(1, dtype=torch.int32, device=keys.device)
        check(lib().vx_cuckoo_set_insert(self.h, _ptr(keys), keys.numel(), _ptr(failed), _stream(stream)),
              "vx_cuckoo_set_insert")
        return failed

    def contains_batch(self, keys: torch.Tensor, stream=None) -> torch.Tensor:
        found = torch.empty(keys.numel(), dtype=torch.uint8, device=keys.device)
        check(lib().vx_cuckoo_set_contains(self.h, _ptr(keys), keys.numel(), _ptr(found), _stream(stream)),
              "vx_cuckoo_set_contains")
        return found
