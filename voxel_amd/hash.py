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
    "vx_slab_set_u32_create": (_i32, [_u32, _u32, C.POINTER(_P)]),
    "vx_slab_set_u32_destroy": (_i32, [_P]),
    "vx_slab_set_u32_clear": (_i32, [_P, _P]),
    "vx_slab_set_u32_insert": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_slab_set_u32_contains": (_i32, [_P, _P, _u32, _P, _P, _P]),
    "vx_slab_set_u32_erase": (_i32, [_P, _P, _u32, _P, _P]),
    "vx_slab_set_u32_slabs_used": (_i32, [_P, C.POINTER(_u32)]),
    "vx_slab_alloc_create": (_i32, [_u32, C.POINTER(_P)]),
    "vx_slab_alloc_destroy": (_i32, [_P]),
    "vx_slab_alloc_clear": (_i32, [_P, _P]),
    "vx_slab_alloc_allocate": (_i32, [_P, _u32, _P, _P]),
    "vx_slab_alloc_deallocate": (_i32, [_P, _P, _u32, _P]),
    "vx_slab_alloc_used": (_i32, [_P, C.POINTER(_u32)]),
    "vx_slab_alloc_view_get": (_i32, [_P, _P]),
    "vx_accel_set_create": (_i32, [_u32, _u32, _u32, C.POINTER(_P)]),
    "vx_accel_set_destroy": (_i32, [_P]),
    "vx_accel_set_clear": (_i32, [_P, _P]),
    "vx_accel_set_find": (_i32, [_P, _P, _u32, _P, _P, _P]),
    "vx_accel_set_insert_unique_unchecked": (_i32, [_P, _P, _u32, _P, _P, _P]),
    "vx_accel_set_read": (_i32, [_P, _P, _u32, _P, _P, _P]),
    "vx_accel_set_view": (_i32, [_P, _P]),
    "vx_accel_set_used": (_i32, [_P, C.POINTER(_u32)]),
    "vx_accel_map_create": (_i32, [_u32, _u32, _u32, C.POINTER(_P)]),
    "vx_accel_map_destroy": (_i32, [_P]),
    "vx_accel_map_clear": (_i32, [_P, _P]),
    "vx_accel_map_find": (_i32, [_P, _P, _u32, _P, _P, _P, _P]),
    "vx_accel_map_insert_unique_unchecked": (_i32, [_P, _P, _P, _u32, _P, _P, _P]),
    "vx_accel_map_read": (_i32, [_P, _P, _u32, _P, _P, _P, _P]),
    "vx_accel_map_view": (_i32, [_P, _P]),
    "vx_accel_map_used": (_i32, [_P, C.POINTER(_u32)]),
})


class AccelView(C.Structure):
    """vx_accel_view: the device layout a caller's kernel works on."""
    _fields_ = [("slots", C.c_void_p), ("slot_mask", C.c_uint64), ("items", C.c_void_p), ("values", C.c_void_p),
                ("used", C.c_void_p), ("capacity", C.c_uint32), ("words", C.c_uint32)]


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


class SlabHashSetU32(_Owned):
    """SlabHashSetU32 (slab_hash/set/host.cuh:15-107): u32 keys, 0xffffffff and
    0xfffffffe reserved; overflow slabs from the slab allocator."""
    _destroy = "vx_slab_set_u32_destroy"

    def __init__(self, bucket_count: int, slab_capacity: int):
        self.h = C.c_void_p()
        check(lib().vx_slab_set_u32_create(bucket_count, slab_capacity, C.byref(self.h)), "vx_slab_set_u32_create")

    def clear(self, stream=None):
        check(lib().vx_slab_set_u32_clear(self.h, _stream(stream)), "vx_slab_set_u32_clear")

    def insert_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None):
        check(lib().vx_slab_set_u32_insert(self.h, _ptr(keys), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_set_u32_insert")

    def contains_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        found = torch.empty(keys.numel(), dtype=torch.uint8, device=keys.device)
        check(lib().vx_slab_set_u32_contains(self.h, _ptr(keys), keys.numel(), _ptr(found), _ptr(statuses),
                                             _stream(stream)), "vx_slab_set_u32_contains")
        return found

    def erase_batch(self, keys: torch.Tensor, statuses: torch.Tensor | None = None, stream=None):
        check(lib().vx_slab_set_u32_erase(self.h, _ptr(keys), keys.numel(), _ptr(statuses), _stream(stream)),
              "vx_slab_set_u32_erase")

    def slabs_used(self) -> int:
        v = C.c_uint32()
        check(lib().vx_slab_set_u32_slabs_used(self.h, C.byref(v)), "vx_slab_set_u32_slabs_used")
        return v.value


class SlabAlloc(_Owned):
    """SlabAlloc (cuda/hash_table/slab_alloc.cuh:26-251): reusable slab
    ordinals; allocate returns one per request (NULL_SLAB when none is
    left), deallocate frees them."""
    _destroy = "vx_slab_alloc_destroy"
    NULL_SLAB = 0xFFFFFFFF

    def __init__(self, capacity: int):
        self.h = C.c_void_p()
        check(lib().vx_slab_alloc_create(capacity, C.byref(self.h)), "vx_slab_alloc_create")

    def clear(self, stream=None):
        check(lib().vx_slab_alloc_clear(self.h, _stream(stream)), "vx_slab_alloc_clear")

    def allocate(self, n: int, device="cuda", stream=None) -> torch.Tensor:
        out = torch.empty(n, dtype=torch.int32, device=device)
        check(lib().vx_slab_alloc_allocate(self.h, n, _ptr(out), _stream(stream)), "vx_slab_alloc_allocate")
        return out

    def deallocate(self, ordinals: torch.Tensor, stream=None):
        check(lib().vx_slab_alloc_deallocate(self.h, _ptr(ordinals), ordinals.numel(), _stream(stream)),
              "vx_slab_alloc_deallocate")

    def used(self) -> int:
        v = C.c_uint32()
        check(lib().vx_slab_alloc_used(self.h, C.byref(v)), "vx_slab_alloc_used")
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


class AccelerationHashSet32(_Owned):
    """AccelerationHashSet32<W> (acceleration_hash/set/host.cuh:15-117): a
    content-addressed set of W-word items; find_batch /
    insert_unique_unchecked_batch / read_items with handles (0 = none) and
    status codes (FOUND / NOT_FOUND, INSERTED / OVERFLOW)."""
    _destroy = "vx_accel_set_destroy"
    _map = False

    def __init__(self, item_words: int, bucket_count: int, overflow_slab_capacity: int):
        self.h = C.c_void_p()
        self.item_words = item_words
        prefix = "vx_accel_map" if self._map else "vx_accel_set"
        self._p = prefix
        check(getattr(lib(), prefix + "_create")(item_words, bucket_count, overflow_slab_capacity, C.byref(self.h)),
              prefix + "_create")

    def clear(self, stream=None):
        check(getattr(lib(), self._p + "_clear")(self.h, _stream(stream)), self._p + "_clear")

    def view(self) -> AccelView:
        v = AccelView()
        check(getattr(lib(), self._p + "_view")(self.h, C.byref(v)), self._p + "_view")
        return v

    def used(self) -> int:
        n = C.c_uint32()
        check(getattr(lib(), self._p + "_used")(self.h, C.byref(n)), self._p + "_used")
        return n.value

    def _n(self, items: torch.Tensor) -> int:
        if items.dim() != 2 or items.shape[1] != self.item_words or items.element_size() != 4:
            raise ValueError("items are an (n, item_words) 32-bit tensor")
        return items.shape[0]

    def find_batch(self, items: torch.Tensor, stream=None):
        """(handles, status) of every item; a map also returns the values."""
        n = self._n(items)
        ptrs = torch.zeros(n, dtype=torch.int32, device=items.device)
        st = torch.zeros(n, dtype=torch.int32, device=items.device)
        if self._map:
            vals = torch.zeros(n, dtype=torch.int32, device=items.device)
            check(lib().vx_accel_map_find(self.h, _ptr(items), n, _ptr(ptrs), _ptr(vals), _ptr(st), _stream(stream)),
                  "vx_accel_map_find")
            return ptrs, vals, st
        check(lib().vx_accel_set_find(self.h, _ptr(items), n, _ptr(ptrs), _ptr(st), _stream(stream)),
              "vx_accel_set_find")
        return ptrs, st

    def read_items(self, ptrs: torch.Tensor, stream=None):
        """(items, ok) behind each handle; a map returns (items, values, ok)."""
        n = ptrs.numel()
        items = torch.zeros((n, self.item_words), dtype=torch.int32, device=ptrs.device)
        ok = torch.zeros(n, dtype=torch.int32, device=ptrs.device)
        if self._map:
            vals = torch.zeros(n, dtype=torch.int32, device=ptrs.device)
            check(lib().vx_accel_map_read(self.h, _ptr(ptrs), n, _ptr(items), _ptr(vals), _ptr(ok), _stream(stream)),
                  "vx_accel_map_read")
            return items, vals, ok
        check(lib().vx_accel_set_read(self.h, _ptr(ptrs), n, _ptr(items), _ptr(ok), _stream(stream)),
              "vx_accel_set_read")
        return items, ok

    def insert_unique_unchecked_batch(self, items: torch.Tensor, stream=None):
        n = self._n(items)
        ptrs = torch.zeros(n, dtype=torch.int32, device=items.device)
        st = torch.zeros(n, dtype=torch.int32, device=items.device)
        check(lib().vx_accel_set_insert_unique_unchecked(self.h, _ptr(items), n, _ptr(ptrs), _ptr(st),
                                                         _stream(stream)), "vx_accel_set_insert_unique_unchecked")
        return ptrs, st


class AccelerationHashMap32(AccelerationHashSet32):
    """AccelerationHashMap32<W> (acceleration_hash/map/host.cuh:15-116): the
    set with one u32 value per stored item."""
    _destroy = "vx_accel_map_destroy"
    _map = True

    def insert_unique_unchecked_batch(self, items: torch.Tensor, values: torch.Tensor, stream=None):
        n = self._n(items)
        if values.numel() != n:
            raise ValueError("one value per item")
        ptrs = torch.zeros(n, dtype=torch.int32, device=items.device)
        st = torch.zeros(n, dtype=torch.int32, device=items.device)
        check(lib().vx_accel_map_insert_unique_unchecked(self.h, _ptr(items), _ptr(values), n, _ptr(ptrs), _ptr(st),
                                                         _stream(stream)), "vx_accel_map_insert_unique_unchecked")
        return ptrs, st
