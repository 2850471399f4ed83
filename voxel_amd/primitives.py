"""Host mirror of the reference's primitive API, over the C ABI.

Names and argument meaning follow ``algo::cuda::scan`` (cuda/scan/scan.cuh:79-161)
and ``algo::cuda::sort`` (cuda/sort/sort.cuh:100-167): operations are in place
on device buffers, take an optional caller workspace and a stream, and raise
``InvalidValue`` where the reference returns ``cudaErrorInvalidValue``.
Device buffers are torch CUDA tensors (plumbing for HIP memory and streams).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, lib


def _stream(stream):
    if stream is None:
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, torch.cuda.Stream):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(stream)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _workspace(nbytes: int, workspace, device):
    if nbytes == 0:
        return None, 0
    if workspace is None:
        workspace = torch.empty(nbytes, dtype=torch.uint8, device=device)
    return workspace, workspace.numel() * workspace.element_size()


# ---- scan ---------------------------------------------------------------------
def required_workspace_size(count: int, elem_bytes: int = 4) -> int:
    """Bytes of workspace the scan needs for `count` elements of 4 or 8 bytes
    (required_workspace_size<T>, cuda/scan/scan.cuh:153-161; 0 = none)."""
    n = C.c_size_t()
    if elem_bytes == 8:
        check(lib().vx_scan_workspace_size_u64(count, C.byref(n)), "vx_scan_workspace_size_u64")
    else:
        check(lib().vx_scan_workspace_size(count, C.byref(n)), "vx_scan_workspace_size")
    return n.value


def _scan(name, data: torch.Tensor, workspace, stream):
    if data.dtype in (torch.int32, torch.uint32):
        eb = 4
    elif data.dtype in (torch.int64, torch.uint64):
        eb = 8
    else:
        raise TypeError("scan supports 32- and 64-bit unsigned data")
    fn = getattr(lib(), f"vx_{name}_u{8 * eb}")
    n = data.numel()
    ws, wsb = _workspace(required_workspace_size(n, eb), workspace, data.device)
    check(fn(_ptr(data), n, _ptr(ws), wsb, _stream(stream)), fn.__name__)
    return data


def exclusive_sum(data, workspace=None, stream=None):
    return _scan("exclusive_sum", data, workspace, stream)


def inclusive_sum(data, workspace=None, stream=None):
    return _scan("inclusive_sum", data, workspace, stream)


def exclusive_max(data, workspace=None, stream=None):
    return _scan("exclusive_max", data, workspace, stream)


def inclusive_max(data, workspace=None, stream=None):
    return _scan("inclusive_max", data, workspace, stream)


def compact(values: torch.Tensor, flags: torch.Tensor, workspace=None, stream=None):
    """Keep values[i] where flags[i] != 0, in order. Returns (out, device count)."""
    n = values.numel()
    out = torch.empty_like(values)
    count = torch.zeros(1, dtype=torch.int32, device=values.device)
    ws, wsb = _workspace(required_workspace_size(n), workspace, values.device)
    check(lib().vx_compact_u32(_ptr(values), _ptr(flags), n, _ptr(out), _ptr(count), _ptr(ws), wsb, _stream(stream)),
          "vx_compact_u32")
    return out, count


# ---- sort ---------------------------------------------------------------------
def _key_info(keys: torch.Tensor):
    if keys.element_size() == 4:
        return 4, lib().vx_sort_pairs_u32
    if keys.element_size() == 8:
        return 8, lib().vx_sort_pairs_u64
    raise TypeError("sort keys must be 32- or 64-bit")


def required_sort_workspace_size(count: int, key_bytes: int, key_bits: int | None = None,
                                 with_values: bool = True) -> int:
    n = C.c_size_t()
    kb = key_bits if key_bits is not None else key_bytes * 8
    check(lib().vx_sort_workspace_size(count, key_bytes, kb, int(with_values), C.byref(n)), "vx_sort_workspace_size")
    return n.value


class ValueArray(C.Structure):
    _fields_ = [("data", C.c_void_p), ("elem_bytes", C.c_uint64)]


def _elem_bytes(t: torch.Tensor) -> int:
    """Bytes of one element of a value array: a row of an (n, ...) tensor."""
    k = 1
    for d in t.shape[1:]:
        k *= d
    return t.element_size() * k


def _key_layout(keys: torch.Tensor):
    """(key bytes, count): 1-D 32/64-bit keys, or (n, W) int32 words (UIntKey<W>)."""
    if keys.dim() == 1:
        if keys.element_size() not in (4, 8):
            raise TypeError("sort keys must be 32- or 64-bit")
        return keys.element_size(), keys.shape[0]
    if keys.dim() == 2 and keys.element_size() == 4:
        return 4 * keys.shape[1], keys.shape[0]
    raise TypeError("multi-word keys are (n, W) 32-bit tensors")


def _arrays(values, n):
    if values is None:
        return []
    vs = [values] if isinstance(values, torch.Tensor) else list(values)
    for v in vs:
        if v.shape[0] != n or not v.is_contiguous():
            raise ValueError("value arrays must be contiguous with one element per key")
    return vs


def required_sort_by_key_workspace_size(count: int, key_bytes: int, key_bits: int, elem_bytes=()) -> int:
    arr = (ValueArray * max(1, len(elem_bytes)))(*[ValueArray(None, e) for e in elem_bytes])
    n = C.c_size_t()
    check(lib().vx_sort_by_key_workspace_size(count, key_bytes, key_bits, arr, len(elem_bytes), C.byref(n)),
          "vx_sort_by_key_workspace_size")
    return n.value


def sort_by_key(keys: torch.Tensor, values=None, key_bits: int | None = None, workspace=None, stream=None):
    """Stable in-place sort of keys by their low key_bits bits, permuting every
    value array with them (sort_by_key<Key, Values...>, cuda/sort/sort.cuh:118-138).
    keys: 1-D u32 / u64 tensor, or (n, W) u32 words (UIntKey<W>, word 0 least
    significant). values: None, one tensor, or a sequence of tensors whose
    first axis is n (an (n, k) tensor is one array of k-element records)."""
    kbytes, n = _key_layout(keys)
    kb = key_bits if key_bits is not None else kbytes * 8
    vs = _arrays(values, n)
    if kbytes in (4, 8) and keys.dim() == 1 and (not vs or (len(vs) == 1 and _elem_bytes(vs[0]) == 4)):
        fn = lib().vx_sort_pairs_u32 if kbytes == 4 else lib().vx_sort_pairs_u64
        v = vs[0] if vs else None
        ws, wsb = _workspace(required_sort_workspace_size(n, kbytes, kb, v is not None), workspace, keys.device)
        check(fn(_ptr(keys), _ptr(v), n, kb, _ptr(ws), wsb, _stream(stream)), fn.__name__)
        return keys, values
    arr = (ValueArray * max(1, len(vs)))(*[ValueArray(v.data_ptr(), _elem_bytes(v)) for v in vs])
    need = required_sort_by_key_workspace_size(n, kbytes, kb, [_elem_bytes(v) for v in vs])
    ws, wsb = _workspace(need, workspace, keys.device)
    check(lib().vx_sort_by_key(_ptr(keys), kbytes, n, kb, arr, len(vs), _ptr(ws), wsb, _stream(stream)),
          "vx_sort_by_key")
    return keys, values


def sort_keys(keys, key_bits=None, workspace=None, stream=None):
    return sort_by_key(keys, None, key_bits, workspace, stream)[0]


SORT_KEYS_ONLY = 0x100  # or'ed into key_bytes: the tuning row of sorts without values


def set_sort_tuning(key_bytes: int, radix_bits: int, keys_per_thread: int, block: int) -> None:
    check(lib().vx_sort_set_tuning(key_bytes, radix_bits, keys_per_thread, block), "vx_sort_set_tuning")


def get_sort_tuning(key_bytes: int):
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(lib().vx_sort_get_tuning(key_bytes, C.byref(a), C.byref(b), C.byref(c)), "vx_sort_get_tuning")
    return a.value, b.value, c.value


def set_scan_tuning(block: int, vectors_per_lane: int) -> None:
    check(lib().vx_scan_set_tuning(block, vectors_per_lane), "vx_scan_set_tuning")


def get_scan_tuning():
    a, b = C.c_uint32(), C.c_uint32()
    check(lib().vx_scan_get_tuning(C.byref(a), C.byref(b)), "vx_scan_get_tuning")
    return a.value, b.value


# ---- generators ---------------------------------------------------------------
def gen_u32(seed, stream_id, n, lo, hi, device="cuda"):
    out = torch.empty(n, dtype=torch.int32, device=device)
    check(lib().vx_gen_u32(seed, stream_id, n, lo, hi, _ptr(out), _stream(None)), "vx_gen_u32")
    return out


def gen_morton(seed, stream_id, n, coord_bits, device="cuda"):
    out = torch.empty(n, dtype=torch.int64, device=device)
    check(lib().vx_gen_morton(seed, stream_id, n, coord_bits, _ptr(out), _stream(None)), "vx_gen_morton")
    return out


def gen_unique_u64(seed, stream_id, first, n, device="cuda"):
    out = torch.empty(n, dtype=torch.int64, device=device)
    check(lib().vx_gen_unique_u64(seed, stream_id, first, n, _ptr(out), _stream(None)), "vx_gen_unique_u64")
    return out


def iota(n, device="cuda"):
    out = torch.empty(n, dtype=torch.int32, device=device)
    check(lib().vx_iota_u32(_ptr(out), n, _stream(None)), "vx_iota_u32")
    return out
