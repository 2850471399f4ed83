"""Content-addressed set / map of multi-word items (AccelerationHashSet32 /
Map32) on the GPU: the reference's tests re-expressed
(test/cuda/test_cuda_acceleration_hash.cu:275-729) — empty table, large
items, calls from a caller's own kernel, chains beyond one bucket, overflow,
full-payload compares, reading items back, the host API, map values and
duplicate inserts — plus a randomized check against set truth at 2^20 items."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NOT_FOUND, INSERTED, OVERFLOW, FOUND = 0, 1, 4, 6
TEST_KERNELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libvx_test_kernels.so")


def fill_items(count, words, seed=0):
    """fill_items (test_cuda_acceleration_hash.cu:146-154)."""
    i = np.arange(count, dtype=np.uint64)[:, None]
    w = np.arange(words, dtype=np.uint64)[None, :]
    return (0x10000 + seed * 97 + i * 131 + w * 17).astype(np.uint32)


def _dev(a, gpu):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(gpu)


def _h(t):
    return t.cpu().numpy().view(np.uint32)


def _set(words, buckets, overflow):
    from voxel_amd.hash import AccelerationHashSet32
    return AccelerationHashSet32(words, buckets, overflow)


def _map(words, buckets, overflow):
    from voxel_amd.hash import AccelerationHashMap32
    return AccelerationHashMap32(words, buckets, overflow)


def test_empty_table_finds_nothing(gpu):
    t = _set(2, 8, 1)
    ptrs, st = t.find_batch(_dev(fill_items(4, 2), gpu))
    assert _h(st).tolist() == [NOT_FOUND] * 4 and _h(ptrs).tolist() == [0] * 4


def test_insert_and_find_large_items(gpu):
    t = _set(6, 8, 2)
    items = _dev(fill_items(16, 6), gpu)
    ip, st = t.insert_unique_unchecked_batch(items)
    assert (_h(st) == INSERTED).all() and (_h(ip) != 0).all()
    fp, st = t.find_batch(items)
    assert (_h(st) == FOUND).all() and np.array_equal(_h(fp), _h(ip))


def _kernels():
    L = C.CDLL(TEST_KERNELS)
    P, u32 = C.c_void_p, C.c_uint32
    L.vxt_custom_insert.argtypes = [P, P, P, u32, P, P]
    L.vxt_custom_find.argtypes = [P, P, u32, P, P, P]
    L.vxt_custom_read.argtypes = [P, P, u32, P, P]
    return L


@pytest.mark.parametrize("kind", ["set", "map"])
def test_device_api_from_custom_kernel(gpu, kind):
    """A caller's kernel (96-thread blocks: a partial wave) inserts and finds
    through the device API of include/vx_accel_device.hpp."""
    L = _kernels()
    words, n = (10, 64) if kind == "set" else (8, 64)
    t = _set(words, 16, 4) if kind == "set" else _map(words, 16, 4)
    v = t.view()
    items = _dev(fill_items(n, words, 0 if kind == "set" else 7), gpu)
    values = _dev(5000 + 3 * np.arange(n), gpu) if kind == "map" else None
    ip = torch.zeros(n, dtype=torch.int32, device=gpu)
    st = torch.zeros_like(ip)
    vp = C.c_void_p(values.data_ptr()) if values is not None else None
    assert L.vxt_custom_insert(C.byref(v), C.c_void_p(items.data_ptr()), vp, n, C.c_void_p(ip.data_ptr()),
                               C.c_void_p(st.data_ptr())) == 0
    assert (_h(st) == INSERTED).all()
    fp, fv = torch.zeros_like(ip), torch.zeros_like(ip)
    assert L.vxt_custom_find(C.byref(v), C.c_void_p(items.data_ptr()), n, C.c_void_p(fp.data_ptr()),
                             C.c_void_p(fv.data_ptr()), C.c_void_p(st.data_ptr())) == 0
    assert (_h(st) == FOUND).all() and np.array_equal(_h(fp), _h(ip))
    if kind == "map":
        assert np.array_equal(_h(fv), _h(values))
    out = torch.zeros((n, words), dtype=torch.int32, device=gpu)
    ok = torch.zeros_like(ip)
    assert L.vxt_custom_read(C.byref(v), C.c_void_p(fp.data_ptr()), n, C.c_void_p(out.data_ptr()),
                             C.c_void_p(ok.data_ptr())) == 0
    assert (_h(ok) == 1).all() and torch.equal(out, items)


def test_items_beyond_one_bucket(gpu):
    """GrowsCollisionChain: one bucket plus two overflow slabs hold 36 items
    (more than one bucket's 31)."""
    t = _set(2, 1, 2)
    items = _dev(fill_items(36, 2), gpu)
    ip, st = t.insert_unique_unchecked_batch(items)
    assert (_h(st) == INSERTED).all() and t.used() == 36
    fp, st = t.find_batch(items)
    assert (_h(st) == FOUND).all() and np.array_equal(_h(fp), _h(ip))


def test_reports_overflow_when_full(gpu):
    """One bucket, no overflow slab: 31 items fit, the 32nd overflows."""
    t = _set(2, 1, 0)
    ip, st = t.insert_unique_unchecked_batch(_dev(fill_items(32, 2), gpu))
    s = _h(st)
    assert (s == INSERTED).sum() == 31 and (s == OVERFLOW).sum() == 1
    assert _h(ip)[s == OVERFLOW].tolist() == [0]


def test_find_compares_full_payload(gpu):
    t = _set(2, 4, 1)
    _, st = t.insert_unique_unchecked_batch(_dev([[1, 2]], gpu))
    assert _h(st).tolist() == [INSERTED]
    ptrs, st = t.find_batch(_dev([[1, 3]], gpu))
    assert _h(st).tolist() == [NOT_FOUND] and _h(ptrs).tolist() == [0]


def test_reads_payload_from_returned_handle(gpu):
    t = _set(6, 4, 2)
    items = _dev(fill_items(8, 6), gpu)
    ptrs, _ = t.insert_unique_unchecked_batch(items)
    out, ok = t.read_items(ptrs)
    assert (_h(ok) == 1).all() and torch.equal(out, items)
    bad = torch.tensor([0, 9, 1 << 20], dtype=torch.int32, device=gpu)
    _, ok = t.read_items(bad)
    assert _h(ok).tolist() == [0, 0, 0]


def test_host_api_insert_and_find(gpu):
    t = _set(2, 4, 2)
    items = _dev(fill_items(5, 2), gpu)
    ip, st = t.insert_unique_unchecked_batch(items)
    assert (_h(st) == INSERTED).all()
    fp, st = t.find_batch(items)
    assert (_h(st) == FOUND).all() and np.array_equal(_h(fp), _h(ip))
    t.clear()
    _, st = t.find_batch(items)
    assert (_h(st) == NOT_FOUND).all() and t.used() == 0


def test_map_insert_and_find_values(gpu):
    t = _map(8, 8, 2)
    items = _dev(fill_items(16, 8), gpu)
    values = _dev(1000 + np.arange(16), gpu)
    ip, st = t.insert_unique_unchecked_batch(items, values)
    assert (_h(st) == INSERTED).all() and (_h(ip) != 0).all()
    fp, fv, st = t.find_batch(items)
    assert (_h(st) == FOUND).all() and np.array_equal(_h(fp), _h(ip)) and torch.equal(fv, values)
    out, rv, ok = t.read_items(fp)
    assert (_h(ok) == 1).all() and torch.equal(out, items) and torch.equal(rv, values)


def test_map_duplicate_insert_finds_one_representative(gpu):
    t = _map(8, 1, 1)
    item = 0xABC000 + 17 * np.arange(8)
    items = _dev(np.tile(item, (4, 1)), gpu)
    values = _dev(10 + np.arange(4), gpu)
    _, st = t.insert_unique_unchecked_batch(items, values)
    assert (_h(st) == INSERTED).all()
    fp, fv, st = t.find_batch(items)
    assert (_h(st) == FOUND).all()
    assert len(set(_h(fp).tolist())) == 1 and len(set(_h(fv).tolist())) == 1
    assert 10 <= _h(fv)[0] <= 13


@pytest.mark.parametrize("words", [1, 3, 9, 16])
def test_randomized_against_set_truth(gpu, words):
    """2^20 distinct random items (half inserted), queries of all of them:
    found exactly for the inserted half, handles read back to the items."""
    rng = np.random.default_rng(words)
    n = 1 << 20
    items = rng.integers(0, 1 << 32, (n, words), dtype=np.uint64).astype(np.uint32)
    items[:, 0] = (np.arange(n, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(np.uint32)  # distinct (odd factor)
    ins = items[: n // 2]
    t = _set(words, (n // 2 + 30) // 31, 0)
    ip, st = t.insert_unique_unchecked_batch(_dev(ins, gpu))
    assert (_h(st) == INSERTED).all()
    q = np.concatenate([items[n // 2:], ins])[rng.permutation(n)]
    fp, st = t.find_batch(_dev(q, gpu))
    hit = _h(st) == FOUND
    truth = np.isin(q[:, 0], ins[:, 0])
    assert np.array_equal(hit, truth)
    out, ok = t.read_items(fp)
    assert (_h(ok)[hit] == 1).all() and np.array_equal(_h(out)[hit].reshape(-1, words), q[hit])
