"""u64 scans and the general sort_by_key (multi-word keys, several value
arrays of any element size) on the GPU vs the CPU oracle, bit-exact.

Fixed cases re-express the reference's tests: WarpShuffleHandlesUint64Input
and DefaultExclusiveMaxUint64MatchesReference (test/cuda/test_cuda_scan.cu:716-813);
the (u32 ids, u16 labels) value arrays, UIntKey<3> / UIntKey<16> keys and
u64 keys with values (test/cuda/test_cuda_sort.cu:380-440, 855-945,
1099-1170). The DAG's own call shape — u32 leaf keys with a u64 voxel-mask
payload (svt/hash_dag_gpu/detail/build_leaf_masks/pipeline.cuh:94-96) — runs
at 2^20 records.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, dev):
    a = np.ascontiguousarray(a)
    view = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[a.dtype.itemsize]
    return torch.from_numpy(a.view(view).copy()).to(dev)


def _host(t, dtype):
    return t.cpu().numpy().view(dtype)


ODD = np.arange(1, 34, 2, dtype=np.uint64)  # {1, 3, ..., 33}


@pytest.mark.parametrize("fn", ["exclusive_sum", "inclusive_sum", "exclusive_max", "inclusive_max"])
@pytest.mark.parametrize("n", [0, 1, 17, 2049, 4095, 4096, 4097, 8193, 100_003, 1 << 21])
def test_u64_scan_matches_oracle(gpu, fn, n):
    import voxel_amd.primitives as P
    if n == 17:
        a = ODD.copy()
    elif n == 2049:
        a = (np.arange(n, dtype=np.uint64) * 37) % 1021
    else:
        lo = O.gen_u32(31, O.STREAM_SCAN, n, 0, 0xFFFFFFFF).astype(np.uint64)
        hi = O.gen_u32(32, O.STREAM_SCAN, n, 0, 0xFFFFFFFF).astype(np.uint64)
        a = (hi << np.uint64(32)) | lo  # full-range values: sums wrap modulo 2^64
    d = _dev(a, gpu)
    getattr(P, fn)(d)
    assert np.array_equal(_host(d, np.uint64), getattr(O, fn + "_u64")(a))


def test_u64_scan_reference_known_answers(gpu):
    import voxel_amd.primitives as P
    d = _dev(ODD, gpu)
    P.inclusive_sum(d)
    assert _host(d, np.uint64).tolist() == [(k + 1) ** 2 for k in range(17)]
    a = (np.arange(2049, dtype=np.uint64) * 37) % 1021
    d = _dev(a, gpu)
    P.exclusive_max(d)
    got = _host(d, np.uint64)
    assert got[0] == 0 and got[28] == max(a[:28]) and got.max() == 1020


def test_u64_scan_unaligned_view(gpu):
    """An 8- but not 16-byte aligned array takes the element-wise path."""
    import voxel_amd.primitives as P
    a = (np.arange(50_001, dtype=np.uint64) * 0x9E3779B97F4A7C15) & np.uint64(0xFFFFFFFFFFFF)
    base = _dev(np.concatenate([[np.uint64(5)], a]), gpu)
    P.inclusive_sum(base[1:])
    got = _host(base, np.uint64)
    assert got[0] == 5 and np.array_equal(got[1:], O.inclusive_sum_u64(a))


def test_u64_scan_rejects_short_workspace(gpu):
    import voxel_amd.primitives as P
    from voxel_amd import InvalidValue
    n = 1 << 16
    need = P.required_workspace_size(n, 8)
    assert need > 0
    d = torch.zeros(n, dtype=torch.int64, device=gpu)
    with pytest.raises(InvalidValue):
        P.exclusive_sum(d, workspace=torch.empty(need - 1, dtype=torch.uint8, device=gpu))


def test_sort_ids_and_labels(gpu):
    """u32 keys with a u32 id array and a u16 label array, stable."""
    import voxel_amd.primitives as P
    keys = np.array([0, 42, 0, 42, 7, 7, 7, 1, 0, 999, 42, 3], np.uint32)
    ids = np.arange(100, 112, dtype=np.uint32)
    labels = np.array([9, 8, 7, 6, 5, 4, 3, 2, 1, 10, 11, 12], np.uint16)
    dk, di, dl = _dev(keys, gpu), _dev(ids, gpu), _dev(labels, gpu)
    P.sort_by_key(dk, [di, dl])
    ek, (ei, el) = O.sort_by_key(keys, [ids, labels], 32)
    assert np.array_equal(_host(dk, np.uint32), ek)
    assert np.array_equal(_host(di, np.uint32), ei) and np.array_equal(_host(dl, np.uint16), el)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ei, ids[order]) and np.array_equal(el, labels[order])


def test_sort_three_word_fixed_keys(gpu):
    import voxel_amd.primitives as P
    keys = np.array([[0, 0, 1], [3, 0, 0], [0, 2, 0], [0xFFFFFFFF] * 3, [0, 0, 0], [1, 0, 1]], np.uint32)
    vals = np.arange(6, dtype=np.uint32)
    dk, dv = _dev(keys, gpu), _dev(vals, gpu)
    P.sort_by_key(dk, dv)
    got = _host(dk, np.uint32).reshape(-1, 3)
    assert got.tolist() == [[0, 0, 0], [3, 0, 0], [0, 2, 0], [0, 0, 1], [1, 0, 1], [0xFFFFFFFF] * 3]
    assert _host(dv, np.uint32).tolist() == [4, 1, 2, 0, 5, 3]


@pytest.mark.parametrize("words,n,bits", [(3, 1025, 96), (16, 257, 512), (3, 70_001, 80), (5, 9000, 33)])
def test_sort_multiword_random_keys(gpu, words, n, bits):
    import voxel_amd.primitives as P
    keys = O.gen_u32(words * 1000 + n, O.STREAM_KEYS, n * words, 0, 0xFFFFFFFF).reshape(n, words)
    keys[::3, -1] = 7  # repeated high words
    keys[::5] = keys[0]  # exact duplicates: stability
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = _dev(keys, gpu), _dev(vals, gpu)
    P.sort_by_key(dk, dv, key_bits=bits)
    ek, (ev,) = O.sort_by_key(keys, [vals], bits)
    assert np.array_equal(_host(dk, np.uint32).reshape(n, words), ek)
    assert np.array_equal(_host(dv, np.uint32), ev)
    masked = [keys[:, w] & np.uint32(min(0xFFFFFFFF, (1 << max(0, min(32, bits - 32 * w))) - 1))
              for w in range(words)]
    assert np.array_equal(ev, np.lexsort([vals] + masked))  # numpy's lexsort: last key primary


def test_sort_u64_keys_with_values(gpu):
    import voxel_amd.primitives as P
    keys = np.array([0x100000000, 3, 0xFFFFFFFFFFFFFF, 0, 0x8000000000000000, 0x100000001], np.uint64)
    vals = np.arange(6, dtype=np.uint32)
    labels = np.arange(60, 0, -10, dtype=np.uint8)
    dk, dv, dl = _dev(keys, gpu), _dev(vals, gpu), _dev(labels, gpu)
    P.sort_by_key(dk, [dv, dl])
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(_host(dk, np.uint64), keys[order])
    assert np.array_equal(_host(dv, np.uint32), vals[order]) and np.array_equal(_host(dl, np.uint8), labels[order])


@pytest.mark.parametrize("n", [2, 257, 4097, 1 << 20])
def test_sort_leaf_keys_with_u64_masks(gpu, n):
    """The reference DAG's leaf-mask sort: u32 leaf keys (24 bits) with a u64
    one-hot voxel mask per record."""
    import voxel_amd.primitives as P
    keys = O.gen_u32(77, O.STREAM_KEYS, n, 0, (1 << 24) - 1)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # duplicated leaves
    bits = O.gen_u32(78, O.STREAM_VALUES, n, 0, 63).astype(np.uint64)
    masks = np.uint64(1) << bits
    dk, dm = _dev(keys, gpu), _dev(masks, gpu)
    P.sort_by_key(dk, dm, key_bits=24)
    ek, (em,) = O.sort_by_key(keys, [masks], 24)
    assert np.array_equal(_host(dk, np.uint32), ek) and np.array_equal(_host(dm, np.uint64), em)


@pytest.mark.parametrize("elem", [1, 2, 3, 6, 12, 16, 24, 36])
def test_sort_struct_payloads(gpu, elem):
    """Records of `elem` bytes (odd sizes take byte / half-word gathers)."""
    import voxel_amd.primitives as P
    n = 30_011
    keys = O.gen_u32(elem, O.STREAM_KEYS, n, 0, 4095)
    rec = O.gen_u32(elem + 100, O.STREAM_VALUES, n * elem, 0, 255).astype(np.uint8).reshape(n, elem)
    dk = _dev(keys, gpu)
    dr = torch.from_numpy(rec.copy()).to(gpu)
    P.sort_by_key(dk, dr, key_bits=12)
    order = np.argsort(keys & np.uint32(4095), kind="stable")
    assert np.array_equal(_host(dk, np.uint32), keys[order])
    assert np.array_equal(dr.cpu().numpy(), rec[order])


def test_sort_general_empty_single_and_invalid(gpu):
    import ctypes as C

    import voxel_amd.primitives as P
    from voxel_amd import InvalidValue
    from voxel_amd._lib import lib
    for n in (0, 1):
        k = torch.zeros((n, 3), dtype=torch.int32, device=gpu)
        v = torch.arange(n, dtype=torch.int16, device=gpu)
        P.sort_by_key(k, [v])
        assert P.required_sort_by_key_workspace_size(n, 12, 96, [2]) == 0
    k = torch.zeros((64, 3), dtype=torch.int32, device=gpu)
    need = P.required_sort_by_key_workspace_size(64, 12, 96, [4])
    arr = (P.ValueArray * 1)(P.ValueArray(None, 4))
    ws = torch.empty(need, dtype=torch.uint8, device=gpu)
    assert lib().vx_sort_by_key(C.c_void_p(k.data_ptr()), 12, 64, 96, arr, 1, C.c_void_p(ws.data_ptr()), need,
                                None) == 1  # NULL value array
    v = torch.zeros(64, dtype=torch.int32, device=gpu)
    arr = (P.ValueArray * 1)(P.ValueArray(v.data_ptr(), 4))
    assert lib().vx_sort_by_key(C.c_void_p(k.data_ptr()), 12, 64, 96, arr, 1, C.c_void_p(ws.data_ptr()), need - 1,
                                None) == 1  # short workspace
    assert lib().vx_sort_by_key(C.c_void_p(k.data_ptr()), 12, 64, 97, arr, 1, C.c_void_p(ws.data_ptr()), need,
                                None) == 1  # key_bits above the key
    assert lib().vx_sort_by_key(C.c_void_p(k.data_ptr()), 10, 64, 64, arr, 1, C.c_void_p(ws.data_ptr()), need,
                                None) == 1  # not whole words
    with pytest.raises(InvalidValue):
        P.sort_by_key(k, [v], workspace=torch.empty(need - 1, dtype=torch.uint8, device=gpu))
