"""Scan, compaction and onesweep sort on the GPU vs the CPU oracle, bit-exact.

Fixed vectors are the reference's own (test/cuda/test_cuda_scan.cu:252-308,
test/cuda/test_cuda_sort.cu:485-553, 1184-1198); random cases follow its
seeds/size pattern with the counter-based generator.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev_u32(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(dev)


def _host_u32(t):
    return t.cpu().numpy().view(np.uint32)


FIXED = [3, 1, 4, 1, 5, 9, 2, 6, 5]


@pytest.mark.parametrize("fn,ref", [("exclusive_sum", O.exclusive_sum), ("inclusive_sum", O.inclusive_sum),
                                    ("exclusive_max", O.exclusive_max), ("inclusive_max", O.inclusive_max)])
@pytest.mark.parametrize("n", [0, 1, 9, 1000, 4095, 4096, 4097, 65537, 1 << 20, 3_000_001])
def test_scan_matches_oracle(gpu, fn, ref, n):
    import voxel_amd.primitives as P
    a = np.array(FIXED, np.uint32) if n == 9 else O.gen_u32(123, O.STREAM_SCAN, n, 0, 9)
    d = _dev_u32(a, gpu)
    getattr(P, fn)(d)
    torch.cuda.synchronize()
    assert np.array_equal(_host_u32(d), ref(a))


def test_scan_fixed_vector_known_answer(gpu):
    import voxel_amd.primitives as P
    d = _dev_u32(FIXED, gpu)
    P.exclusive_sum(d)
    assert _host_u32(d).tolist() == [0, 3, 4, 8, 9, 14, 23, 25, 31]


def test_scan_wraps_like_u32(gpu):
    import voxel_amd.primitives as P
    a = np.full(100_000, 0xFFFF_FFF0, np.uint32)
    d = _dev_u32(a, gpu)
    P.inclusive_sum(d)
    assert np.array_equal(_host_u32(d), O.inclusive_sum(a))


def test_scan_rejects_short_workspace(gpu):
    import voxel_amd.primitives as P
    from voxel_amd import InvalidValue
    n = 1 << 16
    need = P.required_workspace_size(n)
    d = torch.zeros(n, dtype=torch.int32, device=gpu)
    ws = torch.empty(need - 1, dtype=torch.uint8, device=gpu)
    with pytest.raises(InvalidValue):
        P.exclusive_sum(d, workspace=ws)


def test_scan_unaligned_pointer_uses_generic_path(gpu):
    import voxel_amd.primitives as P
    a = O.gen_u32(5, O.STREAM_SCAN, 100_001, 0, 15)
    base = _dev_u32(np.concatenate([[7], a]), gpu)
    view = base[1:]
    P.exclusive_sum(view)
    assert np.array_equal(_host_u32(view), O.exclusive_sum(a))
    assert _host_u32(base)[0] == 7


@pytest.mark.parametrize("n", [1, 100, 5000, 1 << 20])
def test_compact_matches_oracle(gpu, n):
    import voxel_amd.primitives as P
    v = O.gen_u32(9, O.STREAM_VALUES, n, 0, 0xFFFFFFFF)
    f = O.gen_u32(9, O.STREAM_KEYS, n, 0, 2)
    out, cnt = P.compact(_dev_u32(v, gpu), _dev_u32(f, gpu))
    k = int(cnt.item())
    exp = v[f != 0]
    assert k == exp.size
    assert np.array_equal(_host_u32(out)[:k], exp)


def _sort_check(gpu, keys, key_bits=None, with_vals=True):
    import voxel_amd.primitives as P
    keys = np.asarray(keys)
    n = keys.size
    vals = np.arange(n, dtype=np.uint32)
    if keys.dtype == np.uint64:
        dk = torch.from_numpy(keys.view(np.int64).copy()).to(gpu)
    else:
        dk = _dev_u32(keys, gpu)
    dv = _dev_u32(vals, gpu) if with_vals else None
    P.sort_by_key(dk, dv, key_bits=key_bits)
    torch.cuda.synchronize()
    ek, ev = O.sort_pairs(keys, vals, key_bits)
    got = dk.cpu().numpy().view(keys.dtype)
    assert np.array_equal(got, ek)
    if with_vals:
        assert np.array_equal(_host_u32(dv), ev)


def test_sort_reference_fixed_vectors(gpu):
    _sort_check(gpu, np.array([3, 1, 4, 1, 5, 9, 2, 6, 5], np.uint32))
    _sort_check(gpu, np.array([9, 7, 5, 3, 1, 8, 6, 4, 2, 0], np.uint32))
    _sort_check(gpu, np.array([0, 42, 0, 42, 7, 7, 7, 1, 0, 999, 42, 3], np.uint32))


def test_sort_lower_bits_only(gpu):
    _sort_check(gpu, np.array([0x30, 0x21, 0x12, 0x03, 0x34, 0x25, 0x16, 0x07], np.uint32), key_bits=4)
    _sort_check(gpu, np.array([0x1ff, 0x281, 0x142, 0x003, 0x3c0, 0x224, 0x165, 0x026], np.uint32), key_bits=8)


def test_sort_stability_equal_keys(gpu):
    _sort_check(gpu, np.full(4097, 7, np.uint32))
    _sort_check(gpu, np.full(100_003, 7, np.uint64), key_bits=42)


@pytest.mark.parametrize("n", [2, 257, 1025, 4095, 4096, 4097, 100_000, 1 << 20])
@pytest.mark.parametrize("kind", ["u32_4096", "u32_full", "u64_morton42", "u64_full"])
def test_sort_random_matches_oracle(gpu, n, kind):
    if kind == "u32_4096":
        _sort_check(gpu, O.gen_u32(123, O.STREAM_KEYS, n, 0, 4095))
    elif kind == "u32_full":
        _sort_check(gpu, O.gen_u32(456, O.STREAM_KEYS, n, 0, 0xFFFFFFFF))
    elif kind == "u64_morton42":
        _sort_check(gpu, O.gen_morton(2, O.STREAM_KEYS, n, 14), key_bits=42)
    else:
        _sort_check(gpu, O.gen_unique_u64(1234, O.STREAM_KEYS, n), key_bits=64)


def test_sort_keys_only(gpu):
    _sort_check(gpu, O.gen_u32(321, O.STREAM_KEYS, 50_000, 0, 0xFFFFFFFF), with_vals=False)


@pytest.mark.parametrize("rb,kpt,block", [(4, 8, 256), (4, 15, 512), (8, 8, 1024), (8, 23, 256), (8, 15, 512),
                                          (9, 8, 1024), (9, 8, 512), (9, 15, 512), (9, 23, 512), (9, 15, 1024)])
def test_sort_tunings_agree(gpu, rb, kpt, block):
    """Every tile shape sorts like the oracle; a row whose tile does not fit
    LDS is left out."""
    import voxel_amd.primitives as P
    from voxel_amd import VxError
    from voxel_amd._lib import SORT_KEYS_ONLY
    rows = (4, 8, 4 | SORT_KEYS_ONLY, 8 | SORT_KEYS_ONLY)
    old = [P.get_sort_tuning(r) for r in rows]
    try:
        built = []
        for r in rows:
            try:
                P.set_sort_tuning(r, rb, kpt, block)
                built.append(r)
            except VxError:
                pass
        if not built:
            pytest.skip("tile does not fit LDS")
        if 4 in built:
            _sort_check(gpu, O.gen_u32(654, O.STREAM_KEYS, 300_001, 0, 0xFFFFFFFF))
        if 8 in built:
            _sort_check(gpu, O.gen_morton(987, O.STREAM_KEYS, 200_001, 14), key_bits=42)
            _sort_check(gpu, O.gen_unique_u64(989, O.STREAM_KEYS, 100_003), key_bits=64)
        if 4 | SORT_KEYS_ONLY in built:
            _sort_check(gpu, O.gen_u32(655, O.STREAM_KEYS, 300_001, 0, 0xFFFFFFFF), with_vals=False)
            _sort_check(gpu, O.gen_u32(656, O.STREAM_KEYS, 70_001, 0, 0xFFFFFFFF), key_bits=24, with_vals=False)
        if 8 | SORT_KEYS_ONLY in built:
            _sort_check(gpu, O.gen_morton(988, O.STREAM_KEYS, 200_001, 12), key_bits=36, with_vals=False)
            _sort_check(gpu, O.gen_morton(990, O.STREAM_KEYS, 50_001, 10), key_bits=30, with_vals=False)
    finally:
        for r, t in zip(rows, old):
            P.set_sort_tuning(r, *t)


def test_sort_rejects_bad_arguments(gpu):
    import voxel_amd.primitives as P
    from voxel_amd import InvalidValue
    k = torch.zeros(1000, dtype=torch.int32, device=gpu)
    with pytest.raises(InvalidValue):
        P.sort_by_key(k, None, key_bits=0)
    with pytest.raises(InvalidValue):
        P.sort_by_key(k, None, key_bits=33)
    need = P.required_sort_workspace_size(1000, 4, 32, False)
    with pytest.raises(InvalidValue):
        P.sort_by_key(k, None, workspace=torch.empty(need - 1, dtype=torch.uint8, device=gpu))


def test_device_generators_match_oracle(gpu):
    import voxel_amd.primitives as P
    assert np.array_equal(_host_u32(P.gen_u32(2, 1, 10_000, 0, 15)), O.gen_u32(2, 1, 10_000, 0, 15))
    assert np.array_equal(P.gen_morton(2, 1, 10_000, 14).cpu().numpy().view(np.uint64), O.gen_morton(2, 1, 10_000, 14))
    assert np.array_equal(P.gen_unique_u64(2, 4, 0, 10_000).cpu().numpy().view(np.uint64),
                          O.gen_unique_u64(2, 4, 10_000))


@pytest.mark.parametrize("block,vec", [(256, 4), (256, 8), (256, 16), (512, 4), (512, 8), (1024, 4)])
def test_scan_tile_shapes_agree(gpu, block, vec):
    import voxel_amd.primitives as P
    old = P.get_scan_tuning()
    try:
        P.set_scan_tuning(block, vec)
        for n in (5000, 70_001, 1 << 20):
            a = O.gen_u32(11, O.STREAM_SCAN, n, 0, 15)
            d = _dev_u32(a, gpu)
            P.inclusive_sum(d)
            assert np.array_equal(_host_u32(d), O.inclusive_sum(a))
    finally:
        P.set_scan_tuning(*old)


def test_sort_deep_lookback_many_tiles(gpu):
    """Enough tiles that predecessor walks cross several look-back windows."""
    _sort_check(gpu, O.gen_u32(77, O.STREAM_KEYS, 3_000_017, 0, 0xFFFFFFFF))
    _sort_check(gpu, O.gen_u32(78, O.STREAM_KEYS, 2_000_003, 0, 3))
