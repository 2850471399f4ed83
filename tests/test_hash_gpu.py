"""Slab set / slab map / cuckoo set on the GPU: results are exact against
ground truth by construction (inserted vs disjoint keys), plus the
reference's behaviours (test/cuda/test_cuda_slab_hash*.cu, test_cuda_cuckoo_hash.cu):
duplicates, erase, reserved keys, chain growth, overflow."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _d64(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint64).view(np.int64)).to(dev)


def _d32(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(dev)


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 200_000])
def test_slab_set_insert_contains_truth(gpu, n):
    from voxel_amd.hash import ALREADY_PRESENT, INSERTED, SlabHashSetU64, slab_buckets_for
    keys = O.gen_unique_u64(2, O.STREAM_HASH, 2 * n)
    ins, miss = keys[:n], keys[n:]
    s = SlabHashSetU64(slab_buckets_for(n), max(16, n // 20))
    st = torch.zeros(n, dtype=torch.int32, device=gpu)
    s.insert_batch(_d64(ins, gpu), st)
    assert (st.cpu().numpy() == INSERTED).all()
    q = np.concatenate([ins, miss])
    perm = np.random.default_rng(0).permutation(q.size)
    q = q[perm]
    found = s.contains_batch(_d64(q, gpu)).cpu().numpy()
    assert np.array_equal(found, O.set_contains(ins, q))
    s.insert_batch(_d64(ins[: min(n, 100)], gpu), st[: min(n, 100)])
    assert (st[: min(n, 100)].cpu().numpy() == ALREADY_PRESENT).all()


def test_slab_set_duplicates_in_one_batch(gpu):
    from voxel_amd.hash import ALREADY_PRESENT, INSERTED, SlabHashSetU64
    keys = np.array([5, 7, 5, 5, 9, 7], np.uint64)
    s = SlabHashSetU64(4, 4)
    st = torch.zeros(keys.size, dtype=torch.int32, device=gpu)
    s.insert_batch(_d64(keys, gpu), st)
    st = st.cpu().numpy()
    for k in (5, 7, 9):
        sel = st[keys == k]
        assert (sel == INSERTED).sum() == 1 and ((sel == INSERTED) | (sel == ALREADY_PRESENT)).all()


def test_slab_set_chain_growth_and_overflow(gpu):
    from voxel_amd.hash import INSERTED, OVERFLOW, SlabHashSetU64
    keys = O.gen_unique_u64(3, O.STREAM_HASH, 63 * 3)
    s = SlabHashSetU64(1, 2)  # one bucket, two overflow slabs = 189 slots
    st = torch.zeros(keys.size, dtype=torch.int32, device=gpu)
    s.insert_batch(_d64(keys, gpu), st)
    assert (st.cpu().numpy() == INSERTED).all()
    assert s.slabs_used() == 3
    extra = O.gen_unique_u64(4, O.STREAM_HASH, 10)
    st2 = torch.zeros(10, dtype=torch.int32, device=gpu)
    s.insert_batch(_d64(extra, gpu), st2)
    assert (st2.cpu().numpy() == OVERFLOW).all()
    assert s.contains_batch(_d64(keys, gpu)).cpu().numpy().all()


def test_slab_set_erase_and_reserved_keys(gpu):
    from voxel_amd.hash import ERASED, INVALID_KEY, NOT_FOUND, SlabHashSetU64
    keys = O.gen_unique_u64(5, O.STREAM_HASH, 5000)
    s = SlabHashSetU64(100, 100)
    s.insert_batch(_d64(keys, gpu))
    st = torch.zeros(2500, dtype=torch.int32, device=gpu)
    s.erase_batch(_d64(keys[:2500], gpu), st)
    assert (st.cpu().numpy() == ERASED).all()
    s.erase_batch(_d64(keys[:2500], gpu), st)
    assert (st.cpu().numpy() == NOT_FOUND).all()
    f = s.contains_batch(_d64(keys, gpu)).cpu().numpy()
    assert not f[:2500].any() and f[2500:].all()
    bad = np.array([2**64 - 1, 2**64 - 2], np.uint64)
    st = torch.zeros(2, dtype=torch.int32, device=gpu)
    s.insert_batch(_d64(bad, gpu), st)
    assert (st.cpu().numpy() == INVALID_KEY).all()


def test_slab_map_insert_replace_find_erase(gpu):
    from voxel_amd.hash import ERASED, FOUND, INSERTED, NOT_FOUND, REPLACED, SlabHashMapU32
    keys = np.unique(O.gen_u32(6, O.STREAM_HASH, 20_000, 0, 0xFFFFFFFD))
    vals = O.gen_u32(6, O.STREAM_VALUES, keys.size, 0, 0xFFFFFFFF)
    m = SlabHashMapU32(keys.size // 40 + 1, keys.size // 40 + 64)
    st = torch.zeros(keys.size, dtype=torch.int32, device=gpu)
    m.insert_or_replace_batch(_d32(keys, gpu), _d32(vals, gpu), st)
    assert (st.cpu().numpy() == INSERTED).all()
    got = m.find_batch(_d32(keys, gpu), st).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, vals) and (st.cpu().numpy() == FOUND).all()
    vals2 = vals ^ 0x5A5A5A5A
    m.insert_or_replace_batch(_d32(keys, gpu), _d32(vals2, gpu), st)
    assert (st.cpu().numpy() == REPLACED).all()
    assert np.array_equal(m.find_batch(_d32(keys, gpu)).cpu().numpy().view(np.uint32), vals2)
    half = keys.size // 2
    st_h = torch.zeros(half, dtype=torch.int32, device=gpu)
    m.erase_batch(_d32(keys[:half], gpu), st_h)
    assert (st_h.cpu().numpy() == ERASED).all()
    got = m.find_batch(_d32(keys, gpu), st).cpu().numpy().view(np.uint32)
    s = st.cpu().numpy()
    assert (s[:half] == NOT_FOUND).all() and (got[:half] == 0xFFFFFFFF).all()
    assert (s[half:] == FOUND).all() and np.array_equal(got[half:], vals2[half:])


@pytest.mark.parametrize("n,load", [(1000, 0.5), (300_000, 0.8)])
def test_cuckoo_set_truth(gpu, n, load):
    from voxel_amd.hash import CuckooSetU64
    keys = O.gen_unique_u64(7, O.STREAM_HASH, 2 * n)
    ins, miss = keys[:n], keys[n:]
    c = CuckooSetU64(int(n / load) + 1)
    failed = c.insert_batch(_d64(ins, gpu))
    assert int(failed.item()) == 0
    q = np.concatenate([ins, miss])
    assert np.array_equal(c.contains_batch(_d64(q, gpu)).cpu().numpy(), O.set_contains(ins, q))


def test_cuckoo_forced_eviction_chain(gpu):
    """A table barely larger than the key set forces long eviction chains
    (test_cuda_cuckoo_hash.cu forces one by colliding keys)."""
    from voxel_amd.hash import CuckooSetU64
    keys = O.gen_unique_u64(8, O.STREAM_HASH, 900)
    c = CuckooSetU64(1000, 4096)
    assert int(c.insert_batch(_d64(keys, gpu)).item()) == 0
    assert c.contains_batch(_d64(keys, gpu)).cpu().numpy().all()
