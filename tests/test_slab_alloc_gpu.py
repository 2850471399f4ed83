"""Reusable slab allocator (SlabAlloc) and the u32-key slab set
(SlabHashSetU32) on the GPU: the reference's tests re-expressed
(test/cuda/test_cuda_slab_alloc.cu:128-260, test/cuda/test_cuda_slab_hash_set.cu:129-380).
The slab sets here hold 63 keys per slab (a 512-byte slab per 64-lane wave)
where the reference's hold 31, so the chain and overflow cases use 63."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NOT_FOUND, INSERTED, ERASED, OVERFLOW, INVALID_KEY, FOUND, ALREADY_PRESENT = 0, 1, 3, 4, 5, 6, 7
KEYS_PER_SLAB = 63
NULL = 0xFFFFFFFF


def _h(t):
    return t.cpu().numpy().view(np.uint32)


def _dev(a, gpu):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(gpu)


@pytest.mark.parametrize("capacity", [0, 17, 1031])
def test_fresh_allocator_is_empty(gpu, capacity):
    from voxel_amd.hash import SlabAlloc
    a = SlabAlloc(capacity)
    assert a.used() == 0
    if capacity == 0:
        assert _h(a.allocate(3)).tolist() == [NULL] * 3


def test_unique_ordinals_until_exhausted(gpu):
    from voxel_amd.hash import SlabAlloc
    a = SlabAlloc(65)
    got = _h(a.allocate(65 + 17))
    ok = got[got != NULL]
    assert ok.size == 65 and np.unique(ok).size == 65 and ok.max() < 65
    assert (got == NULL).sum() == 17 and a.used() == 65


def test_allocates_beyond_one_word_and_block(gpu):
    """1024 + 3 ordinals all come back, unique, some past the first 1024."""
    from voxel_amd.hash import SlabAlloc
    a = SlabAlloc(1024 + 3)
    got = _h(a.allocate(1024 + 3))
    assert (got != NULL).all() and np.unique(got).size == got.size and got.max() >= 1024


def test_deallocation_allows_reuse_and_double_free_is_harmless(gpu):
    from voxel_amd.hash import SlabAlloc
    a = SlabAlloc(1)
    first = a.allocate(1)
    assert _h(first).tolist() == [0]
    assert _h(a.allocate(1)).tolist() == [NULL]
    a.deallocate(first)
    a.deallocate(first)
    assert _h(a.allocate(1)).tolist() == [0]


def test_out_of_range_deallocation_is_a_no_op(gpu):
    from voxel_amd.hash import SlabAlloc
    a = SlabAlloc(8)
    got = a.allocate(8)
    a.deallocate(_dev([8, 13, NULL], gpu))
    assert a.used() == 8 and _h(a.allocate(1)).tolist() == [NULL]
    a.deallocate(got[:3])
    assert a.used() == 5
    again = _h(a.allocate(4))
    assert sorted(again[again != NULL].tolist()) == sorted(_h(got[:3]).tolist()) and (again == NULL).sum() == 1
    a.clear()
    assert a.used() == 0


def test_many_rounds_of_allocate_and_free(gpu):
    """Random frees and re-allocations: every live ordinal unique, the count
    of live ordinals always the model's."""
    from voxel_amd.hash import SlabAlloc
    cap = 5000
    a = SlabAlloc(cap)
    rng = np.random.default_rng(3)
    live = set()
    for r in range(12):
        got = _h(a.allocate(int(rng.integers(100, 2000))))
        new = got[got != NULL].tolist()
        assert len(new) == len(set(new)) and not (set(new) & live)
        live |= set(new)
        assert len(new) == min(got.size, cap - (len(live) - len(new)))
        drop = rng.choice(sorted(live), size=len(live) // 3, replace=False) if live else []
        a.deallocate(_dev(np.array(drop, np.uint32), gpu))
        live -= set(int(x) for x in drop)
        assert a.used() == len(live)


def test_u32_set_insert_contains(gpu):
    from voxel_amd.hash import SlabHashSetU32
    t = SlabHashSetU32(8, 2)
    keys = _dev(10 + 7 * np.arange(4), gpu)
    st = torch.zeros(4, dtype=torch.int32, device=gpu)
    t.insert_batch(keys, st)
    assert (_h(st) == INSERTED).all()
    found = t.contains_batch(keys, st)
    assert (_h(st) == FOUND).all() and found.cpu().tolist() == [1] * 4
    found = t.contains_batch(_dev([9999], gpu), st[:1])
    assert _h(st)[0] == NOT_FOUND and found.cpu().tolist() == [0]


def test_u32_set_insert_contains_erase(gpu):
    from voxel_amd.hash import SlabHashSetU32
    t = SlabHashSetU32(4, 2)
    keys = _dev([7, 11, 13], gpu)
    st = torch.zeros(3, dtype=torch.int32, device=gpu)
    t.insert_batch(keys, st)
    assert (_h(st) == INSERTED).all()
    t.erase_batch(keys[:1], st[:1])
    assert _h(st)[0] == ERASED
    found = t.contains_batch(keys, st)
    assert _h(st).tolist() == [NOT_FOUND, FOUND, FOUND] and found.cpu().tolist() == [0, 1, 1]
    t.insert_batch(keys[:1], st[:1])  # the erased slot is taken again
    assert _h(st)[0] == INSERTED and t.slabs_used() == 4


def test_u32_set_grows_chain_and_clear_frees_slabs(gpu):
    """One bucket with two overflow slabs holds a chain beyond one slab; clear
    returns the slabs to the allocator and the keys go in again."""
    from voxel_amd.hash import SlabHashSetU32
    n = KEYS_PER_SLAB + 5
    t = SlabHashSetU32(1, 2)
    keys = _dev(np.arange(1, n + 1), gpu)
    st = torch.zeros(n, dtype=torch.int32, device=gpu)
    for round_ in range(2):
        t.insert_batch(keys, st)
        assert (_h(st) == INSERTED).all(), round_
        assert t.slabs_used() == 2
        t.contains_batch(keys, st)
        assert (_h(st) == FOUND).all()
        t.clear()
        assert t.slabs_used() == 1
        t.contains_batch(keys, st)
        assert (_h(st) == NOT_FOUND).all()


def test_u32_set_duplicates_and_reserved_keys(gpu):
    from voxel_amd.hash import SlabHashSetU32
    t = SlabHashSetU32(4, 2)
    st = torch.zeros(8, dtype=torch.int32, device=gpu)
    t.insert_batch(_dev([99] * 8, gpu), st)
    s = _h(st)
    assert (s == INSERTED).sum() == 1 and (s == ALREADY_PRESENT).sum() == 7
    t.insert_batch(_dev([0xFFFFFFFF, 0xFFFFFFFE], gpu), st[:2])
    assert _h(st)[:2].tolist() == [INVALID_KEY, INVALID_KEY]


def test_u32_set_reports_overflow_when_pool_is_full(gpu):
    from voxel_amd.hash import SlabHashSetU32
    t = SlabHashSetU32(1, 0)
    st = torch.zeros(KEYS_PER_SLAB + 1, dtype=torch.int32, device=gpu)
    t.insert_batch(_dev(np.arange(1, KEYS_PER_SLAB + 2), gpu), st)
    s = _h(st)
    assert (s == INSERTED).sum() == KEYS_PER_SLAB and (s == OVERFLOW).sum() == 1


def test_u32_set_against_truth(gpu):
    """2^20 distinct keys inserted at load 0.8, lookups with 50% hits."""
    from voxel_amd.hash import SlabHashSetU32, slab_buckets_for
    rng = np.random.default_rng(5)
    keys = rng.choice(np.uint32(0xFFFFFFF0), size=1 << 21, replace=False).astype(np.uint32)
    ins, other = keys[: 1 << 20], keys[1 << 20:]
    b = slab_buckets_for(ins.size)
    t = SlabHashSetU32(b, b // 2)
    st = torch.zeros(ins.size, dtype=torch.int32, device=gpu)
    t.insert_batch(_dev(ins, gpu), st)
    assert (_h(st) == INSERTED).all()
    q = np.concatenate([ins[::2], other[: ins.size // 2]])
    found = t.contains_batch(_dev(q, gpu)).cpu().numpy()
    assert np.array_equal(found.astype(bool), np.isin(q, ins))
