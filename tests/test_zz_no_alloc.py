"""The allocation-free exclusive sum (vx_exclusive_sum_u32_noalloc, through
the C ABI) against the workspace entry point vx_exclusive_sum_u32: bit-exact
on two calls in a row, the second smaller (its static look-back state still
holds the first call's records), on an unaligned view, and the size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _both(a, dev, offset=0):
    """(no-alloc result, workspace-entry result) for host array a."""
    from voxel_amd._lib import lib
    import voxel_amd.primitives as P
    n = a.size
    pad = np.concatenate([np.full(offset, 5, np.uint32), a])
    x = torch.from_numpy(pad.view(np.int32).copy()).to(dev)
    y = x.clone()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().vx_exclusive_sum_u32_noalloc(C.c_void_p(x[offset:].data_ptr()), n, s) == 0
    need = P.required_workspace_size(n)
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
    assert lib().vx_exclusive_sum_u32(C.c_void_p(y[offset:].data_ptr()), n, C.c_void_p(ws.data_ptr()), need, s) == 0
    torch.cuda.synchronize()
    xs, ys = x.cpu().numpy().view(np.uint32), y.cpu().numpy().view(np.uint32)
    assert (xs[:offset] == 5).all()
    return xs[offset:], ys[offset:]


def test_noalloc_matches_workspace_entry_twice_smaller_second(gpu):
    first = O.gen_u32(41, O.STREAM_SCAN, 3_000_001, 0, 15)
    got, want = _both(first, gpu)
    assert np.array_equal(got, want) and np.array_equal(got, O.exclusive_sum(first))
    second = O.gen_u32(42, O.STREAM_SCAN, 100_003, 0, 0xFFFFFFFF)
    got, want = _both(second, gpu)
    assert np.array_equal(got, want) and np.array_equal(got, O.exclusive_sum(second))


def test_noalloc_unaligned_and_single_tile(gpu):
    a = O.gen_u32(43, O.STREAM_SCAN, 50_001, 0, 9)
    got, want = _both(a, gpu, offset=1)
    assert np.array_equal(got, want)
    b = O.gen_u32(44, O.STREAM_SCAN, 1000, 0, 9)
    got, want = _both(b, gpu)
    assert np.array_equal(got, want)


def test_noalloc_limits(gpu):
    from voxel_amd._lib import lib
    x = torch.zeros(16, dtype=torch.int32, device=gpu)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().vx_exclusive_sum_u32_noalloc(C.c_void_p(x.data_ptr()), (1 << 28) + 1, s) == 1
    assert lib().vx_exclusive_sum_u32_noalloc(None, 5, s) == 1
    assert lib().vx_exclusive_sum_u32_noalloc(None, 0, s) == 0
