"""hbt_direct_potential_noalloc (the entry point that allocates nothing)
through the C-ABI against hbt_direct_potential on the same inputs: results
bit for bit in every mode, on two calls in a row, the second smaller than the
first (the scratch array keeps the first call's table); more groups than the
scratch holds and descending offsets are errors."""
import ctypes

import numpy as np
import pytest

from hbtplus_amd import abi

pytestmark = pytest.mark.gpu


def _inputs(sizes, seed):
    rng = np.random.default_rng(seed)
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(seg[-1])
    pos = (rng.normal(size=(n, 3)) * 30.0 + 5.0e4).astype(np.float32)
    m = rng.uniform(1e-5, 2e-4, n).astype(np.float32)
    return seg, pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), m


def _noalloc(U, seg, x, y, z, m, mode):
    t = U.torch
    d = lambda a: U._dev(np.ascontiguousarray(a))  # noqa: E731
    sd, xd, yd, zd, md = d(seg), d(x), d(y), d(z), d(m)
    out = t.full((max(len(x), 1),), np.nan, dtype=t.float64, device=f"cuda:{U.device}")
    t.cuda.synchronize(U.device)
    rc = U.lib.hbt_direct_potential_noalloc(U.ctx, len(seg) - 1, sd.data_ptr(), xd.data_ptr(), yd.data_ptr(),
                                            zd.data_ptr(), md.data_ptr(), mode, out.data_ptr())
    return rc, out.cpu().numpy()[:len(x)]


def test_noalloc_equals_direct_on_two_calls(gpu):
    U = gpu(abi.make_params(G=43007.1, softening=0.5, periodic=1, box_size=1.0e5))
    # the second call has fewer groups and particles than the first
    for sizes, seed in (([3000, 17, 1, 0, 513, 2000, 256, 257], 1), ([64, 5, 300], 2)):
        seg, x, y, z, m = _inputs(sizes, seed)
        for mode in (abi.HBT_DIRECT_FP32, abi.HBT_DIRECT_FP32_KAHAN, abi.HBT_DIRECT_FP64):
            want = U.direct_potential(seg, x, y, z, m, mode)
            rc, got = _noalloc(U, seg, x, y, z, m, mode)
            assert rc == abi.HBT_OK, U.lib.hbt_last_error()
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (sizes, mode)


def test_noalloc_limits(gpu):
    U = gpu(abi.make_params())
    seg, x, y, z, m = _inputs([1] * 65537, 3)
    rc, _ = _noalloc(U, seg, x, y, z, m, abi.HBT_DIRECT_FP64)
    assert rc == abi.HBT_ERR_CAPACITY and b"65536" in U.lib.hbt_last_error()
    _, x, y, z, m = _inputs([10, 20], 4)
    seg = np.array([0, 25, 20, 30], np.int64)  # the second group descends
    rc, _ = _noalloc(U, seg, x, y, z, m, abi.HBT_DIRECT_FP64)
    assert rc == abi.HBT_ERR_INVALID_ARG and b"ascending" in U.lib.hbt_last_error()
    # the entry point still works after the rejected calls
    seg, x, y, z, m = _inputs([100, 50], 5)
    rc, got = _noalloc(U, seg, x, y, z, m, abi.HBT_DIRECT_FP64)
    assert rc == abi.HBT_OK
    assert np.array_equal(got, U.direct_potential(seg, x, y, z, m, abi.HBT_DIRECT_FP64))
