"""mbg_sa_batch_dev_noalloc (include/mbg.h): SA locate whose only scratch, the kernel's work-queue
counter, lives in static device memory, so the call allocates nothing.  Its output must equal the
existing device entry point's (mbg_sa_batch_dev, same kernel with the context's pool) and the host
entry point's (mbg_sa_batch, pinned to the reference's mb_bwt_sa_batch by the SA tests), bit for bit,
on two calls in a row, the second smaller, so that a stale counter from the first would show."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_sa_noalloc_equals_sa_batch_dev():
    from minibwa_amd import mbg
    with mbg.Context(0) as ctx:
        ctx.idx_load(GOLDEN + "/chrM", 10)
        n_rows = int(ctx.info()["seq_len"]) + 1
        rng = np.random.default_rng(7)
        for n in (250000, 40000):
            print(f"[no_alloc] call with {n} rows", flush=True)
            rows = rng.integers(0, n_rows, n, dtype=np.uint64)
            rows[:3] = [0, n_rows - 1, ctx.info()["primary"]]
            d_a, d_b = ctx.to_device(rows), ctx.to_device(rows)
            mbg.check(ctx.L.mbg_sa_batch_dev(ctx.h, None, n, C.c_void_p(d_a.ptr)), "mbg_sa_batch_dev")
            mbg.check(ctx.L.mbg_sa_batch_dev_noalloc(ctx.h, None, n, C.c_void_p(d_b.ptr)), "mbg_sa_batch_dev_noalloc")
            a = ctx.to_host(d_a.ptr, np.uint64, n)
            b = ctx.to_host(d_b.ptr, np.uint64, n)
            assert np.array_equal(a, b)
            assert np.array_equal(b, ctx.bwt_sa_batch(rows))


def test_sa_noalloc_limits():
    """Sizes past the stated maximum are refused before any device work; a row outside the index comes
    back as 2^64-1 from both device entry points instead of being walked."""
    from minibwa_amd import mbg
    with mbg.Context(0) as ctx:
        ctx.idx_load(GOLDEN + "/chrM", 10)
        small = ctx.to_device(np.zeros(16, np.uint64))       # never touched: the size is refused first
        rc = ctx.L.mbg_sa_batch_dev_noalloc(ctx.h, None, (1 << 31) + 1, C.c_void_p(small.ptr))
        assert rc == -7                                      # MBG_EUNSUPPORTED
        seq_len = int(ctx.info()["seq_len"])
        rows = np.array([5, seq_len + 1, 17, seq_len + 1000], np.uint64)
        for fn in (ctx.L.mbg_sa_batch_dev, ctx.L.mbg_sa_batch_dev_noalloc):
            d = ctx.to_device(rows)
            mbg.check(fn(ctx.h, None, len(rows), C.c_void_p(d.ptr)), "sa")
            got = ctx.to_host(d.ptr, np.uint64, len(rows))
            assert got[1] == got[3] == np.uint64(2 ** 64 - 1)
            assert np.array_equal(got[[0, 2]], ctx.bwt_sa_batch(rows[[0, 2]]))
