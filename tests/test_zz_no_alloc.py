"""The no-allocation compaction entry point against fpdb_filter_compact:
identical outputs on two calls in a row (the second smaller, so rows of the
first stay in the shared scratch), nullable columns, the output-capacity
bound, the row maximum, and no growth of the context's allocations."""
import ctypes

import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd.device import DeviceColumn

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 28


def _noalloc(ctx, cols, n, bm, outs, d_cnt):
    C.check(C.lib().fpdb_filter_compact_noalloc(ctx.handle, C.columns_array([c.c() for c in cols]), len(cols), n,
                                                bm.ptr, C.columns_array([o.c() for o in outs]), d_cnt.ptr))


def _reference(ctx, cols, n, bm, outs):
    cnt = ctypes.c_int64()
    C.check(C.lib().fpdb_filter_compact(ctx.handle, C.columns_array([c.c() for c in cols]), len(cols), n, bm.ptr,
                                        C.columns_array([o.c() for o in outs]), ctypes.byref(cnt)))
    return cnt.value


@pytest.mark.parametrize("sizes", [(3_000_017, 1_000_003), (40_000_000, 9_001)])
def test_noalloc_matches_filter_compact(ctx, sizes):
    rng = np.random.default_rng(sizes[1])
    d_cnt = ctx.zeros(1, np.int64)
    for n in sizes:
        sel = rng.random(n) < 0.37
        a = rng.integers(-2 ** 62, 2 ** 62, n, dtype=np.int64)
        b = rng.integers(-1000, 1000, n).astype(np.int32)
        bvalid = rng.random(n) < 0.8
        cols = [DeviceColumn.from_numpy(ctx, a), DeviceColumn.from_numpy(ctx, b, bvalid)]
        bm = ctx.upload(np.packbits(sel, bitorder="little"))
        want = [DeviceColumn.empty(ctx, n, np.int64), DeviceColumn.empty(ctx, n, np.int32, nullable=True)]
        got = [DeviceColumn.empty(ctx, n, np.int64), DeviceColumn.empty(ctx, n, np.int32, nullable=True)]
        ref = _reference(ctx, cols, n, bm, want)
        ctx.synchronize()
        used_before = ctx.memory()[0]
        _noalloc(ctx, cols, n, bm, got, d_cnt)
        ctx.synchronize()
        assert ctx.memory()[0] == used_before          # nothing allocated through the context
        cnt = int(d_cnt.numpy()[0])
        assert cnt == ref == int(sel.sum())
        assert np.array_equal(got[0].numpy(cnt), want[0].numpy(cnt))
        assert np.array_equal(got[1].numpy(cnt)[got[1].valid_mask(cnt)], want[1].numpy(cnt)[want[1].valid_mask(cnt)])
        assert np.array_equal(got[1].valid_mask(cnt), want[1].valid_mask(cnt))
        assert np.array_equal(got[0].numpy(cnt), a[sel])


def test_noalloc_capacity_and_limits(ctx):
    n = 200_000
    rng = np.random.default_rng(3)
    sel = rng.random(n) < 0.5
    a = np.arange(n, dtype=np.int64)
    src = DeviceColumn.from_numpy(ctx, a)
    bm = ctx.upload(np.packbits(sel, bitorder="little"))
    need = int(sel.sum())
    buf = ctx.upload(np.full(need, -1, np.int64))
    short = DeviceColumn(buf, None, C.TYPE_INT64, need - 500)
    d_cnt = ctx.zeros(1, np.int64)
    _noalloc(ctx, [src], n, bm, [short], d_cnt)
    ctx.synchronize()
    assert int(d_cnt.numpy()[0]) == need
    got = buf.numpy()
    assert np.array_equal(got[:need - 500], a[sel][:need - 500]) and (got[need - 500:] == -1).all()
    # zero rows: the count is written, nothing else
    _noalloc(ctx, [src], 0, bm, [short], d_cnt)
    ctx.synchronize()
    assert int(d_cnt.numpy()[0]) == 0
    # beyond the scratch's maximum: rejected before anything runs
    big = DeviceColumn(src.values, None, C.TYPE_INT64, MAX_ROWS + 1)
    st = C.lib().fpdb_filter_compact_noalloc(ctx.handle, C.columns_array([big.c()]), 1, MAX_ROWS + 1, bm.ptr,
                                             C.columns_array([big.c()]), d_cnt.ptr)
    assert st == C.FPDB_ERR_INVALID and b"maximum" in C.lib().fpdb_last_error()
