"""Device-resident segment cache against the oracle's LRU model: the usage
queue, sizes, hit / miss counts and the cached data after every operation of
seeded store/load sequences; pins keep evicted segments readable."""
import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import ops
from flexpushdowndb_amd.device import DeviceColumn
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DTYPES = [np.int64, np.int32, np.float64, np.int16, np.int8]


def _segment(key, n):
    # deterministic content per key, so a re-store carries the same data
    seed = abs(hash(key)) % (2 ** 32)
    rng = np.random.default_rng(seed)
    dt = DTYPES[seed % len(DTYPES)]
    return rng.integers(-100, 100, n).astype(dt)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cache_matches_lru_model(ctx, seed):
    print(f"cache sequence seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    cap = 64 * 1024
    cache = ops.SegmentCache(ctx, cap)
    model = O.LruCacheModel(cap)
    sizes = {}
    for step in range(300):
        key = ("part%d" % rng.integers(3), "col%d" % rng.integers(5))
        if key not in sizes:
            sizes[key] = int(rng.integers(0, 4000))
        data = _segment(key, sizes[key])
        nbytes = data.nbytes
        if rng.random() < 0.45:
            on_device = rng.random() < 0.5
            src = DeviceColumn.from_numpy(ctx, data) if on_device else data
            assert cache.store(key[0], key[1], 0, nbytes, src) == model.store(key, nbytes)
        else:
            col = cache.load(key[0], key[1])
            assert (col is not None) == model.load(key)
            if col is not None:
                assert np.array_equal(col.numpy(), data)
                cache.release(col)
        info = cache.info()
        assert cache.layout() == model.layout()
        assert info["segments"] == len(model.queue)
        assert info["free"] == model.free and info["used"] == cap - model.free
        assert (info["hits"], info["misses"]) == (model.hits, model.misses)
    cache.close()


def test_cache_oversize_and_remove(ctx):
    cache = ops.SegmentCache(ctx, 1000)
    big = np.arange(200, dtype=np.int64)           # 1600 B > capacity
    assert not cache.store("p", "big", 0, 1, big)
    assert cache.load("p", "big") is None
    assert cache.store("p", "a", 0, 1, np.arange(50, dtype=np.int64))
    assert cache.store("p", "b", 0, 1, np.arange(50, dtype=np.int32))
    assert cache.store("q", "a", 0, 1, np.arange(10, dtype=np.int8))
    assert cache.remove("p") == 2
    assert cache.layout() == [("q", "a")]
    assert cache.info()["used"] == 10
    cache.clear()
    assert cache.info() == {"segments": 0, "used": 0, "free": 1000, "hits": 0, "misses": 0}
    cache.close()


def test_cache_pinned_segment_survives_eviction(ctx):
    cache = ops.SegmentCache(ctx, 8000)
    a = np.arange(900, dtype=np.int64)
    assert cache.store("p", "a", 0, 1, a)
    col = cache.load("p", "a")                      # pinned
    b = np.arange(900, dtype=np.int64) * 3
    assert cache.store("p", "b", 0, 1, b)           # evicts a from the cache
    assert cache.layout() == [("p", "b")]
    assert cache.load("p", "a") is None
    # the pinned buffers stay valid and unchanged while other work allocates
    filler = [DeviceColumn.from_numpy(ctx, np.full(900, 7, np.int64)) for _ in range(4)]
    assert np.array_equal(col.numpy(), a)
    used_pinned = ctx.memory()[0]
    cache.release(col)
    assert ctx.memory()[0] < used_pinned
    with pytest.raises(C.FpdbError):
        cache.release(col)                           # no longer pinned
    del filler
    cache.close()


def test_cache_validity_and_offsets(ctx):
    cache = ops.SegmentCache(ctx, 1 << 20)
    n = 1001
    rng = np.random.default_rng(4)
    vals = rng.integers(0, 1000, n + 13).astype(np.int32)
    valid = rng.random(n + 13) < 0.6
    # host source with a validity mask
    assert cache.store("p", "host", 0, 1, vals[:n], valid[:n])
    col = cache.load("p", "host")
    assert np.array_equal(col.valid_mask(), valid[:n])
    assert np.array_equal(col.numpy()[valid[:n]], vals[:n][valid[:n]])
    cache.release(col)
    # device source at an element offset of 13 (validity not byte aligned)
    dc = DeviceColumn.from_numpy(ctx, vals, valid)
    dc.offset, dc.length = 13, n
    assert cache.store("p", "dev", 0, 1, dc)
    col = cache.load("p", "dev")
    assert np.array_equal(col.valid_mask(), valid[13:])
    assert np.array_equal(col.numpy()[valid[13:]], vals[13:][valid[13:]])
    cache.release(col)
    cache.close()
