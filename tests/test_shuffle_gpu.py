"""Hash shuffle on the GPU against the oracle: partition offsets and every
output column (values and validity) must match the stable partition order of
ShuffleKernel2::shuffle exactly."""
import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import ops
from flexpushdowndb_amd.device import DeviceColumn
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _valid_bits(mask):
    return np.packbits(np.asarray(mask, bool), bitorder="little")


def _table(n, seed, key_kind):
    rng = np.random.default_rng(seed)
    k64 = rng.integers(-2 ** 40, 2 ** 40, n, dtype=np.int64)
    k32 = rng.integers(-1000, 1000, n, dtype=np.int64).astype(np.int32)
    f64 = np.round(rng.uniform(900, 105000, n), 2)
    i8 = rng.integers(-128, 128, n, dtype=np.int64).astype(np.int8)
    i16 = rng.integers(-2 ** 15, 2 ** 15, n, dtype=np.int64).astype(np.int16)
    nul = rng.integers(0, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    nul_valid = rng.random(n) < 0.7
    key_valid = rng.random(n) < 0.8
    host = {"k64": (k64, None), "k32": (k32, None), "f64": (f64, None), "i8": (i8, None),
            "i16": (i16, None), "nul": (nul, nul_valid)}
    if key_kind == "int64":
        keys = [("k64", None)]
    elif key_kind == "int32":
        keys = [("k32", None)]
    elif key_kind == "two":
        keys = [("k64", None), ("k32", None)]
    else:   # nullable int64 key
        keys = [("k64", key_valid)]
    return host, keys


def _check(ctx, n, num_slots, key_kind, seed=1):
    host, keys = _table(n, seed, key_kind)
    dev = {name: DeviceColumn.from_numpy(ctx, a, v) for name, (a, v) in host.items()}
    key_dev = [DeviceColumn.from_numpy(ctx, host[name][0], kv) if kv is not None else dev[name]
               for name, kv in keys]
    names = list(host)
    out, offsets = ops.shuffle(ctx, key_dev, [dev[k] for k in names], n, num_slots)
    hashes = O.arrow6_hash([(host[name][0], None if kv is None else _valid_bits(kv)) for name, kv in keys])
    order, want_offsets = O.shuffle_order(hashes, num_slots)
    assert np.array_equal(offsets, want_offsets)
    for name, o in zip(names, out):
        vals, valid = host[name]
        got = o.numpy(n)
        if valid is None:
            assert np.array_equal(got, vals[order]), name
        else:
            want_valid = valid[order]
            assert np.array_equal(o.valid_mask(n), want_valid), name
            assert np.array_equal(got[want_valid], vals[order][want_valid]), name


@pytest.mark.parametrize("num_slots", [1, 2, 3, 64, 257, 4096])
@pytest.mark.parametrize("n", [0, 1, 100, 4095, 4096, 4097, 70_000])
def test_shuffle_int64_key(ctx, n, num_slots):
    print(f"shuffle int64 key n={n} slots={num_slots}", flush=True)
    _check(ctx, n, num_slots, "int64")


@pytest.mark.parametrize("key_kind", ["int32", "two", "nullable"])
@pytest.mark.parametrize("num_slots", [5, 64])
def test_shuffle_key_kinds(ctx, key_kind, num_slots):
    print(f"shuffle {key_kind} keys slots={num_slots}", flush=True)
    _check(ctx, 50_001, num_slots, key_kind, seed=num_slots)


def test_shuffle_large(ctx):
    # 2M rows, 64 slots: the partition sizes of a uniform hash
    print("shuffle 2M rows 64 slots", flush=True)
    _check(ctx, 2_000_000, 64, "int64", seed=9)


def test_shuffle_input_offsets(ctx):
    # columns that start at an element offset (Arrow slices), validity included
    n, off = 10_000, 37
    rng = np.random.default_rng(3)
    k = rng.integers(0, 10 ** 9, n + off, dtype=np.int64)
    v = rng.integers(0, 100, n + off, dtype=np.int64).astype(np.int32)
    valid = rng.random(n + off) < 0.5
    kc = DeviceColumn.from_numpy(ctx, k)
    kc.offset, kc.length = off, n
    vc = DeviceColumn.from_numpy(ctx, v, valid)
    vc.offset, vc.length = off, n
    out, offsets = ops.shuffle(ctx, [kc], [kc, vc], n, 7)
    order, want = O.shuffle_order(O.arrow6_hash([(k[off:], None)]), 7)
    assert np.array_equal(offsets, want)
    assert np.array_equal(out[0].numpy(n), k[off:][order])
    wv = valid[off:][order]
    assert np.array_equal(out[1].valid_mask(n), wv)
    assert np.array_equal(out[1].numpy(n)[wv], v[off:][order][wv])


def test_shuffle_rejects_bad_arguments(ctx):
    k = DeviceColumn.from_numpy(ctx, np.arange(10, dtype=np.int64))
    for bad in (0, 4097):
        with pytest.raises(C.FpdbError) as e:
            ops.shuffle(ctx, [k], [k], 10, bad)
        assert e.value.code == C.FPDB_ERR_INVALID
    # a nullable input needs an output validity bitmap
    kv = DeviceColumn.from_numpy(ctx, np.arange(10, dtype=np.int64), np.ones(10, bool))
    out = [DeviceColumn.empty(ctx, 10, np.int64)]
    offs = np.zeros(3, np.int64)
    st = C.lib().fpdb_shuffle(ctx.handle, C.columns_array([k.c()]), 1, C.columns_array([kv.c()]), 1, 10, 2,
                              C.columns_array([o.c() for o in out]), None, offs.ctypes.data_as(C._PI64))
    assert st == C.FPDB_ERR_INVALID


def test_shuffle_many_slots_long_input(ctx):
    # 4096 slots over 17M rows: the per-(slot, tile) bookkeeping would pass
    # its bound, so workgroups take runs of several tiles; same output
    n, slots = 17_000_000, 4096
    print(f"shuffle {n} rows {slots} slots", flush=True)
    rng = np.random.default_rng(21)
    k = rng.integers(0, 2 ** 62, n, dtype=np.int64)
    rid = np.arange(n, dtype=np.int32)
    kc, rc = DeviceColumn.from_numpy(ctx, k), DeviceColumn.from_numpy(ctx, rid)
    out, offsets = ops.shuffle(ctx, [kc], [kc, rc], n, slots)
    order, want = O.shuffle_order(O.arrow6_hash([(k, None)]), slots)
    assert np.array_equal(offsets, want)
    assert np.array_equal(out[1].numpy(n), rid[order])
    assert np.array_equal(out[0].numpy(n), k[order])
