"""utf8 columns and the key dictionary on the GPU: take, comparison with a
constant, compaction and concatenation of utf8 columns against Python's own
string operations; key-dictionary ids against the oracle's first-appearance
factorisation (itself pinned against pandas.factorize), including nulls,
mixed integer widths, -0.0, forced hash collisions and several batches."""
import ctypes

import numpy as np
import pandas as pd
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import ops
from flexpushdowndb_amd.device import DeviceColumn, Utf8Column
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _strings(rng, n, null_frac=0.1, alphabet="abcdé€xyz", maxlen=12):
    out = []
    for _ in range(n):
        if rng.random() < null_frac:
            out.append(None)
        else:
            k = int(rng.integers(0, maxlen + 1))
            out.append("".join(rng.choice(list(alphabet), k)))
    return out


def test_utf8_roundtrip_and_take(ctx):
    rng = np.random.default_rng(1)
    s = _strings(rng, 5000)
    col = Utf8Column.from_strings(ctx, s)
    assert col.to_list() == s
    rows = rng.integers(-1, len(s), 7000)
    got = ops.utf8_take(ctx, col, DeviceColumn.from_numpy(ctx, rows.astype(np.int64)), len(rows)).to_list()
    assert got == [s[r] if r >= 0 else None for r in rows]
    # int32 row numbers with nulls, and an input at an element offset
    rv = rng.random(len(rows)) < 0.8
    sub = Utf8Column(col.offsets, col.data, col.data_bytes, col.validity, len(s) - 100, offset=100)
    r32 = rng.integers(0, len(s) - 100, len(rows)).astype(np.int32)
    got = ops.utf8_take(ctx, sub, DeviceColumn.from_numpy(ctx, r32, rv), len(rows)).to_list()
    assert got == [s[100 + r] if v else None for r, v in zip(r32, rv)]
    # empty take
    assert ops.utf8_take(ctx, col, DeviceColumn.from_numpy(ctx, np.zeros(0, np.int64)), 0).to_list() == []


def test_utf8_take_errors(ctx):
    col = Utf8Column.from_strings(ctx, ["abc", "de", None, "fghij"])
    rows = DeviceColumn.from_numpy(ctx, np.array([3, 0, 3], np.int64))
    small = Utf8Column.empty(ctx, 3, 5, nullable=True)
    oc, cin, rc = small.c(), col.c(), rows.c()
    with pytest.raises(C.FpdbError) as e:
        C.check(C.lib().fpdb_utf8_take(ctx.handle, ctypes.byref(cin), ctypes.byref(rc), 3, ctypes.byref(oc)))
    assert e.value.code == C.FPDB_ERR_CAPACITY
    bad = DeviceColumn.from_numpy(ctx, np.array([0, 4], np.int64))
    with pytest.raises(C.FpdbError) as e:
        ops.utf8_take(ctx, col, bad, 2)
    assert e.value.code == C.FPDB_ERR_INVALID
    assert ops.utf8_take(ctx, col, rows, 3).to_list() == ["fghij", "abc", "fghij"]


@pytest.mark.parametrize("op,fn", [(C.CMP_LT, lambda a, b: a < b), (C.CMP_LE, lambda a, b: a <= b),
                                   (C.CMP_GT, lambda a, b: a > b), (C.CMP_GE, lambda a, b: a >= b),
                                   (C.CMP_EQ, lambda a, b: a == b), (C.CMP_NE, lambda a, b: a != b)])
def test_compare_utf8(ctx, op, fn):
    rng = np.random.default_rng(2)
    s = _strings(rng, 3001, maxlen=4, alphabet="abé")
    col = Utf8Column.from_strings(ctx, s)
    for const in ["ab", "", "b", "éa", "abab"]:
        bm = ops.compare_bitmap_utf8(ctx, col, op, const, len(s))
        got = np.unpackbits(bm.numpy(), bitorder="little")[:len(s)].astype(bool)
        cb = const.encode()
        want = np.array([v is not None and fn(v.encode(), cb) for v in s])
        assert np.array_equal(got, want), const


def test_compact_and_concatenate_with_utf8(ctx):
    rng = np.random.default_rng(3)
    n = 20_000
    s = _strings(rng, n)
    k = rng.integers(0, 100, n).astype(np.int64)
    cols = [DeviceColumn.from_numpy(ctx, k), Utf8Column.from_strings(ctx, s)]
    bm, cnt = ops.compare_bitmap(ctx, cols[0], C.CMP_LT, 37, n)
    out, m = ops.filter_compact(ctx, cols, n, bm)
    sel = k < 37
    assert m == sel.sum()
    assert np.array_equal(out[0].numpy(m), k[sel])
    assert out[1].to_list(m) == [v for v, t in zip(s, sel) if t]
    # concatenation: both halves of the compacted column after each other
    a, b = Utf8Column.from_strings(ctx, s[:700]), Utf8Column.from_strings(ctx, s[700:1500])
    dst = Utf8Column.empty(ctx, 1500, a.data_bytes + b.data_bytes, nullable=True)
    dc = C.columns_array([dst.c()])
    C.check(C.lib().fpdb_copy_rows(ctx.handle, C.columns_array([a.c()]), 1, 700, dc, 0))
    C.check(C.lib().fpdb_copy_rows(ctx.handle, C.columns_array([b.c()]), 1, 800, dc, 700))
    assert dst.to_list() == s[:1500]
    tight = Utf8Column.empty(ctx, 1500, a.data_bytes, nullable=True)
    with pytest.raises(C.FpdbError):
        tc = C.columns_array([tight.c()])
        C.check(C.lib().fpdb_copy_rows(ctx.handle, C.columns_array([a.c()]), 1, 700, tc, 0))
        C.check(C.lib().fpdb_copy_rows(ctx.handle, C.columns_array([b.c()]), 1, 800, tc, 700))


def _ids(col, n):
    v = col.numpy(n).astype(np.int64)
    return np.where(col.valid_mask(n), v, -1)


def test_keydict_groups_strings_ints_and_nulls(ctx):
    rng = np.random.default_rng(4)
    n = 30_000
    a = rng.integers(-5, 5, n).astype(np.int32)
    av = rng.random(n) > 0.05
    s = _strings(rng, n, null_frac=0.05, maxlen=3, alphabet="xyz")
    pa = [int(x) if v else None for x, v in zip(a, av)]
    d = ops.KeyDict(ctx, [C.TYPE_INT32, C.TYPE_UTF8])
    keys = [DeviceColumn.from_numpy(ctx, a, av), Utf8Column.from_strings(ctx, s)]
    half = n // 2
    # two batches: ids continue across them
    ids1 = d.encode([keys[0], keys[1]], half, insert=True, null_is_key=True)
    b2 = [DeviceColumn.from_numpy(ctx, a[half:], av[half:]), Utf8Column.from_strings(ctx, s[half:])]
    ids2 = d.encode(b2, n - half, insert=True, null_is_key=True)
    want, dictionary = O.factorize_tuples([pa, s])
    got = np.concatenate([_ids(ids1, half), _ids(ids2, n - half)])
    assert np.array_equal(got, want)
    assert d.size() == len(dictionary)
    # decode: the stored tuples in id order
    allids = DeviceColumn.from_numpy(ctx, np.arange(d.size(), dtype=np.int32))
    k0 = d.decode(0, allids, d.size())
    k1 = d.decode(1, allids, d.size())
    dec0 = [int(x) if v else None for x, v in zip(k0.numpy(), k0.valid_mask())]
    assert list(zip(dec0, k1.to_list())) == dictionary
    # lookup with a wider integer type: absent tuples and null keys give -1
    q = rng.integers(-7, 7, 5000).astype(np.int64)
    qs = _strings(rng, 5000, null_frac=0.05, maxlen=3, alphabet="xyzw")
    ids = d.encode([DeviceColumn.from_numpy(ctx, q), Utf8Column.from_strings(ctx, qs)], 5000, insert=False)
    assert np.array_equal(_ids(ids, 5000), O.lookup_tuples(dictionary, [q.tolist(), qs]))
    assert d.size() == len(dictionary)


def test_keydict_join_keys_exclude_nulls_and_floats(ctx):
    x = np.array([1.5, 0.0, -0.0, np.nan, 1.5, 2.0], np.float64)
    xv = np.array([1, 1, 1, 1, 1, 0], bool)
    d = ops.KeyDict(ctx, [C.TYPE_FLOAT64])
    ids = _ids(d.encode([DeviceColumn.from_numpy(ctx, x, xv)], 6, insert=True, null_is_key=False), 6)
    assert ids[0] == ids[4] and ids[1] == ids[2] and ids[5] == -1 and len(set(ids[:4])) == 3
    f32 = np.array([0.0, 1.5, 3.0], np.float32)
    with pytest.raises(C.FpdbError):   # inserting needs the dictionary's exact type
        d.encode([DeviceColumn.from_numpy(ctx, f32)], 3, insert=True)
    look = _ids(d.encode([DeviceColumn.from_numpy(ctx, f32)], 3, insert=False), 3)
    assert look.tolist() == [ids[1], ids[0], -1]


@pytest.mark.parametrize("hash_bits", [1, 3])
def test_keydict_forced_hash_collisions(ctx, hash_bits):
    """With only `hash_bits` bits of hash every distinct tuple shares its hash
    with many others: ids must still be one per distinct tuple."""
    rng = np.random.default_rng(6)
    n = 4000
    s = [f"k{v}" for v in rng.integers(0, 150, n)]
    d = ops.KeyDict(ctx, [C.TYPE_UTF8], hash_bits=hash_bits)
    ids = _ids(d.encode([Utf8Column.from_strings(ctx, s)], n, insert=True, null_is_key=True), n)
    codes, uniq = pd.factorize(pd.Series(s))
    assert d.size() == len(uniq)
    # a bijection between ids and strings, dense in [0, size)
    pairs = set(zip(ids.tolist(), s))
    assert len(pairs) == len(uniq) and set(ids.tolist()) == set(range(len(uniq)))
    dec = d.decode(0, DeviceColumn.from_numpy(ctx, ids.astype(np.int32)), n).to_list()
    assert dec == s
    # lookups through the collided chains
    q = [f"k{v}" for v in rng.integers(0, 300, 2000)]
    look = _ids(d.encode([Utf8Column.from_strings(ctx, q)], 2000, insert=False), 2000)
    id_of = dict(zip(s, ids.tolist()))
    assert look.tolist() == [id_of.get(v, -1) for v in q]


def test_keydict_large_against_pandas(ctx):
    rng = np.random.default_rng(7)
    n = 2_000_000
    k1 = rng.integers(0, 50_000, n).astype(np.int64)
    k2 = rng.integers(0, 3, n).astype(np.int16)
    d = ops.KeyDict(ctx, [C.TYPE_INT64, C.TYPE_INT16])
    ids = _ids(d.encode([DeviceColumn.from_numpy(ctx, k1), DeviceColumn.from_numpy(ctx, k2)], n), n)
    codes, _ = pd.factorize(pd.Series(k1 * 4 + k2))
    assert np.array_equal(ids, codes)
