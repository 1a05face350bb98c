"""Hash join on the GPU against the oracle: sorted join-output multisets must
match exactly (the reference's join output order is unspecified)."""
import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import datagen, ops
from flexpushdowndb_amd.device import DeviceColumn
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(cols, n, masks=None):
    arrs = [c.numpy(n) for c in cols]
    if masks is not None:
        arrs = [np.where(m, a, 0) if m is not None else a for a, m in zip(arrs, masks)]
    return sorted(zip(*[a.tolist() for a in arrs]))


@pytest.mark.parametrize("n_build,n_probe,seed", [(200_000, 1_200_000, 5), (1_000_000, 6_000_000, 1)])
def test_c1_inner_join_closed_form(ctx, n_build, n_probe, seed):
    """The join workload (the full 1M-row build / 6M-row probe case with the
    benchmark's seed, and a fifth of it): the build keys are a permutation, so
    every probe row matches exactly one build row and the output is fixed by
    direct indexing -- compared column by column in probe-row order."""
    w = datagen.join_workload(n_build=n_build, n_probe=n_probe, seed=seed)
    bk = DeviceColumn.from_numpy(ctx, w["build_key"])
    bp = DeviceColumn.from_numpy(ctx, w["build_pay"])
    t = ops.HashJoinTable(ctx, [bk], len(w["build_key"]), payload=bp)
    assert t.info()["unique"] and t.info()["inserted"] == n_build
    pf = DeviceColumn.from_numpy(ctx, w["probe_fk"])
    pq = DeviceColumn.from_numpy(ctx, w["probe_qty"])
    pp = DeviceColumn.from_numpy(ctx, w["probe_price"])
    n = len(w["probe_fk"])
    out, cnt = t.probe(C.JOIN_INNER, [pf], n, [pf, pq, pp], [bk, "payload"], capacity=n)
    assert cnt == n
    # output is in probe-row order and every probe row has exactly one match
    inv = np.empty(n_build, np.int64)
    inv[w["build_key"]] = np.arange(n_build)
    assert np.array_equal(out[0].numpy(), w["probe_fk"])
    assert np.array_equal(out[1].numpy(), w["probe_qty"])
    assert np.array_equal(out[2].numpy(), w["probe_price"])
    assert np.array_equal(out[3].numpy(), w["probe_fk"])
    assert np.array_equal(out[4].numpy(), w["build_pay"][inv[w["probe_fk"]]])


def test_inner_join_duplicates_misses_vs_oracle(ctx):
    rng = np.random.default_rng(2)
    bk = rng.integers(0, 5000, 20_000).astype(np.int64)
    bv = rng.integers(-9, 9, 20_000).astype(np.int32)
    pk = rng.integers(0, 8000, 30_001).astype(np.int64)
    pv = rng.random(30_001)
    dbk, dbv = DeviceColumn.from_numpy(ctx, bk), DeviceColumn.from_numpy(ctx, bv)
    dpk, dpv = DeviceColumn.from_numpy(ctx, pk), DeviceColumn.from_numpy(ctx, pv)
    t = ops.HashJoinTable(ctx, [dbk], len(bk))
    assert not t.info()["unique"]
    op, ob = O.join_pairs(bk, pk)
    assert t.probe_count(C.JOIN_INNER, [dpk], len(pk)) == len(op)
    out, cnt = t.probe(C.JOIN_INNER, [dpk], len(pk), [dpk, dpv], [dbk, dbv])
    assert cnt == len(op)
    ref = sorted(zip(pk[op].tolist(), pv[op].tolist(), bk[ob].tolist(), bv[ob].tolist()))
    assert _rows(out, cnt) == ref
    # rows come out grouped in probe order
    assert (np.diff(np.searchsorted(np.sort(op), op)) >= 0).all()
    # capacity error reports the needed count
    with pytest.raises(C.FpdbError) as e:
        t.probe(C.JOIN_INNER, [dpk], len(pk), [dpk], [], capacity=10)
    assert e.value.code == C.FPDB_ERR_CAPACITY


def test_null_keys_never_match_and_int32_keys(ctx):
    rng = np.random.default_rng(3)
    bk = rng.integers(-100, 100, 3000).astype(np.int32)
    bvalid = rng.random(3000) > 0.2
    pk = rng.integers(-150, 150, 4000).astype(np.int64)
    pvalid = rng.random(4000) > 0.2
    dbk = DeviceColumn.from_numpy(ctx, bk, bvalid)
    dpk = DeviceColumn.from_numpy(ctx, pk, pvalid)
    t = ops.HashJoinTable(ctx, [dbk], len(bk))
    assert t.info()["inserted"] == bvalid.sum()
    op, ob = O.join_pairs(bk.astype(np.int64), pk, np.packbits(bvalid, bitorder="little"),
                          np.packbits(pvalid, bitorder="little"))
    rows = DeviceColumn.from_numpy(ctx, np.arange(4000, dtype=np.int64))
    brows = DeviceColumn.from_numpy(ctx, np.arange(3000, dtype=np.int64))
    out, cnt = t.probe(C.JOIN_INNER, [dpk], len(pk), [rows], [brows, dbk])
    assert sorted(zip(out[0].numpy(cnt).tolist(), out[1].numpy(cnt).tolist())) == sorted(zip(op.tolist(), ob.tolist()))
    # the (nullable, int32) build key as an output column: the matched key, always valid
    got = out[1].numpy(cnt)
    assert np.array_equal(out[2].numpy(cnt), bk[got])
    assert out[2].valid_mask()[:cnt].all()


def test_two_column_keys(ctx):
    rng = np.random.default_rng(8)
    b1 = rng.integers(0, 50, 4000).astype(np.int32)
    b2 = rng.integers(0, 50, 4000).astype(np.int32)
    p1 = rng.integers(0, 60, 5000).astype(np.int32)
    p2 = rng.integers(0, 60, 5000).astype(np.int32)
    t = ops.HashJoinTable(ctx, [DeviceColumn.from_numpy(ctx, b1), DeviceColumn.from_numpy(ctx, b2)], 4000)
    bidx = DeviceColumn.from_numpy(ctx, np.arange(4000, dtype=np.int64))
    pidx = DeviceColumn.from_numpy(ctx, np.arange(5000, dtype=np.int64))
    out, cnt = t.probe(C.JOIN_INNER, [DeviceColumn.from_numpy(ctx, p1), DeviceColumn.from_numpy(ctx, p2)], 5000,
                       [pidx], [bidx])
    comb_b = b1.astype(np.int64) * 1000 + b2
    comb_p = p1.astype(np.int64) * 1000 + p2
    op, ob = O.join_pairs(comb_b, comb_p)
    assert sorted(zip(out[0].numpy(cnt).tolist(), out[1].numpy(cnt).tolist())) == sorted(zip(op.tolist(), ob.tolist()))


def test_outer_and_semi_joins(ctx):
    rng = np.random.default_rng(4)
    bk = rng.integers(0, 300, 500).astype(np.int64)
    pk = rng.integers(100, 400, 700).astype(np.int64)
    dbk, dpk = DeviceColumn.from_numpy(ctx, bk), DeviceColumn.from_numpy(ctx, pk)
    bidx = DeviceColumn.from_numpy(ctx, np.arange(500, dtype=np.int64))
    pidx = DeviceColumn.from_numpy(ctx, np.arange(700, dtype=np.int64))
    op, ob, bmatched = O.join_pairs(bk, pk, want_matched=True)
    pmatched = np.zeros(700, bool)
    pmatched[op] = True
    inner = sorted(zip(op.tolist(), ob.tolist()))

    # RIGHT: probe side kept, unmatched probe rows with a null build side
    t = ops.HashJoinTable(ctx, [dbk], 500)
    out, cnt = t.probe(C.JOIN_RIGHT, [dpk], 700, [pidx], [bidx])
    assert cnt == len(op) + (~pmatched).sum()
    p_out, b_out, b_valid = out[0].numpy(cnt), out[1].numpy(cnt), out[1].valid_mask(cnt)
    assert sorted(zip(p_out[b_valid].tolist(), b_out[b_valid].tolist())) == inner
    assert sorted(p_out[~b_valid].tolist()) == np.flatnonzero(~pmatched).tolist()

    # RIGHT_SEMI: probe rows with a match, in probe order
    out, cnt = t.probe(C.JOIN_RIGHT_SEMI, [dpk], 700, [pidx], [])
    assert out[0].numpy(cnt).tolist() == np.flatnonzero(pmatched).tolist()

    # LEFT: inner pairs + unmatched build rows
    t2 = ops.HashJoinTable(ctx, [dbk], 500, track_build_matches=True)
    out, cnt = t2.probe(C.JOIN_LEFT, [dpk], 700, [pidx], [bidx])
    assert _rows(out, cnt) == inner
    rows, nr = t2.build_side_rows(matched=False)
    assert rows.numpy(nr).tolist() == np.flatnonzero(~bmatched).tolist()

    # FULL: both sides' unmatched rows
    t3 = ops.HashJoinTable(ctx, [dbk], 500, track_build_matches=True)
    out, cnt = t3.probe(C.JOIN_FULL, [dpk], 700, [pidx], [bidx])
    assert cnt == len(op) + (~pmatched).sum()
    rows, nr = t3.build_side_rows(matched=False)
    assert rows.numpy(nr).tolist() == np.flatnonzero(~bmatched).tolist()
    g = ops.gather(ctx, [bidx], rows, nr)
    assert g[0].numpy(nr).tolist() == np.flatnonzero(~bmatched).tolist()

    # LEFT_SEMI: build rows with a match, in build order
    t4 = ops.HashJoinTable(ctx, [dbk], 500, track_build_matches=True)
    _, cnt = t4.probe(C.JOIN_LEFT_SEMI, [dpk], 700, [], [])
    assert cnt == 0
    rows, nr = t4.build_side_rows(matched=True)
    assert rows.numpy(nr).tolist() == np.flatnonzero(bmatched).tolist()


def test_empty_sides(ctx):
    dpk = DeviceColumn.from_numpy(ctx, np.arange(100, dtype=np.int64))
    empty = DeviceColumn.from_numpy(ctx, np.zeros(0, np.int64))
    t = ops.HashJoinTable(ctx, [empty], 0)
    out, cnt = t.probe(C.JOIN_INNER, [dpk], 100, [dpk], [])
    assert cnt == 0
    out, cnt = t.probe(C.JOIN_RIGHT, [dpk], 100, [dpk], [empty])
    assert cnt == 100 and not out[1].valid_mask(100).any()
    t2 = ops.HashJoinTable(ctx, [dpk], 100)
    out, cnt = t2.probe(C.JOIN_INNER, [empty], 0, [empty], [dpk])
    assert cnt == 0


def test_attic_join_vector(ctx):
    # fpdb-executor/ATTIC/test/JoinTest.cpp:82-123: a join b on aa = ba -> 4 rows
    a = np.array([[10, 13, 16], [10, 14, 17], [12, 15, 18]], np.int64)
    b = np.array([[10, 23, 26], [11, 24, 27], [10, 25, 28]], np.int64)
    bc = [DeviceColumn.from_numpy(ctx, b[:, i].copy()) for i in range(3)]
    ac = [DeviceColumn.from_numpy(ctx, a[:, i].copy()) for i in range(3)]
    t = ops.HashJoinTable(ctx, [bc[0]], 3)
    out, cnt = t.probe(C.JOIN_INNER, [ac[0]], 3, ac, bc)
    assert cnt == 4
    assert _rows(out, cnt) == sorted([(10, 13, 16, 10, 23, 26), (10, 13, 16, 10, 25, 28),
                                      (10, 14, 17, 10, 23, 26), (10, 14, 17, 10, 25, 28)])


def _bits(mask):
    return np.packbits(mask, bitorder="little")


def _check_all_join_types(ctx, bk, pk, bvalid, pvalid, want_form):
    """Every join type of a build/probe pair against the oracle's hash join."""
    nb, npr = len(bk), len(pk)
    dbk = DeviceColumn.from_numpy(ctx, bk, bvalid)
    dpk = DeviceColumn.from_numpy(ctx, pk, pvalid)
    bidx = DeviceColumn.from_numpy(ctx, np.arange(nb, dtype=np.int64))
    pidx = DeviceColumn.from_numpy(ctx, np.arange(npr, dtype=np.int64))
    op, ob, bmatched = O.join_pairs(bk.astype(np.int64), pk.astype(np.int64),
                                    None if bvalid is None else _bits(bvalid),
                                    None if pvalid is None else _bits(pvalid), want_matched=True)
    pmatched = np.zeros(npr, bool)
    pmatched[op] = True
    inner = sorted(zip(op.tolist(), ob.tolist()))
    t = ops.HashJoinTable(ctx, [dbk], nb)
    assert t.form() == want_form
    out, cnt = t.probe(C.JOIN_INNER, [dpk], npr, [pidx], [bidx, dbk])
    assert sorted(zip(out[0].numpy(cnt).tolist(), out[1].numpy(cnt).tolist())) == inner
    assert np.array_equal(out[2].numpy(cnt), bk[out[1].numpy(cnt)])   # build key column = matched key
    # probe order: output rows follow the probe rows
    assert (np.diff(out[0].numpy(cnt)) >= 0).all()
    out, cnt = t.probe(C.JOIN_RIGHT, [dpk], npr, [pidx], [bidx])
    assert cnt == len(op) + (~pmatched).sum()
    p_out, b_out, b_valid = out[0].numpy(cnt), out[1].numpy(cnt), out[1].valid_mask(cnt)
    assert sorted(zip(p_out[b_valid].tolist(), b_out[b_valid].tolist())) == inner
    assert sorted(p_out[~b_valid].tolist()) == np.flatnonzero(~pmatched).tolist()
    out, cnt = t.probe(C.JOIN_RIGHT_SEMI, [dpk], npr, [pidx], [])
    assert out[0].numpy(cnt).tolist() == np.flatnonzero(pmatched).tolist()
    for jt in (C.JOIN_LEFT, C.JOIN_FULL, C.JOIN_LEFT_SEMI):
        tt = ops.HashJoinTable(ctx, [dbk], nb, track_build_matches=True)
        assert tt.form() == want_form
        out, cnt = tt.probe(jt, [dpk], npr, [pidx], [bidx])
        if jt == C.JOIN_LEFT:
            assert _rows(out, cnt) == inner
        elif jt == C.JOIN_FULL:
            assert cnt == len(op) + (~pmatched).sum()
        else:
            assert cnt == 0
        rows, nr = tt.build_side_rows(matched=False)
        assert rows.numpy(nr).tolist() == np.flatnonzero(~bmatched).tolist()
        tt.close()


@pytest.mark.parametrize("dtype", [np.int64, np.int32])
def test_dense_table_all_join_types(ctx, dtype):
    """Distinct keys with gaps over a range < 2x the rows, negative keys, nulls
    on both sides, probe keys inside and outside the range: the dense table."""
    rng = np.random.default_rng(21)
    nb = 30_000
    bk = rng.choice(np.arange(-25_000, 25_000), nb, replace=False).astype(dtype)
    pk = rng.integers(-30_000, 30_000, 70_001).astype(dtype)
    _check_all_join_types(ctx, bk, pk, rng.random(nb) > 0.1, rng.random(len(pk)) > 0.1, "dense")
    _check_all_join_types(ctx, bk, pk, None, None, "dense")


def test_dense_table_extreme_keys(ctx):
    """Keys at the int64 limits: slot arithmetic must not wrap into the range."""
    lo = np.iinfo(np.int64).min
    hi = np.iinfo(np.int64).max
    bk = np.concatenate([np.arange(hi - 999, hi + 1, dtype=np.int64)])
    pk = np.array([hi, hi - 1, hi - 1000, lo, lo + 1, 0, -1, hi - 500] * 100, np.int64)
    _check_all_join_types(ctx, bk, pk, None, None, "dense")
    bk2 = np.arange(lo, lo + 1000, dtype=np.int64)
    pk2 = np.array([lo, lo + 999, lo + 1000, hi, 0, -1, lo + 3] * 100, np.int64)
    _check_all_join_types(ctx, bk2, pk2, None, None, "dense")


def test_dense_falls_back_to_hash(ctx):
    """A duplicate key or a range wider than twice the rows: the hash table,
    with the same results."""
    rng = np.random.default_rng(22)
    bk = rng.permutation(20_000).astype(np.int64)
    bk[7] = bk[19_000]                       # one duplicate
    pk = rng.integers(-10, 20_010, 50_000).astype(np.int64)
    _check_all_join_types(ctx, bk, pk, None, None, "hash")
    sparse = rng.choice(100_000, 20_000, replace=False).astype(np.int64)   # span 5x the rows
    _check_all_join_types(ctx, sparse, pk * 5, None, None, "hash")
    # all keys null: nothing inserted, hash form
    _check_all_join_types(ctx, bk, pk, np.zeros(len(bk), bool), None, "hash")


def test_dense_and_hash_same_output(ctx):
    """The same join through both forms (a duplicate forces the hash table;
    its extra match is removed from the comparison): identical pairs."""
    rng = np.random.default_rng(23)
    n = 100_000
    bk = rng.permutation(n).astype(np.int64)
    pay = rng.integers(0, 2 ** 31, n).astype(np.int32)
    pk = rng.integers(0, n, 300_000).astype(np.int64)
    dbk, dpay, dpk = (DeviceColumn.from_numpy(ctx, a) for a in (bk, pay, pk))
    td = ops.HashJoinTable(ctx, [dbk], n, payload=dpay)
    assert td.form() == "dense"
    out_d, cnt_d = td.probe(C.JOIN_INNER, [dpk], len(pk), [dpk], [dbk, "payload"], capacity=len(pk))
    bk2 = np.concatenate([bk, [n + 5]])      # range 2x + a key outside: still distinct, but spread
    bk2[-1] = 10 * n
    pay2 = np.concatenate([pay, [7]]).astype(np.int32)
    th = ops.HashJoinTable(ctx, [DeviceColumn.from_numpy(ctx, bk2)], n + 1, payload=DeviceColumn.from_numpy(ctx, pay2))
    assert th.form() == "hash"
    out_h, cnt_h = th.probe(C.JOIN_INNER, [dpk], len(pk), [dpk], [DeviceColumn.from_numpy(ctx, bk2), "payload"],
                            capacity=len(pk))
    assert cnt_d == cnt_h == len(pk)
    for a, b in zip(out_d, out_h):
        assert np.array_equal(a.numpy(cnt_d), b.numpy(cnt_h))


@pytest.mark.parametrize("spread", [1, 3])
def test_probe_longer_than_one_scan_chunk(ctx, spread):
    """10M probe rows are 9,766 tiles of 1,024 rows: the tile-count scan takes
    two chunks of 8,192 tiles.  Dense form (keys a permutation) and hash form
    (keys spread 3x); build keys k, probe keys uniform over twice the key
    range, so about half the probe rows miss.  Closed form: the output rows
    are the hitting probe rows in probe order, the payload that of the build
    row holding the key."""
    rng = np.random.default_rng(31 + spread)
    nb, npr = 500_000, 10_000_000
    perm = rng.permutation(nb).astype(np.int64)
    bk = perm * spread + (spread > 1)
    pay = rng.integers(-2 ** 31, 2 ** 31, nb).astype(np.int32)
    raw = rng.integers(0, 2 * nb, npr).astype(np.int64)
    pk = raw * spread + (spread > 1)
    dbk, dpay, dpk = (DeviceColumn.from_numpy(ctx, a) for a in (bk, pay, pk))
    t = ops.HashJoinTable(ctx, [dbk], nb, payload=dpay)
    assert t.form() == ("dense" if spread == 1 else "hash")
    out, cnt = t.probe(C.JOIN_INNER, [dpk], npr, [dpk], [dbk, "payload"], capacity=npr)
    hit = raw < nb
    assert cnt == int(hit.sum())
    inv = np.empty(nb, np.int64)
    inv[perm] = np.arange(nb)
    assert np.array_equal(out[0].numpy(cnt), pk[hit])
    assert np.array_equal(out[1].numpy(cnt), pk[hit])
    assert np.array_equal(out[2].numpy(cnt), pay[inv[raw[hit]]])


@pytest.mark.parametrize("nb,dup", [(2, True), (97, True), (3000, False), (4096, True), (4096, False),
                                    (4097, True)])
def test_small_hash_tables(ctx, nb, dup):
    """Small build sides (tables of 64 to 16,384 slots, the sizes of filtered
    dimension tables).  Keys spread over +-2^40 (never the dense form),
    duplicate or distinct build keys, nulls on both sides, probe keys that hit
    and miss; every join type against the oracle."""
    rng = np.random.default_rng(50 + nb + dup)
    if dup:
        bk = rng.choice(rng.integers(-2 ** 40, 2 ** 40, max(1, nb // 2)), nb).astype(np.int64)
    else:
        bk = rng.permutation(np.unique(rng.integers(-2 ** 40, 2 ** 40, 2 * nb))[:nb]).astype(np.int64)
    pk = np.concatenate([rng.choice(bk, 150_000), rng.integers(-2 ** 40, 2 ** 40, 50_001)]).astype(np.int64)
    rng.shuffle(pk)
    _check_all_join_types(ctx, bk, pk, rng.random(nb) > 0.1, rng.random(len(pk)) > 0.1, "hash")
    _check_all_join_types(ctx, bk, pk, None, None, "hash")


@pytest.mark.parametrize("extra", [0, 1])
def test_dense_probe_at_scan_chunk_boundary(ctx, extra):
    """8,192 probe tiles of 1,024 rows fill exactly one chunk of the tile-count
    scan; one more row starts a second chunk."""
    rng = np.random.default_rng(80 + extra)
    nb, npr = 100_000, 8192 * 1024 + extra
    bk = rng.permutation(nb).astype(np.int64)
    pay = rng.integers(0, 2 ** 31, nb).astype(np.int32)
    pk = rng.integers(0, 2 * nb, npr).astype(np.int64)
    dbk, dpay, dpk = (DeviceColumn.from_numpy(ctx, a) for a in (bk, pay, pk))
    t = ops.HashJoinTable(ctx, [dbk], nb, payload=dpay)
    assert t.form() == "dense"
    out, cnt = t.probe(C.JOIN_INNER, [dpk], npr, [dpk], ["payload"], capacity=npr)
    hit = pk < nb
    inv = np.empty(nb, np.int64)
    inv[bk] = np.arange(nb)
    assert cnt == int(hit.sum())
    assert np.array_equal(out[0].numpy(cnt), pk[hit])
    assert np.array_equal(out[1].numpy(cnt), pay[inv[pk[hit]]])


def test_objects_outliving_their_context():
    """Closing a context before the tables made on it: the wrappers' later
    close (or garbage collection) must not touch the destroyed context."""
    from flexpushdowndb_amd.device import Context
    c2 = Context(0)
    keys = DeviceColumn.from_numpy(c2, np.arange(1000, dtype=np.int64))
    t = ops.HashJoinTable(c2, [keys], 1000)
    g = ops.GroupBy(c2, [C.TYPE_INT64], [(C.AGG_COUNT_ALL, None, False)], 100)
    cache = ops.SegmentCache(c2, 1 << 20)
    c2.close()
    t.close()
    g.close()
    cache.close()
    del t, g, cache, keys
