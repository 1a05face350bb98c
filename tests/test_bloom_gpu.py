"""Bloom filters and filter compaction on the GPU, bit-exact against the oracle.

Blocked filter: bit arrays and selection bitmaps are compared with the oracle's
restatement of the vendored Arrow BlockedBloomFilter fed with the oracle's
restatement of the Arrow 6 key hash (parity with the reference binary's hash is
unpinned; the filter arithmetic given the hash is pinned).  Legacy filter:
fully pinned, compared bit for bit.
"""
import ctypes
import types

import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import datagen, ops
from flexpushdowndb_amd.device import DeviceColumn
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits(bm_dev, n):
    return np.unpackbits(bm_dev.numpy((n + 7) // 8), bitorder="little")[:n].astype(bool)


def _valid_bits(mask):
    return np.packbits(mask, bitorder="little")


@pytest.mark.parametrize("dtype", [np.int64, np.int32])
def test_hash_keys_match_oracle(ctx, dtype):
    rng = np.random.default_rng(1)
    v = rng.integers(-2 ** 31, 2 ** 31 - 1, 10_001).astype(dtype)
    valid = rng.random(10_001) > 0.1
    col = DeviceColumn.from_numpy(ctx, v, valid)
    got = ops.hash_keys(ctx, [col], len(v)).numpy()
    assert np.array_equal(got, O.arrow6_hash([(v, _valid_bits(valid))]))
    w = rng.integers(0, 1000, 10_001).astype(np.int32)
    got2 = ops.hash_keys(ctx, [col, DeviceColumn.from_numpy(ctx, w)], len(v)).numpy()
    assert np.array_equal(got2, O.arrow6_hash([(v, _valid_bits(valid)), (w, None)]))


@pytest.mark.parametrize("n_build,n_probe", [(0, 100), (1, 1), (63, 64), (64, 65), (1000, 4097),
                                             (70_000, 200_003)])
def test_blocked_build_find_bit_exact(ctx, n_build, n_probe):
    rng = np.random.default_rng(n_build + 7)
    keys = rng.integers(0, 2 ** 40, n_build).astype(np.int64)
    probe = np.concatenate([keys[: n_probe // 2], rng.integers(0, 2 ** 40, n_probe - min(n_probe // 2, n_build))])
    probe = probe[:n_probe].astype(np.int64)
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], n_build)
    lnb, ref_blocks = O.blocked_build(O.arrow6_hash([(keys, None)]) if n_build else np.zeros(0, np.uint32),
                                      n_capacity=n_build)
    assert f.log_num_blocks == lnb
    assert np.array_equal(f.blocks_numpy(), ref_blocks)
    bm, cnt = f.find([DeviceColumn.from_numpy(ctx, probe)], n_probe)
    ref_bm = O.blocked_find(lnb, ref_blocks, O.arrow6_hash([(probe, None)]))
    got_bytes = bm.numpy((n_probe + 7) // 8)
    assert np.array_equal(got_bytes, ref_bm)
    assert cnt == int(np.unpackbits(ref_bm, bitorder="little").sum())


@pytest.mark.parametrize("n_build,lnb", [(100_000, 14), (200_000, 15), (400_000, 16), (600_000, 17)])
def test_blocked_find_filter_slices(ctx, n_build, lnb):
    """Filters of 1, 2 and 4 LDS slices (and the 2^16-reachable-block cap):
    find bitmap, count and fused compaction match the oracle."""
    rng = np.random.default_rng(lnb)
    keys = rng.integers(0, 2 ** 40, n_build).astype(np.int64)
    n = 1_000_003
    probe = np.concatenate([rng.choice(keys, n // 2), rng.integers(0, 2 ** 40, n - n // 2)]).astype(np.int64)
    rng.shuffle(probe)
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], n_build)
    assert f.log_num_blocks == lnb
    ref_bm = O.blocked_find(lnb, f.blocks_numpy(), O.arrow6_hash([(probe, None)]))
    dp = DeviceColumn.from_numpy(ctx, probe)
    bm, cnt = f.find([dp], n)
    assert np.array_equal(bm.numpy((n + 7) // 8), ref_bm)
    sel = np.unpackbits(ref_bm, bitorder="little")[:n].astype(bool)
    assert cnt == int(sel.sum())
    out, c2, bm2 = f.filter([dp], [0], n)
    assert c2 == cnt and np.array_equal(out[0].numpy(), probe[sel])
    assert np.array_equal(bm2.numpy((n + 7) // 8), ref_bm)


@pytest.mark.parametrize("n_build", [9_000, 200_000, 400_000])
def test_blocked_find_slices_nullable_int32(ctx, n_build):
    """Generic key path (int32 with nulls) through 1, 2 and 4 slices."""
    rng = np.random.default_rng(n_build)
    keys = rng.integers(-2 ** 31, 2 ** 31 - 1, n_build).astype(np.int32)
    n = 300_017
    probe = np.concatenate([rng.choice(keys, n // 2), rng.integers(-2 ** 31, 2 ** 31 - 1, n - n // 2)])
    probe = probe.astype(np.int32)
    valid = rng.random(n) > 0.05
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], n_build)
    h = O.arrow6_hash([(probe, _valid_bits(valid))])
    ref_bm = O.blocked_find(f.log_num_blocks, f.blocks_numpy(), h)
    bm, cnt = f.find([DeviceColumn.from_numpy(ctx, probe, valid)], n)
    assert np.array_equal(bm.numpy((n + 7) // 8), ref_bm)
    assert cnt == int(np.unpackbits(ref_bm, bitorder="little")[:n].sum())


def test_blocked_more_than_2_16_blocks_quirk(ctx):
    # 32-bit hashes address at most 2^16 blocks: a filter sized for > 524,288
    # keys never sets blocks >= 2^16 (block_id uses hash bits 16..31)
    n = 600_000
    keys = np.arange(n, dtype=np.int64) * 7919
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], n)
    blocks = f.blocks_numpy()
    assert f.log_num_blocks == 17
    assert not blocks[1 << 16:].any()
    _, ref = O.blocked_build(O.arrow6_hash([(keys, None)]), n_capacity=n)
    assert np.array_equal(blocks, ref)


def test_blocked_with_nulls_and_select_and_int32(ctx):
    rng = np.random.default_rng(4)
    n = 50_000
    k32 = rng.integers(-1000, 1000, n).astype(np.int32)
    valid = rng.random(n) > 0.2
    select = rng.random(n) > 0.5
    kc = DeviceColumn.from_numpy(ctx, k32, valid)
    sel = ctx.upload(np.packbits(select, bitorder="little"))
    f = ops.BlockedBloomFilter.build(ctx, [kc], n, select=sel, capacity=int(select.sum()))
    h = O.arrow6_hash([(k32, _valid_bits(valid))])
    lnb, ref = O.blocked_build(h[select], n_capacity=int(select.sum()))
    assert np.array_equal(f.blocks_numpy(), ref)
    bm, _ = f.find([kc], n)
    assert np.array_equal(bm.numpy((n + 7) // 8), O.blocked_find(lnb, ref, h))


def test_blocked_insert_is_or_merge(ctx):
    # partitioned builds OR together into the single-build bit array
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 2 ** 40, 30_000).astype(np.int64)
    whole = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], len(keys))
    parts = ops.BlockedBloomFilter(ctx, len(keys))
    for chunk in np.array_split(keys, 5):
        parts.insert([DeviceColumn.from_numpy(ctx, chunk)], len(chunk))
    assert np.array_equal(whole.blocks_numpy(), parts.blocks_numpy())


@pytest.mark.parametrize("n", [0, 1, 100, 2048, 2049, 100_003])
def test_fused_bloom_filter_compaction(ctx, n):
    rng = np.random.default_rng(n + 3)
    keys = rng.integers(0, 2 ** 20, 5000).astype(np.int64)
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, keys)], len(keys))
    probe = rng.integers(0, 2 ** 21, n).astype(np.int64)
    pay = rng.integers(-50, 50, n).astype(np.int32)
    dbl = rng.random(n)
    pvalid = rng.random(n) > 0.3
    cols = [DeviceColumn.from_numpy(ctx, probe), DeviceColumn.from_numpy(ctx, pay, pvalid),
            DeviceColumn.from_numpy(ctx, dbl)]
    out, cnt, bm = f.filter(cols, [0], n)
    ref_bm = O.blocked_find(f.log_num_blocks, f.blocks_numpy(), O.arrow6_hash([(probe, None)]))
    sel = np.unpackbits(ref_bm, bitorder="little")[:n].astype(bool) if n else np.zeros(0, bool)
    assert cnt == int(sel.sum())
    assert np.array_equal(bm.numpy((n + 7) // 8), ref_bm)
    assert np.array_equal(out[0].numpy(), probe[sel])
    assert np.array_equal(out[2].numpy(), dbl[sel])
    assert np.array_equal(out[1].valid_mask(), pvalid[sel])
    assert np.array_equal(out[1].numpy()[pvalid[sel]], pay[sel][pvalid[sel]])
    # no false negatives for inserted keys
    if n:
        _, c2, _ = f.filter([DeviceColumn.from_numpy(ctx, keys)], [0], len(keys))
        assert c2 == len(keys)


def test_filter_compact_plain(ctx):
    rng = np.random.default_rng(12)
    n = 77_777
    sel = rng.random(n) > 0.6
    a = rng.integers(0, 1 << 62, n).astype(np.int64)
    b = rng.integers(0, 255, n).astype(np.uint8)
    out, cnt = ops.filter_compact(ctx, [DeviceColumn.from_numpy(ctx, a), DeviceColumn.from_numpy(ctx, b)], n,
                                  ctx.upload(np.packbits(sel, bitorder="little")))
    assert cnt == sel.sum()
    assert np.array_equal(out[0].numpy(), a[sel]) and np.array_equal(out[1].numpy(), b[sel])


@pytest.mark.parametrize("misalign", [1, 3, 8])
def test_filter_compact_unaligned_bitmap(ctx, misalign):
    """A selection bitmap that does not start on an 8-byte boundary: the count
    and copy passes assemble each 64-row word from bytes."""
    rng = np.random.default_rng(40 + misalign)
    n = 300_001
    sel = rng.random(n) > 0.45
    a = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    packed = np.packbits(sel, bitorder="little")
    buf = ctx.upload(np.concatenate([np.full(misalign, 0xFF, np.uint8), packed, np.full(8, 0xFF, np.uint8)]))
    out, cnt = ops.filter_compact(ctx, [DeviceColumn.from_numpy(ctx, a)], n,
                                  types.SimpleNamespace(ptr=buf.ptr + misalign))
    assert cnt == int(sel.sum())
    assert np.array_equal(out[0].numpy(cnt), a[sel])


def test_legacy_build_find_bit_exact(ctx):
    rng = np.random.default_rng(21)
    n = 40_000
    keys = rng.integers(-2 ** 40, 2 ** 40, n).astype(np.int64)
    lf = ops.LegacyBloomFilter(ctx, n, 0.01, seed=1234)
    assert lf.k == 7 and lf.num_bits == O.legacy_m(n, 0.01)
    assert np.array_equal(lf.hash_functions, O.legacy_params(1234, lf.num_bits, 7))
    lf.build([DeviceColumn.from_numpy(ctx, keys)], n)
    okeys = O.legacy_keys([keys])
    ref_words = O.legacy_build(lf.hash_functions, lf.num_bits, okeys)
    assert np.array_equal(lf.words.numpy(), ref_words)
    probe = np.concatenate([keys[:20_000], rng.integers(-2 ** 40, 2 ** 40, 30_001)]).astype(np.int64)
    bm, cnt = lf.find([DeviceColumn.from_numpy(ctx, probe)], len(probe))
    ref_bm = O.legacy_find(lf.hash_functions, n, ref_words, O.legacy_keys([probe]))
    assert np.array_equal(bm.numpy(len(ref_bm)), ref_bm)
    sel = np.unpackbits(ref_bm, bitorder="little")[:len(probe)]
    assert cnt == sel.sum() and sel[:20_000].all()
    # the false-positive rate lands near the 1% target
    fpr = sel[20_000:].mean()
    assert fpr < 0.02


def test_legacy_two_columns_and_capacity_zero(ctx):
    rng = np.random.default_rng(5)
    a = rng.integers(-100, 100, 5000).astype(np.int32)
    b = rng.integers(0, 10 ** 12, 5000).astype(np.int64)
    lf = ops.LegacyBloomFilter(ctx, 5000, 0.001, seed=7)
    lf.build([DeviceColumn.from_numpy(ctx, a), DeviceColumn.from_numpy(ctx, b)], 5000)
    ref = O.legacy_build(lf.hash_functions, lf.num_bits, O.legacy_keys([a, b]))
    assert np.array_equal(lf.words.numpy(), ref)
    empty = ops.LegacyBloomFilter(ctx, 0, 0.01)
    assert empty.k == 0
    empty.build([DeviceColumn.from_numpy(ctx, b)], 0)
    bm, cnt = empty.find([DeviceColumn.from_numpy(ctx, b)], 5000)
    assert cnt == 0


def test_bloom_workload_full_size_exact(ctx):
    """Full single-edge predicate-transfer workload: 2M keys filtered to ~400k,
    60M probe keys.  The 2^16-block bit array and all 60M bits of the selection
    bitmap equal the oracle's; the compacted column holds every selected key
    at its rank."""
    w = datagen.bloom_workload()
    kept = w["kept"]
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, kept)], len(kept))
    lnb, ref = O.blocked_build(O.arrow6_hash([(kept, None)]), n_capacity=len(kept))
    assert f.log_num_blocks == lnb == 16
    assert np.array_equal(f.blocks_numpy(), ref)
    probe = w["probe"]
    pc = DeviceColumn.from_numpy(ctx, probe)
    bm, cnt = f.find([pc], len(probe))
    got = bm.numpy((len(probe) + 7) // 8)
    ref_bm = O.blocked_find(lnb, ref, O.arrow6_hash([(probe, None)]))
    assert np.array_equal(got, ref_bm)
    sel = np.unpackbits(got, bitorder="little")[:len(probe)].astype(bool)
    assert cnt == int(sel.sum())
    out, c2, bm2 = f.filter([pc], [0], len(probe))
    assert c2 == cnt and np.array_equal(bm2.numpy(len(got)), got)
    # 14,649 compaction tiles (two scan chunks): every kept key at its rank
    assert np.array_equal(out[0].numpy(c2), probe[sel])
    assert sel[np.isin(probe, kept)].all()   # no false negatives


def test_legacy_filter_c2_size_exact(ctx):
    """The universal-hash filter at the workload's size: the ~400k kept keys
    at a 1% target false-positive rate (k = 7, m = ceil(n |ln p| / ln^2 2))
    and 60M probe keys.  Bit array words and all 60M selection bits equal the
    oracle's (BloomFilter.cpp:62-91, UniversalHashFunction.cpp:51-57)."""
    w = datagen.bloom_workload()
    kept, probe = w["kept"], w["probe"]
    n = len(kept)
    lf = ops.LegacyBloomFilter(ctx, n, 0.01, seed=2024)
    assert lf.k == 7 and lf.num_bits == O.legacy_m(n, 0.01)
    lf.build([DeviceColumn.from_numpy(ctx, kept)], n)
    ref_words = O.legacy_build(lf.hash_functions, lf.num_bits, O.legacy_keys([kept]))
    assert np.array_equal(lf.words.numpy(), ref_words)
    bm, cnt = lf.find([DeviceColumn.from_numpy(ctx, probe)], len(probe))
    ref_bm = O.legacy_find(lf.hash_functions, n, ref_words, O.legacy_keys([probe]))
    assert np.array_equal(bm.numpy(len(ref_bm)), ref_bm)
    sel = np.unpackbits(ref_bm, bitorder="little")[:len(probe)].astype(bool)
    assert cnt == int(sel.sum())
    member = np.isin(probe, kept)
    assert sel[member].all()
    assert sel[~member].mean() < 0.012   # false positives near the 1% target


def test_filter_compact_capacity(ctx):
    """Outputs shorter than the selected count: the rows that fit are written,
    nothing past them, and the call reports FPDB_ERR_CAPACITY with the count."""
    n = 100_000
    rng = np.random.default_rng(90)
    sel = rng.random(n) < 0.5
    a = np.arange(n, dtype=np.int64) + 11
    need = int(sel.sum())
    cap = need - 1000
    src = DeviceColumn.from_numpy(ctx, a)
    bmd = ctx.upload(np.packbits(sel, bitorder="little"))
    buf = ctx.upload(np.full(need + 64, -7, np.int64))
    dst = DeviceColumn(buf, None, C.TYPE_INT64, cap)
    nout = ctypes.c_int64()
    st = C.lib().fpdb_filter_compact(ctx.handle, C.columns_array([src.c()]), 1, n, bmd.ptr,
                                     C.columns_array([dst.c()]), ctypes.byref(nout))
    assert st == C.FPDB_ERR_CAPACITY and nout.value == need
    got = buf.numpy()
    assert np.array_equal(got[:cap], a[sel][:cap])
    assert (got[cap:] == -7).all()
    with pytest.raises(C.FpdbError) as e:
        ops.filter_compact(ctx, [src], n, bmd, capacity=cap)
    assert e.value.code == C.FPDB_ERR_CAPACITY
    out, cnt = ops.filter_compact(ctx, [src], n, bmd, capacity=need)
    assert cnt == need and np.array_equal(out[0].numpy(cnt), a[sel])


@pytest.mark.parametrize("extra", [-1, 0, 1])
def test_filter_compact_at_scan_chunk_boundary(ctx, extra):
    """8,192 compaction tiles of 4,096 rows fill exactly one chunk of the
    tile-count scan; one more row starts a second chunk, one fewer leaves the
    last tile partial."""
    n = 8192 * 4096 + extra
    rng = np.random.default_rng(70 + extra)
    sel = rng.random(n) < 0.3
    a = np.arange(n, dtype=np.int64) * 7 - 5
    out, cnt = ops.filter_compact(ctx, [DeviceColumn.from_numpy(ctx, a)], n,
                                  ctx.upload(np.packbits(sel, bitorder="little")))
    assert cnt == int(sel.sum())
    assert np.array_equal(out[0].numpy(cnt), a[sel])
