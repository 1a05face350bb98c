"""Wire forms on the GPU: selection bitmaps to and from the int64 words of
the reference's bitmap pushdown (against the oracle's setBit/getBit), and
bloom filters through Arrow IPC streams and back, bit for bit."""
import ctypes

import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import ops, wire
from flexpushdowndb_amd.device import DeviceColumn
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sizes", [[1], [64], [63, 1, 64, 130], [1000, 0, 7, 4095, 513], [100_003, 250_000]])
def test_bitmap_words_round_trip(ctx, sizes):
    print(f"bitmap words sizes={sizes}", flush=True)
    rng = np.random.default_rng(len(sizes) + sum(sizes))
    sels = [rng.random(n) < 0.3 for n in sizes]
    total = sum(sizes)
    nwords = (total + 63) // 64
    d_words = ctx.zeros(max(1, nwords), np.int64)
    row = 0
    for sel in sels:
        bm = np.zeros(max(8, (len(sel) + 63) // 64 * 8), np.uint8)
        packed = np.packbits(sel, bitorder="little")
        bm[:len(packed)] = packed
        d_bm = ctx.upload(bm)
        wire.bitmap_to_words(ctx, d_bm.ptr, len(sel), row, d_words.ptr)
        row += len(sel)
    words = d_words.numpy(nwords)
    assert np.array_equal(words, O.bitmap_words(sels))
    row = 0
    for sel in sels:
        n = len(sel)
        out = ctx.zeros(max(8, (n + 63) // 64 * 8), np.uint8)
        wire.bitmap_from_words(ctx, d_words.ptr, nwords, row, n, out.ptr)
        got = np.unpackbits(out.numpy((n + 7) // 8), bitorder="little")[:n].astype(bool)
        assert np.array_equal(got, sel)
        row += n
    with pytest.raises(C.FpdbError):
        wire.bitmap_from_words(ctx, d_words.ptr, nwords, nwords * 64 - 3, 10, d_words.ptr)


def test_blocked_filter_through_ipc(ctx):
    rng = np.random.default_rng(8)
    keys = rng.integers(0, 10 ** 9, 150_000, dtype=np.int64)
    kc = DeviceColumn.from_numpy(ctx, keys)
    f = ops.BlockedBloomFilter.build(ctx, [kc], len(keys))
    batches = wire.blocked_filter_record_batches(ctx, f.blocks.ptr, f.log_num_blocks)
    assert [b.num_rows for b in batches] == [136, 8 << f.log_num_blocks]
    data = wire.ipc_bytes(batches)
    lnb, d_blocks = wire.blocked_filter_from_record_batches(ctx, wire.ipc_read(data))
    assert lnb == f.log_num_blocks
    assert np.array_equal(d_blocks.numpy(1 << lnb), f.blocks_numpy())
    # the received filter selects exactly what the original does
    probe = np.concatenate([keys[:5000], rng.integers(0, 10 ** 9, 20_000, dtype=np.int64)])
    pc = DeviceColumn.from_numpy(ctx, probe)
    n = len(probe)
    bm1 = ctx.zeros((n + 63) // 64 * 8, np.uint8)
    bm2 = ctx.zeros((n + 63) // 64 * 8, np.uint8)
    a = C.columns_array([pc.c()])
    C.check(C.lib().fpdb_bloom_blocked_find(ctx.handle, f.blocks.ptr, lnb, a, 1, n, bm1.ptr, None))
    C.check(C.lib().fpdb_bloom_blocked_find(ctx.handle, d_blocks.ptr, lnb, a, 1, n, bm2.ptr, None))
    assert np.array_equal(bm1.numpy(), bm2.numpy())
    # masks other than BloomFilterMasks() are refused
    masks = O.bloom_masks()
    masks[3] ^= 1
    bad = [wire._batch(masks), batches[1]]
    with pytest.raises(C.FpdbError):
        wire.blocked_filter_from_record_batches(ctx, bad)


def test_legacy_filter_words_through_ipc(ctx):
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 10 ** 6, 20_000, dtype=np.int64)
    lf = ops.LegacyBloomFilter(ctx, len(keys), 0.01, seed=3)
    lf.build([DeviceColumn.from_numpy(ctx, keys)], len(keys))
    words = lf.words.numpy()
    back = wire.record_batch_to_bitmap(wire.ipc_read(wire.ipc_bytes([wire.bitmap_record_batch(words)]))[0])
    assert np.array_equal(back, words)
