"""Wire forms of filters and selection bitmaps as Arrow record batches.

The reference ships them to its store with Flight DoPut
(BloomFilterCreatePOp::putBloomFilterToStore,
fpdb-executor/src/physical/bloomfilter/BloomFilterCreatePOp.cpp:139-170;
FilterPOp::putBitmapToFPDBStore, fpdb-executor/src/physical/filter/FilterPOp.cpp:537-591):
  * a blocked bloom filter: two batches of one uint8 column "bitmap", the 136
    mask bytes then the block bytes (ArrowBloomFilter.cpp:97-116);
  * a legacy bloom filter or a filter's selection bitmap: one batch of one
    int64 column "bitmap" (ArrowSerializer.cpp:330-348), row r at bit r % 64
    of word r / 64.
The device side (filter blocks, bitmap words) goes through the C ABI; the
framing uses pyarrow, imported only here.
"""
import ctypes

import numpy as np

from . import _capi as C

BITMAP_FIELD_NAME = "bitmap"


def _pa():
    import pyarrow as pa
    return pa


def _batch(values):
    pa = _pa()
    arr = pa.array(values)
    return pa.RecordBatch.from_arrays([arr], schema=pa.schema([pa.field(BITMAP_FIELD_NAME, arr.type)]))


def blocked_filter_record_batches(ctx, d_blocks, log_num_blocks):
    """ArrowBloomFilter::makeBitmapRecordBatches for device blocks."""
    masks = np.zeros(136, np.uint8)
    blocks = np.zeros(8 << log_num_blocks, np.uint8)
    C.check(C.lib().fpdb_bloom_blocked_export(ctx.handle, d_blocks, log_num_blocks, masks.ctypes.data,
                                              blocks.ctypes.data))
    return [_batch(masks), _batch(blocks)]


def blocked_filter_from_record_batches(ctx, batches):
    """ArrowBloomFilter::saveBitmapRecordBatches: (log_num_blocks, device blocks)."""
    if len(batches) != 2:
        raise ValueError("RecordBatch stream for ArrowBloomFilter's bitmap should contain two recordBatches")
    masks = np.ascontiguousarray(batches[0].column(0).to_numpy(zero_copy_only=False), dtype=np.uint8)
    blocks = np.ascontiguousarray(batches[1].column(0).to_numpy(zero_copy_only=False), dtype=np.uint8)
    lnb = ctypes.c_int()
    C.check(C.lib().fpdb_bloom_blocked_import(ctx.handle, masks.ctypes.data, blocks.ctypes.data, blocks.nbytes,
                                              ctypes.byref(lnb), None))
    d = ctx.empty(blocks.nbytes // 8, np.uint64)
    C.check(C.lib().fpdb_bloom_blocked_import(ctx.handle, masks.ctypes.data, blocks.ctypes.data, blocks.nbytes,
                                              ctypes.byref(lnb), d.ptr))
    return lnb.value, d


def bitmap_record_batch(words):
    """ArrowSerializer::bitmap_to_recordBatch."""
    return _batch(np.ascontiguousarray(words, dtype=np.int64))


def record_batch_to_bitmap(batch):
    """ArrowSerializer::recordBatch_to_bitmap."""
    pa = _pa()
    idx = batch.schema.get_field_index(BITMAP_FIELD_NAME)
    if idx < 0:
        raise ValueError("Bitmap array not found in the recordBatch")
    if batch.schema.field(idx).type != pa.int64():
        raise ValueError("Type of bitmap array is not int64")
    return batch.column(idx).to_numpy().astype(np.int64)


def ipc_bytes(batches):
    """The Arrow IPC stream of the batches (what a Flight DoPut carries)."""
    pa = _pa()
    sink = pa.BufferOutputStream()
    with pa.ipc.new_stream(sink, batches[0].schema) as w:
        for b in batches:
            w.write_batch(b)
    return sink.getvalue().to_pybytes()


def ipc_read(data):
    pa = _pa()
    return list(pa.ipc.open_stream(pa.BufferReader(data)))


def bitmap_to_words(ctx, d_bitmap, n, row_offset, d_words):
    """FilterPOp::bufferBitMap on the device (fpdb_bitmap_to_words)."""
    C.check(C.lib().fpdb_bitmap_to_words(ctx.handle, d_bitmap, int(n), int(row_offset), d_words))


def bitmap_from_words(ctx, d_words, num_words, row_offset, n, d_bitmap):
    """FilterPOp::makeSelectionVector on the device (fpdb_bitmap_from_words)."""
    C.check(C.lib().fpdb_bitmap_from_words(ctx.handle, d_words, int(num_words), int(row_offset), int(n), d_bitmap))
