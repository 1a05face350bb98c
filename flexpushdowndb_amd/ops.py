"""Kernel-level operators over the C ABI (device columns in, device columns out).

These are the seams the reference's kernels expose
(HashJoinArrowKernel, BloomFilterCreate*Kernel / BloomFilterUseKernel,
GroupArrowKernel; fpdb-executor/src/physical/...), driven from Python for tests
and benchmarks.  The C++ operator layer (flexpushdowndb_amd/pop) drives the same
C ABI from the reference's PhysicalOp interface.
"""
import ctypes

import numpy as np

from . import _capi as C
from .device import DeviceColumn, Utf8Column


def _cols(cols, n=None):
    return C.columns_array([c.c(n) for c in cols])


def hash_keys(ctx, key_cols, n):
    """Arrow 6 HashMultiColumn hashes (RecordBatchHasher.cpp:69-104)."""
    out = ctx.empty(max(1, n), np.uint32)
    C.check(C.lib().fpdb_hash_keys(ctx.handle, _cols(key_cols), len(key_cols), n, out.ptr))
    return out


def bloom_masks():
    out = np.zeros(136, np.uint8)
    C.check(C.lib().fpdb_bloom_masks(out.ctypes.data))
    return out


class BlockedBloomFilter:
    """The default (Arrow) blocked bloom filter, device-resident
    (fpdb-tuple/src/arrow/exec/BloomFilter.cpp, ArrowBloomFilter.cpp)."""

    MAX_INPUT_SIZE = 20_000_000   # BLOOM_FILTER_MAX_INPUT_SIZE, physical/Globals.h:55

    def __init__(self, ctx, capacity):
        self.ctx = ctx
        self.capacity = int(capacity)
        self.valid = self.capacity <= self.MAX_INPUT_SIZE
        self.log_num_blocks = int(C.lib().fpdb_bloom_blocked_log_num_blocks(self.capacity))
        self.blocks = ctx.zeros(1 << self.log_num_blocks, np.uint64) if self.valid else None

    @classmethod
    def build(cls, ctx, key_cols, n, select=None, capacity=None):
        f = cls(ctx, n if capacity is None else capacity)
        if f.valid:
            C.check(C.lib().fpdb_bloom_blocked_build(ctx.handle, _cols(key_cols), len(key_cols), n,
                                                     select.ptr if select is not None else None,
                                                     f.log_num_blocks, f.blocks.ptr))
        return f

    def insert(self, key_cols, n, select=None):
        C.check(C.lib().fpdb_bloom_blocked_insert(self.ctx.handle, _cols(key_cols), len(key_cols), n,
                                                  select.ptr if select is not None else None,
                                                  self.log_num_blocks, self.blocks.ptr))

    def find(self, key_cols, n):
        """Selection bitmap (device, (n+7)/8 bytes) and selected count."""
        bm = self.ctx.zeros(max(8, ((n + 63) // 64) * 8), np.uint8)
        cnt = self.ctx.zeros(1, np.int64)
        C.check(C.lib().fpdb_bloom_blocked_find(self.ctx.handle, self.blocks.ptr, self.log_num_blocks,
                                                _cols(key_cols), len(key_cols), n, bm.ptr, cnt.ptr))
        return bm, int(cnt.numpy()[0])

    def filter(self, cols, key_index, n, want_bitmap=True):
        """Fused Find + compaction of all columns (BloomFilterUseKernel.cpp:180-207);
        with utf8 columns, Find and then a compaction that carries them."""
        if any(isinstance(c, Utf8Column) for c in cols):
            bm, _ = self.find([cols[i] for i in key_index], n)
            out, kept = filter_compact(self.ctx, cols, n, bm)
            return out, kept, bm if want_bitmap else None
        out = [DeviceColumn.empty(self.ctx, n, c.dtype, c.validity is not None, c.type_id) for c in cols]
        bm = self.ctx.zeros(max(8, ((n + 63) // 64) * 8), np.uint8) if want_bitmap else None
        ki = (ctypes.c_int * len(key_index))(*key_index)
        nout = ctypes.c_int64()
        C.check(C.lib().fpdb_bloom_blocked_filter(self.ctx.handle, self.blocks.ptr, self.log_num_blocks,
                                                  _cols(cols), len(cols), ki, len(key_index), n,
                                                  bm.ptr if bm is not None else None,
                                                  C.columns_array([o.c() for o in out]), ctypes.byref(nout)))
        for o in out:
            o.length = nout.value
        return out, nout.value, bm

    def blocks_numpy(self):
        return self.blocks.numpy()


class LegacyBloomFilter:
    """The in-tree universal-hash bloom filter with a target false-positive
    rate (fpdb-executor/src/physical/bloomfilter/BloomFilter.cpp)."""

    MAX_INPUT_SIZE = 20_000_000

    def __init__(self, ctx, capacity, fpr, hash_functions=None, seed=0):
        self.ctx = ctx
        self.capacity = int(capacity)
        self.fpr = float(fpr)
        self.valid = self.capacity <= self.MAX_INPUT_SIZE
        k, m = ctypes.c_int64(), ctypes.c_int64()
        C.check(C.lib().fpdb_bloom_legacy_params(self.capacity, self.fpr, ctypes.byref(k), ctypes.byref(m)))
        self.num_bits = m.value
        if self.capacity == 0:
            self.k = 0      # BloomFilter::makeHashFunctions, BloomFilter.cpp:209-213
        else:
            self.k = k.value
        if hash_functions is None:
            hf = (C.UniversalHash * max(1, self.k))()
            if self.k:
                C.check(C.lib().fpdb_bloom_legacy_make_hash_functions(seed, self.num_bits, self.k, hf))
            self.hash_functions = np.array([[h.a, h.b, h.m, h.p] for h in hf[:self.k]], np.int64).reshape(-1, 4)
        else:
            self.hash_functions = np.asarray(hash_functions, np.int64).reshape(-1, 4)
            self.k = len(self.hash_functions)
        self.words = ctx.zeros(self.num_bits // 64 + 1, np.int64)

    def _hf(self):
        hf = (C.UniversalHash * max(1, self.k))()
        for i, (a, b, m, p) in enumerate(self.hash_functions):
            hf[i] = C.UniversalHash(int(a), int(b), int(m), int(p))
        return hf

    def build(self, key_cols, n, select=None):
        C.check(C.lib().fpdb_bloom_legacy_build(self.ctx.handle, _cols(key_cols), len(key_cols), n,
                                                select.ptr if select is not None else None, self._hf(),
                                                self.k, self.num_bits, self.words.ptr))
        return self

    def find(self, key_cols, n):
        bm = self.ctx.zeros(max(8, ((n + 63) // 64) * 8), np.uint8)
        cnt = self.ctx.zeros(1, np.int64)
        C.check(C.lib().fpdb_bloom_legacy_find(self.ctx.handle, self._hf(), self.k, self.capacity,
                                               self.num_bits, self.words.ptr, _cols(key_cols), len(key_cols), n, bm.ptr,
                                               cnt.ptr))
        return bm, int(cnt.numpy()[0])


_CMP_OPS = {"<": C.CMP_LT, "<=": C.CMP_LE, ">": C.CMP_GT, ">=": C.CMP_GE, "==": C.CMP_EQ, "!=": C.CMP_NE}


def compare_bitmap(ctx, col, op, constant, n, combine=None, combine_or=False, out=None):
    """Selection bitmap of `col op constant` (nulls unselected), AND-ed with the
    bitmap `combine` (OR-ed when combine_or); `out` may be `combine` itself.
    Returns (bitmap, selected count)."""
    code = _CMP_OPS[op] if isinstance(op, str) else int(op)
    if combine_or:
        code |= C.CMP_OR
    bm = out if out is not None else ctx.empty(max(8, (n + 63) // 64 * 8), np.uint8)
    cnt = ctx.zeros(1, np.int64)
    if isinstance(col, Utf8Column):
        b = constant.encode() if isinstance(constant, str) else bytes(constant)
        cc = col.c()
        C.check(C.lib().fpdb_compare_bitmap_utf8(ctx.handle, ctypes.byref(cc), code, b, len(b),
                                                 combine.ptr if combine is not None else None, n, bm.ptr, cnt.ptr))
        return bm, int(cnt.numpy()[0])
    is_f = np.dtype(col.dtype).kind == "f"
    C.check(C.lib().fpdb_compare_bitmap(ctx.handle, ctypes.byref(col.c()), code, 0 if is_f else int(constant),
                                        float(constant) if is_f else 0.0, combine.ptr if combine is not None else None,
                                        n, bm.ptr, cnt.ptr))
    return bm, int(cnt.numpy()[0])


def filter_compact(ctx, cols, n, bitmap, capacity=None):
    """Order-preserving compaction of the rows whose bit is set into output
    columns of `capacity` rows (default n); raises FpdbError with
    FPDB_ERR_CAPACITY when more rows are selected than fit."""
    cap = n if capacity is None else capacity
    out = [Utf8Column.empty(ctx, cap, c.data_bytes, c.validity is not None) if isinstance(c, Utf8Column)
           else DeviceColumn.empty(ctx, cap, c.dtype, c.validity is not None, c.type_id, validity_rows=n)
           for c in cols]
    nout = ctypes.c_int64()
    C.check(C.lib().fpdb_filter_compact(ctx.handle, _cols(cols), len(cols), n, bitmap.ptr,
                                        C.columns_array([o.c() for o in out]), ctypes.byref(nout)))
    for o in out:
        o.length = nout.value
    return out, nout.value


def shuffle(ctx, key_cols, cols, n, num_slots):
    """Hash shuffle (ShuffleKernel2::shuffle, ShuffleKernel2.cpp:28-126): the
    columns reordered partition by partition, rows of partition p (key hash %
    num_slots == p) in input order at [offsets[p], offsets[p + 1]).
    Returns (output columns, offsets as an int64 numpy array)."""
    out = [DeviceColumn.empty(ctx, n, c.dtype, c.validity is not None, c.type_id) for c in cols]
    offsets = np.zeros(num_slots + 1, np.int64)
    C.check(C.lib().fpdb_shuffle(ctx.handle, _cols(key_cols), len(key_cols), _cols(cols), len(cols), n, num_slots,
                                 C.columns_array([o.c() for o in out]), None,
                                 offsets.ctypes.data_as(C._PI64)))
    return out, offsets


class _CacheBuffer:
    """A cache-owned device buffer (not freed by Python)."""

    def __init__(self, ctx, ptr, nbytes, dtype, length):
        self.ctx, self.ptr, self.nbytes, self.dtype, self.length = ctx, ptr, nbytes, np.dtype(dtype), length

    def numpy(self, n=None):
        return self.ctx.download(self.ptr, self.length if n is None else n, self.dtype)


class SegmentCache:
    """Device-resident segment cache (SegmentCache + LRUCachingPolicy,
    fpdb-cache/src/SegmentCache.cpp, fpdb-cache/src/policy/LRUCachingPolicy.cpp)."""

    def __init__(self, ctx, capacity_bytes, policy=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        C.check(C.lib().fpdb_cache_create(ctx.handle, int(capacity_bytes), policy, ctypes.byref(h)))
        self.handle = h

    def close(self):
        # after its context is closed the native object is gone with it
        if self.handle and self.ctx.handle:
            C.lib().fpdb_cache_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def store(self, partition, column, start, finish, data, valid=None):
        """data: a DeviceColumn, or a numpy array (host memory, copied in
        directly) with an optional boolean validity mask."""
        if isinstance(data, DeviceColumn):
            col = data.c()
        else:
            arr = np.ascontiguousarray(data)
            self._host_keepalive = arr
            vb = None
            if valid is not None:
                vb = np.packbits(np.asarray(valid, bool), bitorder="little")
                self._host_valid_keepalive = vb
            col = C.Column(C.NP_TO_TYPE[arr.dtype], 0, len(arr), 0, arr.ctypes.data,
                           vb.ctypes.data if vb is not None else None)
        stored = ctypes.c_int()
        C.check(C.lib().fpdb_cache_store(self.handle, partition.encode(), column.encode(), int(start), int(finish),
                                         ctypes.byref(col), ctypes.byref(stored)))
        return bool(stored.value)

    def load(self, partition, column, start=0, finish=0):
        """A pinned DeviceColumn over the cached buffers, or None on a miss."""
        out = C.Column()
        hit = ctypes.c_int()
        C.check(C.lib().fpdb_cache_load(self.handle, partition.encode(), column.encode(), int(start), int(finish),
                                        ctypes.byref(out), ctypes.byref(hit)))
        if not hit.value:
            return None
        dtype = C.TYPE_TO_NP[out.type]
        vals = _CacheBuffer(self.ctx, out.values, out.length * dtype.itemsize, dtype, out.length)
        valid = _CacheBuffer(self.ctx, out.validity, (out.length + 7) // 8, np.uint8, (out.length + 7) // 8) \
            if out.validity else None
        return DeviceColumn(vals, valid, out.type, out.length)

    def release(self, col):
        C.check(C.lib().fpdb_cache_release(self.handle, col.values.ptr))

    def remove(self, partition, column=None):
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_cache_remove(self.handle, partition.encode(),
                                          column.encode() if column is not None else None, ctypes.byref(n)))
        return n.value

    def clear(self):
        C.check(C.lib().fpdb_cache_clear(self.handle))

    def info(self):
        ns, hits, misses = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        used, free = ctypes.c_uint64(), ctypes.c_uint64()
        C.check(C.lib().fpdb_cache_info(self.handle, ctypes.byref(ns), ctypes.byref(used), ctypes.byref(free),
                                        ctypes.byref(hits), ctypes.byref(misses)))
        return {"segments": ns.value, "used": used.value, "free": free.value, "hits": hits.value,
                "misses": misses.value}

    def layout(self):
        """Keys in the usage queue, most recently used first."""
        need = ctypes.c_uint64()
        C.check(C.lib().fpdb_cache_layout(self.handle, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        C.check(C.lib().fpdb_cache_layout(self.handle, buf, need.value, ctypes.byref(need)))
        return [tuple(line.split("\t")) for line in buf.value.decode().splitlines()]


class HashJoinTable:
    """Device hash table of the build side (HashJoinArrowKernel build phase)."""

    def __init__(self, ctx, key_cols, n, payload=None, track_build_matches=False):
        self.ctx = ctx
        self.handle = ctypes.c_void_p()
        C.check(C.lib().fpdb_hj_build(ctx.handle, _cols(key_cols), len(key_cols), n,
                                      ctypes.byref(payload.c()) if payload is not None else None,
                                      1 if track_build_matches else 0, ctypes.byref(self.handle)))
        self.payload_type = payload.type_id if payload is not None else None

    def info(self):
        b, i, cap, u = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        C.check(C.lib().fpdb_hj_table_info(self.handle, ctypes.byref(b), ctypes.byref(i), ctypes.byref(cap),
                                           ctypes.byref(u)))
        return {"build_rows": b.value, "inserted": i.value, "capacity": cap.value, "unique": bool(u.value)}

    def form(self):
        """'dense' (direct-index table) or 'hash' (open addressing)."""
        f = ctypes.c_int()
        C.check(C.lib().fpdb_hj_table_form(self.handle, ctypes.byref(f)))
        return "dense" if f.value == 1 else "hash"

    def close(self):
        # after its context is closed the native object is gone with it
        if self.handle and self.ctx.handle:
            C.lib().fpdb_hj_table_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def probe_count(self, join_type, key_cols, n):
        cnt = ctypes.c_int64()
        C.check(C.lib().fpdb_hj_probe_count(self.ctx.handle, self.handle, join_type, _cols(key_cols),
                                            len(key_cols), n, ctypes.byref(cnt)))
        return cnt.value

    def probe(self, join_type, key_cols, n, probe_out, build_out, capacity=None, out=None):
        """build_out entries: DeviceColumn, or the string 'payload' for the inline payload.
        utf8 columns on either side go through the probe as row numbers and
        are taken afterwards."""
        outer_build = join_type in (C.JOIN_RIGHT, C.JOIN_FULL)
        if join_type in (C.JOIN_RIGHT_SEMI, C.JOIN_LEFT_SEMI):
            build_out = []
        if join_type == C.JOIN_LEFT_SEMI:
            probe_out = []
        if out is None and any(isinstance(c, Utf8Column) for c in list(probe_out) + list(build_out)):
            rows_p = iota(self.ctx, n)
            nb = self.info()["build_rows"]
            rows_b = iota(self.ctx, nb)
            rp = DeviceColumn(rows_p, None, C.TYPE_INT64, n)
            rb = DeviceColumn(rows_b, None, C.TYPE_INT64, nb)
            po = [rp if isinstance(c, Utf8Column) else c for c in probe_out]
            bo = [rb if isinstance(c, Utf8Column) else c for c in build_out]
            res, m = self.probe(join_type, key_cols, n, po, bo, capacity)
            src = list(probe_out) + list(build_out)
            for i, c in enumerate(src):
                if isinstance(c, Utf8Column):
                    res[i] = utf8_take(self.ctx, c, res[i], m)
            return res, m
        if capacity is None:
            capacity = self.probe_count(join_type, key_cols, n)
        bcols = []
        for b in build_out:
            if isinstance(b, str):
                bcols.append(C.Column(self.payload_type, 0, 0, 0, None, None))
            else:
                bcols.append(b.c())
        if out is None:
            out = []
            for p in probe_out:
                out.append(DeviceColumn.empty(self.ctx, capacity, p.dtype, p.validity is not None, p.type_id))
            for b in build_out:
                if isinstance(b, str):
                    out.append(DeviceColumn.empty(self.ctx, capacity, C.TYPE_TO_NP[self.payload_type],
                                                  outer_build, self.payload_type))
                else:
                    out.append(DeviceColumn.empty(self.ctx, capacity, b.dtype,
                                                  outer_build or b.validity is not None, b.type_id))
        nout = ctypes.c_int64()
        C.check(C.lib().fpdb_hj_probe(self.ctx.handle, self.handle, join_type, _cols(key_cols), len(key_cols), n,
                                      _cols(probe_out), len(probe_out), C.columns_array(bcols), len(bcols),
                                      C.columns_array([o.c() for o in out]), capacity, ctypes.byref(nout)))
        for o in out:
            o.length = nout.value
        return out, nout.value

    def build_side_rows(self, matched):
        b = self.info()["build_rows"]
        rows = self.ctx.empty(max(1, b), np.int64)
        nout = ctypes.c_int64()
        C.check(C.lib().fpdb_hj_build_side_rows(self.ctx.handle, self.handle, 1 if matched else 0, rows.ptr,
                                                max(1, b), ctypes.byref(nout)))
        return rows, nout.value


def column_min_max(ctx, col, n):
    """(min, max, valid count) of an integer device column."""
    lo, hi, cnt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    C.check(C.lib().fpdb_column_min_max(ctx.handle, ctypes.byref(col.c()), n, ctypes.byref(lo), ctypes.byref(hi),
                                        ctypes.byref(cnt)))
    return lo.value, hi.value, cnt.value


def packing_for(ranges):
    """(key_min, key_bits) per key from (min, max) ranges, or None when the
    packed key would need more than 64 bits."""
    out = []
    for lo, hi in ranges:
        span = int(hi) - int(lo)
        out.append((int(lo), max(1, span.bit_length())))
    return out if sum(b for _, b in out) <= 64 else None


def gather(ctx, cols, rows, n):
    out = [DeviceColumn.empty(ctx, n, c.dtype, True, c.type_id) for c in cols]
    C.check(C.lib().fpdb_gather(ctx.handle, _cols(cols, 0), len(cols), rows.ptr, n,
                                C.columns_array([o.c() for o in out])))
    return out


class GroupBy:
    """Accumulating hash group-by (GroupArrowKernel's aggregate node)."""

    def __init__(self, ctx, key_types, aggs, max_groups, packed=None):
        """aggs: list of (kind, input_type, nullable).  packed: (key_min, key_bits)
        per key column, for up to four keys packed into one 64-bit key."""
        self.ctx = ctx
        self.key_types = list(key_types)
        self.specs = [C.AggSpec(k, t if t is not None else 0, 1 if nl else 0, 0) for k, t, nl in aggs]
        kt = (ctypes.c_int32 * max(1, len(key_types)))(*key_types)
        sp = (C.AggSpec * max(1, len(aggs)))(*self.specs)
        self.handle = ctypes.c_void_p()
        if packed is None:
            C.check(C.lib().fpdb_groupby_create(ctx.handle, kt, len(key_types), sp, len(aggs), max_groups,
                                                ctypes.byref(self.handle)))
        else:
            mins = (ctypes.c_int64 * len(key_types))(*[int(m) for m, _ in packed])
            bits = (ctypes.c_int32 * len(key_types))(*[int(b) for _, b in packed])
            C.check(C.lib().fpdb_groupby_create_packed(ctx.handle, kt, len(key_types), mins, bits, sp, len(aggs),
                                                       max_groups, ctypes.byref(self.handle)))

    def close(self):
        # after its context is closed the native object is gone with it
        if self.handle and self.ctx.handle:
            C.lib().fpdb_groupby_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def consume(self, key_cols, value_cols, n):
        vals = [v.c() if v is not None else C.Column(0, 0, 0, 0, None, None) for v in value_cols]
        C.check(C.lib().fpdb_groupby_consume(self.ctx.handle, self.handle, _cols(key_cols),
                                             C.columns_array(vals), n))

    def num_groups(self):
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_groupby_num_groups(self.ctx.handle, self.handle, ctypes.byref(n)))
        return n.value

    def finalize(self):
        ng = self.num_groups()
        keys = [DeviceColumn.empty(self.ctx, ng, C.TYPE_TO_NP[t], False, t) for t in self.key_types]
        aggs = []
        for s in self.specs:
            ot = C.lib().fpdb_groupby_output_type(s)
            aggs.append(DeviceColumn.empty(self.ctx, ng, C.TYPE_TO_NP[ot], True, ot))
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_groupby_finalize(self.ctx.handle, self.handle, C.columns_array([k.c() for k in keys]),
                                              C.columns_array([a.c() for a in aggs]), ctypes.byref(n)))
        for c in keys + aggs:
            c.length = n.value
        return keys, aggs


# ---- utf8 columns and composite keys -------------------------------------------------

def _rows_col(rows, n):
    """A row-number column: a DeviceColumn (INT64/INT32) or a DeviceArray of int64."""
    if isinstance(rows, DeviceColumn):
        return rows.c(n)
    return C.Column(C.TYPE_INT64, 0, int(n), 0, rows.ptr, None)


def utf8_take(ctx, col, rows, n):
    """Rows `rows` (-1 or null = null row) of a utf8 column, as a new Utf8Column."""
    cin = col.c()
    rc = _rows_col(rows, n)
    nbytes = ctypes.c_int64()
    C.check(C.lib().fpdb_utf8_take_bytes(ctx.handle, ctypes.byref(cin), ctypes.byref(rc), n, ctypes.byref(nbytes)))
    out = Utf8Column.empty(ctx, n, nbytes.value, nullable=True)
    oc = out.c()
    C.check(C.lib().fpdb_utf8_take(ctx.handle, ctypes.byref(cin), ctypes.byref(rc), n, ctypes.byref(oc)))
    return out


def iota(ctx, n, start=0):
    out = ctx.empty(max(1, n), np.int64)
    C.check(C.lib().fpdb_iota_int64(ctx.handle, out.ptr, n, start))
    return out


def compare_bitmap_utf8(ctx, col, op, constant, n, combine=None, combine_or=False, out=None):
    """Selection bitmap of `col op constant` over a utf8 column (bytewise order)."""
    return compare_bitmap(ctx, col, op, constant, n, combine, combine_or, out)[0]


class KeyDict:
    """Dense int32 ids for key tuples of any width and type (utf8 included),
    in HBM: the key encoder the join and group-by operators use for keys their
    integer kernels do not take directly."""

    def __init__(self, ctx, key_types, hash_bits=64):
        self.ctx = ctx
        self.key_types = list(key_types)
        types = (ctypes.c_int32 * len(self.key_types))(*self.key_types)
        h = ctypes.c_void_p()
        C.check(C.lib().fpdb_keydict_create(ctx.handle, types, len(self.key_types), hash_bits, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            C.lib().fpdb_keydict_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_keydict_size(self.handle, ctypes.byref(n)))
        return n.value

    def encode(self, key_cols, n, insert=True, null_is_key=False):
        """INT32 ids of the rows (-1 and a cleared validity bit: absent, or a
        null key with null_is_key False)."""
        ids = DeviceColumn.empty(self.ctx, n, np.int32, nullable=True)
        oc = ids.c()
        total = ctypes.c_int64()
        C.check(C.lib().fpdb_keydict_encode(self.ctx.handle, self.handle, _cols(key_cols, n), len(key_cols), n,
                                            1 if insert else 0, 1 if null_is_key else 0, ctypes.byref(oc),
                                            ctypes.byref(total)))
        return ids

    def decode(self, key, ids, n):
        """Values of key column `key` for the given ids (a DeviceColumn)."""
        t = self.key_types[key]
        rc = ids.c(n)
        if t == C.TYPE_UTF8:
            nbytes = ctypes.c_int64()
            C.check(C.lib().fpdb_keydict_decode_bytes(self.ctx.handle, self.handle, key, ctypes.byref(rc), n,
                                                      ctypes.byref(nbytes)))
            out = Utf8Column.empty(self.ctx, n, nbytes.value, nullable=True)
        else:
            out = DeviceColumn.empty(self.ctx, n, C.TYPE_TO_NP[t], nullable=True, type_id=t)
        oc = out.c()
        C.check(C.lib().fpdb_keydict_decode(self.ctx.handle, self.handle, key, ctypes.byref(rc), n,
                                            ctypes.byref(oc)))
        return out
