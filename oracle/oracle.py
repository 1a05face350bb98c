"""numpy/ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It restates FlexPushdownDB's hot-path operators on the CPU so the
HIP path can be checked against it; it is never the thing measured on the GPU
and the product library never imports it.

Each function names the reference file:line it restates (paths relative to the
FlexPushdownDB source tree).  The C++ core is fpdb_oracle.cpp next to this file.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, u32, dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_double
        L.orc_bloom_masks.argtypes = [P]
        L.orc_bloom_masks.restype = i32
        L.orc_mt19937_raw.argtypes = [i32, P, i64]
        L.orc_uniform_u64_draws.argtypes = [P, i64]
        L.orc_blocked_log_num_blocks.argtypes = [i64]
        L.orc_blocked_log_num_blocks.restype = i32
        L.orc_blocked_insert.argtypes = [P, P, i32, P, i64]
        L.orc_blocked_find.argtypes = [P, P, i32, P, i64, P]
        L.orc_arrow6_hash_column.argtypes = [P, i32, P, i64, P, i32]
        L.orc_legacy_k.argtypes = [dbl]
        L.orc_legacy_k.restype = i64
        L.orc_legacy_m.argtypes = [i64, dbl]
        L.orc_legacy_m.restype = i64
        L.orc_legacy_keys.argtypes = [P, P, i32, i64, P]
        L.orc_legacy_add.argtypes = [P, i32, P, P, i64]
        L.orc_legacy_find.argtypes = [P, i32, i64, P, P, i64, P]
        L.orc_is_prime.argtypes = [ctypes.c_uint64]
        L.orc_is_prime.restype = i32
        L.orc_next_prime.argtypes = [ctypes.c_uint64]
        L.orc_next_prime.restype = ctypes.c_uint64
        L.orc_legacy_make_params.argtypes = [u32, i64, i32, P]
        L.orc_join_pairs.argtypes = [P, P, i64, P, P, i64, P, P, P]
        L.orc_join_pairs.restype = i64
        L.orc_groupby.argtypes = [P, i64, P, i32, P, i32, i64, P, P, P, P, P]
        L.orc_groupby.restype = i64
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --- blocked bloom filter ---------------------------------------------------

def bloom_masks():
    """136-byte mask bit vector (fpdb-tuple/src/arrow/exec/BloomFilter.cpp:16-79)."""
    out = np.zeros(136, np.uint8)
    lib().orc_bloom_masks(_p(out))
    return out


def mask_table_u64(masks=None):
    """The 1024 57-bit masks expanded to uint64 (BloomFilter.h:52-61)."""
    masks = bloom_masks() if masks is None else masks
    buf = np.zeros(144, np.uint8)
    buf[:136] = masks
    out = np.zeros(1024, np.uint64)
    for i in range(1024):
        w = int.from_bytes(bytes(buf[i // 8:i // 8 + 8]), "little")
        out[i] = (w >> (i % 8)) & ((1 << 57) - 1)
    return out


def blocked_log_num_blocks(n):
    return int(lib().orc_blocked_log_num_blocks(int(n)))


def arrow6_hash(columns):
    """uint32 key hash per row over fixed-width columns.

    columns: list of (values ndarray, validity bitmap ndarray or None).
    Restates Arrow 6 Hashing::HashMultiColumn as called at
    fpdb-tuple/src/RecordBatchHasher.cpp:97-100 (parity unpinned: the Arrow 6
    source is not in the reference tree).
    """
    n = len(columns[0][0])
    out = np.zeros(n, np.uint32)
    for i, (vals, valid) in enumerate(columns):
        vals = np.ascontiguousarray(vals)
        lib().orc_arrow6_hash_column(_p(vals), vals.dtype.itemsize, _p(valid), n, _p(out), 1 if i else 0)
    return out


def blocked_build(hashes, n_capacity=None):
    """BlockedBloomFilter built by BloomFilterBuilder_SingleThreaded
    (ArrowBloomFilter.cpp:29-75, BloomFilter.cpp:87-136)."""
    hashes = np.ascontiguousarray(hashes, dtype=np.uint32)
    cap = len(hashes) if n_capacity is None else n_capacity
    lnb = blocked_log_num_blocks(cap)
    blocks = np.zeros(1 << lnb, np.uint64)
    masks = bloom_masks()
    lib().orc_blocked_insert(_p(masks), _p(blocks), lnb, _p(hashes), len(hashes))
    return lnb, blocks


def blocked_find(lnb, blocks, hashes):
    """Packed selection bitmap (BloomFilter.cpp:138-170)."""
    hashes = np.ascontiguousarray(hashes, dtype=np.uint32)
    n = len(hashes)
    out = np.zeros((n + 7) // 8, np.uint8)
    masks = bloom_masks()
    lib().orc_blocked_find(_p(masks), _p(blocks), lnb, _p(hashes), n, _p(out))
    return out


# --- legacy bloom filter ----------------------------------------------------

def legacy_k(p):
    return int(lib().orc_legacy_k(p))


def legacy_m(n, p):
    return int(lib().orc_legacy_m(int(n), p))


def legacy_keys(cols):
    """hashCombine over std::hash of each column (ArrayHasher.cpp:12-19, Util.cpp:106-113)."""
    cols = [np.ascontiguousarray(c) for c in cols]
    ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    widths = (ctypes.c_int * len(cols))(*[c.dtype.itemsize for c in cols])
    out = np.zeros(len(cols[0]), np.int64)
    lib().orc_legacy_keys(ptrs, widths, len(cols), len(cols[0]), _p(out))
    return out


def legacy_params(seed, m, k):
    out = np.zeros((k, 4), np.int64)
    lib().orc_legacy_make_params(seed, m, k, _p(out))
    return out


def legacy_build(params, num_bits, keys):
    words = np.zeros(num_bits // 64 + 1, np.int64)     # BloomFilter.cpp:234-237
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    params = np.ascontiguousarray(params, dtype=np.int64)
    lib().orc_legacy_add(_p(params), len(params), _p(words), _p(keys), len(keys))
    return words


def legacy_find(params, capacity, words, keys):
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    params = np.ascontiguousarray(params, dtype=np.int64)
    out = np.zeros((len(keys) + 7) // 8, np.uint8)
    lib().orc_legacy_find(_p(params), len(params), int(capacity), _p(words), _p(keys), len(keys), _p(out))
    return out


def is_prime(n):
    return bool(lib().orc_is_prime(int(n)))


def next_prime(n):
    return int(lib().orc_next_prime(int(n)))


# --- hash join ----------------------------------------------------------------

def join_pairs(bkeys, pkeys, bvalid=None, pvalid=None, want_matched=False):
    """(probe_row, build_row) pairs of an inner equi-join in probe order
    (HashJoinArrowKernel.cpp:126-206 semantics; TupleSetIndex.cpp:44-88)."""
    bkeys = np.ascontiguousarray(bkeys, dtype=np.int64)
    pkeys = np.ascontiguousarray(pkeys, dtype=np.int64)
    L = lib()
    matched = np.zeros(len(bkeys), np.uint8) if want_matched else None
    cnt = L.orc_join_pairs(_p(bkeys), _p(bvalid), len(bkeys), _p(pkeys), _p(pvalid), len(pkeys),
                           None, None, None)
    op = np.zeros(cnt, np.int64)
    ob = np.zeros(cnt, np.int64)
    L.orc_join_pairs(_p(bkeys), _p(bvalid), len(bkeys), _p(pkeys), _p(pvalid), len(pkeys),
                     _p(op), _p(ob), _p(matched))
    if want_matched:
        return op, ob, matched.astype(bool)
    return op, ob


# --- group-by -------------------------------------------------------------------

def groupby(keys_u64, sums_i64=(), sums_f64=(), max_groups=None):
    """Grouped COUNT(*), SUM(int64) and SUM(float64) in first-appearance group
    order (GroupArrowKernel.cpp:202-243, Sum.cpp:63-74, Count.cpp:91-104)."""
    keys_u64 = np.ascontiguousarray(keys_u64, dtype=np.uint64)
    n = len(keys_u64)
    mg = int(max_groups or max(1, n))
    si = [np.ascontiguousarray(v, dtype=np.int64) for v in sums_i64]
    sf = [np.ascontiguousarray(v, dtype=np.float64) for v in sums_f64]
    pi = (ctypes.c_void_p * max(1, len(si)))(*[v.ctypes.data for v in si])
    pf = (ctypes.c_void_p * max(1, len(sf)))(*[v.ctypes.data for v in sf])
    ok = np.zeros(mg, np.uint64)
    oc = np.zeros(mg, np.int64)
    osi = np.zeros((max(1, len(si)), mg), np.int64)
    osf = np.zeros((max(1, len(sf)), mg), np.float64)
    ofr = np.zeros(mg, np.int64)
    ng = lib().orc_groupby(_p(keys_u64), n, pi, len(si), pf, len(sf), mg,
                           _p(ok), _p(oc), _p(osi), _p(osf), _p(ofr))
    if ng < 0:
        raise ValueError("more groups than max_groups")
    return {
        "keys": ok[:ng].copy(), "count": oc[:ng].copy(),
        "sum_i64": [osi[j, :ng].copy() for j in range(len(si))],
        "sum_f64": [osf[j, :ng].copy() for j in range(len(sf))],
        "first_row": ofr[:ng].copy(),
    }


# --- hash shuffle -----------------------------------------------------------------

def shuffle_order(hashes, num_slots):
    """Row order and slot offsets of ShuffleKernel2::shuffle
    (fpdb-executor/src/physical/shuffle/ShuffleKernel2.cpp:66-97): row r joins
    partition hashes[r] % num_slots (:68), partitions keep row order and are
    concatenated in slot order (:85-89).  Returns (row order, offsets)."""
    part = np.asarray(hashes, dtype=np.uint32) % np.uint32(num_slots)
    order = np.argsort(part, kind="stable")
    counts = np.bincount(part, minlength=num_slots)
    offsets = np.zeros(num_slots + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    return order, offsets


# --- segment cache (LRU policy) -------------------------------------------------

class LruCacheModel:
    """SegmentCache::store/load (fpdb-cache/src/SegmentCache.cpp:21-45) with
    LRUCachingPolicy (fpdb-cache/src/policy/LRUCachingPolicy.cpp:18-73): a
    usage queue, most recently used first; a store evicts from the back until
    the segment fits and declines a segment larger than the capacity; a load
    moves a present key to the front.  Keys are (partition, column).  A store
    of a key already present refreshes it (the reference would queue it
    twice; the device cache and this model both refresh)."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.free = capacity
        self.queue = []          # [(key, size)], front = most recent
        self.hits = 0
        self.misses = 0

    def _find(self, key):
        for i, (k, _) in enumerate(self.queue):
            if k == key:
                return i
        return -1

    def store(self, key, size):
        i = self._find(key)
        if i >= 0:
            self.queue.insert(0, self.queue.pop(i))
            return True
        if size > self.capacity:
            return False
        while self.free < size:
            _, s = self.queue.pop()
            self.free += s
        self.queue.insert(0, (key, size))
        self.free -= size
        return True

    def load(self, key):
        i = self._find(key)
        if i < 0:
            self.misses += 1
            return False
        self.queue.insert(0, self.queue.pop(i))
        self.hits += 1
        return True

    def remove(self, partition, column=None):
        keep = [(k, s) for k, s in self.queue if not (k[0] == partition and (column is None or k[1] == column))]
        removed = len(self.queue) - len(keep)
        self.free += sum(s for k, s in self.queue) - sum(s for k, s in keep)
        self.queue = keep
        return removed

    def layout(self):
        return [k for k, _ in self.queue]


# --- predicate-transfer plan ----------------------------------------------------

def pred_trans_plan(order, tables, joins):
    """Connection steps of the predicate-transfer orderings
    (fpdb-executor/src/physical/transform/pred-trans/SmallToLargePredTransOrder.cpp:37-214,
    BFSPredTransOrder.cpp:40-227), with insertion order where the reference
    iterates hash sets.

    order: "s2l", "bfs" or "yan"; tables: [(name, rows)]; joins:
    [(left, right, left_columns, right_columns, type)].  Returns (steps, ends)
    with steps = [(transfer_op, create_op, src, tgt, src_attached, tgt_attached)]
    and ends = {table: op its rows leave from}, or raises ValueError on a cycle
    (s2l)."""
    names = [t for t, _ in tables]
    rows = dict(tables)
    curr = {t: "scan_" + t for t in names}
    steps = []

    def pred(a, b):
        return "-".join(x + "-" + y for x, y in zip(a, b))

    def connect(create, use, src, tgt):
        steps.append((use, create, src, tgt, curr[src], curr[tgt]))
        curr[tgt] = use

    if order == "s2l":
        fw, bw = [], []
        need = {"F": {t: 0 for t in names}, "B": {t: 0 for t in names}}
        for i, (l, r, lc, rc, ty) in enumerate([j for j in joins if j[4] != "FULL"]):
            p = pred(lc, rc)
            if ty != "RIGHT":
                fw.append(("BloomFilterCreate(F)<%d>-%s" % (i, p), "BloomFilterUse(F)<%d>-%s" % (i, p), l, r))
                need["F"][r] += 1
            if ty != "LEFT":
                bw.append(("BloomFilterCreate(B)<%d>-%s" % (i, p), "BloomFilterUse(B)<%d>-%s" % (i, p), r, l))
                need["B"][l] += 1
        for tag, nodes in (("F", fw), ("B", bw)):
            done = {t: 0 for t in names}
            queue = [n for n in nodes if need[tag][n[2]] == 0]
            seen = set(queue)
            while queue:
                create, use, src, tgt = queue.pop(0)
                connect(create, use, src, tgt)
                done[tgt] += 1
                if done[tgt] == need[tag][tgt]:
                    for n in nodes:
                        if n[2] == tgt and n not in seen:
                            queue.append(n)
                            seen.add(n)
            if len(seen) != len(nodes):
                raise ValueError("cycle (%s predicate transfer)" % ("forward" if tag == "F" else "backward"))
        return steps, curr

    # BFS from the largest table (Yannakakis: the first one)
    nbrs = {t: [] for t in names}
    in_graph = []
    for l, r, lc, rc, ty in joins:
        if ty == "FULL":
            continue
        ld, rd = {"LEFT": ("fw", "bw"), "RIGHT": ("bw", "fw")}.get(ty, ("both", "both"))
        nbrs[l].append((r, ld, lc, rc))
        nbrs[r].append((l, rd, rc, lc))
        for t in (l, r):
            if t not in in_graph:
                in_graph.append(t)
    todo = [t for t in names if t in in_graph]
    stack = []
    while todo:
        if order == "yan":
            root = todo[0]
        else:
            root = todo[0]
            for t in todo:
                if rows[t] > rows[root]:
                    root = t
        todo.remove(root)
        q = [root]
        while q:
            u = q.pop(0)
            for v, d, own, other in nbrs[u]:
                if v not in todo:
                    continue
                stack.append((v, u, other, own, d in ("both", "bw"), d in ("both", "fw")))
                q.append(v)
                todo.remove(v)
    ident = [0]

    def transfer(src, tgt, sc, tc, tag):
        p = pred(sc, tc)
        i = ident[0]
        ident[0] += 1
        if order == "yan":
            name = "HashJoinArrow(%s)<%d>-%s" % (tag, i, p)
            steps.append((name, name, src, tgt, curr[src], curr[tgt]))
            curr[tgt] = name
        else:
            connect("BloomFilterCreate(%s)<%d>-%s" % (tag, i, p), "BloomFilterUse(%s)<%d>-%s" % (tag, i, p), src, tgt)

    back = []
    while stack:
        src, tgt, sc, tc, fwd, bwd = stack.pop()
        if fwd:
            transfer(src, tgt, sc, tc, "F")
        back.append((tgt, src, tc, sc, bwd, fwd))
    while back:
        src, tgt, sc, tc, fwd, _ = back.pop()
        if fwd:
            transfer(src, tgt, sc, tc, "B")
    return steps, curr


# --- wire forms -------------------------------------------------------------------

def bitmap_words(batches):
    """int64 words of a filter's selection over consecutive batches
    (FilterPOp::bufferBitMap, fpdb-executor/src/physical/filter/FilterPOp.cpp:264-279,
    with fpdb::util::setBit, fpdb-util/src/Util.cpp:115-117): row r of the
    concatenation is bit r % 64 of word r / 64; ceil(rows / 64) words.
    batches: list of boolean selection arrays."""
    total = sum(len(b) for b in batches)
    words = [0] * ((total + 63) // 64)
    row = 0
    for b in batches:
        for i in np.nonzero(np.asarray(b, bool))[0]:
            r = row + int(i)
            words[r // 64] |= 1 << (r % 64)
        row += len(b)
    return np.array([w - (1 << 64) if w >= 1 << 63 else w for w in words], dtype=np.int64)


def selection_from_words(words, start, n):
    """FilterPOp::makeSelectionVector (FilterPOp.cpp:280-297): rows
    [start, start + n) as a boolean array (fpdb::util::getBit, Util.cpp:123-125)."""
    w = [int(x) & ((1 << 64) - 1) for x in words]
    return np.array([(w[r // 64] >> (r % 64)) & 1 for r in range(start, start + n)], dtype=bool)


def factorize_tuples(columns, null_is_key=True):
    """Dense ids of key tuples in order of first appearance -- the group ids
    Arrow's Grouper hands the acero aggregate node (GroupArrowKernel.cpp:245-291)
    and that the group-by's first-appearance group order follows.  `columns`:
    sequences of equal length whose None entries are nulls.  Integers compare by
    value, floats by value (0.0 == -0.0), strings by their characters.  With
    null_is_key False a tuple holding a null gets id -1 (a join key never
    matches).  Pure Python: small inputs only."""
    ids = {}
    out = []
    n = len(columns[0]) if columns else 0
    for r in range(n):
        t = tuple(c[r] for c in columns)
        if not null_is_key and any(v is None for v in t):
            out.append(-1)
            continue
        t = tuple(0.0 if isinstance(v, float) and v == 0.0 else v for v in t)
        if t not in ids:
            ids[t] = len(ids)
        out.append(ids[t])
    return np.array(out, np.int64), list(ids.keys())


def lookup_tuples(dictionary, columns):
    """Ids of the tuples in `dictionary` (the key list factorize_tuples
    returned), -1 for tuples absent from it or holding a null."""
    index = {t: i for i, t in enumerate(dictionary)}
    n = len(columns[0]) if columns else 0
    out = []
    for r in range(n):
        t = tuple(c[r] for c in columns)
        if any(v is None for v in t):
            out.append(-1)
            continue
        t = tuple(0.0 if isinstance(v, float) and v == 0.0 else v for v in t)
        out.append(index.get(t, -1))
    return np.array(out, np.int64)
