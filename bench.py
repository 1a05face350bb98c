#!/usr/bin/env python3
"""Benchmark of the MI355X hot path of FlexPushdownDB.

Headline (one JSON line on stdout): hash-join probe throughput on the join
workload -- build 1,000,000 rows {int64 key (permutation), int32 pay}, probe
6,000,000 rows {int64 fk, int32 qty, float64 price}, inner join producing 5
columns (probe fk, qty, price, build key, pay).  One step = build the table +
probe and materialise all 6M output rows, inputs resident in HBM.
value = probe rows / step time (G rows/s).

Also reported (same line): bloom-filter use on the predicate-transfer edge
(60,000,000 int64 probe keys against a filter built from ~400k kept keys) and
hash group-by over 60,000,000 Zipf(1.1) rows, each with its kernel's fraction
of HBM peak, plus the CPU oracle's join timed on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "join-probe Grows/s + % of HBM peak; batched vs plain speedup; p99 query ms"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(ctx, name):
    from flexpushdowndb_amd import _capi as C
    t, n = ctypes.c_double(), ctypes.c_int64()
    C.check(C.lib().fpdb_profile_kernel(ctx.handle, name.encode(), ctypes.byref(t), ctypes.byref(n)))
    return (t.value / n.value if n.value else float("nan")), n.value


def load_traffic(name):
    """HBM bytes per launch of a kernel group, measured with rocprofv3 --pmc
    (the L2's memory-side read and write requests by size) by
    `python -m flexpushdowndb_amd.profiling` and committed under profiles/ as a
    small JSON summary."""
    p = os.path.join(ROOT, "profiles", f"{name}.traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def bench_join(ctx, steps, warmup, spread=1, w=None, n_build=1_000_000):
    """spread = 1: the workload as specified (build keys a permutation of
    0..n-1, so the table takes its dense form); spread = 3: every key k
    becomes 3k + 1 (same rows, same matches), so the keys span 3x the build
    rows and the table takes its open-addressing hash form.  n_build: build
    rows of a generated workload (the probe keys stay uniform over them)."""
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import datagen
    from flexpushdowndb_amd.device import DeviceColumn
    L = C.lib()
    n_probe = 6_000_000
    if w is None:
        log("join: generating inputs")
        w = datagen.join_workload(n_build, n_probe)
    n_build = len(w["build_key"])
    if spread != 1:
        w = dict(w)
        w["build_key"] = w["build_key"] * spread + 1
        w["probe_fk"] = w["probe_fk"] * spread + 1
    bk = DeviceColumn.from_numpy(ctx, w["build_key"])
    bp = DeviceColumn.from_numpy(ctx, w["build_pay"])
    pf = DeviceColumn.from_numpy(ctx, w["probe_fk"])
    pq = DeviceColumn.from_numpy(ctx, w["probe_qty"])
    pp = DeviceColumn.from_numpy(ctx, w["probe_price"])
    outs = [DeviceColumn.empty(ctx, n_probe, np.int64), DeviceColumn.empty(ctx, n_probe, np.int32),
            DeviceColumn.empty(ctx, n_probe, np.float64), DeviceColumn.empty(ctx, n_probe, np.int64),
            DeviceColumn.empty(ctx, n_probe, np.int32)]
    d_count = ctx.zeros(1, np.int64)
    # argument blocks built once; the timed loop only calls the C ABI
    a_bkeys = C.columns_array([bk.c()])
    a_pay = bp.c()
    a_pkeys = C.columns_array([pf.c()])
    a_pout = C.columns_array([pf.c(), pq.c(), pp.c()])
    a_bout = C.columns_array([bk.c(), C.Column(C.TYPE_INT32, 0, 0, 0, None, None)])
    a_out = C.columns_array([o.c() for o in outs])
    table = ctypes.c_void_p()

    form = ctypes.c_int(-1)

    def step():
        C.check(L.fpdb_hj_build(ctx.handle, a_bkeys, 1, n_build, ctypes.byref(a_pay), 0, ctypes.byref(table)))
        C.check(L.fpdb_hj_probe_async(ctx.handle, table, C.JOIN_INNER, a_pkeys, 1, n_probe, a_pout, 3,
                                      a_bout, 2, a_out, n_probe, d_count.ptr))
        C.check(L.fpdb_hj_table_form(table, ctypes.byref(form)))
        L.fpdb_hj_table_destroy(table)

    for _ in range(warmup):
        step()
    ctx.synchronize()
    L.fpdb_profile_reset(ctx.handle)
    L.fpdb_context_set_profiling(ctx.handle, 1)
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        step()
    ms_total = ctx.timer_stop()
    ctx.synchronize()
    probe_ms, probe_n = kernel_ms(ctx, "hj_probe")
    build_ms, _ = kernel_ms(ctx, "hj_build")
    L.fpdb_context_set_profiling(ctx.handle, 0)
    # validate the last step's output: closed form of the permutation build side
    cnt = int(d_count.numpy()[0])
    inv = np.empty(n_build, np.int64)
    inv[(w["build_key"] - 1) // spread if spread != 1 else w["build_key"]] = np.arange(n_build)
    pidx = (w["probe_fk"] - 1) // spread if spread != 1 else w["probe_fk"]
    ok = (cnt == n_probe and np.array_equal(outs[0].numpy(), w["probe_fk"])
          and np.array_equal(outs[4].numpy(), w["build_pay"][inv[pidx]])
          and np.array_equal(outs[2].numpy(), w["probe_price"])
          and np.array_equal(outs[3].numpy(), w["probe_fk"]))
    if not ok:
        raise SystemExit("join output does not match the closed form")
    ms_step = ms_total / steps
    alg_bytes = n_probe * (8 + 4 + 8) + n_probe * 12 + n_probe * 32   # read + build gather + write
    return {
        "ms_per_step": ms_step,
        "value": n_probe / (ms_step * 1e-3) / 1e9,
        "probe_ms": probe_ms, "build_ms": build_ms, "probe_launches": probe_n,
        "alg_bytes": alg_bytes, "inputs": w, "table_form": "dense" if form.value == 1 else "hash",
    }


def bench_bloom(ctx, steps):
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import datagen, ops
    from flexpushdowndb_amd.device import DeviceColumn
    from oracle import oracle as O
    L = C.lib()
    log("bloom: generating inputs (60M probe keys)")
    w = datagen.bloom_workload()
    kept = w["kept"]
    kc = DeviceColumn.from_numpy(ctx, kept)
    f = ops.BlockedBloomFilter.build(ctx, [kc], len(kept))
    n = len(w["probe"])
    pc = DeviceColumn.from_numpy(ctx, w["probe"])
    bm = ctx.zeros(((n + 63) // 64) * 8, np.uint8)
    cnt = ctx.zeros(1, np.int64)
    a_keys = C.columns_array([pc.c()])
    out = [DeviceColumn.empty(ctx, n, np.int64)]
    a_out = C.columns_array([out[0].c()])
    ki = (ctypes.c_int * 1)(0)
    nout = ctypes.c_int64()
    for _ in range(2):
        C.check(L.fpdb_bloom_blocked_find(ctx.handle, f.blocks.ptr, f.log_num_blocks, a_keys, 1, n, bm.ptr, cnt.ptr))
    ctx.synchronize()
    L.fpdb_profile_reset(ctx.handle)
    L.fpdb_context_set_profiling(ctx.handle, 1)
    for _ in range(steps):
        C.check(L.fpdb_bloom_blocked_find(ctx.handle, f.blocks.ptr, f.log_num_blocks, a_keys, 1, n, bm.ptr, cnt.ptr))
    for _ in range(max(3, steps // 4)):
        C.check(L.fpdb_bloom_blocked_filter(ctx.handle, f.blocks.ptr, f.log_num_blocks, a_keys, 1, ki, 1, n,
                                            None, a_out, ctypes.byref(nout)))
        C.check(L.fpdb_bloom_blocked_insert(ctx.handle, C.columns_array([kc.c()]), 1, len(kept), None,
                                            f.log_num_blocks, f.blocks.ptr))
    ctx.synchronize()
    find_ms, _ = kernel_ms(ctx, "bloom_blocked_find")
    filt_ms, _ = kernel_ms(ctx, "bloom_blocked_filter")
    ins_ms, _ = kernel_ms(ctx, "bloom_blocked_insert")
    L.fpdb_context_set_profiling(ctx.handle, 0)
    # bit array and a slice of the bitmap against the oracle
    _, ref = O.blocked_build(O.arrow6_hash([(kept, None)]), n_capacity=len(kept))
    ok = np.array_equal(f.blocks_numpy(), ref)
    head = O.blocked_find(f.log_num_blocks, ref, O.arrow6_hash([(w["probe"][:1_000_000], None)]))
    ok = ok and np.array_equal(bm.numpy(125_000), head) and int(cnt.numpy()[0]) == nout.value
    if not ok:
        raise SystemExit("bloom filter output does not match the oracle")
    find_bytes = n * 8 + (n + 7) // 8
    filt_bytes = n * 8 + nout.value * 8
    # the same probe column against a filter of 100,000 of the kept keys
    # (2^14 blocks, 128 KiB): every workgroup holds the whole filter in LDS
    small = kept[:100_000]
    fs = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, small)], len(small))
    ctx.synchronize()
    L.fpdb_profile_reset(ctx.handle)
    L.fpdb_context_set_profiling(ctx.handle, 1)
    for _ in range(steps):
        C.check(L.fpdb_bloom_blocked_find(ctx.handle, fs.blocks.ptr, fs.log_num_blocks, a_keys, 1, n, bm.ptr, cnt.ptr))
    ctx.synchronize()
    small_ms, _ = kernel_ms(ctx, "bloom_blocked_find")
    L.fpdb_context_set_profiling(ctx.handle, 0)
    _, ref_s = O.blocked_build(O.arrow6_hash([(small, None)]), n_capacity=len(small))
    if not (np.array_equal(fs.blocks_numpy(), ref_s) and np.array_equal(
            bm.numpy(125_000), O.blocked_find(fs.log_num_blocks, ref_s, O.arrow6_hash([(w["probe"][:1_000_000], None)])))):
        raise SystemExit("bloom filter (LDS-resident) output does not match the oracle")
    # CPU: the oracle's hash + FindImp over the same 60M probe keys, on one
    # host thread and on P threads over P slices of the probe column
    m = n
    t0 = time.perf_counter()
    O.blocked_find(f.log_num_blocks, ref, O.arrow6_hash([(w["probe"][:m], None)]))
    cpu_s = time.perf_counter() - t0
    p = host_cores()
    slices = np.array_split(w["probe"], p)
    cpu_p = run_parallel(lambda sl: O.blocked_find(f.log_num_blocks, ref, O.arrow6_hash([(sl, None)])), slices)
    legacy = bench_bloom_legacy(ctx, steps, w, pc, a_keys, bm, cnt)
    return {
        "legacy_filter": legacy,
        "cpu_baseline": {"value": m / cpu_p / 1e9, "unit": "Gkeys/s", "cores": p, "kind": "port",
                         "sample": f"the whole workload: oracle Arrow 6 hash + blocked-filter Find (C++) over all "
                                   f"{m} probe keys against the same 2^{f.log_num_blocks}-block filter, {p} host "
                                   f"threads over {p} slices, {cpu_p:.2f} s",
                         "single_core": {"value": m / cpu_s / 1e9, "unit": "Gkeys/s", "cores": 1, "kind": "port",
                                         "sample": f"the same on one host thread, {cpu_s:.2f} s"}},
        "workload": f"blocked bloom use: filter from {len(kept)} keys (2^{f.log_num_blocks} blocks), "
                    f"{n} int64 probe keys",
        "find_kernel_ms": find_ms, "find_gkeys_per_s": n / (find_ms * 1e-3) / 1e9,
        "find_hbm_frac": find_bytes / (find_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "find_traffic_bytes": load_traffic("bloom_blocked_find"),
        "filter_kernel_ms": filt_ms, "filter_hbm_frac": filt_bytes / (filt_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "selected": nout.value, "insert_kernel_ms": ins_ms,
        "lds_filter": {"workload": f"same {n} probe keys, filter from 100000 keys (2^{fs.log_num_blocks} blocks, "
                                   "held whole in every workgroup's LDS)",
                       "find_kernel_ms": small_ms, "find_hbm_frac": find_bytes / (small_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }


def bench_bloom_legacy(ctx, steps, w, pc, a_keys, bm, cnt):
    """The universal-hash filter on the same edge: built from the kept keys
    at a 1% target false-positive rate (k = 7), applied to the 60M probe
    keys.  Bit array and the whole selection bitmap checked against the
    oracle."""
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import ops
    from flexpushdowndb_amd.device import DeviceColumn
    from oracle import oracle as O
    L = C.lib()
    kept, probe = w["kept"], w["probe"]
    n = len(probe)
    lf = ops.LegacyBloomFilter(ctx, len(kept), 0.01, seed=2024)
    kc = DeviceColumn.from_numpy(ctx, kept)
    a_kept = C.columns_array([kc.c()])
    hf = lf._hf()

    def find():
        C.check(L.fpdb_bloom_legacy_find(ctx.handle, hf, lf.k, lf.capacity, lf.num_bits, lf.words.ptr, a_keys, 1, n,
                                         bm.ptr, cnt.ptr))

    lf.build([kc], len(kept))
    find()
    ctx.synchronize()
    L.fpdb_profile_reset(ctx.handle)
    L.fpdb_context_set_profiling(ctx.handle, 1)
    for _ in range(steps):
        find()
        C.check(L.fpdb_bloom_legacy_build(ctx.handle, a_kept, 1, len(kept), None, hf, lf.k, lf.num_bits,
                                          lf.words.ptr))
    ctx.synchronize()
    find_ms, _ = kernel_ms(ctx, "bloom_legacy_find")
    build_ms, _ = kernel_ms(ctx, "bloom_legacy_build")
    L.fpdb_context_set_profiling(ctx.handle, 0)
    find()
    okeys = O.legacy_keys([probe])
    ref_words = O.legacy_build(lf.hash_functions, lf.num_bits, O.legacy_keys([kept]))
    t0 = time.perf_counter()
    ref_bm = O.legacy_find(lf.hash_functions, len(kept), ref_words, okeys)
    cpu_s = time.perf_counter() - t0
    if not (np.array_equal(lf.words.numpy(), ref_words) and np.array_equal(bm.numpy(len(ref_bm)), ref_bm)):
        raise SystemExit("legacy bloom filter output does not match the oracle")
    find_bytes = n * 8 + (n + 7) // 8
    return {"workload": f"legacy universal-hash filter: {len(kept)} kept keys, p = 0.01 (k = {lf.k}, "
                        f"m = {lf.num_bits} bits), {n} int64 probe keys",
            "find_kernel_ms": find_ms, "find_gkeys_per_s": n / (find_ms * 1e-3) / 1e9,
            "find_hbm_frac": find_bytes / (find_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "build_kernel_ms": build_ms,
            "selected": int(np.unpackbits(ref_bm, bitorder="little")[:n].sum()),
            "cpu_baseline": {"value": n / cpu_s / 1e9, "unit": "Gkeys/s", "cores": 1, "kind": "port",
                             "sample": f"the whole workload: oracle legacy find (hashCombine key, k universal "
                                       f"hashes with 128-bit products) over all {n} probe keys, keys hashed "
                                       f"beforehand, {cpu_s:.2f} s"}}


def bench_groupby(ctx, steps):
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import datagen, ops
    from flexpushdowndb_amd.device import DeviceColumn
    L = C.lib()
    log("groupby: generating inputs (60M rows)")
    w = datagen.groupby_workload()
    n = len(w["key"])
    kc = DeviceColumn.from_numpy(ctx, w["key"])
    vals = [None, DeviceColumn.from_numpy(ctx, w["qty"]), DeviceColumn.from_numpy(ctx, w["price"]),
            DeviceColumn.from_numpy(ctx, w["disc"])]
    aggs = [(C.AGG_COUNT_ALL, None, False), (C.AGG_SUM, C.TYPE_INT64, False), (C.AGG_SUM, C.TYPE_INT64, False),
            (C.AGG_SUM, C.TYPE_FLOAT64, False)]
    g = None
    wall = []
    for i in range(steps + 1):
        if g is not None:
            g.close()
        g = ops.GroupBy(ctx, [C.TYPE_INT32], aggs, 200_000)
        ctx.synchronize()
        if i == 1:
            L.fpdb_profile_reset(ctx.handle)
            L.fpdb_context_set_profiling(ctx.handle, 1)
        t0 = time.perf_counter()
        g.consume([kc], vals, n)
        ctx.synchronize()
        if i >= 1:
            wall.append((time.perf_counter() - t0) * 1e3)
    # consume = sample, count pass + offsets, main pass, head merge + compact pass
    sample_ms, _ = kernel_ms(ctx, "groupby_sample")
    count_ms, _ = kernel_ms(ctx, "groupby_count")
    main_ms, _ = kernel_ms(ctx, "groupby_main")
    compact_ms, _ = kernel_ms(ctx, "groupby_compact")
    cons_ms = sample_ms + count_ms + main_ms + compact_ms
    t0 = time.perf_counter()
    keys, out = g.finalize()
    finalize_ms = (time.perf_counter() - t0) * 1e3
    L.fpdb_context_set_profiling(ctx.handle, 0)
    k = keys[0].numpy()
    c = out[0].numpy()
    key = w["key"].astype(np.int64)
    ok = c.sum() == n and np.array_equal(c, np.bincount(key, minlength=100_000)[k])
    ok = ok and np.array_equal(out[1].numpy(), np.bincount(key, weights=w["qty"], minlength=100_000)[k].astype(np.int64))
    if not ok:
        raise SystemExit("group-by output does not match the oracle's closed form")
    alg = n * (4 + 8 + 8 + 8)
    # CPU: the oracle's hash group-by (C++ unordered_map, first-appearance
    # order) over the same 60M rows, on one host thread and on P threads over
    # P key partitions (each a partial group-by of its partition's groups);
    # pyarrow's acero group-by on all host cores when present
    from oracle import oracle as O
    m = n
    t0 = time.perf_counter()
    O.groupby(w["key"][:m].astype(np.uint64), sums_i64=(w["qty"][:m], w["price"][:m]), sums_f64=(w["disc"][:m],),
              max_groups=100_000)
    cpu_s = time.perf_counter() - t0
    p = host_cores()
    parts = [(w["key"][i].astype(np.uint64), w["qty"][i], w["price"][i], w["disc"][i]) for i in hash_parts(w["key"], p)]
    cpu_p = run_parallel(lambda q: O.groupby(q[0], sums_i64=(q[1], q[2]), sums_f64=(q[3],), max_groups=100_000), parts)
    del parts
    cpu = {"value": m / cpu_p / 1e9, "unit": "Grows/s", "cores": p, "kind": "port",
           "sample": f"the whole workload: oracle group-by (COUNT, SUM int64 x2, SUM float64) over all {m} rows, "
                     f"{p} host threads over {p} key partitions, {cpu_p:.2f} s",
           "single_core": {"value": m / cpu_s / 1e9, "unit": "Grows/s", "cores": 1, "kind": "port",
                           "sample": f"the same on one host thread, {cpu_s:.2f} s"}}
    try:
        import pyarrow as pa
        tbl = pa.table({"key": w["key"], "qty": w["qty"], "price": w["price"], "disc": w["disc"]})
        t0 = time.perf_counter()
        res = tbl.group_by("key", use_threads=True).aggregate(
            [("key", "count"), ("qty", "sum"), ("price", "sum"), ("disc", "sum")])
        ta = time.perf_counter() - t0
        assert res.num_rows == len(k)
        cpu["acero"] = {"value": m / ta / 1e9, "unit": "Grows/s", "cores": pa.cpu_count(), "kind": "acero",
                        "sample": f"pyarrow {pa.__version__} Table.group_by(...).aggregate (acero hash aggregate, "
                                  f"COUNT, SUM x3, use_threads) over all {m} rows, {ta:.2f} s"}
        del tbl, res
    except ImportError:
        pass
    return {"cpu_baseline": cpu,
            "workload": f"hash group-by: {n} rows, int32 Zipf(1.1) key over 100000 values, "
                        "COUNT, SUM int64 x2, SUM float64", "groups": int(len(k)),
            "consume_kernels_ms": cons_ms, "sample_ms": sample_ms, "count_ms": count_ms, "main_ms": main_ms,
            "compact_ms": compact_ms, "consume_wall_ms": statistics.median(wall),
            "grows_per_s": n / (cons_ms * 1e-3) / 1e9, "finalize_ms": finalize_ms,
            "hbm_frac": alg / (cons_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic_bytes": load_traffic("groupby_consume")}


def bench_star(ctx, steps, device_name):
    """Batched (persistent, graph-replayed, autotuned) vs plain per-partition
    execution of the 64-partition star join with full predicate transfer."""
    from flexpushdowndb_amd import datagen
    from flexpushdowndb_amd.starjoin import StarJoin, autotune
    log("star: generating inputs (60M fact rows, 64 partitions)")
    w = datagen.star_workload()
    exp = datagen.star_expected(w)
    sj = StarJoin(ctx, w)
    best, entry = autotune(sj, device_name)
    log(f"star: autotuned {best} ({entry['ms']:.3f} ms)")
    sj.capture()
    sj.replay()
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        sj.replay()
    batched_ms = ctx.timer_stop() / steps
    got = sj.result()
    ok = all(np.array_equal(got[k], exp[k]) for k in ("g_a", "g_b", "count", "sum"))
    plain = []
    for _ in range(3):
        t0 = time.perf_counter()
        sj.run_plain()
        ctx.synchronize()
        plain.append((time.perf_counter() - t0) * 1e3)
    got2 = sj.result()
    ok = ok and all(np.array_equal(got2[k], exp[k]) for k in ("g_a", "g_b", "count", "sum"))
    if not ok:
        raise SystemExit("star join result does not match the expected join + group-by")
    plain_ms = statistics.median(plain)
    stage_bytes = star_stage_bytes(w, sj.default_order())
    # CPU: the same query over all fact rows with the oracle's hash joins
    # (one per dimension, filtered dimension rows as the build side) and its
    # group-by
    from oracle import oracle as O
    m = len(w["measure"])
    t0 = time.perf_counter()
    keep = np.ones(m, bool)
    grp = {}
    for d, dim in enumerate(w["dims"]):
        rows = np.nonzero(dim["attr"] < dim["keep_below"])[0]
        op, ob = O.join_pairs(dim["key"][rows], w["fact_fk"][d][:m])
        hit = np.zeros(m, bool)
        hit[op] = True
        keep &= hit
        if d in (0, 3):
            g = np.zeros(m, np.int64)
            g[op] = dim["group"][rows[ob]]
            grp[d] = g
    key = (grp[0][keep].astype(np.uint64) << np.uint64(32)) | grp[3][keep].astype(np.uint64)
    O.groupby(key, sums_i64=(w["measure"][:m][keep],))
    cpu_s = time.perf_counter() - t0
    return {"cpu_baseline": {"value": m / cpu_s / 1e9, "unit": "G fact rows/s", "cores": 1, "kind": "port",
                             "sample": f"the whole workload, all {m} fact rows (no predicate transfer): 4 oracle "
                                       f"hash joins (C++ unordered_map) + oracle group-by, {cpu_s:.2f} s"},
            "workload": "star join, 60,000,000 fact rows in 64 lognormal(0,1) partitions, 4 dimensions "
                        "(30,000/200,000/2,000,000/2,556 rows kept at 20%/4%/1%/14%), forward + backward "
                        "blocked bloom filters, 4 joins, group-by on 2 dimension attributes",
            "batched_ms": batched_ms, "plain_ms": plain_ms, "speedup": plain_ms / batched_ms,
            "groups": int(len(got["count"])), "survivors": exp["survivors"], "tuning": list(best),
            "alg_bytes": stage_bytes["total"], "alg_bytes_by_stage": stage_bytes,
            "hbm_frac": stage_bytes["total"] / (batched_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def star_stage_bytes(w, order):
    """Algorithmic bytes of one batched star join, stage by stage from the
    workload's actual row counts: each dimension's key, filter and group
    columns read once; the fact pass reads the first filter's key column in
    full and each later filter's key column only for the rows still passing
    (the forward filters' real pass counts, false positives included, from the
    oracle's filters); survivors are written once (row, keys, measure) and read
    back by the backward-filter build and by the joins and group-by."""
    from oracle import oracle as O
    n = len(w["measure"])
    nd = len(w["dims"])
    dims = sum(len(d["key"]) * (8 + 4 + (4 if "group" in d else 0)) for d in w["dims"])
    passing = np.ones(n, bool)
    fact = 0
    for i, d in enumerate(order):
        dim = w["dims"][d]
        kept = dim["key"][dim["attr"] < dim["keep_below"]]
        lnb, blocks = O.blocked_build(O.arrow6_hash([(kept, None)]) if len(kept) else np.zeros(0, np.uint32),
                                      n_capacity=len(kept))
        fact += 8 * (n if i == 0 else int(passing.sum()))
        bits = O.blocked_find(lnb, blocks, O.arrow6_hash([(w["fact_fk"][d], None)]))
        passing &= np.unpackbits(bits, bitorder="little")[:n].astype(bool)
    surv = int(passing.sum())
    surv_write = surv * (8 + 8 * nd + 8) + surv * 8          # row, keys, measure written; measure read
    backward = surv * 8 * nd                                   # keys read for the backward filters
    joins = surv * (8 * nd + 8 + 8)                             # keys, measure and row read by the joins
    out = {"dimensions": dims, "fact_pass": fact, "survivors_written": surv_write, "backward_filters": backward,
           "joins_and_groupby": joins, "survivors": surv}
    out["total"] = dims + fact + surv_write + backward + joins
    return out


SERVE_RATES = ((50, 150), (200, 400), (800, 800))   # (queries/s, queries) per arrival rate


def bench_serve(ctx, rates=SERVE_RATES):
    """The serving loop over the resident 60M-row star schema: Poisson
    arrivals at each rate, replayed in real time by the native loop; p50/p99
    latency (arrival to completion) and throughput per rate."""
    from flexpushdowndb_amd import datagen
    from flexpushdowndb_amd.serve import Server, poisson_arrivals
    log("serve: generating the resident schema (60M fact rows)")
    w = datagen.serve_workload()
    srv = Server(ctx, w, max_active=8)
    # warm-up, and two queries checked against direct indexing
    warm = datagen.serve_queries(8, seed=99)
    ids = [srv.submit(q) for q in warm]
    srv.drain(ids)
    for q, i in list(zip(warm, ids))[:2]:
        got, exp = srv.result(i), datagen.serve_expected(w, q)
        if not all(np.array_equal(got[k], exp[k]) for k in ("g0", "g1", "count", "sum")):
            raise SystemExit("serving result does not match the expected join + group-by")
    out = {"workload": "serving: star schema resident in HBM (60,000,000 fact rows, 4 dimensions of "
                       "30,000/200,000/2,000,000/2,556 rows); each query filters 1-4 dimensions at log-uniform "
                       "selectivity in [0.1%, 50%] and groups by 1-2 dimension attributes; up to 8 queries "
                       "share a circular scan of the fact table in 8 chunks, admitted at any chunk",
           "rates": []}
    for k, (rate, n) in enumerate(rates):
        qs = datagen.serve_queries(n, seed=100 + k)
        arr = poisson_arrivals(n, rate, seed=200 + k)
        ids, lat, makespan = srv.run_trace(qs, arr)
        lat_ms = lat / 1e6
        out["rates"].append({"arrival_qps": rate, "queries": n, "p50_ms": float(np.percentile(lat_ms, 50)),
                             "p99_ms": float(np.percentile(lat_ms, 99)), "mean_ms": float(lat_ms.mean()),
                             "throughput_qps": n / (makespan / 1e9)})
        log(f"serve: {rate} q/s: p50 {out['rates'][-1]['p50_ms']:.3f} ms, p99 {out['rates'][-1]['p99_ms']:.3f} ms")
    srv.close()
    # CPU: the same kind of query answered on the host by direct indexing of
    # the dense dimension keys (numpy, one thread), one query after another
    qs = datagen.serve_queries(12, seed=300)
    t0 = time.perf_counter()
    for q in qs:
        datagen.serve_expected(w, q)
    cpu_s = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": len(qs) / cpu_s, "unit": "queries/s", "cores": 1, "kind": "port",
                           "sample": f"{len(qs)} serving queries over the 60M-row fact table answered one after "
                                     f"another by numpy direct indexing of the dense dimension keys (filters, join "
                                     f"by key index, group-by by first appearance), {cpu_s:.2f} s"}
    return out


def bench_shuffle(ctx, steps, n=60_000_000, num_slots=64):
    """Hash shuffle of n rows {int64 key (uniform), int32 row id, float64} on
    the key into num_slots slots (ShuffleKernel2::shuffle).  Checked in full
    through size-independent properties: every output row's key hashes to its
    slot, row ids ascend within each slot (stable), and the row ids are a
    permutation carrying their key and float64 values along."""
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import ops
    from flexpushdowndb_amd.device import DeviceColumn
    from oracle import oracle as O
    L = C.lib()
    log(f"shuffle: generating inputs ({n} rows)")
    rng = np.random.default_rng(2024)
    key = rng.integers(0, 2 ** 62, n, dtype=np.int64)
    rid = np.arange(n, dtype=np.int32)
    val = np.round(rng.uniform(900, 105000, n), 2)
    cols = [DeviceColumn.from_numpy(ctx, a) for a in (key, rid, val)]
    out = [DeviceColumn.empty(ctx, n, a.dtype) for a in (key, rid, val)]
    a_keys, a_in = C.columns_array([cols[0].c()]), C.columns_array([c.c() for c in cols])
    a_out = C.columns_array([o.c() for o in out])
    d_off = ctx.empty(num_slots + 1, np.int64)
    offsets = np.zeros(num_slots + 1, np.int64)
    poff = offsets.ctypes.data_as(C._PI64)
    for _ in range(2):
        C.check(L.fpdb_shuffle(ctx.handle, a_keys, 1, a_in, 3, n, num_slots, a_out, d_off.ptr, poff))
    L.fpdb_profile_reset(ctx.handle)
    L.fpdb_context_set_profiling(ctx.handle, 1)
    wall = []
    for _ in range(steps):
        t0 = time.perf_counter()
        C.check(L.fpdb_shuffle(ctx.handle, a_keys, 1, a_in, 3, n, num_slots, a_out, d_off.ptr, poff))
        wall.append((time.perf_counter() - t0) * 1e3)
    shuffle_ms, _ = kernel_ms(ctx, "shuffle")
    scatter_ms, _ = kernel_ms(ctx, "shuffle_scatter")
    L.fpdb_context_set_profiling(ctx.handle, 0)
    ok_key, ok_rid, ok_val = out[0].numpy(n), out[1].numpy(n), out[2].numpy(n)
    slot = np.repeat(np.arange(num_slots, dtype=np.uint32), np.diff(offsets))
    ok = offsets[-1] == n and np.array_equal(O.arrow6_hash([(ok_key, None)]) % np.uint32(num_slots), slot)
    ok = ok and np.array_equal(np.sort(ok_rid), rid)
    ok = ok and np.array_equal(key[ok_rid], ok_key) and np.array_equal(val[ok_rid], ok_val)
    d = np.diff(ok_rid.astype(np.int64))
    inside = np.ones(n - 1, bool)
    inside[offsets[1:-1] - 1] = False     # steps across a slot boundary may go down
    ok = ok and bool(np.all(d[inside] > 0))
    if not ok:
        raise SystemExit("shuffle output does not match the oracle's properties")
    alg = n * 20 * 2
    scatter_alg = n * 20 * 2 + n        # columns read and written, plus one partition-id byte a row
    scatter_gbs = scatter_alg / (scatter_ms * 1e-3) / 1e9
    # CPU: the oracle's stable partition + take on a 6,000,000-row slice
    m = 6_000_000
    t0 = time.perf_counter()
    order, _ = O.shuffle_order(O.arrow6_hash([(key[:m], None)]), num_slots)
    _ = (key[:m][order], rid[:m][order], val[:m][order])
    cpu_s = time.perf_counter() - t0
    return {"workload": f"hash shuffle: {n} rows {{int64 key, int32, float64}} on the key into {num_slots} slots",
            "kernels_ms": shuffle_ms, "scatter_kernel_ms": scatter_ms, "wall_ms": statistics.median(wall),
            "grows_per_s": n / (shuffle_ms * 1e-3) / 1e9,
            "alg_bytes": alg, "hbm_frac": alg / (shuffle_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "scatter_hbm_frac": alg / (scatter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "roofline": {"bound": "hbm", "kernel": "shuffle_scatter", "achieved": round(scatter_gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(scatter_gbs / HBM_PEAK_GBS, 4),
                         "traffic": load_traffic("shuffle_scatter"), "alg_bytes_per_launch": scatter_alg},
            "cpu_baseline": {"value": m / cpu_s / 1e9, "unit": "Grows/s", "cores": 1, "kind": "port",
                             "sample": f"oracle stable partition (numpy argsort of hash % {num_slots}) + take "
                                       f"of the three columns over the first {m} rows, {cpu_s:.2f} s"}}


def bench_cache(ctx, steps, w=None):
    """The join workload's probe side served through the device segment cache:
    a cold step stores the three probe columns from host memory (PCIe) into
    the cache and then builds, probes and materialises; a warm step loads
    them from the cache (a hit hands out the HBM buffers) and does the same
    join.  The same output either way (checked)."""
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd import datagen, ops
    from flexpushdowndb_amd.device import DeviceColumn
    L = C.lib()
    n_build, n_probe = 1_000_000, 6_000_000
    if w is None:
        log("cache: generating inputs")
        w = datagen.join_workload(n_build, n_probe)
    host = {"lo_fk": np.ascontiguousarray(w["probe_fk"]), "lo_qty": np.ascontiguousarray(w["probe_qty"]),
            "lo_price": np.ascontiguousarray(w["probe_price"])}
    bk = DeviceColumn.from_numpy(ctx, w["build_key"])
    bp = DeviceColumn.from_numpy(ctx, w["build_pay"])
    outs = [DeviceColumn.empty(ctx, n_probe, np.int64), DeviceColumn.empty(ctx, n_probe, np.int32),
            DeviceColumn.empty(ctx, n_probe, np.float64), DeviceColumn.empty(ctx, n_probe, np.int64),
            DeviceColumn.empty(ctx, n_probe, np.int32)]
    d_count = ctx.zeros(1, np.int64)
    a_bkeys = C.columns_array([bk.c()])
    a_pay = bp.c()
    a_bout = C.columns_array([bk.c(), C.Column(C.TYPE_INT32, 0, 0, 0, None, None)])
    a_out = C.columns_array([o.c() for o in outs])
    table = ctypes.c_void_p()
    cache = ops.SegmentCache(ctx, 1 << 30)

    def join(cols):
        a_pkeys = C.columns_array([cols[0].c()])
        a_pout = C.columns_array([c.c() for c in cols])
        C.check(L.fpdb_hj_build(ctx.handle, a_bkeys, 1, n_build, ctypes.byref(a_pay), 0, ctypes.byref(table)))
        C.check(L.fpdb_hj_probe_async(ctx.handle, table, C.JOIN_INNER, a_pkeys, 1, n_probe, a_pout, 3,
                                      a_bout, 2, a_out, n_probe, d_count.ptr))
        L.fpdb_hj_table_destroy(table)

    def cached_step():
        cols = []
        for name in host:
            col = cache.load("lineorder/part0", name)
            if col is None:
                cache.store("lineorder/part0", name, 0, n_probe, host[name])
                col = cache.load("lineorder/part0", name)
            cols.append(col)
        join(cols)
        ctx.synchronize()
        for col in cols:
            cache.release(col)

    times = {"cold": [], "warm": []}
    for i in range(steps + 1):
        cache.clear()
        for kind in ("cold", "warm"):
            t0 = time.perf_counter()
            cached_step()
            if i:
                times[kind].append((time.perf_counter() - t0) * 1e3)
    info = cache.info()
    ok = int(d_count.numpy()[0]) == n_probe and np.array_equal(outs[0].numpy(), w["probe_fk"]) \
        and np.array_equal(outs[2].numpy(), w["probe_price"])
    if not ok:
        raise SystemExit("join through the segment cache does not match the closed form")
    cache.close()
    cold, warm = statistics.median(times["cold"]), statistics.median(times["warm"])
    probe_bytes = sum(a.nbytes for a in host.values())
    return {"workload": "join workload (1,000,000-row build, 6,000,000-row probe) with the probe columns "
                        "{int64 fk, int32 qty, float64 price} served through the device segment cache",
            "cold_step_ms": cold, "warm_step_ms": warm, "speedup": cold / warm,
            "segment_bytes": probe_bytes, "host_to_hbm_gbs": probe_bytes / ((cold - warm) * 1e-3) / 1e9,
            "segments": info["segments"], "hits": info["hits"], "misses": info["misses"]}


def bench_wire(ctx, steps, n=60_000_000):
    """Wire forms: a 60M-row selection bitmap into the int64 words of bitmap
    pushdown (in 100,000-row batches at running offsets, as FilterPOp buffers
    it, and in one call) and back per batch; a 2^16-block filter exported as
    its two record batches, framed as an Arrow IPC stream, parsed and
    imported.  Checked against the oracle's words / the original blocks."""
    from flexpushdowndb_amd import datagen, ops, wire
    from flexpushdowndb_amd.device import DeviceColumn
    from oracle import oracle as O
    rng = np.random.default_rng(12)
    sel = rng.random(n) < 0.25
    packed = np.packbits(sel, bitorder="little")
    bm = np.zeros((n + 63) // 64 * 8, np.uint8)
    bm[:len(packed)] = packed
    d_bm = ctx.upload(bm)
    nwords = (n + 63) // 64
    d_words = ctx.zeros(nwords, np.int64)
    out = ctx.zeros((n + 63) // 64 * 8, np.uint8)
    wire.bitmap_to_words(ctx, d_bm.ptr, n, 0, d_words.ptr)
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        wire.bitmap_to_words(ctx, d_bm.ptr, n, 0, d_words.ptr)
    to_ms = ctx.timer_stop() / steps
    ctx.timer_start()
    for _ in range(steps):
        wire.bitmap_from_words(ctx, d_words.ptr, nwords, 0, n, out.ptr)
    from_ms = ctx.timer_stop() / steps
    words = d_words.numpy()
    ok = np.array_equal(words.view(np.uint8)[:len(packed)], packed) and np.array_equal(out.numpy(len(packed)), packed)
    # batch by batch at running offsets (FilterPOp::bufferBitMap)
    chunk = 100_000
    d_words2 = ctx.zeros(nwords, np.int64)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s0 in range(0, n, chunk):
        wire.bitmap_to_words(ctx, d_bm.ptr + s0 // 8, min(chunk, n - s0), s0, d_words2.ptr)
    ctx.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3
    ok = ok and np.array_equal(d_words2.numpy(), words)
    head = O.bitmap_words([sel[:64 * 1562]])      # whole words only
    ok = ok and np.array_equal(words[:len(head)], head)
    # the 2^16-block filter of the bloom workload through an IPC stream
    w = datagen.bloom_workload()
    f = ops.BlockedBloomFilter.build(ctx, [DeviceColumn.from_numpy(ctx, w["kept"])], len(w["kept"]))
    times = []
    for _ in range(max(3, steps // 4)):
        t0 = time.perf_counter()
        data = wire.ipc_bytes(wire.blocked_filter_record_batches(ctx, f.blocks.ptr, f.log_num_blocks))
        lnb, d_blocks = wire.blocked_filter_from_record_batches(ctx, wire.ipc_read(data))
        ctx.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ok = ok and lnb == f.log_num_blocks and np.array_equal(d_blocks.numpy(1 << lnb), f.blocks_numpy())
    if not ok:
        raise SystemExit("wire forms do not match the oracle")
    alg = nwords * 8 * 2
    return {"workload": f"{n}-row selection bitmap <-> int64 bitmap words; 2^{f.log_num_blocks}-block filter "
                        f"through an Arrow IPC stream ({len(data)} bytes)",
            "to_words_ms": to_ms, "from_words_ms": from_ms, "to_words_batched_ms": batched_ms,
            "to_words_hbm_frac": alg / (to_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "filter_round_trip_ms": statistics.median(times), "ipc_bytes": len(data)}


def bench_copy_calibration(ctx, steps, n=64 * 1024 * 1024):
    """A plain copy of n int64 values (8n bytes read, 8n written) with
    copy_rows_kernel: the reference point for the FETCH_SIZE / WRITE_SIZE
    counters of the profiled kernels."""
    from flexpushdowndb_amd import _capi as C
    from flexpushdowndb_amd.device import DeviceColumn
    L = C.lib()
    src = DeviceColumn.empty(ctx, n, np.int64)
    dst = DeviceColumn.empty(ctx, n, np.int64)
    ctx.fill(src.values.ptr, 1, src.values.nbytes)
    a_in, a_out = C.columns_array([src.c()]), C.columns_array([dst.c()])
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(steps):
        C.check(L.fpdb_copy_rows(ctx.handle, a_in, 1, n, a_out, 0))
    ms = ctx.timer_stop() / steps
    return {"workload": f"copy_rows_kernel over {n} int64 values", "bytes_read": 8 * n, "bytes_written": 8 * n,
            "ms": ms, "gbs": 16 * n / (ms * 1e-3) / 1e9}


def host_cores():
    """Host CPUs this process may run on (the GPU box's share, not the whole machine)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def run_parallel(fn, parts):
    """Runs fn over the parts on one host thread each (the oracle's C++ calls
    release the GIL) and returns the wall time in seconds."""
    import concurrent.futures
    with concurrent.futures.ThreadPoolExecutor(len(parts)) as ex:
        t0 = time.perf_counter()
        list(ex.map(fn, parts))
        return time.perf_counter() - t0


def hash_parts(key, p):
    """Row indices of each of p partitions of key % p (the reference's
    PARALLEL_DEGREE operator instances each get the rows of one partition)."""
    slot = (np.asarray(key).astype(np.int64) % p).astype(np.int64)
    order = np.argsort(slot, kind="stable")
    bounds = np.searchsorted(slot[order], np.arange(p + 1))
    return [order[bounds[i]:bounds[i + 1]] for i in range(p)]


def cpu_baseline_join(w, runs=3):
    """The oracle's hash join (C++ unordered_map build + probe, pairs only) on
    the whole join workload: one host thread, P host threads over P key
    partitions (P = host cores, inputs partitioned beforehand, as the
    reference's parallel operator instances receive shuffled inputs), and
    pyarrow's acero hash join on all host cores when pyarrow is present."""
    from oracle import oracle as O
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        op, _ = O.join_pairs(w["build_key"], w["probe_fk"])
        times.append(time.perf_counter() - t0)
        assert len(op) == len(w["probe_fk"])
    t = statistics.median(times)
    n = len(w["probe_fk"])
    p = host_cores()
    bparts, pparts = hash_parts(w["build_key"], p), hash_parts(w["probe_fk"], p)
    parts = [(w["build_key"][b], w["probe_fk"][q]) for b, q in zip(bparts, pparts)]
    tp = statistics.median(run_parallel(lambda bp: O.join_pairs(bp[0], bp[1]), parts) for _ in range(runs))
    out = {"value": n / tp / 1e9, "unit": "Grows/s", "cores": p, "kind": "port",
           "sample": f"whole join workload (1,000,000-row build, 6,000,000-row probe), oracle C++ unordered_map "
                     f"hash join producing (probe,build) row pairs, {p} host threads each joining one of {p} key "
                     f"partitions; median of {runs} runs, {tp:.3f} s each",
           "single_core": {"value": n / t / 1e9, "unit": "Grows/s", "cores": 1, "kind": "port",
                           "sample": f"the same oracle join on one host thread, {t:.3f} s"}}
    acero = acero_join(w, runs)
    if acero:
        out["acero"] = acero
    return out


def acero_join(w, runs=3):
    """pyarrow's acero hash join (the successor of the reference's Arrow 6
    delegate) on the join workload, all host cores; None without pyarrow."""
    try:
        import pyarrow as pa
    except ImportError:
        return None
    build = pa.table({"key": w["build_key"], "pay": w["build_pay"]})
    probe = pa.table({"fk": w["probe_fk"], "qty": w["probe_qty"], "price": w["probe_price"]})
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        res = probe.join(build, keys="fk", right_keys="key", join_type="inner", use_threads=True)
        times.append(time.perf_counter() - t0)
        assert res.num_rows == len(w["probe_fk"])
    t = statistics.median(times)
    return {"value": len(w["probe_fk"]) / t / 1e9, "unit": "Grows/s", "cores": pa.cpu_count(), "kind": "acero",
            "sample": f"pyarrow {pa.__version__} Table.join (acero hash join, inner, use_threads) over the whole "
                      f"join workload; median of {runs} runs, {t:.3f} s each"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", choices=["join", "bloom", "groupby", "shuffle", "cache", "wire", "star", "serve",
                                       "calib"],
                    default=None,
                    help="run one workload only (for profiler runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this implementation runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)
    from flexpushdowndb_amd.device import Context
    ctx = Context(0)
    info = ctx.device_info()
    log(f"device: {info}")
    res = {}
    if args.only == "calib":
        print(json.dumps(bench_copy_calibration(ctx, args.steps)))
        return
    if args.only in (None, "join"):
        res["join"] = bench_join(ctx, args.steps, args.warmup)
        log(f"join: {res['join']['ms_per_step']:.4f} ms/step, probe kernel {res['join']['probe_ms']:.4f} ms "
            f"({res['join']['table_form']} table)")
        hj = bench_join(ctx, args.steps, args.warmup, spread=3, w=res["join"]["inputs"])
        res["join_hash"] = {k: v for k, v in hj.items() if k != "inputs"}
        log(f"join (keys spread 3x): {hj['ms_per_step']:.4f} ms/step, probe kernel {hj['probe_ms']:.4f} ms "
            f"({hj['table_form']} table)")
        hs = bench_join(ctx, args.steps, args.warmup, spread=3, n_build=4_000)
        res["join_small_hash"] = {k: v for k, v in hs.items() if k != "inputs"}
        log(f"join (4,000 build rows, keys spread 3x): {hs['ms_per_step']:.4f} ms/step, probe kernel "
            f"{hs['probe_ms']:.4f} ms ({hs['table_form']} table)")
    if args.only in (None, "bloom"):
        res["bloom"] = bench_bloom(ctx, args.steps)
        log(f"bloom: find {res['bloom']['find_kernel_ms']:.4f} ms")
    if args.only in (None, "groupby"):
        res["groupby"] = bench_groupby(ctx, max(2, args.steps // 5))
        gb = res["groupby"]
        log(f"groupby: consume {gb['consume_kernels_ms']:.4f} ms (sample {gb['sample_ms']:.3f}, count {gb['count_ms']:.3f}, "
            f"main {gb['main_ms']:.3f}, "
            f"compact {gb['compact_ms']:.3f}; wall {gb['consume_wall_ms']:.3f})")
    if args.only in (None, "shuffle"):
        res["shuffle"] = bench_shuffle(ctx, args.steps)
        log(f"shuffle: {res['shuffle']['kernels_ms']:.4f} ms (scatter {res['shuffle']['scatter_kernel_ms']:.4f})")
    if args.only in (None, "cache"):
        res["cache"] = bench_cache(ctx, max(3, args.steps // 4), res["join"]["inputs"] if "join" in res else None)
        log(f"cache: cold {res['cache']['cold_step_ms']:.3f} ms, warm {res['cache']['warm_step_ms']:.3f} ms")
    if args.only in (None, "wire"):
        res["wire"] = bench_wire(ctx, args.steps)
        log(f"wire: to words {res['wire']['to_words_ms']:.4f} ms, filter round trip "
            f"{res['wire']['filter_round_trip_ms']:.3f} ms")
    if args.only in (None, "star"):
        res["star"] = bench_star(ctx, max(3, args.steps // 2), f"gfx950-{info['num_cus']}cu")
        log(f"star: batched {res['star']['batched_ms']:.3f} ms, plain {res['star']['plain_ms']:.3f} ms")
    if args.only in (None, "serve"):
        res["serve"] = bench_serve(ctx)
    if args.only not in (None, "join"):
        print(json.dumps({k: v for k, v in res[args.only].items() if k != "inputs"}))
        return
    j = res["join"]
    achieved = j["alg_bytes"] / (j["probe_ms"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(j["value"], 4),
        "unit": "Grows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(j["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic: seeded counter-based SplitMix64 columns (flexpushdowndb_amd/datagen.py)",
        "config": {"workload": "inner hash join: build 1,000,000 rows {int64 key = permutation, int32 pay}, "
                               "probe 6,000,000 rows {int64 fk, int32 qty, float64 price}; 5 output columns; "
                               "step = build + probe + materialise"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("hj_probe"),
                     "kernel": "hj_probe", "kernel_ms": round(j["probe_ms"], 5),
                     "alg_bytes_per_launch": j["alg_bytes"]},
        "build_kernel_ms": round(j["build_ms"], 5),
        "table_form": j["table_form"],
    }
    if "join_hash" in res:
        h = res["join_hash"]
        line["join_hash_table"] = {
            "workload": "the same join with every key k replaced by 3k + 1 (keys span 3x the build rows, so the "
                        "build takes the open-addressing hash form)",
            "value_grows_per_s": round(h["value"], 4), "ms_per_step": round(h["ms_per_step"], 5),
            "probe_kernel_ms": round(h["probe_ms"], 5), "build_kernel_ms": round(h["build_ms"], 5),
            "probe_hbm_frac": round(h["alg_bytes"] / (h["probe_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "probe_traffic": load_traffic("hj_probe_hash"),
            "table_form": h["table_form"]}
    if "join_small_hash" in res:
        h = res["join_small_hash"]
        line["join_small_hash_table"] = {
            "workload": "build 4,000 rows {int64 key 3k + 1 over a permutation, int32 pay}, the same 6,000,000 "
                        "probe rows shape (keys uniform over the build keys); an 8,192-slot (128 KiB) hash table",
            "value_grows_per_s": round(h["value"], 4), "ms_per_step": round(h["ms_per_step"], 5),
            "probe_kernel_ms": round(h["probe_ms"], 5), "build_kernel_ms": round(h["build_ms"], 5),
            "probe_hbm_frac": round(h["alg_bytes"] / (h["probe_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "table_form": h["table_form"]}
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle join on the host")
        line["cpu_baseline"] = cpu_baseline_join(j["inputs"])
    if "bloom" in res:
        line["bloom_use"] = res["bloom"]
    if "groupby" in res:
        line["groupby"] = res["groupby"]
    if "shuffle" in res:
        line["hash_shuffle"] = res["shuffle"]
    if "cache" in res:
        line["segment_cache"] = res["cache"]
    if "wire" in res:
        line["wire_forms"] = res["wire"]
    if "star" in res:
        line["batched_star_join"] = res["star"]
    if "serve" in res:
        line["serving"] = res["serve"]
    line["device"] = info
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
