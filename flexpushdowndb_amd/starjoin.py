"""Batched star join with predicate transfer over the C ABI (fpdb_star_*),
and the autotuner that picks its launch shape.

The autotuner times the batched form for each candidate tuning on the actual
inputs, keeps the fastest, and caches it per (workload signature, device) in a
JSON file (FPDB_AUTOTUNE_CACHE, default flexpushdowndb_amd/_build/autotune.json).
"""
import ctypes
import itertools
import json
import os

import numpy as np

from . import _capi as C
from .device import DeviceColumn

_PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(_PKG, "_build", "autotune.json")

# the autotuner's search, knob by knob: launch shape (workgroups per CU x
# hash-table load shift), then fact rows per thread per step, then the order of
# the forward filters over the fact rows
SHAPES = [(bpc, shift) for bpc in (1, 2, 4, 8) for shift in (1, 2)]
ITEMS = (4, 8, 16)
MAX_ORDERS = 24


class StarJoin:
    """One prepared star-join plan over device-resident columns."""

    def __init__(self, ctx, workload, group_dims=(0, 3), filter_order=None, tuning=None, max_groups=4096,
                 barrier_timeout_us=0, split_partitions=False):
        """split_partitions: hand each partition's fact columns over in
        device buffers of their own (one column per partition and dimension)
        instead of one contiguous fact table."""
        self.ctx = ctx
        self.barrier_timeout_us = int(barrier_timeout_us)
        self.w = workload
        dims = workload["dims"]
        self.cols = []          # keep device columns alive
        q = C.StarQuery()
        q.num_dims = len(dims)
        q.group_dims[0], q.group_dims[1] = group_dims
        order = filter_order if filter_order is not None else self.default_order()
        for i, d in enumerate(order):
            q.filter_order[i] = d
        for d, dim in enumerate(dims):
            kc = DeviceColumn.from_numpy(ctx, dim["key"])
            ac = DeviceColumn.from_numpy(ctx, dim["attr"])
            fc = DeviceColumn.from_numpy(ctx, workload["fact_fk"][d])
            self.cols += [kc, ac, fc]
            sd = C.StarDimension()
            sd.key, sd.filter_attr, sd.fact_fk = kc.c(), ac.c(), fc.c()
            sd.filter_op = C.CMP_LT
            sd.filter_constant = int(dim["keep_below"])
            if "group" in dim:
                gc = DeviceColumn.from_numpy(ctx, dim["group"])
                self.cols.append(gc)
                sd.has_group_attr = 1
                sd.group_attr = gc.c()
            q.dims[d] = sd
        mc = DeviceColumn.from_numpy(ctx, workload["measure"])
        self.cols.append(mc)
        q.fact_measure = mc.c()
        q.fact_rows = len(workload["measure"])
        offs = np.ascontiguousarray(workload["offsets"], dtype=np.int64)
        self._offs = offs
        q.num_partitions = len(offs) - 1
        q.partition_offsets = offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        if split_partitions:
            nd, npart = len(dims), len(offs) - 1
            fk_arr = (C.Column * (npart * nd))()
            m_arr = (C.Column * npart)()
            for p in range(npart):
                lo, hi = int(offs[p]), int(offs[p + 1])
                for d in range(nd):
                    pc = DeviceColumn.from_numpy(ctx, np.ascontiguousarray(workload["fact_fk"][d][lo:hi]))
                    self.cols.append(pc)
                    fk_arr[p * nd + d] = pc.c()
                pm = DeviceColumn.from_numpy(ctx, np.ascontiguousarray(workload["measure"][lo:hi]))
                self.cols.append(pm)
                m_arr[p] = pm.c()
            self._split = (fk_arr, m_arr)
            q.partition_fact_fk = ctypes.cast(fk_arr, ctypes.POINTER(C.Column))
            q.partition_measure = ctypes.cast(m_arr, ctypes.POINTER(C.Column))
        q.max_groups = max_groups
        self.q = q
        self.max_groups = max_groups
        self.handle = None
        self.prepare(tuning)

    def default_order(self):
        """Most selective dimension first (kept fraction x key-range coverage)."""
        sel = [dim["keep_below"] / 100.0 for dim in self.w["dims"]]
        return list(np.argsort(sel, kind="stable"))

    def set_order(self, order):
        for i, d in enumerate(order):
            self.q.filter_order[i] = int(d)

    def order(self):
        return [int(self.q.filter_order[i]) for i in range(self.q.num_dims)]

    def prepare(self, tuning=None):
        """tuning: (blocks per CU, hash-table load shift[, items per thread[,
        forward filter order]]) or None for the defaults."""
        self.close()
        t = None
        if tuning is not None and len(tuning) > 3:
            self.set_order(tuning[3])
        if tuning is not None or self.barrier_timeout_us:
            bpc, shift = tuning[:2] if tuning is not None else (0, 1)
            items = tuning[2] if tuning is not None and len(tuning) > 2 else 0
            t = C.StarTuning(int(bpc), self.barrier_timeout_us, int(shift), int(items))
        h = ctypes.c_void_p()
        C.check(C.lib().fpdb_star_prepare(self.ctx.handle, ctypes.byref(self.q), ctypes.byref(t) if t else None,
                                          ctypes.byref(h)))
        self.handle = h
        self.tuning = tuning

    def close(self):
        # after its context is closed the native object is gone with it
        if self.handle and self.ctx.handle:
            C.lib().fpdb_star_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_plain(self):
        C.check(C.lib().fpdb_star_run_plain(self.ctx.handle, self.handle))

    def run_batched(self):
        C.check(C.lib().fpdb_star_run_batched(self.ctx.handle, self.handle))

    def capture(self):
        C.check(C.lib().fpdb_star_graph_capture(self.ctx.handle, self.handle))

    def replay(self):
        C.check(C.lib().fpdb_star_graph_launch(self.ctx.handle, self.handle))

    def result(self):
        m = self.max_groups
        ga, gb = np.zeros(m, np.int32), np.zeros(m, np.int32)
        cnt, sm = np.zeros(m, np.int64), np.zeros(m, np.int64)
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_star_result(self.ctx.handle, self.handle, ga.ctypes.data, gb.ctypes.data,
                                         cnt.ctypes.data, sm.ctypes.data, m, ctypes.byref(n)))
        k = n.value
        return {"g_a": ga[:k].astype(np.int64), "g_b": gb[:k].astype(np.int64), "count": cnt[:k], "sum": sm[:k]}

    def filter(self, dim, backward):
        lnb, rows = ctypes.c_int(), ctypes.c_int64()
        C.check(C.lib().fpdb_star_filter(self.ctx.handle, self.handle, dim, 1 if backward else 0, None, 0,
                                         ctypes.byref(lnb), ctypes.byref(rows)))
        blocks = np.zeros(1 << lnb.value, np.uint64)
        C.check(C.lib().fpdb_star_filter(self.ctx.handle, self.handle, dim, 1 if backward else 0,
                                         blocks.ctypes.data, len(blocks), ctypes.byref(lnb), ctypes.byref(rows)))
        return lnb.value, blocks, rows.value

    def signature(self):
        dims = self.w["dims"]
        return "star:" + ",".join(f"{len(d['key'])}/{d['keep_below']}" for d in dims) + \
               f":fact={len(self.w['measure'])}:parts={len(self._offs) - 1}"


def _time_batched(sj, reps):
    sj.ctx.synchronize()
    sj.ctx.timer_start()
    for _ in range(reps):
        sj.run_batched()
    return sj.ctx.timer_stop() / reps


def autotune(sj, device_name, reps=3, cache_path=None):
    """Tunes the batched form on the prepared plan, one knob at a time, each
    candidate timed over `reps` runs: the launch shape (blocks per CU x
    hash-table load shift), then fact rows per thread per step, then the
    forward filters' order (every order for up to four dimensions, else the
    default and its neighbours by one swap).  The result (blocks per CU, load
    shift, items, order) is cached per (signature, device)."""
    path = cache_path or os.environ.get("FPDB_AUTOTUNE_CACHE", DEFAULT_CACHE)
    key = f"{sj.signature()}@{device_name}"
    cache = {}
    if os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if key in cache and len(cache[key]["tuning"]) == 4:
        t = cache[key]["tuning"]
        best = (t[0], t[1], t[2], tuple(t[3]))
        sj.prepare(best)
        return best, cache[key]
    timings = {}

    def measure(cand):
        if cand not in timings:
            sj.prepare(cand)
            sj.run_batched()          # warm
            timings[cand] = _time_batched(sj, reps)
        return timings[cand]

    base_order = tuple(sj.order())
    shape = min(SHAPES, key=lambda c: measure((c[0], c[1], 8, base_order)))
    items = min(ITEMS, key=lambda i: measure((shape[0], shape[1], i, base_order)))
    nd = len(base_order)
    if nd <= 4:
        orders = list(itertools.permutations(base_order))[:MAX_ORDERS]
    else:
        orders = [base_order] + [tuple(base_order[:i] + (base_order[i + 1], base_order[i]) + base_order[i + 2:])
                                 for i in range(nd - 1)]
    order = min(orders, key=lambda o: measure((shape[0], shape[1], items, tuple(o))))
    best = (shape[0], shape[1], items, tuple(order))
    sj.prepare(best)
    entry = {"tuning": [best[0], best[1], best[2], list(best[3])], "ms": timings[best],
             "candidates": {"/".join(str(x) for x in c[:3]) + "/" + "".join(str(d) for d in c[3]): ms
                            for c, ms in timings.items()}}
    cache[key] = entry
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass
    return best, entry
