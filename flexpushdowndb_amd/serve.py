"""Serving loop over a star schema resident in HBM: queries are submitted,
admitted in FIFO order at any tick of a circular shared scan of the fact
table (continuous batching), and finished once they have seen every chunk
(fpdb_serve_* entry points)."""
import ctypes

import numpy as np

from . import _capi as C
from .device import DeviceColumn


QUEUED, SCANNING, FINISHED, FAILED = 0, 1, 2, 3


class Server:
    def __init__(self, ctx, w, max_active=8, workspace_bytes=0, chunks=0, data_ctx=None):
        """data_ctx: the context holding the schema's columns (default ctx)."""
        self.ctx = ctx
        self.w = w
        self._cols = []
        n = len(w["measure"])
        dctx = data_ctx if data_ctx is not None else ctx

        def dc(a):
            c = DeviceColumn.from_numpy(dctx, np.ascontiguousarray(a))
            self._cols.append(c)
            return c.c()

        S = C.ServeSchema()
        S.num_dims = len(w["dims"])
        for d, dim in enumerate(w["dims"]):
            D = S.dims[d]
            D.key = dc(dim["key"])
            D.filter_attr = dc(dim["attr"])
            D.group_attr = dc(dim["group"])
            D.group_domain = dim["domain"]
            D.fact_fk = dc(w["fact_fk"][d])
        S.fact_measure = dc(w["measure"])
        S.fact_rows = n
        cfg = C.ServeConfig(max_active, chunks, workspace_bytes)
        self.handle = ctypes.c_void_p()
        C.check(C.lib().fpdb_serve_create(ctx.handle, ctypes.byref(S), ctypes.byref(cfg), ctypes.byref(self.handle)))

    def close(self):
        # after its context is closed the native object is gone with it
        if self.handle and self.ctx.handle:
            C.lib().fpdb_serve_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def query_struct(q):
        Q = C.ServeQuery()
        Q.num_filters = len(q["filters"])
        for i, (d, c) in enumerate(q["filters"]):
            Q.filter_dim[i] = d
            Q.filter_op[i] = C.CMP_LT
            Q.filter_constant[i] = c
        Q.num_group = len(q["group"])
        for i, d in enumerate(q["group"]):
            Q.group_dim[i] = d
        return Q

    def submit(self, q):
        qid = ctypes.c_int64()
        C.check(C.lib().fpdb_serve_submit(self.handle, ctypes.byref(self.query_struct(q)), ctypes.byref(qid)))
        return qid.value

    def step(self):
        n = ctypes.c_int()
        C.check(C.lib().fpdb_serve_step(self.handle, ctypes.byref(n)))
        return n.value

    def status(self, qid):
        st = ctypes.c_int()
        C.check(C.lib().fpdb_serve_status(self.handle, qid, ctypes.byref(st)))
        return st.value

    def drain(self, ids, max_steps=1_000_000):
        """Steps until every query in ids is finished or failed; returns the steps taken."""
        steps = 0
        while any(self.status(i) < FINISHED for i in ids):
            self.step()
            steps += 1
            if steps > max_steps:
                raise RuntimeError("serving queries did not finish")
        return steps

    def result(self, qid, capacity=4096):
        g0 = np.zeros(capacity, np.int32)
        g1 = np.zeros(capacity, np.int32)
        cnt = np.zeros(capacity, np.int64)
        sm = np.zeros(capacity, np.int64)
        n = ctypes.c_int64()
        C.check(C.lib().fpdb_serve_result(self.handle, qid, g0.ctypes.data, g1.ctypes.data, cnt.ctypes.data,
                                          sm.ctypes.data, capacity, ctypes.byref(n)))
        k = n.value
        return {"g0": g0[:k].astype(np.int64), "g1": g1[:k].astype(np.int64), "count": cnt[:k], "sum": sm[:k]}

    def run_trace(self, queries, arrival_ns):
        n = len(queries)
        Qs = (C.ServeQuery * max(1, n))(*[self.query_struct(q) for q in queries])
        arr = np.ascontiguousarray(arrival_ns, dtype=np.int64)
        ids = np.zeros(n, np.int64)
        lat = np.zeros(n, np.int64)
        mk = ctypes.c_int64()
        C.check(C.lib().fpdb_serve_run_trace(self.handle, Qs, arr.ctypes.data_as(C._PI64), n,
                                             ids.ctypes.data_as(C._PI64), lat.ctypes.data_as(C._PI64),
                                             ctypes.byref(mk)))
        return ids, lat, mk.value


def poisson_arrivals(n, rate_per_s, seed):
    """Arrival times (ns) of a Poisson process with the given rate."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=n)
    return (np.cumsum(gaps) * 1e9).astype(np.int64)
