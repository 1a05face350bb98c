"""Serving loop: every query's result equals the direct-indexing expectation,
and is identical when the query runs alone and in a mix of concurrently
admitted queries (continuous batching over shared scans)."""
import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import datagen
from flexpushdowndb_amd.serve import Server, poisson_arrivals

pytestmark = pytest.mark.gpu

W = {}


def _w():
    if not W:
        W["w"] = datagen.serve_workload(n_fact=2_000_000, seed=11)
        # widen some filters so results are not all empty
        qs = datagen.serve_queries(40, seed=12)
        W["q"] = [{"filters": [(d, min(99_999, c * (8 if i % 2 else 1))) for d, c in q["filters"]],
                   "group": q["group"]} for i, q in enumerate(qs)]
    return W["w"], W["q"]


def _same(a, b):
    for k in ("g0", "g1", "count", "sum"):
        assert np.array_equal(a[k], b[k]), k


def test_mix_alone_and_expected(ctx):
    w, qs = _w()
    mix = Server(ctx, w, max_active=8)
    ids = [mix.submit(q) for q in qs]
    steps = 0
    while any(mix.status(i) != 2 for i in ids):
        assert mix.step() > 0
        steps += 1
    assert steps == 5                      # 40 queries, 8 per shared scan
    alone = Server(ctx, w, max_active=1)
    nonempty = 0
    for q, i in zip(qs, ids):
        got = mix.result(i)
        _same(got, datagen.serve_expected(w, q))
        j = alone.submit(q)
        assert alone.step() == 1
        _same(alone.result(j), got)
        nonempty += len(got["count"]) > 0
    assert nonempty >= 10
    mix.close()
    alone.close()


def test_trace_and_workspace_budget(ctx):
    w, qs = _w()
    # a budget that holds one query's workspace at a time: one query per scan
    one = 256 * 2000 * 16
    srv = Server(ctx, w, max_active=8, workspace_bytes=one)
    arr = poisson_arrivals(len(qs), 2000.0, seed=1)
    ids, lat, makespan = srv.run_trace(qs, arr)
    assert (lat > 0).all() and makespan >= arr[-1]
    for q, i in zip(qs, ids):
        _same(srv.result(int(i)), datagen.serve_expected(w, q))
    srv.close()


def test_errors(ctx):
    w, qs = _w()
    srv = Server(ctx, w)
    with pytest.raises(C.FpdbError):
        srv.submit({"filters": [(7, 10)], "group": [0]})          # no dimension 7
    with pytest.raises(C.FpdbError):
        srv.submit({"filters": [(0, 10)], "group": [1, 1]})       # same group dimension twice
    qid = srv.submit(qs[0])
    with pytest.raises(C.FpdbError):
        srv.result(qid)                                           # not finished
    srv.close()
    bad = dict(w)
    bad["dims"] = [dict(d) for d in w["dims"]]
    bad["dims"][1]["key"] = bad["dims"][1]["key"][::-1].copy()    # keys not dense
    with pytest.raises(C.FpdbError):
        Server(ctx, bad)


def test_full_size_queries_exact(ctx):
    """The serving workload at its full size (60M fact rows, four dimensions
    of 30,000 / 200,000 / 2,000,000 / 2,556 rows): a mix of queries admitted
    into shared scans, each result equal to the direct-indexing expectation."""
    w = datagen.serve_workload()
    qs = datagen.serve_queries(6, seed=31)
    srv = Server(ctx, w, max_active=8)
    ids = [srv.submit(q) for q in qs]
    while any(srv.status(i) != 2 for i in ids):
        assert srv.step() > 0
    for q, i in zip(qs, ids):
        _same(srv.result(i), datagen.serve_expected(w, q))
    srv.close()
