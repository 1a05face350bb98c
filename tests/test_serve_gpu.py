"""Serving loop: every query's result equals the direct-indexing expectation,
and is identical when the query runs alone and in a mix of concurrently
admitted queries (continuous batching over a circular shared scan): queries
admitted at any chunk, a query joining mid-pass, and one failing query
leaving the others' results unchanged."""
import numpy as np
import pytest

from flexpushdowndb_amd import _capi as C
from flexpushdowndb_amd import datagen
from flexpushdowndb_amd.device import Context
from flexpushdowndb_amd.serve import FAILED, FINISHED, SCANNING, Server, poisson_arrivals

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
    steps = mix.drain(ids)
    # 40 queries, 8 in the scan at once, one pass of 8 chunks each: five
    # passes of ticks (the finalizes of the last ones may take a step more)
    assert 40 <= steps <= 42
    alone = Server(ctx, w, max_active=1, chunks=3)
    nonempty = 0
    for q, i in zip(qs, ids):
        got = mix.result(i)
        _same(got, datagen.serve_expected(w, q))
        j = alone.submit(q)
        alone.drain([j])
        _same(alone.result(j), got)
        nonempty += len(got["count"]) > 0
    assert nonempty >= 10
    mix.close()
    alone.close()


def test_trace_and_workspace_budget(ctx):
    w, qs = _w()
    # a budget that holds one query's workspace at a time: one query per scan
    one = 257 * 2000 * 24
    srv = Server(ctx, w, max_active=8, workspace_bytes=one)
    arr = poisson_arrivals(len(qs), 2000.0, seed=1)
    ids, lat, makespan = srv.run_trace(qs, arr)
    assert (lat > 0).all() and makespan >= arr[-1]
    for q, i in zip(qs, ids):
        _same(srv.result(int(i)), datagen.serve_expected(w, q))
    srv.close()


def test_query_joining_mid_pass(ctx):
    """A query submitted while another is partway through its pass joins the
    scan at the current chunk, stays for one full pass (wrapping around the
    end of the fact table) and finishes after the first one, both exact."""
    w, qs = _w()
    srv = Server(ctx, w, max_active=8, chunks=8)
    a = srv.submit(qs[0])
    for _ in range(3):
        srv.step()
    assert srv.status(a) == SCANNING
    b = srv.submit(qs[1])
    order = []
    while len(order) < 2:
        srv.step()
        for i in (a, b):
            if i not in order and srv.status(i) == FINISHED:
                order.append(i)
    # b left the scan three ticks after a, and finalizes run in order
    assert order == [a, b]
    _same(srv.result(a), datagen.serve_expected(w, qs[0]))
    _same(srv.result(b), datagen.serve_expected(w, qs[1]))
    srv.close()


def test_failing_query_leaves_others_unchanged():
    """The server's context gets an HBM budget too small for a query with
    2,000 groups (its partials and totals: 257 x 2,000 groups x 24 B = 12.3 MB)
    but large enough for the others; that query fails at admission with the
    budget error while the queries around it in the queue finish exactly as
    they do alone."""
    w, qs = _w()
    data = Context(0)
    # the server's own buffers: one query byte per dimension row and a 4-byte
    # copy of each fact foreign-key column; then 8 MiB for the queries
    fixed = sum(len(d["key"]) for d in w["dims"]) + 4 * len(w["measure"]) * len(w["dims"])
    srv_ctx = Context(0, hbm_budget_bytes=fixed + 8 * 2**20)
    big = {"filters": [(0, 50_000)], "group": [0, 2]}        # 40 x 50 groups
    assert 257 * 40 * 50 * 24 > 8 * 2**20
    small = [q for q in qs if np.prod([w["dims"][d]["domain"] for d in q["group"]]) <= 50][:6]
    mix = small[:3] + [big] + small[3:]
    srv = Server(srv_ctx, w, max_active=8, data_ctx=data)
    ids = [srv.submit(q) for q in mix]
    srv.drain(ids)
    assert srv.status(ids[3]) == FAILED
    with pytest.raises(C.FpdbError) as e:
        srv.result(ids[3])
    assert e.value.code == C.FPDB_ERR_BUDGET
    for q, i in zip(mix, ids):
        if i == ids[3]:
            continue
        assert srv.status(i) == FINISHED
        _same(srv.result(i), datagen.serve_expected(w, q))
    # the server goes on serving after the failure
    j = srv.submit(small[0])
    srv.drain([j])
    _same(srv.result(j), datagen.serve_expected(w, small[0]))
    srv.close()
    srv_ctx.close()
    data.close()


def test_servers_one_after_another(ctx):
    """A server created after another was destroyed may get the same pinned
    result pages back; its results must still be its own."""
    w, qs = _w()
    for k in range(3):
        srv = Server(ctx, w, max_active=4, chunks=2)
        mine = qs[5 * k:5 * k + 5]
        ids = [srv.submit(q) for q in mine]
        srv.drain(ids)
        for q, i in zip(mine, ids):
            _same(srv.result(i), datagen.serve_expected(w, q))
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
    srv.drain(ids)
    for q, i in zip(qs, ids):
        _same(srv.result(i), datagen.serve_expected(w, q))
    srv.close()
