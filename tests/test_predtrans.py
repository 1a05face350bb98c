"""The predicate-transfer plan builder (flexpushdowndb_amd/pop/PredTrans.cpp,
run through tests/cpp/pop_test's plan-only mode, which touches no device)
against the oracle's restatement of SmallToLargePredTransOrder and
BFSPredTransOrder: the same transfers, connected in the same order, to the
same operators, for star, snowflake, chain, multi-column, outer-join,
disconnected and cyclic join graphs."""
import os
import subprocess

import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "pop_test")

GRAPHS = {
    "star": ([("F", 6e6), ("D1", 1000), ("D2", 500)],
             [("D1", "F", ["k1"], ["fk1"], "INNER"), ("D2", "F", ["k2"], ["fk2"], "INNER")]),
    "ssb": ([("lineorder", 6e6), ("date", 2556), ("customer", 30000), ("supplier", 2000), ("part", 200000)],
            [("date", "lineorder", ["d_datekey"], ["lo_orderdate"], "INNER"),
             ("customer", "lineorder", ["c_custkey"], ["lo_custkey"], "INNER"),
             ("supplier", "lineorder", ["s_suppkey"], ["lo_suppkey"], "INNER"),
             ("part", "lineorder", ["p_partkey"], ["lo_partkey"], "INNER")]),
    "snowflake": ([("F", 1e6), ("D1", 1e4), ("D11", 100), ("D2", 50)],
                  [("D11", "D1", ["a"], ["b"], "INNER"), ("D1", "F", ["k1"], ["fk1"], "INNER"),
                   ("D2", "F", ["k2"], ["fk2"], "INNER")]),
    "chain": ([("A", 10), ("B", 100), ("C", 1000), ("D", 10000)],
              [("A", "B", ["a"], ["b"], "INNER"), ("B", "C", ["b"], ["c"], "INNER"),
               ("C", "D", ["c"], ["d"], "INNER")]),
    "multicol": ([("F", 1e5), ("D", 10)], [("D", "F", ["x", "y"], ["fx", "fy"], "INNER")]),
    "outer": ([("A", 10), ("B", 20), ("C", 30), ("E", 5)],
              [("A", "B", ["a"], ["b"], "LEFT"), ("B", "C", ["b"], ["c"], "RIGHT"),
               ("A", "C", ["a"], ["c"], "FULL"), ("E", "C", ["e"], ["c"], "INNER")]),
    "disconnected": ([("A", 10), ("B", 20), ("C", 30), ("D", 40)],
                     [("A", "B", ["a"], ["b"], "INNER"), ("C", "D", ["c"], ["d"], "INNER")]),
    "cycle": ([("A", 1), ("B", 2), ("C", 3)],
              [("A", "B", ["x"], ["y"], "INNER"), ("B", "C", ["y"], ["z"], "INNER"),
               ("C", "A", ["z"], ["x"], "INNER")]),
}


def _cpp_plan(order, tables, joins):
    assert os.path.exists(BIN), "tests/cpp/pop_test not built (python -m flexpushdowndb_amd.build)"
    tspec = ",".join("%s=%r" % (t, float(n)) for t, n in tables)
    jspec = ";".join("%s.%s>%s.%s@%s" % (l, "+".join(lc), r, "+".join(rc), ty) for l, r, lc, rc, ty in joins)
    out = subprocess.run([BIN, "plan", order, tspec, jspec], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stdout
    steps, ends, error = [], {}, None
    for line in out.stdout.splitlines():
        if line.startswith("STEP "):
            steps.append(tuple(line[5:].split("|")))
        elif line.startswith("END "):
            t, op = line[4:].split("|")
            ends[t] = op
        elif line.startswith("ERROR "):
            error = line[6:]
    return steps, ends, error


@pytest.mark.parametrize("order", ["s2l", "bfs", "yan"])
@pytest.mark.parametrize("graph", sorted(GRAPHS))
def test_plan_matches_oracle(graph, order):
    tables, joins = GRAPHS[graph]
    steps, ends, error = _cpp_plan(order, tables, joins)
    try:
        want_steps, want_ends = O.pred_trans_plan(order, tables, joins)
    except ValueError as e:
        assert error is not None and "cycles" in error, (error, e)
        return
    assert error is None, error
    assert steps == want_steps
    assert ends == want_ends


def test_every_table_gets_every_transfer_it_can():
    # in the SSB star every dimension filters lineorder (forward) and is
    # filtered by the reduced lineorder (backward), whichever ordering
    tables, joins = GRAPHS["ssb"]
    for order in ("s2l", "bfs"):
        steps, _, _ = _cpp_plan(order, tables, joins)
        pairs = sorted((s[2], s[3]) for s in steps)
        dims = ["customer", "date", "part", "supplier"]
        assert pairs == sorted([(d, "lineorder") for d in dims] + [("lineorder", d) for d in dims])
