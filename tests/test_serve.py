"""Serving workload generator and the expected-result function it is checked
against (brute force over a tiny schema)."""
import numpy as np

from flexpushdowndb_amd import datagen


def _brute(w, q):
    groups = {}
    for r in range(len(w["measure"])):
        if all(w["dims"][d]["attr"][w["fact_fk"][d][r]] < c for d, c in q["filters"]):
            g = tuple(int(w["dims"][d]["group"][w["fact_fk"][d][r]]) for d in q["group"])
            if len(g) == 1:
                g = (g[0], 0)
            if g not in groups:
                groups[g] = [r, 0, 0]
            groups[g][1] += 1
            groups[g][2] += int(w["measure"][r])
    return sorted(((v[0], k, v[1], v[2]) for k, v in groups.items()))


def test_expected_matches_brute_force():
    w = datagen.serve_workload(n_fact=3000, seed=1, dims=(50, 70, 400, 9))
    qs = datagen.serve_queries(20, seed=2)
    # widen the filters so the tiny schema keeps rows
    qs = [{"filters": [(d, min(100000, c * 40)) for d, c in q["filters"]], "group": q["group"]} for q in qs]
    nonempty = 0
    for q in qs:
        e = datagen.serve_expected(w, q)
        b = _brute(w, q)
        assert [(k[0], k[1]) for _, k, _, _ in b] == list(zip(e["g0"].tolist(), e["g1"].tolist()))
        assert [c for _, _, c, _ in b] == e["count"].tolist()
        assert [s for _, _, _, s in b] == e["sum"].tolist()
        nonempty += len(b) > 0
    assert nonempty >= 5


def test_query_generator_ranges():
    qs = datagen.serve_queries(500, seed=3)
    for q in qs:
        dims = [d for d, _ in q["filters"]]
        assert 1 <= len(dims) <= 4 and len(set(dims)) == len(dims)
        assert all(100 <= c <= 50_000 for _, c in q["filters"])   # 0.1 % .. 50 % of [0, 100000)
        assert 1 <= len(q["group"]) <= 2 and len(set(q["group"])) == len(q["group"])
    sels = np.array([c for q in qs for _, c in q["filters"]]) / 100_000
    # log-uniform: about half the selectivities lie below the geometric midpoint
    mid = np.sqrt(0.001 * 0.5)
    assert 0.4 < (sels < mid).mean() < 0.6
