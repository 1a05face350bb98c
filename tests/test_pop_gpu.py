"""The drop-in physical operators (C++, flexpushdowndb_amd/pop) driven by the
operator-level test program tests/cpp/pop_test, plus a predicate-transfer
star join + group-by plan checked row for row against numpy."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "pop_test")


def _run(args, timeout=240):
    assert os.path.exists(BIN), "tests/cpp/pop_test not built (python -m flexpushdowndb_amd.build)"
    return subprocess.run([BIN, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=timeout)


def test_operator_unit_tests():
    r = _run(["unit"])
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout


M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(x):
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _uni(seed, col, n, lo, hi):
    with np.errstate(over="ignore"):
        base = np.uint64(seed) ^ (np.uint64(col) * np.uint64(0x9E3779B97F4A7C15))
        u = _splitmix(base ^ np.arange(n, dtype=np.uint64))
    return (u % np.uint64(hi - lo + 1)).astype(np.int64) + lo


def _star_expected(n):
    seed = 77
    fk1, fk2, m = _uni(seed, 1, n, 0, 999), _uni(seed, 2, n, 0, 499), _uni(seed, 3, n, 1, 100)
    a1, g1 = _uni(seed, 4, 1000, 0, 99), _uni(seed, 5, 1000, 0, 9)
    a2, g2 = _uni(seed, 6, 500, 0, 99), _uni(seed, 7, 500, 0, 4)
    keep = (a1[fk1] < 30) & (a2[fk2] < 10)
    kg1, kg2, km = g1[fk1][keep], g2[fk2][keep], m[keep]
    # groups in order of first appearance in the joined stream (fact order)
    key = kg1 * 1000 + kg2
    _, first = np.unique(key, return_index=True)
    order = np.sort(first)
    exp = []
    for i in order:
        sel = key == key[i]
        exp.append((kg1[i], kg2[i], km[sel].sum(), sel.sum()))
    return np.array(exp, dtype=np.int64).reshape(-1, 4)


@pytest.mark.parametrize("n", [1000, 250_000])
def test_predicate_transfer_star_join(tmp_path, n):
    out = tmp_path / "star.csv"
    r = _run(["starjoin", str(n), str(out)])
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    got = np.loadtxt(out, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
    assert np.array_equal(got, _star_expected(n))


@pytest.mark.parametrize("order", ["s2l", "bfs", "yan"])
def test_planned_predicate_transfer_star_join(tmp_path, order):
    # the same query with its filters (or Yannakakis semi-joins) laid out by
    # the plan builder: the same groups in the same order
    n = 250_000
    out = tmp_path / f"star_{order}.csv"
    r = _run(["starjoin-planned", order, str(n), str(out)])
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    got = np.loadtxt(out, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
    assert np.array_equal(got, _star_expected(n))
