"""Batched star join with full predicate transfer: the per-partition (plain)
and persistent (batched, also replayed from a captured graph) executions must
both produce exactly the join + group-by result, and their forward and
backward bloom filters must be bit-exact with the oracle's."""
import numpy as np
import pytest

from flexpushdowndb_amd import datagen
from flexpushdowndb_amd.device import Context
from flexpushdowndb_amd.starjoin import StarJoin, autotune
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same(got, exp):
    for k in ("g_a", "g_b", "count", "sum"):
        assert np.array_equal(got[k], exp[k]), k


def _oracle_filters(w, order):
    fwd = {}
    for d, dim in enumerate(w["dims"]):
        kept = dim["key"][dim["attr"] < dim["keep_below"]]
        fwd[d] = O.blocked_build(O.arrow6_hash([(kept, None)]) if len(kept) else np.zeros(0, np.uint32),
                                 n_capacity=len(kept))
    surv = np.ones(len(w["measure"]), bool)
    for d in order:
        lnb, blocks = fwd[d]
        bits = O.blocked_find(lnb, blocks, O.arrow6_hash([(w["fact_fk"][d], None)]))
        surv &= np.unpackbits(bits, bitorder="little")[:len(surv)].astype(bool)
    bwd = {}
    for d in range(len(w["dims"])):
        keys = w["fact_fk"][d][surv]
        bwd[d] = O.blocked_build(O.arrow6_hash([(keys, None)]) if len(keys) else np.zeros(0, np.uint32),
                                 n_capacity=len(keys))
    return fwd, bwd, int(surv.sum())


@pytest.mark.parametrize("n_fact,n_parts", [(2_000_000, 8), (300_000, 64)])
def test_star_join_plain_batched_graph(ctx, n_fact, n_parts):
    w = datagen.star_workload(n_fact=n_fact, n_parts=n_parts, seed=9,
                              keep_below=(30, 20, 10, 40))
    exp = datagen.star_expected(w)
    assert len(exp["count"]) > 10
    sj = StarJoin(ctx, w)
    fwd, bwd, nsurv = _oracle_filters(w, sj.default_order())
    sj.run_plain()
    _same(sj.result(), exp)
    for d in range(4):
        lnb, blocks, rows = sj.filter(d, backward=False)
        assert lnb == fwd[d][0] and np.array_equal(blocks, fwd[d][1])
        lnb, blocks, rows = sj.filter(d, backward=True)
        assert rows == nsurv and lnb == bwd[d][0] and np.array_equal(blocks, bwd[d][1])
    sj.run_batched()
    _same(sj.result(), exp)
    for d in range(4):
        lnb, blocks, rows = sj.filter(d, backward=False)
        assert np.array_equal(blocks, fwd[d][1])
        lnb, blocks, rows = sj.filter(d, backward=True)
        assert rows == nsurv and np.array_equal(blocks, bwd[d][1])
    sj.capture()
    for _ in range(3):
        sj.replay()
        _same(sj.result(), exp)
        sj.run_plain()             # eager work between replays
        _same(sj.result(), exp)


@pytest.mark.parametrize("domain0", [300, 60])
def test_star_join_many_groups(ctx, domain0):
    """Group counts on both sides of one workgroup's worth (1,024), where the
    emit ranks groups by counting smaller first rows or by a bitonic sort."""
    w = datagen.star_workload(n_fact=2_000_000, n_parts=16, seed=13, keep_below=(60, 50, 40, 80))
    rng = np.random.default_rng(14)
    w["dims"][0]["group"] = rng.integers(0, domain0, len(w["dims"][0]["key"])).astype(np.int32)
    exp = datagen.star_expected(w)
    assert (len(exp["count"]) > 1024) == (domain0 == 300)
    sj = StarJoin(ctx, w)
    sj.run_batched()
    _same(sj.result(), exp)
    sj.run_plain()
    _same(sj.result(), exp)


def test_star_join_empty_dimension_and_partitions(ctx):
    w = datagen.star_workload(n_fact=100_000, n_parts=16, seed=3, keep_below=(30, 0, 10, 40))
    w["offsets"] = np.array([0, 0, 50_000, 50_000, 100_000], np.int64)
    sj = StarJoin(ctx, w)
    sj.run_plain()
    assert len(sj.result()["count"]) == 0
    sj.run_batched()
    assert len(sj.result()["count"]) == 0


@pytest.mark.parametrize("n_fact,n_parts,empty", [(1_000_000, 16, False), (200_000, 8, True)])
def test_star_join_partitions_in_separate_buffers(ctx, n_fact, n_parts, empty):
    """Each partition's fact columns in device buffers of their own (the
    reference's one TupleSet per partition): plain, batched and replayed runs
    give the same exact result as the contiguous table, empty partitions
    included."""
    w = datagen.star_workload(n_fact=n_fact, n_parts=n_parts, seed=5, keep_below=(30, 20, 10, 40))
    if empty:
        offs = w["offsets"].copy()
        offs[2] = offs[1]                       # partition 1 empty
        offs[-2] = offs[-1]                     # the last partition empty
        w["offsets"] = offs
    exp = datagen.star_expected(w)
    assert len(exp["count"]) > 10
    sj = StarJoin(ctx, w, split_partitions=True)
    fwd, bwd, nsurv = _oracle_filters(w, sj.default_order())
    sj.run_plain()
    _same(sj.result(), exp)
    sj.run_batched()
    _same(sj.result(), exp)
    for d in range(4):
        lnb, blocks, rows = sj.filter(d, backward=True)
        assert rows == nsurv and np.array_equal(blocks, bwd[d][1])
    sj.capture()
    for _ in range(2):
        sj.replay()
        _same(sj.result(), exp)


def test_star_join_autotune_and_full_size(ctx, tmp_path):
    """The 60M-row, 64-partition workload: exact result from the autotuned
    batched form and from the plain form."""
    w = datagen.star_workload()
    exp = datagen.star_expected(w)
    sj = StarJoin(ctx, w)
    best, entry = autotune(sj, "test-device", reps=1, cache_path=str(tmp_path / "tune.json"))
    t = entry["tuning"]
    assert (t[0], t[1], t[2], tuple(t[3])) == best
    assert sorted(best[3]) == list(range(4)) and best[2] in (4, 8, 16)
    sj.run_batched()
    _same(sj.result(), exp)
    sj.run_plain()
    _same(sj.result(), exp)
    # a second autotune call is a cache hit
    best2, _ = autotune(sj, "test-device", reps=1, cache_path=str(tmp_path / "tune.json"))
    assert best2 == best


def test_barrier_wait_bound_reports_error_and_next_run_is_exact(ctx):
    """A grid-barrier wait bound of 1 us (set from the host) makes the
    persistent kernel's waiting workgroups give up: the launch still ends
    (every workgroup stops at its next barrier) and the result call reports
    FPDB_ERR_INTERNAL.  The next run, with the default bound, is exact."""
    from flexpushdowndb_amd import _capi as C
    w = datagen.star_workload(n_fact=2_000_000, n_parts=8, seed=9, keep_below=(30, 20, 10, 40))
    exp = datagen.star_expected(w)
    tight = StarJoin(ctx, w, barrier_timeout_us=1)
    tight.run_batched()
    with pytest.raises(C.FpdbError) as e:
        tight.result()
    assert e.value.code == C.FPDB_ERR_INTERNAL
    tight.run_plain()                  # the plain form has no grid barrier
    _same(tight.result(), exp)
    tight.close()
    normal = StarJoin(ctx, w)
    normal.run_batched()
    _same(normal.result(), exp)
    normal.run_batched()
    _same(normal.result(), exp)
