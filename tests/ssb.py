"""The 13 Star Schema Benchmark queries at scale 0.01 as star-join plans, run
two ways and compared with the answers the reference printed for them
(resources/result/ssb-sf0.01/ssb/original/*.r, over the parquet data of
fpdb-main/test-resources/fpdb-store-0/flexpushdowndb/ssb-sf0.01):

* `oracle_run`: filters in numpy, then the oracle's hash join
  (orc_join_pairs) and group-by (orc_groupby);
* `device_run`: the library's device path with full predicate transfer as
  SmallToLargePredTransOrder plans it (pred-trans/SmallToLargePredTransOrder.cpp:38-214):
  dimension filters (compare bitmaps + compaction), a forward blocked bloom
  filter per dimension applied to the fact table, backward filters built from
  the surviving fact rows and applied to the dimensions, then one hash join per
  dimension and a hash group-by.

String columns are dictionary-encoded with sorted dictionaries (see
tests/golden/make_ssb_fixtures.py), so string equality and ranges become
integer comparisons.  The measure expressions (lo_extendedprice*lo_discount,
lo_revenue - lo_supplycost) are evaluated in numpy and handed to both paths as
an int64 column: projection arithmetic is not part of the device path.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

TABLES = {"d": "date", "p": "part", "s": "supplier", "c": "customer"}
JOIN_KEYS = {"date": ("lo_orderdate", "d_datekey"), "part": ("lo_partkey", "p_partkey"),
             "supplier": ("lo_suppkey", "s_suppkey"), "customer": ("lo_custkey", "c_custkey")}


def _between(col, lo, hi):
    return [[(col, ">=", lo)], [(col, "<=", hi)]]


# where: conjunction of clauses, each a disjunction of (column, op, constant).
# measure: ("*"|"-"|None, column[, column]).  group: output group columns.
PLANS = {
    "1.1": dict(dims=["date"], where=[[("d_year", "==", 1992)], [("lo_quantity", "<", 25)]]
                + _between("lo_discount", 1, 3), measure=("*", "lo_extendedprice", "lo_discount"), group=[]),
    "1.2": dict(dims=["date"], where=[[("d_yearmonthnum", "==", 199201)]] + _between("lo_discount", 4, 6)
                + _between("lo_quantity", 26, 35), measure=("*", "lo_extendedprice", "lo_discount"), group=[]),
    "1.3": dict(dims=["date"], where=[[("d_weeknuminyear", "==", 3)], [("d_year", "==", 1992)]]
                + _between("lo_discount", 5, 7) + _between("lo_quantity", 26, 35),
                measure=("*", "lo_extendedprice", "lo_discount"), group=[]),
    "2.1": dict(dims=["supplier", "date", "part"], where=[[("p_category", "==", "MFGR#12")],
                                                          [("s_region", "==", "AMERICA")]],
                measure=(None, "lo_revenue"), group=["d_year", "p_brand1"]),
    "2.2": dict(dims=["supplier", "date", "part"], where=_between("p_brand1", "MFGR#2221", "MFGR#2228")
                + [[("s_region", "==", "ASIA")]], measure=(None, "lo_revenue"), group=["d_year", "p_brand1"]),
    "2.3": dict(dims=["supplier", "date", "part"], where=[[("p_brand1", "==", "MFGR#2221")],
                                                          [("s_region", "==", "EUROPE")]],
                measure=(None, "lo_revenue"), group=["d_year", "p_brand1"]),
    "3.1": dict(dims=["supplier", "date", "customer"], where=[[("c_region", "==", "ASIA")], [("s_region", "==", "ASIA")]]
                + _between("d_year", 1992, 1997), measure=(None, "lo_revenue"),
                group=["c_nation", "s_nation", "d_year"]),
    "3.2": dict(dims=["supplier", "date", "customer"], where=[[("c_nation", "==", "CHINA")],
                                                              [("s_nation", "==", "UNITED STATES")]]
                + _between("d_year", 1992, 1997), measure=(None, "lo_revenue"), group=["c_city", "s_city", "d_year"]),
    "3.3": dict(dims=["supplier", "date", "customer"],
                where=[[("c_city", "==", "UNITED KI1"), ("c_city", "==", "UNITED KI5")],
                       [("s_city", "==", "UNITED ST0"), ("s_city", "==", "UNITED ST9")]]
                + _between("d_year", 1992, 1997), measure=(None, "lo_revenue"), group=["c_city", "s_city", "d_year"]),
    "3.4": dict(dims=["supplier", "date", "customer"],
                where=[[("c_city", "==", "UNITED KI1"), ("c_city", "==", "UNITED KI5")],
                       [("s_city", "==", "UNITED ST0"), ("s_city", "==", "UNITED ST9")],
                       [("d_yearmonth", "==", "Jan1992")]],
                measure=(None, "lo_revenue"), group=["c_city", "s_city", "d_year"]),
    "4.1": dict(dims=["supplier", "date", "part", "customer"],
                where=[[("c_region", "==", "AMERICA")], [("s_region", "==", "AMERICA")],
                       [("p_mfgr", "==", "MFGR#1"), ("p_mfgr", "==", "MFGR#2")]],
                measure=("-", "lo_revenue", "lo_supplycost"), group=["d_year", "c_nation"]),
    "4.2": dict(dims=["supplier", "date", "part", "customer"],
                where=[[("c_region", "==", "AMERICA")], [("s_region", "==", "AMERICA")],
                       [("d_year", "==", 1992), ("d_year", "==", 1993)],
                       [("p_mfgr", "==", "MFGR#1"), ("p_mfgr", "==", "MFGR#2")]],
                measure=("-", "lo_revenue", "lo_supplycost"), group=["d_year", "s_nation", "p_category"]),
    "4.3": dict(dims=["supplier", "date", "part", "customer"],
                where=[[("c_region", "==", "AMERICA")], [("s_nation", "==", "UNITED STATES")],
                       [("d_year", "==", 1992), ("d_year", "==", 1993)], [("p_category", "==", "MFGR#14")]],
                measure=("-", "lo_revenue", "lo_supplycost"), group=["d_year", "s_city", "p_brand1"]),
}
QUERIES = list(PLANS)


def table_of(col):
    return "lineorder" if col.startswith("lo_") else TABLES[col.split("_")[0]]


def load():
    z = np.load(os.path.join(GOLDEN, "ssb_sf0.01.npz"))
    cols = {k: (z[k].astype(np.int64) if not k.endswith("__dict") else z[k]) for k in z.files}
    with open(os.path.join(GOLDEN, "ssb_sf0.01_answers.json")) as f:
        answers = json.load(f)
    return cols, answers


def encode(cols, col, op, const):
    """(op, integer constant) on the column's codes; string ranges map through
    the sorted dictionary."""
    if not isinstance(const, str):
        return op, int(const)
    d = cols[col + "__dict"]
    if op == "==":
        hit = np.flatnonzero(d == const)
        return "==", int(hit[0]) if len(hit) else -1
    if op == ">=":
        return ">=", int(np.searchsorted(d, const, "left"))
    if op == "<=":
        return "<=", int(np.searchsorted(d, const, "right")) - 1
    raise ValueError(op)


def decode(cols, col, v):
    key = col + "__dict"
    return str(cols[key][v]) if key in cols else int(v)


def clauses_of(plan, table):
    """Clauses on one table, the disjunctive one (at most one) first, so the
    chain AND/OR over one bitmap evaluates them exactly."""
    cl = [c for c in plan["where"] if table_of(c[0][0]) == table]
    assert all(len({table_of(a[0]) for a in c}) == 1 for c in cl)
    cl.sort(key=lambda c: -len(c))
    assert sum(1 for c in cl if len(c) > 1) <= 1, "one disjunctive clause per table"
    return cl


def measure(cols, plan):
    m = plan["measure"]
    if m[0] is None:
        return cols[m[1]].copy()
    a, b = cols[m[1]], cols[m[2]]
    return a * b if m[0] == "*" else a - b


_NP_OPS = {"<": np.less, "<=": np.less_equal, ">": np.greater, ">=": np.greater_equal, "==": np.equal,
           "!=": np.not_equal}


def _np_select(cols, clauses, n):
    sel = np.ones(n, bool)
    for clause in clauses:
        c = np.zeros(n, bool)
        for col, op, const in clause:
            op2, k = encode(cols, col, op, const)
            c |= _NP_OPS[op2](cols[col], k)
        sel &= c
    return sel


def expected_rows(plan, answer):
    """The answer's rows as (group values in plan order..., aggregate)."""
    names = answer["columns"]
    gi = [names.index(g) for g in plan["group"]]
    ai = [i for i in range(len(names)) if i not in gi]
    assert len(ai) == 1
    return sorted(tuple(r[i] for i in gi) + (r[ai[0]],) for r in answer["rows"])


def oracle_run(cols, plan):
    from oracle import oracle as O
    n = len(cols["lo_orderdate"])
    fact_sel = _np_select(cols, clauses_of(plan, "lineorder"), n)
    fact = {"rows": np.flatnonzero(fact_sel)}
    cur = fact["rows"]                       # fact row ids still joined
    gathered = {}                            # group column -> values aligned with cur
    for t in plan["dims"]:
        fk, pk = JOIN_KEYS[t]
        dn = len(cols[pk])
        dsel = np.flatnonzero(_np_select(cols, clauses_of(plan, t), dn))
        op, ob = O.join_pairs(cols[pk][dsel], cols[fk][cur])
        cur = cur[op]
        gathered = {k: v[op] for k, v in gathered.items()}
        for g in plan["group"]:
            if table_of(g) == t:
                gathered[g] = cols[g][dsel][ob]
    m = measure(cols, plan)[cur]
    if plan["group"]:
        keys = np.zeros(len(cur), np.uint64)
        shift = 0
        for g in plan["group"]:
            v = gathered[g]
            keys |= (v - v.min() if len(v) else v).astype(np.uint64) << np.uint64(shift)
            shift += 20
        res = O.groupby(keys, sums_i64=[m])
        rows = []
        for k, s in zip(res["keys"], res["sum_i64"][0]):
            vals, shift = [], 0
            for g in plan["group"]:
                v = gathered[g]
                code = int((int(k) >> shift) & 0xFFFFF) + int(v.min())
                vals.append(decode(cols, g, code))
                shift += 20
            rows.append(tuple(vals) + (int(s),))
        return sorted(rows)
    if len(cur) == 0:
        return []
    res = O.groupby(np.zeros(len(cur), np.uint64), sums_i64=[m])
    return [(int(res["sum_i64"][0][0]),)]


def _device_select(ctx, ops, dcols, clauses, n):
    """Bitmap of the conjunction of clauses over device columns (or None)."""
    bm = None
    for clause in clauses:
        for j, (col, op, const) in enumerate(clause):
            op2, k = const
            if j == 0:
                bm, _ = ops.compare_bitmap(ctx, dcols[col], op2, k, n, combine=bm)
            else:
                bm, _ = ops.compare_bitmap(ctx, dcols[col], op2, k, n, combine=bm, combine_or=True, out=bm)
    return bm


def device_run(ctx, cols, plan, strings=False):
    """The plan on the device; returns the sorted result rows and the row
    counts after each stage.  With strings=True the string columns go to the
    device as utf8 columns (their dictionary values) and are filtered, joined
    and grouped as strings: string constants compare bytewise on the device
    and string group keys go through the key dictionary."""
    from flexpushdowndb_amd import ops, _capi as C
    from flexpushdowndb_amd.device import DeviceColumn, Utf8Column
    stats = {}

    def is_str(c):
        return strings and c + "__dict" in cols

    def up_col(c):
        if is_str(c):
            d = cols[c + "__dict"]
            return Utf8Column.from_strings(ctx, [str(d[v]) for v in cols[c]])
        return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(cols[c], np.int64))

    def up(a):
        return DeviceColumn.from_numpy(ctx, np.ascontiguousarray(a, np.int64))

    def enc(clauses):
        return [[(c, op, (op, k) if is_str(c) else encode(cols, c, op, k)) for c, op, k in cl] for cl in clauses]

    # fact table: the foreign keys, the measure and the columns fact predicates use
    n = len(cols["lo_orderdate"])
    fcols = [JOIN_KEYS[t][0] for t in plan["dims"]]
    fpred = sorted({a[0] for cl in clauses_of(plan, "lineorder") for a in cl})
    dev = {c: up_col(c) for c in fcols + fpred}
    dev["measure"] = up(measure(cols, plan))
    names = fcols + ["measure"]
    fcl = clauses_of(plan, "lineorder")
    if fcl:
        bm = _device_select(ctx, ops, dev, enc(fcl), n)
        out, n = ops.filter_compact(ctx, [dev[c] for c in names], n, bm)
        fact = dict(zip(names, out))
    else:
        fact = {c: dev[c] for c in names}
    stats["fact_filtered"] = n

    # dimensions: filter, keep key + group columns, forward bloom filter
    dims = {}
    for t in plan["dims"]:
        fk, pk = JOIN_KEYS[t]
        dn = len(cols[pk])
        gcols = [g for g in plan["group"] if table_of(g) == t]
        need = sorted({a[0] for cl in clauses_of(plan, t) for a in cl})
        dd = {c: up_col(c) for c in set([pk] + gcols + need)}
        dcl = clauses_of(plan, t)
        keep = [pk] + gcols
        if dcl:
            bm = _device_select(ctx, ops, dd, enc(dcl), dn)
            out, kn = ops.filter_compact(ctx, [dd[c] for c in keep], dn, bm)
        else:
            out, kn = [dd[c] for c in keep], dn
        d = dict(zip(keep, out))
        d["_n"] = kn
        d["_fwd"] = ops.BlockedBloomFilter.build(ctx, [d[pk]], kn)
        dims[t] = d

    # forward predicate transfer: every dimension's filter reduces the fact table
    for t in plan["dims"]:
        fk, pk = JOIN_KEYS[t]
        if n == 0:
            break
        out, n, _ = dims[t]["_fwd"].filter([fact[c] for c in names], [names.index(fk)], n, want_bitmap=False)
        fact = dict(zip(names, out))
    stats["fact_after_forward"] = n

    # backward predicate transfer: the surviving fact keys reduce each dimension
    for t in plan["dims"]:
        fk, pk = JOIN_KEYS[t]
        d = dims[t]
        bwd = ops.BlockedBloomFilter.build(ctx, [fact[fk]], n)
        keep = [c for c in d if not c.startswith("_")]
        if d["_n"]:
            out, kn, _ = bwd.filter([d[c] for c in keep], [keep.index(pk)], d["_n"], want_bitmap=False)
            d.update(zip(keep, out))
            d["_n"] = kn
        stats["dim_" + t] = d["_n"]

    # joins, one per dimension, carrying the group columns along
    cur = dict(fact)
    order = list(names)
    for t in plan["dims"]:
        fk, pk = JOIN_KEYS[t]
        d = dims[t]
        gcols = [g for g in plan["group"] if table_of(g) == t]
        if n == 0:
            break
        table = ops.HashJoinTable(ctx, [d[pk]], d["_n"])
        probe_names = [c for c in order if c != fk]
        out, n = table.probe(C.JOIN_INNER, [cur[fk]], n, [cur[c] for c in probe_names], [d[g] for g in gcols])
        table.close()
        order = probe_names + gcols
        cur = dict(zip(order, out))
    stats["joined"] = n
    if n == 0:
        return [], stats

    # group-by
    keys = [cur[g] for g in plan["group"]]
    ktypes = [k.type_id for k in keys]
    if any(isinstance(k, Utf8Column) for k in keys):
        # string keys: group by key-dictionary id, then decode the ids
        kd = ops.KeyDict(ctx, ktypes)
        ids = kd.encode(keys, n, insert=True, null_is_key=True)
        ids.validity = None
        gb = ops.GroupBy(ctx, [C.TYPE_INT32], [(C.AGG_SUM, C.TYPE_INT64, False)], max_groups=max(1, n))
        gb.consume([ids], [cur["measure"]], n)
        kcols, aggs = gb.finalize()
        gb.close()
        sums = aggs[0].numpy()
        ng = len(sums)
        vals = []
        for j, k in enumerate(keys):
            dk = kd.decode(j, kcols[0], ng)
            vals.append(dk.to_list() if isinstance(dk, Utf8Column) else [int(v) for v in dk.numpy()])
        kd.close()
        rows = [tuple(v[i] for v in vals) + (int(sums[i]),) for i in range(ng)]
        return sorted(rows), stats
    packed = None
    if len(keys) > 2 or (len(keys) == 2 and any(np.dtype(k.dtype).itemsize > 4 for k in keys)):
        ranges = [ops.column_min_max(ctx, k, n)[:2] for k in keys]
        packed = ops.packing_for(ranges)
        assert packed is not None
    gb = ops.GroupBy(ctx, ktypes, [(C.AGG_SUM, C.TYPE_INT64, False)], max_groups=max(1, n), packed=packed)
    gb.consume(keys, [cur["measure"]], n)
    kcols, aggs = gb.finalize()
    gb.close()
    sums = aggs[0].numpy()
    kvals = [k.numpy() for k in kcols]
    rows = []
    for i in range(len(sums)):
        rows.append(tuple(decode(cols, g, kvals[j][i]) for j, g in enumerate(plan["group"])) + (int(sums[i]),))
    return sorted(rows), stats
