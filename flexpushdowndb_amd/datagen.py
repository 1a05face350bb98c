"""Seeded synthetic inputs for the hot-path workloads.

All columns come from a counter-based generator, SplitMix64 applied to
seed ^ (column * 0x9E3779B97F4A7C15) ^ row, so any row of any column can be
regenerated independently and the CPU and GPU sides see identical data.

Workloads (sizes are the defaults; tests pass smaller ones):
  join      build 1,000,000 rows {key int64 = permutation of 0..n-1,
            pay int32 in [0, 2^31)}; probe 6,000,000 rows {fk int64 uniform
            over the build keys, qty int32 in [1, 50], price float64 =
            k/100 for k in [90000, 10500000]}.
  bloom     2,000,000 rows {key int64 in [0, 2^40), attr int32 in [0, 100)},
            kept where attr < 20; 60,000,000 probe keys, half drawn from the
            kept keys and half uniform in [0, 2^40).
  groupby   60,000,000 rows {key int32 Zipf(s=1.1) over 100,000 values
            (inverse CDF of the rank, rank -> key by a seeded permutation),
            qty int64 in [1, 50], price int64 cents in [90000, 10500000],
            disc float64 = U[0, 1) * 0.10}.
"""
import numpy as np

PHI = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    with np.errstate(over="ignore"):
        z = x + PHI
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def rand_u64(seed, column, n, start=0):
    with np.errstate(over="ignore"):
        base = np.uint64(seed) ^ (np.uint64(column) * PHI)
        rows = np.arange(start, start + n, dtype=np.uint64)
        return splitmix64(base ^ rows)


def uniform_int(seed, column, n, lo, hi, start=0):
    """Integers in [lo, hi] (inclusive)."""
    span = np.uint64(hi - lo + 1)
    return (rand_u64(seed, column, n, start) % span).astype(np.int64) + lo


def uniform_unit(seed, column, n, start=0):
    """Doubles in [0, 1) with 53 random bits."""
    return (rand_u64(seed, column, n, start) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def permutation(seed, column, n):
    keys = rand_u64(seed, column, n)
    return np.argsort(keys, kind="stable").astype(np.int64)


# --- workloads -------------------------------------------------------------

def join_workload(n_build=1_000_000, n_probe=6_000_000, seed=1):
    build_key = permutation(seed, 1, n_build)
    build_pay = uniform_int(seed, 2, n_build, 0, (1 << 31) - 1).astype(np.int32)
    probe_fk = uniform_int(seed, 3, n_probe, 0, n_build - 1)
    probe_qty = uniform_int(seed, 4, n_probe, 1, 50).astype(np.int32)
    probe_price = uniform_int(seed, 5, n_probe, 90_000, 10_500_000).astype(np.float64) / 100.0
    return {"build_key": build_key, "build_pay": build_pay, "probe_fk": probe_fk,
            "probe_qty": probe_qty, "probe_price": probe_price}


def bloom_workload(n_build=2_000_000, n_probe=60_000_000, seed=2, keep_below=20):
    key = uniform_int(seed, 1, n_build, 0, (1 << 40) - 1)
    attr = uniform_int(seed, 2, n_build, 0, 99).astype(np.int32)
    kept = key[attr < keep_below]
    pick = rand_u64(seed, 3, n_probe)
    from_kept = (pick & np.uint64(1)) == np.uint64(0)
    idx = ((pick >> np.uint64(1)) % np.uint64(max(1, len(kept)))).astype(np.int64)
    probe = uniform_int(seed, 4, n_probe, 0, (1 << 40) - 1)
    if len(kept):
        probe[from_kept] = kept[idx[from_kept]]
    return {"key": key, "attr": attr, "kept": kept, "probe": probe}


def zipf_keys(seed, column, n, n_values=100_000, s=1.1):
    ranks = np.arange(1, n_values + 1, dtype=np.float64)
    w = ranks ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = uniform_unit(seed, column, n)
    r = np.searchsorted(cdf, u, side="right")
    r = np.minimum(r, n_values - 1)
    perm = permutation(seed, column + 100, n_values)
    return perm[r].astype(np.int32)


def lognormal_partitions(n_rows, n_parts, seed, column=50):
    """Partition sizes proportional to lognormal(0, 1) draws (Box-Muller on
    the counter RNG), normalised to n_rows by largest remainder.  Returns the
    n_parts + 1 row offsets."""
    u1 = np.maximum(uniform_unit(seed, column, n_parts), 1e-300)
    u2 = uniform_unit(seed, column + 1, n_parts)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    w = np.exp(z)
    exact = w / w.sum() * n_rows
    sizes = np.floor(exact).astype(np.int64)
    rest = n_rows - int(sizes.sum())
    order = np.argsort(-(exact - sizes), kind="stable")
    sizes[order[:rest]] += 1
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


# star join: four dimensions of 30,000 / 200,000 / 2,000,000 / 2,556 rows kept
# at 20% / 4% / 1% / 14%, group attributes on dimensions 0 and 3
STAR_DIMS = (30_000, 200_000, 2_000_000, 2_556)
STAR_KEEP_BELOW = (20, 4, 1, 14)
STAR_GROUP_DOMAIN = {0: 40, 3: 12}


def star_workload(n_fact=60_000_000, n_parts=64, seed=4, dims=STAR_DIMS, keep_below=STAR_KEEP_BELOW):
    out = {"dims": [], "fact_fk": [], "offsets": lognormal_partitions(n_fact, n_parts, seed)}
    for d, rows in enumerate(dims):
        dim = {"key": np.arange(rows, dtype=np.int64),
               "attr": uniform_int(seed, 10 + d, rows, 0, 99).astype(np.int32),
               "keep_below": keep_below[d]}
        if d in STAR_GROUP_DOMAIN:
            dim["group"] = uniform_int(seed, 20 + d, rows, 0, STAR_GROUP_DOMAIN[d] - 1).astype(np.int32)
        out["dims"].append(dim)
        out["fact_fk"].append(uniform_int(seed, 30 + d, n_fact, 0, rows - 1))
    out["measure"] = uniform_int(seed, 40, n_fact, 1, 100)
    return out


def star_expected(w, group_dims=(0, 3)):
    """SELECT g_a, g_b, COUNT(*), SUM(measure) ... in first-appearance order,
    by direct indexing of the dense dimension keys."""
    keep = np.ones(len(w["measure"]), bool)
    for d, dim in enumerate(w["dims"]):
        keep &= (dim["attr"] < dim["keep_below"])[w["fact_fk"][d]]
    ga = w["dims"][group_dims[0]]["group"][w["fact_fk"][group_dims[0]][keep]].astype(np.int64)
    gb = w["dims"][group_dims[1]]["group"][w["fact_fk"][group_dims[1]][keep]].astype(np.int64)
    meas = w["measure"][keep]
    key = ga * (1 << 32) + gb
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    cnt = np.bincount(inv, minlength=len(uniq))
    sm = np.bincount(inv, weights=meas, minlength=len(uniq)).astype(np.int64)
    order = np.argsort(first, kind="stable")
    return {"g_a": ga[first[order]], "g_b": gb[first[order]], "count": cnt[order], "sum": sm[order],
            "survivors": int(keep.sum())}


def groupby_workload(n=60_000_000, seed=3, n_values=100_000):
    return {"key": zipf_keys(seed, 1, n, n_values),
            "qty": uniform_int(seed, 2, n, 1, 50),
            "price": uniform_int(seed, 3, n, 90_000, 10_500_000),
            "disc": uniform_unit(seed, 4, n) * 0.10}


# ---- serving workload: the star schema resident in HBM, random queries ------

SERVE_GROUP_DOMAIN = (40, 25, 50, 12)
SERVE_ATTR_DOMAIN = 100_000


def serve_workload(n_fact=60_000_000, seed=6, dims=STAR_DIMS, group_domain=SERVE_GROUP_DOMAIN):
    """Star schema for the serving loop: dense dimension keys, a filter
    attribute uniform in [0, 100000) and a group attribute per dimension."""
    out = {"dims": [], "fact_fk": []}
    for d, rows in enumerate(dims):
        out["dims"].append({
            "key": np.arange(rows, dtype=np.int64),
            "attr": uniform_int(seed, 10 + d, rows, 0, SERVE_ATTR_DOMAIN - 1).astype(np.int32),
            "group": uniform_int(seed, 20 + d, rows, 0, group_domain[d] - 1).astype(np.int32),
            "domain": group_domain[d]})
        out["fact_fk"].append(uniform_int(seed, 30 + d, n_fact, 0, rows - 1))
    out["measure"] = uniform_int(seed, 40, n_fact, 1, 100)
    return out


def serve_queries(n, seed=7, n_dims=4):
    """Queries of the serving workload: 1-4 filtered dimensions, each kept
    with selectivity log-uniform in [0.1 %, 50 %] (attr < c), grouped by 1-2
    dimensions' group attributes."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(1, n_dims + 1))
        fdims = sorted(rng.choice(n_dims, size=k, replace=False).tolist())
        sel = 10 ** rng.uniform(np.log10(0.001), np.log10(0.5), size=k)
        consts = [max(1, int(round(s * SERVE_ATTR_DOMAIN))) for s in sel]
        g = int(rng.integers(1, 3))
        gdims = rng.choice(n_dims, size=g, replace=False).tolist()
        out.append({"filters": list(zip(fdims, consts)), "group": gdims})
    return out


def serve_expected(w, q):
    """Result of one serving query by direct indexing of the dense keys:
    groups (g0, g1) with COUNT(*) and SUM(measure) in first-appearance order."""
    keep = np.ones(len(w["measure"]), bool)
    for d, c in q["filters"]:
        keep &= (w["dims"][d]["attr"] < c)[w["fact_fk"][d]]
    rows = np.flatnonzero(keep)
    gd = q["group"]
    g0 = w["dims"][gd[0]]["group"][w["fact_fk"][gd[0]][rows]].astype(np.int64)
    g1 = (w["dims"][gd[1]]["group"][w["fact_fk"][gd[1]][rows]].astype(np.int64) if len(gd) == 2
          else np.zeros(len(rows), np.int64))
    key = g0 * (1 << 32) + g1
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    cnt = np.bincount(inv, minlength=len(uniq))
    sm = np.bincount(inv, weights=w["measure"][rows], minlength=len(uniq)).astype(np.int64)
    order = np.argsort(first, kind="stable")
    return {"g0": g0[first[order]], "g1": g1[first[order]], "count": cnt[order], "sum": sm[order]}
