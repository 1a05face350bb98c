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


def groupby_workload(n=60_000_000, seed=3, n_values=100_000):
    return {"key": zipf_keys(seed, 1, n, n_values),
            "qty": uniform_int(seed, 2, n, 1, 50),
            "price": uniform_int(seed, 3, n, 90_000, 10_500_000),
            "disc": uniform_unit(seed, 4, n) * 0.10}
