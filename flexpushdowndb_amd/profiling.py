"""rocprofv3 summaries of the benchmark workloads: per-kernel duration, HBM
traffic and unit utilisation, written under profiles/.

    python -m flexpushdowndb_amd.profiling --workload join [--out-dir profiles]

runs `bench.py --only <workload>` under rocprofv3 four times (the kernel trace,
then three counter passes that each fit the hardware's per-block counter
limits) and writes profiles/<name>.md plus one <group>.traffic.json per kernel
group that bench.py reports a roofline for.  Run it where the GPU is; it only
starts rocprofv3 as child processes and never touches the GPU itself.

HBM bytes come from the L2's memory-side request counters, by request size:
read bytes = 128 TCC_EA0_RDREQ_128B + 64 TCC_EA0_RDREQ_64B + 32
TCC_EA0_RDREQ_32B, write bytes = 64 TCC_EA0_WRREQ_64B + 32 (TCC_EA0_WRREQ -
TCC_EA0_WRREQ_64B).  (On gfx950 the L2 reads memory in 128-byte requests and
TCC_BUBBLE, which rocprofv3's FETCH_SIZE expression takes as the 128-byte
count, reads zero, so FETCH_SIZE counts every read at half its size; a plain
copy kernel of known size reads and writes exactly the bytes above.)
"""
import argparse
import csv
import glob
import json
import os
import subprocess
import sys
from collections import defaultdict

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PASSES = {
    "rd": ["TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_sum"],
    "wr": ["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_HIT_sum", "TCC_MISS_sum"],
    "busy": ["TA_TA_BUSY_sum", "GRBM_GUI_ACTIVE", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CU_CYCLES"],
}

# kernel groups that bench.py reports a roofline for: name -> (workload,
# kernels whose per-dispatch means add up to one launch of the group: a name
# without template arguments matches every instantiation, one with them only
# that one; description)
GROUPS = {
    "hj_probe": ("join", ["hj_dense_count_kernel", "tile_scan_kernel", "hj_dense_emit_kernel"],
                 "dense join probe: count pass + tile scan + emit pass"),
    "hj_probe_hash": ("join", ["hj_match_kernel", "tile_scan_kernel", "hj_emit_kernel"],
                      "hash-table join probe: match pass + tile scan + emit pass"),
    "bloom_blocked_find": ("bloom", ["blocked_find_kernel<1, false, true>", "find_combine_kernel"],
                           "blocked bloom find on the 2^16-block filter: sliced find + combine"),
    "groupby_consume": ("groupby", ["gb_sample_count_kernel", "gb_sample_gather_kernel", "gb_sample_select_kernel",
                                    "gb_count_kernel", "gb_bucket_scan_kernel", "gb_bucket_offsets_kernel",
                                    "gb_main_staged_kernel", "gb_head_merge_kernel", "gb_compact_kernel",
                                    "gb_compact_piece_kernel", "gb_compact_merge_kernel"],
                        "group-by consume: sample + count + offsets + main + head merge + compact (pieces + merge)"),
    "shuffle_scatter": ("shuffle", ["shuffle_scatter_kernel"], "hash shuffle scatter pass"),
}


def short_name(full):
    name = full.split("(")[0]
    if name.startswith("void "):
        name = name[5:]
    return name.replace("fpdb::", "")


def base_name(short):
    return short.split("<")[0]


def run(cmd, log):
    print("[profiling] " + " ".join(cmd), file=sys.stderr, flush=True)
    with open(log, "w") as f:
        r = subprocess.run(cmd, stdout=f, stderr=subprocess.STDOUT, cwd="/tmp",
                           env=dict(os.environ, TMPDIR="/tmp"))
    if r.returncode != 0:
        raise SystemExit(f"profiling: command failed ({r.returncode}), see {log}")


def bench_cmd(workload, steps):
    return ["python3", os.path.join(ROOT, "bench.py"), "--only", workload, "--steps", str(steps), "--warmup", "1",
            "--no-cpu-baseline"]


def find_csv(d, suffix):
    hits = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    if not hits:
        raise SystemExit(f"profiling: no {suffix} under {d}")
    return hits[0]


def collect(workload, raw_dir, steps, rerun=True):
    """Runs the trace and counter passes (or, rerun False, reads a previous
    run's output); returns (trace stats, per-kernel counter means, the
    commands)."""
    os.makedirs(raw_dir, exist_ok=True)
    cmds = []
    d = os.path.join(raw_dir, f"trace_{workload}")
    cmd = ["rocprofv3", "--kernel-trace", "--stats", "-d", d, "-o", "run", "--output-format", "csv", "--"]
    cmd += bench_cmd(workload, steps)
    if rerun:
        run(["timeout", "-s", "KILL", "180"] + cmd, d + ".log")
    cmds.append(cmd)
    stats = {}
    for r in csv.DictReader(open(find_csv(d, "kernel_stats.csv"))):
        stats[short_name(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3)
    counters = defaultdict(lambda: defaultdict(list))
    for tag, names in PASSES.items():
        d = os.path.join(raw_dir, f"pmc_{workload}_{tag}")
        cmd = ["rocprofv3", "--pmc", *names, "-d", d, "-o", "run", "--output-format", "csv", "--"]
        cmd += bench_cmd(workload, 3)
        if rerun:
            run(["timeout", "-s", "KILL", "120"] + cmd, d + ".log")
        cmds.append(cmd)
        for r in csv.DictReader(open(find_csv(d, "counter_collection.csv"))):
            counters[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    means = {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in counters.items()}
    return stats, means, cmds


def traffic(m):
    rd = 128 * m.get("TCC_EA0_RDREQ_128B_sum", 0) + 64 * m.get("TCC_EA0_RDREQ_64B_sum", 0) + \
        32 * m.get("TCC_EA0_RDREQ_32B_sum", 0)
    w64 = m.get("TCC_EA0_WRREQ_64B_sum", 0)
    wr = 64 * w64 + 32 * (m.get("TCC_EA0_WRREQ_sum", 0) - w64)
    return rd, wr


def summarize(workload, name, title, stats, means, cmds, out_dir, num_cus=256, num_xcds=8):
    rows = []
    for k, (calls, us) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        m = means.get(k, {})
        rd, wr = traffic(m)
        gui = m.get("GRBM_GUI_ACTIVE", 0) / num_xcds
        ta = m.get("TA_TA_BUSY_sum", 0) / (num_cus * gui) if gui else float("nan")
        hit = m.get("TCC_HIT_sum", 0)
        miss = m.get("TCC_MISS_sum", 0)
        l2 = hit / (hit + miss) if hit + miss else float("nan")
        rows.append((k, calls, us, rd / 1e6, wr / 1e6, ta, l2))
    lines = [f"# {name}", "", title, "",
             "Made with (from the repository root on the GPU host; raw output under gpurun_out/):", "",
             f"    python -m flexpushdowndb_amd.profiling --workload {workload}", "",
             "which runs:", ""]
    lines += ["    " + " ".join(c) for c in cmds]
    lines += ["", "Read and written MB per dispatch from the L2's memory-side request counters by request size "
                  "(128 x RDREQ_128B + 64 x RDREQ_64B + 32 x RDREQ_32B; 64 x WRREQ_64B + 32 x the other write "
                  "requests); TA busy = TA_TA_BUSY_sum / (CUs x kernel cycles); L2 hit = TCC_HIT / (HIT + MISS).", "",
              "| kernel | calls | average us | read MB | written MB | TA busy | L2 hit |",
              "|---|---|---|---|---|---|---|"]
    for k, calls, us, rd, wr, ta, l2 in rows:
        lines.append(f"| `{k[:90]}` | {calls} | {us:.1f} | {rd:.1f} | {wr:.1f} | {ta:.2f} | {l2:.2f} |")
    path = os.path.join(out_dir, f"{name}.md")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"[profiling] wrote {path}", file=sys.stderr)
    for g, (wl, prefixes, desc) in GROUPS.items():
        if wl != workload:
            continue
        total_rd = total_wr = 0.0
        total_us = 0.0
        found = []
        for k, calls, us, rd, wr, ta, l2 in rows:
            if base_name(k) in prefixes or k in prefixes:
                total_rd += rd * 1e6
                total_wr += wr * 1e6
                total_us += us
                found.append(k)
        if not found:
            continue
        with open(os.path.join(out_dir, f"{g}.traffic.json"), "w") as f:
            json.dump({"kernels": desc, "matched": found, "bytes_per_launch": total_rd + total_wr,
                       "read_bytes": total_rd, "write_bytes": total_wr, "kernel_us_sum": total_us,
                       "source": f"{name}.md"}, f, indent=1)


NAMES = {"join": ("hj_probe_join_1m_build_6m_probe",
                  "Inner hash join, build 1,000,000 rows {int64 key permutation, int32 pay}, probe 6,000,000 rows "
                  "{int64 fk, int32 qty, float64 price}, 5 output columns: the dense-table form, the same join with "
                  "keys spread 3x (hash-table form) and a 4,000-row build side (hash-table form)."),
         "bloom": ("bloom_blocked_find_60m_keys",
                   "Bloom filter create/use: filter from ~400k kept keys (2^16 blocks), 60,000,000 int64 probe keys; "
                   "also the 2^14-block LDS filter, the legacy universal-hash filter and the compaction."),
         "groupby": ("groupby_zipf_60m_rows",
                     "Hash group-by: 60,000,000 rows, int32 Zipf(1.1) key over 100,000 values, COUNT(*), SUM int64 x2, "
                     "SUM float64."),
         "shuffle": ("shuffle_60m_rows_64_slots", "Hash shuffle: 60,000,000 rows {int64 key, int32, float64} into 64 slots."),
         "star": ("star_persistent_60m_fact_64_partitions",
                  "Batched star join: 60M fact rows in 64 partitions, 4 dimensions, forward/backward bloom filters, "
                  "joins and group-by in one persistent launch, and the plain per-partition launches."),
         "serve": ("serve_scan_60m_fact", "Serving loop over the resident 60M-row star schema."),
         "cache": ("segment_cache_join_probe_columns", "Join through the device segment cache."),
         "wire": ("wire_bitmap_words_60m_rows", "Wire forms of bitmaps and filters."),
         "calib": ("copy_calibration", "Plain copy of 64 Mi int64 values: reads and writes 536,870,912 bytes each.")}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", required=True, choices=sorted(NAMES))
    ap.add_argument("--out-dir", default=os.path.join(ROOT, "profiles"))
    ap.add_argument("--raw-dir", default=os.path.join(ROOT, "gpurun_out", "prof"))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--summarize-only", action="store_true", help="summarise the raw output of an earlier run")
    a = ap.parse_args()
    name, title = NAMES[a.workload]
    stats, means, cmds = collect(a.workload, os.path.abspath(a.raw_dir), a.steps, rerun=not a.summarize_only)
    summarize(a.workload, name, title, stats, means, cmds, os.path.abspath(a.out_dir))


if __name__ == "__main__":
    main()
