"""Writes the SSB scale-0.01 fixtures used by the end-to-end star-join tests.

Inputs (data files shipped with the reference, read as data only):
  fpdb-main/test-resources/fpdb-store-0/flexpushdowndb/ssb-sf0.01/parquet/{lineorder,date,part,supplier,customer}.parquet
  resources/result/ssb-sf0.01/ssb/original/{1.1 ... 4.3}.r   (the reference's printed query answers)

Outputs (committed):
  ssb_sf0.01.npz           the columns the 13 SSB queries touch; string columns
                           dictionary-encoded with a sorted dictionary (so range
                           predicates on strings become range predicates on codes)
  ssb_sf0.01_answers.json  the .r answers parsed into {query: {columns, types, rows}}

Run from the repository root:  python tests/golden/make_ssb_fixtures.py [reference_root]
"""
import json
import os
import sys

import numpy as np
import pyarrow.parquet as pq

HERE = os.path.dirname(os.path.abspath(__file__))
QUERIES = ["1.1", "1.2", "1.3", "2.1", "2.2", "2.3", "3.1", "3.2", "3.3", "3.4", "4.1", "4.2", "4.3"]
COLUMNS = {
    "lineorder": ["lo_custkey", "lo_partkey", "lo_suppkey", "lo_orderdate", "lo_quantity", "lo_discount",
                  "lo_extendedprice", "lo_revenue", "lo_supplycost"],
    "date": ["d_datekey", "d_year", "d_yearmonthnum", "d_yearmonth", "d_weeknuminyear"],
    "part": ["p_partkey", "p_mfgr", "p_category", "p_brand1"],
    "supplier": ["s_suppkey", "s_city", "s_nation", "s_region"],
    "customer": ["c_custkey", "c_city", "c_nation", "c_region"],
}


def parse_answer(path):
    """The .r files are tables printed as '| a | b' lines: names, types, rows."""
    lines = [ln.rstrip("\n") for ln in open(path) if ln.startswith("|")]
    cells = [[c.strip() for c in ln[1:].split("|")] for ln in lines]
    names, types, rows = cells[0], cells[1], cells[2:]
    out = []
    for r in rows:
        out.append([int(v) if t == "int64" else float(v) if t == "double" else v for v, t in zip(r, types)])
    return {"columns": names, "types": types, "rows": out}


def main(ref_root):
    pq_dir = os.path.join(ref_root, "fpdb-main/test-resources/fpdb-store-0/flexpushdowndb/ssb-sf0.01/parquet")
    arrays = {}
    for table, cols in COLUMNS.items():
        t = pq.read_table(os.path.join(pq_dir, table + ".parquet"), columns=cols)
        for c in cols:
            col = t.column(c)
            assert col.null_count == 0, c
            v = col.to_numpy()
            if v.dtype == object:                       # string column -> sorted dictionary codes
                dictionary, codes = np.unique(v.astype(str), return_inverse=True)
                arrays[c] = codes.astype(np.int32)
                arrays[c + "__dict"] = dictionary
            else:
                assert v.dtype == np.int64 and np.abs(v).max() < 2**31, c
                arrays[c] = v.astype(np.int32)          # widened back to int64 by the loader
    np.savez_compressed(os.path.join(HERE, "ssb_sf0.01.npz"), **arrays)
    res_dir = os.path.join(ref_root, "resources/result/ssb-sf0.01/ssb/original")
    answers = {q: parse_answer(os.path.join(res_dir, q + ".r")) for q in QUERIES}
    with open(os.path.join(HERE, "ssb_sf0.01_answers.json"), "w") as f:
        json.dump(answers, f, indent=1)
    print("wrote", sorted(k for k in arrays if not k.endswith("__dict")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
