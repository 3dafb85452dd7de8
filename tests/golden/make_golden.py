#!/usr/bin/env python3
"""Write golden fixtures from the reference's own serial CG code.

Runs in a container that has the reference tree: ``make -C oracle ref`` first
builds ``oracle/_ref/libnpbref_<CLASS>.so`` from the reference sources (the
serial ``makea`` and ``conj_grad`` of HIP/CG/cg.cpp, sizes fixed per NPB class
at compile time).  For each case this script calls the reference generator
with the case's seed and the reference's serial inverse-power loop, and stores
only data: sizes, SHA-256 digests of the assembled rowstr/colidx/a bytes, a
few verbatim rows, the bits of ``ratio`` and the per-iteration (zeta, rnorm)
as exact hex floats.  Output: tests/golden/reference_cases.json.

Usage:  python tests/golden/make_golden.py [--jobs 6]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# name, NPB class whose compile-time sizes apply, seed, outer iterations run
CASES = [
    ("S_npb", "S", 314159265.0, 15),
    ("cfg1_n1400_seed12345", "S", 12345.0, 15),
    ("W_npb", "W", 314159265.0, 15),
    ("A_npb", "A", 314159265.0, 15),
    ("B_npb", "B", 314159265.0, 75),
    ("cfg2_n75000_seed271828", "B", 271828.0, 75),
    ("C_npb", "C", 314159265.0, 75),
    ("cfg3_n150000_seed314", "C", 314.0, 3),
]


def sha(arr) -> str:
    return hashlib.sha256(arr.tobytes()).hexdigest()


def run_case(case):
    import numpy as np

    import oracle

    name, cls, seed, niter = case
    ref = oracle.RefClass(cls)
    t0 = time.time()
    m = ref.makea(seed)
    t_gen = time.time() - t0
    t0 = time.time()
    zeta, rnorm, x = ref.power(niter)
    t_pow = time.time() - t0
    n = m.n
    rows = {}
    for j in (0, 1, n // 2, n - 1):
        s, e = int(m.rowstr[j]), int(m.rowstr[j + 1])
        rows[str(j)] = {
            "colidx": [int(c) for c in m.colidx[s:e][:8]],
            "a": [float(v).hex() for v in m.a[s:e][:8]],
            "len": e - s,
        }
    return {
        "name": name,
        "npb_class": cls,
        "n": n,
        "nonzer": ref.nonzer,
        "shift": ref.shift,
        "rcond": ref.rcond,
        "seed": seed,
        "niter": niter,
        "nnz": int(m.rowstr[n]),
        "sha256_rowstr": sha(m.rowstr),
        "sha256_colidx": sha(m.colidx),
        "sha256_a": sha(m.a),
        "ratio_hex": float(np.power(ref.rcond, 1.0 / n)).hex(),
        "rows_head": rows,
        "zeta_hex": [float(v).hex() for v in zeta],
        "rnorm_hex": [float(v).hex() for v in rnorm],
        "x_final_sha256": sha(x),
        "x_final_head_hex": [float(v).hex() for v in x[:4]],
        "seconds_generate": round(t_gen, 2),
        "seconds_power": round(t_pow, 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--out", default=os.path.join(HERE, "reference_cases.json"))
    args = ap.parse_args()
    with ProcessPoolExecutor(max_workers=args.jobs) as ex:
        results = list(ex.map(run_case, CASES))
    doc = {
        "source": "reference serial makea/conj_grad (HIP/CG/cg.cpp) built by oracle/Makefile `ref`",
        "note": "zeta/rnorm per outer iteration of the serial power loop from x = 1; floats as hex",
        "cases": results,
    }
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    for r in results:
        print(r["name"], r["nnz"], r["seconds_generate"], r["seconds_power"], flush=True)


if __name__ == "__main__":
    main()
