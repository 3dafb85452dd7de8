#!/usr/bin/env python3
"""Benchmark of the NPB CG conj_grad hot path on one MI355X.

Workload (one "step"): the plain FP64 inverse-power run at n=75000 — NPB makea
matrix with nonzer=13, rcond=0.1, shift=60, seed 271828 (13,708,444 nnz, int32
CSR) — x reset to ones, then 75 outer iterations of {25-step conj_grad,
zeta = shift + 1/(x.z), x = z/||z||}, all on the device (the timed region of
cg.cu:491-522).  The matrix and vectors are resident in HBM before timing.

value: whole-step algorithmic bytes / wall time (GB/s), with the algorithmic
bytes of one outer iteration = 26 * B_spmv + 25 * 72 n + 56 n, where
B_spmv = 12 nnz + 4 (n+1) + 16 n (matrix streamed once, vector read once,
result written once).  The SpMV kernel's own duration is
sampled with HIP events on the library's stream during the timed steps and
gives the roofline entry.  cpu_baseline: the reference's serial power loop
(its own code, built under oracle/_ref from the reference sources; the
oracle's C restatement where that build is absent), single thread, on the
same matrix.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "npb-gpu_amd"))

import numpy as np  # noqa: E402

METRIC = "CG time to 1e-10 rel. residual (learned+mixed vs plain FP64); SpMV GB/s"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(workload="npb_cg_fp64_power_n75000", n=75000, nonzer=13, shift=60.0,
                rcond=0.1, seed=271828.0, niter=75, cgitmax=25)


def spmv_bytes(n: int, nnz: int, val_bytes: int = 8, vec_bytes: int = 8) -> int:
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * n * vec_bytes


def outer_iter_bytes(n: int, nnz: int, cgitmax: int) -> int:
    # SpMVs: cgitmax of q = A p plus one r = A z; vector traffic per CG step:
    # z,r update (read z,p,r,q; write z,r) 48n + p update 24n; init and the
    # residual/norm pass 56n (z read, x rescaled, q,z,r,p written; x read).
    return (cgitmax + 1) * spmv_bytes(n, nnz) + cgitmax * 72 * n + 56 * n


def hip_device_synchronize():
    lib = ctypes.CDLL("libamdhip64.so")
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def load_traffic():
    path = os.path.join(ROOT, "profiles", "spmv_fp64_n75000_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def golden_zeta():
    path = os.path.join(ROOT, "tests", "golden", "reference_cases.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    for c in doc["cases"]:
        if c["name"] == "cfg2_n75000_seed271828":
            return np.array([float.fromhex(h) for h in c["zeta_hex"]])
    return None


def cpu_baseline(m, wl, outer_iters: int):
    """The reference's serial power loop (conj_grad + zeta + x = z/||z||,
    cg.cu:458-482 / HIP/CG/cg.cpp) on `outer_iters` outer iterations of the
    same matrix, one pinned thread.  It runs the reference build under
    oracle/_ref (compiled from the reference sources by oracle/Makefile) when
    that build is present and its class matches the workload; otherwise the
    oracle's C restatement of the same loop."""
    import oracle

    ref = None
    if oracle.ref_available("B"):
        ref = oracle.RefClass("B")
        if (ref.n, ref.nonzer, ref.shift) != (wl["n"], wl["nonzer"], wl["shift"]):
            ref = None
    same_matrix = None
    if ref is not None:
        rm = ref.makea(wl["seed"])  # the reference's own generator, outside the timing
        same_matrix = bool(np.array_equal(rm.rowstr, m.rowstr)
                           and np.array_equal(rm.colidx, m.colidx)
                           and np.array_equal(rm.a.view(np.int64), m.a.view(np.int64)))
    else:
        om = oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, wl["nonzer"], wl["shift"],
                        wl["rcond"], wl["seed"])
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        if ref is not None:
            zeta, _, _ = ref.power(outer_iters)
        else:
            zeta, _, _ = oracle.power(om, outer_iters, wl["cgitmax"])
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    gbytes = outer_iters * outer_iter_bytes(m.n, m.nnz, wl["cgitmax"]) / 1e9
    what = ("reference HIP/CG/cg.cpp serial power loop (oracle/_ref/libnpbref_B.so, built from "
            "the reference sources)" if ref is not None else
            "oracle/npbcg_oracle.c serial power loop")
    out = {
        "value": round(gbytes / dt, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "reference" if ref is not None else "port",
        "sample": (f"{outer_iters} of {wl['niter']} outer iterations (x {wl['cgitmax']} CG "
                   f"steps) of the same n={m.n} matrix, {what}, one pinned thread, "
                   f"{dt:.2f} s"),
        "seconds": round(dt, 3),
        "outer_iters_per_s": round(outer_iters / dt, 4),
        "cg_iters_per_s": round(outer_iters * wl["cgitmax"] / dt, 2),
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
        "zeta_last": float(zeta[-1]),
    }
    if same_matrix is not None:
        out["reference_matrix_bit_identical"] = same_matrix
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


PROFILE_COPY_BYTES = 256 << 20


def profile_pass():
    """Short run for hardware-counter passes: a copy of known size (the
    calibration kernel), then two outer iterations launched without graphs."""
    import npbcg

    wl = WORKLOAD
    m = npbcg.makea(wl["n"], wl["nonzer"], wl["shift"], wl["seed"], wl["rcond"])
    with npbcg.Context(m, device=0, cgitmax=wl["cgitmax"], use_graph=0) as ctx:
        ctx.copy_bench(PROFILE_COPY_BYTES, 3)
        ctx.fill_x(1.0)
        ctx.power(2, wl["shift"])
        print(json.dumps({"profile_pass": True, "copy_bytes": PROFILE_COPY_BYTES,
                          "kernel_config": ctx.describe()}))


def profile_pass_solve():
    """Short run for counter passes on the n=150000 solve matrix: a copy of
    known size (the calibration kernel), then 10 launches each of the FP64 and
    the FP32 CG SpMV kernels (npbcg_spmv_kernel_bench)."""
    import npbcg

    sw = SOLVE_WORKLOAD
    m = npbcg.makea(sw["n"], sw["nonzer"], sw["shift"], sw["seed"], sw["rcond"])
    with npbcg.Context(m, device=0, use_graph=0) as ctx:
        ctx.copy_bench(PROFILE_COPY_BYTES, 3)
        fp64 = ctx.spmv_kernel_bench(0, 10)
        fp32 = ctx.spmv_kernel_bench(1, 10)
        print(json.dumps({"profile_pass_solve": True, "copy_bytes": PROFILE_COPY_BYTES,
                          "fp64_ms": fp64, "fp32_ms": fp32, "kernel_config": ctx.describe()}))


def profile_pass_batch():
    """Short run for counter passes on the batched SpMM: the 64-system batch
    (layout 3, 2048-row items), 20 SpMM launches in MIXED mode (the FP32
    kernel) and 20 in FP64 mode, every system running, after a copy of known
    size (the calibration kernel)."""
    import npbcg

    bw = BATCH_WORKLOAD
    mats = npbcg.make_many(batch_systems(bw))
    out = {"profile_pass_batch": True, "copy_bytes": PROFILE_COPY_BYTES}
    with npbcg.Context(mats[0], device=0, use_graph=0) as ctx:
        ctx.copy_bench(PROFILE_COPY_BYTES, 3)
    for name, mode in (("fp32", npbcg.SOLVE_MIXED), ("fp64", npbcg.SOLVE_FP64)):
        with npbcg.Batch(mats, nrhs=bw["nrhs"], mode=mode, rows_per_tile=2048, layout=3) as bt:
            out[name + "_spmm_ms"] = round(bt.spmm_bench(20), 4)
    print(json.dumps(out))


def _pmc_rows(path, counter):
    import csv

    out = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name", counter) != counter:
                continue
            name = row.get("Kernel_Name", "")
            val = float(row.get("Counter_Value", "nan"))
            out.setdefault(name, []).append(val)
    return out


def pmc_summary(fetch_csv, write_csv, out_path):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the SpMV kernel, scaled by
    what the counters read for the copy kernel of known traffic."""
    fetch = _pmc_rows(fetch_csv, "FETCH_SIZE")
    write = _pmc_rows(write_csv, "WRITE_SIZE")

    def pick(d, key):
        # the kernel of that name with the most dispatches (the CG SpMV, not
        # the once-per-iteration residual SpMV of the same template)
        ks = sorted((k for k in d if key in k), key=lambda k: -len(d[k]))
        return (ks[0], d[ks[0]]) if ks else (None, [])

    _, cf = pick(fetch, "copy_kernel")
    _, cw = pick(write, "copy_kernel")
    kname, sf = pick(fetch, "spmv_sell_kernel")
    if kname is None:
        kname, sf = pick(fetch, "spmv_csr_kernel")
    _, sw = pick(write, kname.split("<")[0] if kname else "spmv")
    expect_kib = PROFILE_COPY_BYTES / 1024.0
    f_scale = float(np.median(cf)) / expect_kib
    w_scale = float(np.median(cw)) / expect_kib
    spmv_f = float(np.median(sf)) * 1024.0 / f_scale
    spmv_w = float(np.median(sw)) * 1024.0 / w_scale
    doc = {"kernel": kname, "dispatches": len(sf),
           "copy_fetch_kib_per_dispatch": float(np.median(cf)),
           "copy_write_kib_per_dispatch": float(np.median(cw)),
           "copy_bytes_each_way": PROFILE_COPY_BYTES,
           "fetch_scale": f_scale, "write_scale": w_scale,
           "spmv_fetch_bytes": spmv_f, "spmv_write_bytes": spmv_w,
           "bytes_per_launch": spmv_f + spmv_w,
           "workload": "n=75000 nonzer=13 shift=60 seed=271828, FP64, SELL column-blocked CG SpMV",
           "commands": [
               "rocprofv3 --pmc FETCH_SIZE --output-format csv -d prof_fetch -o fetch "
               "-- python3 bench.py --profile-pass",
               "rocprofv3 --pmc WRITE_SIZE --output-format csv -d prof_write -o write "
               "-- python3 bench.py --profile-pass",
               "python3 bench.py --pmc-summary <fetch csv> <write csv> " + out_path],
           "note": "FETCH_SIZE/WRITE_SIZE are L2-to-fabric KiB per dispatch; each is divided by "
                   "what it reads for the copy kernel of known traffic (fetch_scale, write_scale). "
                   "The matrix fits the 256 MB Infinity Cache, so fetched bytes include hits "
                   "there, not only HBM reads."}
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc))


def pmc_summary_solve(fetch_csv, write_csv, out_path):
    """FETCH_SIZE / WRITE_SIZE per dispatch of the FP64 and FP32 CG SpMV
    kernels of the n=150000 counter pass (--profile-pass-solve), scaled as in
    pmc_summary, beside the bytes the layout streams."""
    fetch = _pmc_rows(fetch_csv, "FETCH_SIZE")
    write = _pmc_rows(write_csv, "WRITE_SIZE")

    def rows(d, key, exclude=None):
        vals = []
        for k, v in d.items():
            if key in k and (exclude is None or exclude not in k):
                vals += v
        return vals

    expect_kib = PROFILE_COPY_BYTES / 1024.0
    f_scale = float(np.median(rows(fetch, "copy_kernel"))) / expect_kib
    w_scale = float(np.median(rows(write, "copy_kernel"))) / expect_kib
    doc = {"workload": "n=150000 nonzer=15 shift=110 seed=314, SELL column-blocked CG SpMV "
                       "kernels (FP64 conj_grad/solve kernel, FP32 mixed inner kernel)",
           "fetch_scale": f_scale, "write_scale": w_scale,
           "copy_bytes_each_way": PROFILE_COPY_BYTES}
    for name, key, excl in (("fp64", "spmv_sell_kernel", None), ("fp32", "spmv_sell32_kernel", None)):
        f = rows(fetch, key, excl)
        w = rows(write, key, excl)
        if not f or not w:
            continue
        fb = float(np.median(f)) * 1024.0 / f_scale
        wb = float(np.median(w)) * 1024.0 / w_scale
        doc[name] = {"dispatches": len(f), "fetch_bytes": fb, "write_bytes": wb,
                     "bytes_per_launch": fb + wb}
    doc["commands"] = [
        "rocprofv3 --pmc FETCH_SIZE --output-format csv -d prof_fetch -o fetch "
        "-- python3 bench.py --profile-pass-solve",
        "rocprofv3 --pmc WRITE_SIZE --output-format csv -d prof_write -o write "
        "-- python3 bench.py --profile-pass-solve",
        "python3 bench.py --pmc-summary-solve <fetch csv> <write csv> " + out_path]
    doc["note"] = ("FETCH_SIZE/WRITE_SIZE are L2-to-fabric KiB per dispatch, divided by what they "
                   "read for a copy of known traffic. Fetched bytes include hits in the 256 MB "
                   "memory-side cache.")
    with open(out_path, "w") as fo:
        json.dump(doc, fo, indent=1)
    print(json.dumps(doc))


SOLVE_WORKLOAD = dict(workload="npb_cg_solve_n150000", n=150000, nonzer=15, shift=110.0,
                      rcond=0.1, seed=314.0, rtol=1e-10)


LEARNED_PARAMS = os.path.join(ROOT, "npb-gpu_amd", "npbcg", "learned_precond.npz")


def sell_streamed_bytes(desc: dict, val_bytes: int) -> int:
    """Bytes one SELL SpMV streams from memory: every stored entry (value +
    16-bit column), the lane-row index of every slice (4 bytes per stored row
    slot, one per row and column block), the row sums written (one per row
    and column block) and the multiplied vector read once."""
    n, nblk, stored = desc["n"], desc["col_blocks"], desc["sell_stored"]
    return stored * (val_bytes + 2) + 4 * nblk * n + nblk * n * val_bytes + n * val_bytes


def spmv_kernel_rooflines(ctx, m, reps: int = 50):
    """Event-timed SpMV kernels on one context's matrix: the FP64 CG SpMV and
    the FP32 SpMV of the mixed solve's inner CG, each launched back to back as
    successive CG steps launch them (walk direction alternating)."""
    desc = ctx.describe()
    out = {"n": m.n, "nnz": m.nnz, "sell_stored": desc["sell_stored"],
           "col_blocks": desc["col_blocks"], "workgroups": desc["sell_workgroups"]}
    for name, prec, vb in (("fp64", 0, 8), ("fp32", 1, 4)):
        ms = ctx.spmv_kernel_bench(prec, reps)
        alg = spmv_bytes(m.n, m.nnz, vb, vb)
        streamed = sell_streamed_bytes(desc, vb)
        out[name] = {
            "avg_launch_ms": round(ms, 5),
            "algorithmic_bytes": alg,
            "streamed_bytes": streamed,
            "achieved_gbs_algorithmic": round(alg / (ms * 1e-3) / 1e9, 1),
            "achieved_gbs_streamed": round(streamed / (ms * 1e-3) / 1e9, 1),
            "frac_algorithmic": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "frac_streamed": round(streamed / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "launches": reps,
        }
    return out


def solve_comparison(reps: int, use_graph: int = 1):
    """Time to a 1e-10 FP64-checked relative residual on the n=150000 matrix
    (b = 2u - 1 from the LCG stream seeded 315), each variant the median of
    `reps` solves with b resident: plain FP64 CG, FP64-refined mixed
    precision, and mixed precision preconditioned by the top-K Neumann form
    (g = h = 0) and by the trained network (when its parameter file exists).
    Preconditioner set-up (graph, network inference, operator values) is
    timed separately: it is paid once per matrix, not per solve."""
    import npbcg

    sw = SOLVE_WORKLOAD
    m = npbcg.makea(sw["n"], sw["nonzer"], sw["shift"], sw["seed"], sw["rcond"])
    b = rhs_for(sw["n"], sw["seed"])
    out = {"workload": sw["workload"], "n": sw["n"], "nnz": m.nnz, "rtol": sw["rtol"]}
    variants = [("fp64", npbcg.SOLVE_FP64, npbcg.PRECOND_NONE),
                ("mixed", npbcg.SOLVE_MIXED, npbcg.PRECOND_NONE),
                ("fp64_neumann", npbcg.SOLVE_FP64, npbcg.PRECOND_NEUMANN),
                ("mixed_neumann", npbcg.SOLVE_MIXED, npbcg.PRECOND_NEUMANN)]
    net = None
    if os.path.exists(LEARNED_PARAMS):
        variants += [("fp64_learned", npbcg.SOLVE_FP64, npbcg.PRECOND_LEARNED),
                     ("mixed_learned", npbcg.SOLVE_MIXED, npbcg.PRECOND_LEARNED)]
    with npbcg.Context(m, device=0, use_graph=use_graph) as ctx:
        out["spmv_kernels"] = spmv_kernel_rooflines(ctx, m)
        ctx.set_rhs(b)
        t0 = time.perf_counter()
        ctx.set_preconditioner(npbcg.PRECOND_NEUMANN)
        # builds the row graph, the operator and its values (plus one
        # application with its n-vector transfers, about 0.3 ms of the figure)
        ctx.apply_preconditioner(b)
        out["precond_setup_wall_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        if os.path.exists(LEARNED_PARAMS):
            net = npbcg.Net(seed=1)
            net.load(LEARNED_PARAMS)
            net.apply(ctx)                         # warm
            inf = [net.apply(ctx) for _ in range(5)]
            out["learned_inference_ms_median"] = round(float(np.median(inf)), 4)
        for name, mode, kind in variants:
            ctx.set_preconditioner(kind)
            ctx.solve(mode, sw["rtol"], 500)  # warm (also loads the operator values)
            ev, wall, st = [], [], None
            for _ in range(reps):
                t0 = time.perf_counter()
                st = ctx.solve(mode, sw["rtol"], 500)
                wall.append((time.perf_counter() - t0) * 1e3)
                ev.append(st.time_ms)
            d = st.as_dict()
            d["time_ms_median"] = round(float(np.median(ev)), 4)
            d["wall_ms_median"] = round(float(np.median(wall)), 4)
            out[name] = d
    if net is not None:
        net.close()
    base = out["fp64"]["time_ms_median"]
    out["speedup_vs_fp64"] = {k: round(base / out[k]["time_ms_median"], 3)
                              for k, _, _ in variants if k != "fp64"}
    if "mixed_learned" in out:
        # one solve paying the whole per-matrix preparation: graph and operator
        # set-up plus one network inference
        setup = out["precond_setup_wall_ms"] + out.get("learned_inference_ms_median", 0.0)
        t = out["mixed_learned"]["time_ms_median"]
        out["mixed_learned_incl_setup_ms"] = round(t + setup, 4)
        out["speedup_vs_fp64_incl_setup"] = {
            "mixed_learned": round(base / (t + setup), 4)}
        # solves on the same matrix after which the preparation has paid for
        # itself against plain FP64 (None: never, the solve itself is slower)
        gain = base - t
        out["mixed_learned_breakeven_solves"] = (round(setup / gain, 1) if gain > 0 else None)
    return out


BATCH_WORKLOAD = dict(workload="npb_cg_batch_64x4", systems=64, n_min=2000, n_max=40000,
                      nonzer_min=7, nonzer_max=13, shift=20.0, rcond=0.1, nrhs=4, rtol=1e-10,
                      maxit=3000)


def batch_systems(bw=BATCH_WORKLOAD):
    """(n, nonzer, shift, seed, rcond) of the batch: system s = 1..64 uses
    matrix seed s, and n ~ U{n_min..n_max}, nonzer ~ U{nonzer_min..nonzer_max}
    drawn (in that order) from numpy's PCG64 seeded s."""
    out = []
    for s in range(1, bw["systems"] + 1):
        rng = np.random.default_rng(s)
        n = int(rng.integers(bw["n_min"], bw["n_max"] + 1))
        nz = int(rng.integers(bw["nonzer_min"], bw["nonzer_max"] + 1))
        out.append((n, nz, bw["shift"], float(s), bw["rcond"]))
    return out


def batch_rhs(mats, nrhs):
    """Right-hand sides uniform(-1, 1): numpy PCG64 seeded 1000 + s per system."""
    return [np.random.default_rng(1000 + s).uniform(-1.0, 1.0, (m.n, nrhs))
            for s, m in enumerate(mats, start=1)]


def batch_comparison(reps: int):
    """The 64-system batch with 4 right-hand sides each: every solve to a
    1e-10 FP64-checked relative residual (or maxit), mixed precision and FP64,
    with the SpMM launch shape chosen by the autotuner."""
    import npbcg

    bw = BATCH_WORKLOAD
    t0 = time.perf_counter()
    mats = npbcg.make_many(batch_systems(bw))
    t_gen = time.perf_counter() - t0
    rhs = batch_rhs(mats, bw["nrhs"])
    nnz = int(sum(m.nnz for m in mats))
    rows = int(sum(m.n for m in mats))
    out = {"workload": bw["workload"], "systems": bw["systems"], "nrhs": bw["nrhs"],
           "rows": rows, "nnz": nnz, "generate_seconds": round(t_gen, 2),
           "nonzer_counts": {str(k): sum(1 for m in mats if m.nonzer == k)
                             for k in range(bw["nonzer_min"], bw["nonzer_max"] + 1)}}
    cache = os.path.join(ROOT, "npb-gpu_amd", "npbcg", "tune_records.txt")
    for name, mode in (("mixed", npbcg.SOLVE_MIXED), ("fp64", npbcg.SOLVE_FP64)):
        with npbcg.Batch(mats, nrhs=bw["nrhs"], mode=mode) as bt:
            lpr, rt, ms = bt.tune(cache)
            bt.set_rhs(rhs)
            bt.solve(bw["rtol"], bw["maxit"], check=False)   # warm
            times, st, steps = [], None, None
            for _ in range(reps):
                st, steps, rel = bt.solve(bw["rtol"], bw["maxit"], check=False)
                times.append(st.time_ms)
            val_b = 4 if mode == npbcg.SOLVE_MIXED else 8
            spmm_bytes = nnz * (val_b + 4) + 4 * (rows + 1) + 2 * rows * bw["nrhs"] * val_b
            d = st.as_dict()
            d.update(time_ms_median=round(float(np.median(times)), 3),
                     steps_median=float(np.median(steps)),
                     unconverged_systems=sorted({i + 1 for i in np.nonzero(rel > bw["rtol"])[0]}),
                     tuned_spmm_ms=round(ms, 4),
                     tuned_spmm_gbs=round(spmm_bytes / (ms * 1e-3) / 1e9, 1),
                     tuned_spmm_frac=round(spmm_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
            out[name] = d
    return out


def spmv_variant_sweep(reps: int = 20):
    """Complete q = A p products (npbcg_spmv_bench: SpMV plus, for the
    LDS-staged layout, the combine of its block row sums) for every SpMV
    variant on the n=75000 power matrix and the n=150000 solve matrix, with
    the FP64 and FP32 CG SpMV kernels alone for the LDS-staged layout."""
    import npbcg

    out = {}
    for wl in (WORKLOAD, SOLVE_WORKLOAD):
        m = npbcg.makea(wl["n"], wl["nonzer"], wl["shift"], wl["seed"], wl["rcond"])
        row = {"n": m.n, "nnz": m.nnz, "algorithmic_bytes": spmv_bytes(m.n, m.nnz)}
        for v, name in ((1, "csr_64_lanes"), (2, "csr_32_lanes"), (3, "csr_16_lanes"),
                        (4, "csr_64_lanes_2_in_flight"), (5, "sell_lds_staged")):
            with npbcg.Context(m, device=0, spmv_variant=v) as ctx:
                ms = ctx.spmv_bench(reps)
                row[name] = {"ms": round(ms, 5),
                             "gbs_algorithmic": round(row["algorithmic_bytes"] / ms / 1e6, 1)}
                if v == 5:
                    row["sell_kernels"] = spmv_kernel_rooflines(ctx, m, reps)
            print(f"spmv sweep {wl['workload']}: {name} {row[name]}", file=sys.stderr, flush=True)
        out[wl["workload"]] = row
    return out


def batch_layout_sweep(reps: int = 10):
    """SpMM time of the 64-system, 4-right-hand-side batch for each layout and
    launch shape the autotuner considers (row gathers: lanes per row x rows per
    tile; SELL: one 2048-row window per item), FP32 (mixed) and FP64."""
    import npbcg

    bw = BATCH_WORKLOAD
    mats = npbcg.make_many(batch_systems(bw))
    out = {}
    for name, mode in (("mixed", npbcg.SOLVE_MIXED), ("fp64", npbcg.SOLVE_FP64)):
        rows = []
        shapes = [(1, lpr, rpt) for lpr in (4, 8, 16, 32, 64) for rpt in (256, 1024, 4096)]
        shapes += [(3, 0, 2048)]
        for layout, lpr, rpt in shapes:
            with npbcg.Batch(mats, nrhs=bw["nrhs"], mode=mode, layout=layout,
                             lanes_per_row=lpr, rows_per_tile=rpt) as bt:
                rows.append({"layout": layout, "lanes_per_row": lpr, "rows_per_tile": rpt,
                             "spmm_ms": round(bt.spmm_bench(reps), 4)})
            print(f"batch sweep {name}: {rows[-1]}", file=sys.stderr, flush=True)
        out[name] = rows
    return out


def rhs_for(n: int, seed: float) -> np.ndarray:
    """b_i = 2 u_i - 1, u_i successive draws of the 46-bit LCG (a = 5^13)
    from seed + 1 (the definition the oracle's orc_make_rhs restates)."""
    x = int(seed) + 1
    a, mask = 1220703125, (1 << 46) - 1
    out = np.empty(n)
    for i in range(n):
        x = (a * x) & mask
        out[i] = 2.0 * (x / float(1 << 46)) - 1.0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-outer-iters", type=int, default=75,
                    help="outer iterations in the CPU baseline sample (0 = skip)")
    ap.add_argument("--spmv-variant", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--fuse-p-update", action="store_true")
    ap.add_argument("--cg-form", type=int, default=0,
                    help="0 auto (one-reduction form), 1 the reference's step order, 2 one-reduction")
    ap.add_argument("--solve-reps", type=int, default=10,
                    help="solves per mode in the n=150000 time-to-1e-10 comparison (0 = skip)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="short run for rocprofv3 counter passes")
    ap.add_argument("--profile-pass-solve", action="store_true",
                    help="short n=150000 SpMV run (FP64 and FP32 kernels) for counter passes")
    ap.add_argument("--profile-pass-batch", action="store_true",
                    help="short batched-SpMM run for rocprofv3 counter passes")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="summarise two counter passes into the traffic file")
    ap.add_argument("--pmc-summary-solve", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="summarise two n=150000 counter passes (--profile-pass-solve)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the n=150000 time-to-1e-10 comparison (JSON on stdout)")
    ap.add_argument("--batch-reps", type=int, default=3,
                    help="repeats of the 64-system batched solve (0 = skip)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 64-system batched solve (JSON on stdout)")
    ap.add_argument("--spmv-sweep", action="store_true",
                    help="every SpMV variant at n=75000 and n=150000 (JSON on stdout)")
    ap.add_argument("--batch-sweep", action="store_true",
                    help="batched SpMM time per layout and launch shape (JSON on stdout)")
    args = ap.parse_args()
    if args.spmv_sweep:
        print(json.dumps(spmv_variant_sweep()))
        return
    if args.batch_sweep:
        print(json.dumps(batch_layout_sweep()))
        return
    if args.solve_only:
        print(json.dumps(solve_comparison(max(1, args.solve_reps), 0 if args.no_graph else 1)))
        return
    if args.batch_only:
        print(json.dumps(batch_comparison(max(1, args.batch_reps))))
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.pmc_summary_solve:
        pmc_summary_solve(*args.pmc_summary_solve)
        return
    if args.profile_pass:
        profile_pass()
        return
    if args.profile_pass_solve:
        profile_pass_solve()
        return
    if args.profile_pass_batch:
        profile_pass_batch()
        return

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the solver is not sharded across GPUs",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import npbcg

    wl = WORKLOAD
    t0 = time.perf_counter()
    m = npbcg.makea(wl["n"], wl["nonzer"], wl["shift"], wl["seed"], wl["rcond"])
    t_gen = time.perf_counter() - t0
    ctx = npbcg.Context(m, device=0, cgitmax=wl["cgitmax"], spmv_variant=args.spmv_variant,
                        use_graph=0 if args.no_graph else 1,
                        fuse_p_update=1 if args.fuse_p_update else 0, cg_form=args.cg_form)
    n, nnz = m.n, m.nnz

    for _ in range(args.warmup):
        ctx.fill_x(1.0)
        ctx.power(wl["niter"], wl["shift"])

    # SpMV launches of the first outer iteration of every step are bracketed
    # with HIP events on the library's stream.
    ctx.set_spmv_sampling(1)
    spmv_ms = []
    gpu_ms = []
    hip_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.fill_x(1.0)
        zeta, rnorm = ctx.power(wl["niter"], wl["shift"])
        t = ctx.timing()
        spmv_ms.append(t.spmv_avg_ms)
        gpu_ms.append(t.total_ms)
    hip_device_synchronize()
    wall = time.perf_counter() - t0
    ctx.set_spmv_sampling(0)

    ms_per_step = wall * 1e3 / args.steps
    step_bytes = wl["niter"] * outer_iter_bytes(n, nnz, wl["cgitmax"])
    value = step_bytes / (ms_per_step * 1e-3) / 1e9
    b_spmv = spmv_bytes(n, nnz)
    spmv_avg_ms = float(np.mean(spmv_ms))
    achieved = b_spmv / (spmv_avg_ms * 1e-3) / 1e9

    ref = golden_zeta()
    zeta_err = float(np.max(np.abs(zeta - ref) / np.abs(ref))) if ref is not None else None

    traffic = load_traffic()
    roofline = {
        "bound": "hbm",
        "residency": ("memory-side cache: the 147 MB SELL copy stays in the 256 MB Infinity "
                      "Cache between SpMVs, so achieved can exceed what HBM alone streams; "
                      "roofline_hbm_n150000 has the SpMVs of a matrix that does not fit"),
        "kernel": ("spmv_sell_kernel, the CG SpMV: s = A r with r.s partials (one-reduction "
                   "form; q = A p with p.q in the reference's step order, --cg-form 1)"),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "algorithmic_bytes_per_launch": b_spmv,
        "avg_launch_ms": round(spmv_avg_ms, 5),
        "samples": int(sum(1 for _ in spmv_ms)) * 25,
    }

    # Streaming references: a read-only stream and a copy at the matrix's
    # size (resident in the 256 MB memory-side cache, like the matrix) and at
    # 1 GiB (HBM).
    small = b_spmv // 16 * 16
    half = b_spmv // 2 // 16 * 16
    stream = {
        "read_gbs_at_matrix_size": round(small / ctx.stream_bench(1, small, 10) / 1e6, 1),
        "copy_gbs_at_matrix_size": round(2 * half / ctx.stream_bench(0, half, 10) / 1e6, 1),
        "read_gbs_1GiB": round((1 << 30) / ctx.stream_bench(1, 1 << 30, 5) / 1e6, 1),
        "copy_gbs_1GiB": round(2 * (1 << 30) / ctx.stream_bench(0, 1 << 30, 5) / 1e6, 1),
    }
    roofline["frac_of_read_stream_at_matrix_size"] = round(
        achieved / stream["read_gbs_at_matrix_size"], 4)

    solve = (solve_comparison(args.solve_reps, 0 if args.no_graph else 1)
             if args.solve_reps > 0 else None)
    batch = batch_comparison(args.batch_reps) if args.batch_reps > 0 else None

    cpu = None
    if rank == 0 and args.cpu_outer_iters > 0:
        cpu = cpu_baseline(m, wl, args.cpu_outer_iters)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: NPB makea generator (seeded LCG), generated on the host",
        "config": {k: wl[k] for k in ("workload", "n", "nonzer", "shift", "rcond", "seed",
                                      "niter", "cgitmax")} | {"nnz": nnz,
                                                              "value_basis": "algorithmic bytes of 75 outer iterations / wall time"},
        "cg_iters_per_s": round(wl["niter"] * wl["cgitmax"] / (ms_per_step * 1e-3), 1),
        "gpu_event_ms_per_step": round(float(np.median(gpu_ms)), 4),
        "zeta_final": float(zeta[-1]),
        "zeta_max_rel_err_vs_reference": zeta_err,
        "generate_seconds": round(t_gen, 2),
        "kernel_config": ctx.describe(),
        "roofline": roofline,
        "roofline_hbm_n150000": solve["spmv_kernels"] if solve else None,
        "time_to_1e-10_ms": ({k: solve[k]["time_ms_median"] for k in
                              ("fp64", "mixed", "mixed_neumann", "mixed_learned") if k in solve}
                             | ({"mixed_learned_incl_setup": solve["mixed_learned_incl_setup_ms"]}
                                if "mixed_learned_incl_setup_ms" in solve else {})
                             if solve else None),
        "speedup_mixed_learned_vs_fp64": (solve["speedup_vs_fp64"].get("mixed_learned")
                                          if solve else None),
        "speedup_mixed_learned_vs_fp64_incl_setup": (
            solve.get("speedup_vs_fp64_incl_setup", {}).get("mixed_learned") if solve else None),
        "stream_reference": stream,
        "solve_to_1e-10": solve,
        "batched_solves": batch,
        "cpu_baseline": cpu,
    }
    ctx.close()
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
