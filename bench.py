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
result written once) — see DESIGN.md.  The SpMV kernel's own duration is
sampled with HIP events on the library's stream during the timed steps and
gives the roofline entry.  cpu_baseline: the oracle's serial power loop
(single thread) on a bounded sample of the same workload.
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
    """Serial oracle power loop (single thread, pinned) on `outer_iters`
    outer iterations of the same matrix."""
    import oracle

    om = oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, wl["nonzer"], wl["shift"],
                    wl["rcond"], wl["seed"])
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        zeta, _, _ = oracle.power(om, outer_iters, wl["cgitmax"])
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    gbytes = outer_iters * outer_iter_bytes(m.n, m.nnz, wl["cgitmax"]) / 1e9
    return {
        "value": round(gbytes / dt, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{outer_iters} of {wl['niter']} outer iterations (x {wl['cgitmax']} CG "
                   f"steps) of the same n={m.n} matrix, oracle/npbcg_oracle.c serial power "
                   f"loop, one pinned thread, {dt:.2f} s"),
        "seconds": round(dt, 3),
        "outer_iters_per_s": round(outer_iters / dt, 4),
        "cg_iters_per_s": round(outer_iters * wl["cgitmax"] / dt, 2),
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
        "zeta_last": float(zeta[-1]),
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-outer-iters", type=int, default=20,
                    help="outer iterations in the CPU baseline sample (0 = skip)")
    ap.add_argument("--spmv-variant", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the path is not sharded "
                          "across GPUs (see DESIGN.md for how it would shard)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import npbcg

    wl = WORKLOAD
    t0 = time.perf_counter()
    m = npbcg.makea(wl["n"], wl["nonzer"], wl["shift"], wl["seed"], wl["rcond"])
    t_gen = time.perf_counter() - t0
    ctx = npbcg.Context(m, device=0, cgitmax=wl["cgitmax"], spmv_variant=args.spmv_variant,
                        use_graph=0 if args.no_graph else 1)
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
        "kernel": "spmv_pq_kernel (q = A p, p.q)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "algorithmic_bytes_per_launch": b_spmv,
        "avg_launch_ms": round(spmv_avg_ms, 5),
        "samples": int(sum(1 for _ in spmv_ms)) * 25,
    }

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
        "cpu_baseline": cpu,
    }
    ctx.close()
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
