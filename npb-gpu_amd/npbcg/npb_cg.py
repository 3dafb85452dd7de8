"""NPB CG benchmark driver on npbcg, with the reference's report.

    python npb-gpu_amd/npbcg/npb_cg.py --class B
    python npb-gpu_amd/npbcg/npb_cg.py --n 75000 --nonzer 13 --niter 75 --shift 60 --seed 271828

Runs what the reference's main does (CUDA/CG/cg.cu:396-640): generate the
matrix (makea, seed 314159265 for the NPB classes), one untimed conj_grad +
power step, x reset to ones, then the timed region of NITER outer iterations
(conj_grad, zeta = shift + 1/(x.z), x = z/||z||) — here the device-resident
loop npbcg_power — and prints the same lines: the per-iteration
"it ||r|| zeta" table (cg.cu:515-516), the verification block against the
class's published zeta with epsilon 1e-10 (cg.cu:534-552), and the
print_results summary (c_print_results.cpp:55-98) with Mop/s from the
reference's operation count (cg.cu:553-558).  --json adds one JSON line.
--profile repeats the NITER iterations from x = 1 kernel by kernel
(npbcg_profile_kernels) and prints the per-kernel table of the reference's
PROFILING build (cg.cu:565-589): workgroup size, seconds and share of the
profiled run's time.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import npbcg  # noqa: E402

# (NA, NONZER, NITER, SHIFT, zeta_verify_value) of the NPB classes
# (setparams class table, RCOND 0.1; verification values cg.cu:373-393).
CLASSES = {
    "S": (1400, 7, 15, 10.0, 8.5971775078648),
    "W": (7000, 8, 15, 12.0, 10.362595087124),
    "A": (14000, 11, 15, 20.0, 17.130235054029),
    "B": (75000, 13, 75, 60.0, 22.712745482631),
    "C": (150000, 15, 75, 110.0, 28.973605592845),
    "D": (1500000, 21, 100, 500.0, 52.514532105794),
    "E": (9000000, 26, 100, 1500.0, 77.522164599383),
}
NPB_SEED = 314159265.0
RCOND = 0.1
EPSILON = 1.0e-10


def mops(n: int, nonzer: int, niter: int, seconds: float) -> float:
    """The reference's Mop/s: 2 NITER NA (3 + NZ(NZ+1) + 25 (5 + NZ(NZ+1)) + 3) / t / 1e6."""
    if seconds <= 0.0:
        return 0.0
    w = nonzer * (nonzer + 1)
    return 2.0 * niter * n * (3.0 + w + 25.0 * (5.0 + w) + 3.0) / seconds / 1e6


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def kernel_table(kernels: list, total_ms: float) -> str:
    """The PROFILING table of cg.cu:565-589: kernel, threads per block,
    seconds and percentage of the total (here one row per kernel of the path,
    with its launch count)."""
    lines = [f"{'GPU Kernel':>5s}\t{'Threads Per Block':>25s}\t{'Time in Seconds':>25s}"
             f"\t{'Time in Percentage':>25s}\t{'Launches':>9s}"]
    for k in kernels:
        pct = k["ms"] * 100.0 / total_ms if total_ms > 0 else 0.0
        lines.append(f"{k['name']:>29s}\t{k['threads']:25d}\t{k['ms'] / 1e3:25f}"
                     f"\t{pct:24.2f}%\t{k['calls']:9d}")
    return "\n".join(lines)


def run(cls: str | None, n: int, nonzer: int, niter: int, shift: float, seed: float,
        device: int = 0, out=sys.stdout, profile: bool = False, use_graph: bool = True) -> dict:
    verify = CLASSES[cls][4] if cls else None
    p = lambda *a: print(*a, file=out, flush=True)  # noqa: E731
    p("\n\n NAS Parallel Benchmarks 4.1 npbcg (gfx950 HIP) version - CG Benchmark\n")
    p(f" Size: {n:11d}")
    p(f" Iterations: {niter:5d}")
    t0 = time.perf_counter()
    m = npbcg.makea(n, nonzer, shift, seed, RCOND)
    t_init = time.perf_counter() - t0
    p(f" Initialization time = {t_init:15.3f} seconds")
    with npbcg.Context(m, device=device, use_graph=int(use_graph)) as ctx:
        ctx.fill_x(1.0)
        ctx.power(1, shift)          # untimed warm-up iteration (cg.cu:458-482)
        ctx.fill_x(1.0)              # reset x = 1 (cg.cu:485-487)
        t0 = time.perf_counter()
        zeta, rnorm = ctx.power(niter, shift)
        t = time.perf_counter() - t0
        config = ctx.describe()
        prof = None
        if profile:
            ctx.fill_x(1.0)
            kernels, prof_ms, prof_zeta = ctx.profile_kernels(niter, shift)
            prof = {"kernels": kernels, "total_ms": prof_ms,
                    "same_zeta": bool((prof_zeta == zeta).all())}
    p("\n   iteration           ||r||                 zeta")
    for it in range(niter):
        p(f"    {it + 1:5d}       {rnorm[it]:20.14e}{zeta[it]:20.13e}")
    p(" Benchmark completed")
    z = float(zeta[-1])
    verified = False
    err = None
    if verify is not None:
        err = abs(z - verify) / verify
        if err <= EPSILON:
            verified = True
            p(" VERIFICATION SUCCESSFUL")
            p(f" Zeta is    {z:20.13e}")
            p(f" Error is   {err:20.13e}")
        else:
            p(" VERIFICATION FAILED")
            p(f" Zeta                {z:20.13e}")
            p(f" The correct zeta is {verify:20.13e}")
    else:
        p(" Problem size unknown")
        p(" NO VERIFICATION PERFORMED")
    mop = mops(n, nonzer, niter, t)
    p(f"\n\n CG Benchmark Completed")
    p(f" class_npb       =                        {cls or 'U'}")
    p(f" Size            =             {n:12d}")
    p(f" Iterations      =             {niter:12d}")
    p(f" Time in seconds =             {t:12.2f}")
    p(f" Mop/s total     =             {mop:12.2f}")
    p(f" Operation type  = {'floating point':>24s}")
    if verify is None:
        p(" Verification    =            NOT PERFORMED")
    elif verified:
        p(" Verification    =               SUCCESSFUL")
    else:
        p(" Verification    =             UNSUCCESSFUL")
    p(f" Version         =             {'npbcg ' + str(npbcg.LIB.npbcg_abi_version()):>12s}")
    p("\n Hardware:")
    p(f"    CPU device   = {cpu_model()}")
    p(f"    GPU device   = {config.get('device_name', 'device ' + str(device))}")
    p("\n Software:")
    p(f"    Parameters   = {json.dumps(config)}")
    if prof is not None:
        p(f"\n Profiled run: {niter} iterations from x = 1, kernel by kernel, "
          f"{prof['total_ms'] / 1e3:.6f} seconds of device time; zeta "
          f"{'identical to' if prof['same_zeta'] else 'DIFFERS from'} the timed run")
        p(kernel_table(prof["kernels"], prof["total_ms"]))
    p("")
    res = {"class": cls or "U", "n": n, "nonzer": nonzer, "niter": niter, "shift": shift,
           "seed": seed, "zeta": z, "zeta_error": err, "verified": verified,
           "time_s": t, "mops": mop, "makea_s": t_init}
    if prof is not None:
        res["profile"] = prof
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="NPB CG on npbcg with the reference's report")
    ap.add_argument("--class", dest="cls", choices=sorted(CLASSES), default=None)
    ap.add_argument("--n", type=int)
    ap.add_argument("--nonzer", type=int)
    ap.add_argument("--niter", type=int)
    ap.add_argument("--shift", type=float)
    ap.add_argument("--seed", type=float, default=NPB_SEED)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--json", action="store_true", help="also print one JSON summary line")
    ap.add_argument("--profile", action="store_true",
                    help="also print per-kernel device times of a profiled repeat")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the timed loop kernel by kernel instead of replaying a "
                         "captured graph (for tracing tools)")
    a = ap.parse_args(argv)
    if a.cls:
        n, nonzer, niter, shift, _ = CLASSES[a.cls]
        seed = NPB_SEED
        if any(v is not None for v in (a.n, a.nonzer, a.niter, a.shift)) or a.seed != NPB_SEED:
            ap.error("--class fixes n, nonzer, niter, shift and seed")
    else:
        if None in (a.n, a.nonzer, a.niter, a.shift):
            ap.error("give --class or all of --n --nonzer --niter --shift")
        n, nonzer, niter, shift, seed = a.n, a.nonzer, a.niter, a.shift, a.seed
    res = run(a.cls, n, nonzer, niter, shift, seed, a.device, profile=a.profile,
              use_graph=not a.no_graph)
    if a.json:
        print(json.dumps(res))
    return 0 if (res["verified"] or a.cls is None) else 1


if __name__ == "__main__":
    sys.exit(main())
