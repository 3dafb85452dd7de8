"""Train the learned preconditioner on generated NPB systems and save it.

    python npb-gpu_amd/npbcg/train_precond.py [--systems 2000] [--out PATH]

Defaults: 2000 systems, n uniform in [10000, 75000], 7 to 15 nonzeros per
generator vector, shift (nonzer+1)^2 * U[0.3, 0.5] (every system definite;
nonzer 15 gives shifts in [77, 128]), rcond 0.1, 8 CG residual/correction
pairs per system.  Training runs on the GPU through
npbcg_net_train (the library's own forward/backward kernels and Adam); the
host only runs the matrix generator threads.  The parameters and feature
normalisation go to an .npz file (default: learned_precond.npz next to
this script), which npbcg.Net.load reads.  A JSON summary line is printed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import npbcg  # noqa: E402

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_precond.npz")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    o = npbcg.default_train_opts()
    ap.add_argument("--systems", type=int, default=o.systems)
    ap.add_argument("--n-min", type=int, default=o.n_min)
    ap.add_argument("--n-max", type=int, default=o.n_max)
    ap.add_argument("--nonzer-min", type=int, default=o.nonzer_min)
    ap.add_argument("--nonzer-max", type=int, default=o.nonzer_max)
    ap.add_argument("--shift-min", type=float, default=o.shift_min)
    ap.add_argument("--shift-max", type=float, default=o.shift_max)
    ap.add_argument("--steps-per-system", type=int, default=o.steps_per_system)
    ap.add_argument("--lr", type=float, default=o.lr)
    ap.add_argument("--seed", type=int, default=o.seed)
    ap.add_argument("--net-seed", type=int, default=1)
    ap.add_argument("--gen-threads", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--history", default=None, help="optional .npz with per-system losses")
    ap.add_argument("--chunk", type=int, default=100,
                    help="systems per npbcg_net_train call (a progress line after each)")
    a = ap.parse_args(argv)
    o.systems, o.n_min, o.n_max = a.systems, a.n_min, a.n_max
    o.nonzer_min, o.nonzer_max = a.nonzer_min, a.nonzer_max
    o.shift_min, o.shift_max = a.shift_min, a.shift_max
    o.steps_per_system, o.lr, o.seed = a.steps_per_system, a.lr, a.seed
    o.gen_threads = a.gen_threads
    lh, bh = [], []
    t0 = time.perf_counter()
    wait_ms = 0.0
    skipped = 0
    with npbcg.Net(seed=a.net_seed) as net:
        for first in range(0, a.systems, a.chunk):
            o.first_system = first
            o.systems = min(a.chunk, a.systems - first)
            st, l, b = net.train(o)
            lh.append(l)
            bh.append(b)
            wait_ms += st.wait_ms
            skipped += st.skipped
            print(f"systems {first + o.systems}/{a.systems}: mean loss/base "
                  f"{np.nanmean(l / b):.4f} skipped {st.skipped} "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
        net.save(a.out)
    lh, bh = np.concatenate(lh), np.concatenate(bh)
    if a.history:
        np.savez(a.history, loss=lh, base=bh)
    w = max(1, a.systems // 10)
    summary = dict(systems=int(lh.size), skipped=skipped, out=a.out,
                   ratio_first=float(np.nanmean(lh[:w] / bh[:w])),
                   ratio_last=float(np.nanmean(lh[-w:] / bh[-w:])),
                   generator_wait_s=round(wait_ms / 1e3, 2),
                   total_s=round(time.perf_counter() - t0, 2))
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
