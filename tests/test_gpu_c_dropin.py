"""A compiled C caller of the drop-in (tests/c/dropin_sequence.c, linked to
libnpbcg.so) drives the reference benchmark's own call sequence: the untimed
conj_grad on host arrays through the exact-signature shim (cg.cu:458-482),
then NITER x {conj_grad_gpu, gpu_kernel_ten, gpu_kernel_eleven} (cg.cu:500-519)
through npbcg_conj_grad_device / npbcg_norms / npbcg_scale_x.

The per-iteration zeta must match the reference's own trajectory (the golden
fixture, relative 1e-12) and rnorm within 1e-12 * sqrt(n); the warm-up
iteration's rnorm and norms are checked against the oracle's conj_grad."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "dropin_sequence")


def _run(c):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (make -C tests/c)")
    args = [BIN, str(c["n"]), str(c["nonzer"]), str(c["niter"]), repr(c["shift"]),
            repr(c["seed"])]
    print("running", " ".join(args), flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0, out.stdout + out.stderr
    warm, its = None, []
    for line in out.stdout.splitlines():
        f = line.split()
        if f[0] == "warmup":
            warm = [float.fromhex(v) for v in f[1:4]]
        elif f[0] == "it":
            its.append((int(f[1]), float.fromhex(f[2]), float.fromhex(f[3])))
        elif f[0] == "done":
            assert f[1] == "0"
    return warm, its


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "A_npb"])
def test_c_caller_reference_sequence(gpu, golden, name):
    c = golden[name]
    warm, its = _run(c)
    assert [k for k, _, _ in its] == list(range(1, c["niter"] + 1))
    zref = np.array([float.fromhex(h) for h in c["zeta_hex"]])
    rref = np.array([float.fromhex(h) for h in c["rnorm_hex"]])
    zeta = np.array([z for _, _, z in its])
    rnorm = np.array([r for _, r, _ in its])
    rel = np.abs(zeta - zref) / np.abs(zref)
    print(f"{name}: max rel zeta err {rel.max():.3e}, "
          f"max abs rnorm err {np.max(np.abs(rnorm - rref)):.3e}")
    assert np.all(rel <= 1e-12)
    assert np.all(np.abs(rnorm - rref) <= 1e-12 * np.sqrt(c["n"]))
    if name.endswith("_npb"):
        zn = oracle.NPB_CLASSES[c["npb_class"]][4]
        assert abs(zeta[-1] - zn) / zn <= oracle.NPB_EPSILON
    # the untimed iteration (cg.cu:458-482) against the oracle's conj_grad
    om = oracle.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
    x = np.ones(c["n"] + 2)
    z, _, _, _, rn = oracle.conj_grad(om, x)
    t1 = float(np.dot(x[: c["n"]], z[: c["n"]]))
    t2 = 1.0 / np.sqrt(float(np.dot(z[: c["n"]], z[: c["n"]])))
    assert abs(warm[0] - rn) <= 1e-12 * np.sqrt(c["n"])
    assert abs(warm[1] - t1) <= 1e-12 * abs(t1)
    assert abs(warm[2] - t2) <= 1e-12 * abs(t2)
