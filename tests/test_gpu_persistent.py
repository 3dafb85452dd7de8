"""The persistent conj_grad engine (one cooperative kernel per npbcg_power /
npbcg_conj_grad_device call, grid barriers between its phases) against the
reference's own trajectories and the oracle, with the tolerances of the
kernel engine (tests/test_gpu_fp64.py): zeta relative 1e-12 per outer
iteration, rnorm within 1e-12 * sqrt(n), vectors within 1e-12 * scale.

The CG step is the one-reduction form the kernel engine uses by default; the
reference loop it replaces is cg.cu:803-846 called from :500-520."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ENGINE = dict(engine=2, spmv_variant=5)


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond, m.seed)


def _gen(lib, c):
    return lib.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])


def _zeta_ok(z, ref, label):
    rel = np.abs(np.asarray(z) - ref) / np.abs(ref)
    print(f"{label}: max rel zeta err {rel.max():.3e}")
    assert np.all(rel <= 1e-12), label


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "W_npb", "A_npb",
                                  "B_npb", "cfg2_n75000_seed271828", "C_npb",
                                  "cfg3_n150000_seed314"])
def test_persistent_power_matches_reference(gpu, golden, name):
    c = golden[name]
    m = _gen(gpu, c)
    with gpu.Context(m, **ENGINE) as ctx:
        cfg = ctx.describe()
        assert cfg["engine"] == 2 and cfg["spmv_variant"] == 5 and cfg["cg_form"] == 2
        zeta, rnorm = ctx.power(c["niter"], c["shift"])
        t = ctx.timing()
        xg = ctx.get_vector(gpu.VEC_X)
    assert t.engine == 2
    _zeta_ok(zeta, np.array([float.fromhex(h) for h in c["zeta_hex"]]), f"{name} persistent")
    rref = np.array([float.fromhex(h) for h in c["rnorm_hex"]])
    assert np.all(np.abs(rnorm - rref) <= 1e-12 * np.sqrt(m.n))
    if m.n <= 14000:   # final x = z/||z|| against the oracle's loop
        _, _, ox = oracle.power(_ocsr(m), c["niter"])
        err = float(np.max(np.abs(xg[: m.n] - ox[: m.n])))
        assert err <= 1e-12, err
        assert xg[m.n] == 1.0
    if name.endswith("_npb"):
        zn = oracle.NPB_CLASSES[c["npb_class"]][4]
        assert abs(zeta[-1] - zn) / zn <= oracle.NPB_EPSILON


@pytest.mark.parametrize("name", ["S_npb", "A_npb"])
def test_persistent_dropin_and_device_sequence(gpu, golden, name):
    """npbcg_conj_grad on host arrays (the reference signature) and the
    cg.cu:500-519 sequence conj_grad_gpu + gpu_kernel_ten + gpu_kernel_eleven."""
    c = golden[name]
    m = _gen(gpu, c)
    n = m.n
    rng = np.random.default_rng(11)
    x = np.ones(n + 2)
    x[:n] += 0.25 * rng.standard_normal(n)
    x[n] = 3.0
    om = _ocsr(m)
    with gpu.Context(m, **ENGINE) as ctx:
        z, p, q, r, rn = ctx.conj_grad(x)
        oz, op, oq, orr, orn = oracle.conj_grad(om, x)
        scale = max(1.0, float(np.max(np.abs(x[:n]))))
        for v, ref, lab in [(z, oz, "z"), (p, op, "p"), (q, oq, "q"), (r, orr, "r")]:
            err = float(np.max(np.abs(v[:n] - ref[:n])))
            print(f"{name} {lab}: {err:.3e}")
            assert err <= 1e-12 * max(scale, float(np.max(np.abs(ref[:n])))), lab
        assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n])
        assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 3.0 and p[n] == 3.0
        # the reference's timed loop, one call per step
        ctx.fill_x(1.0)
        zeta = []
        for _ in range(c["niter"]):
            ctx.conj_grad_device()
            t1, t2 = ctx.norms()
            zeta.append(c["shift"] + 1.0 / t1)
            ctx.scale_x(1.0 / np.sqrt(t2))
    _zeta_ok(zeta, np.array([float.fromhex(h) for h in c["zeta_hex"]]), f"{name} sequence")


def test_persistent_matches_kernel_engine_and_repeats(gpu, golden):
    c = golden["cfg2_n75000_seed271828"]
    m = _gen(gpu, c)
    with gpu.Context(m) as k_ctx:
        zk, rk = k_ctx.power(10, c["shift"])
        tk = k_ctx.timing().total_ms
    with gpu.Context(m, **ENGINE) as ctx:
        runs = []
        for _ in range(3):
            ctx.fill_x(1.0)
            runs.append(ctx.power(10, c["shift"]))
        tp = ctx.timing().total_ms
    for zp, rp in runs[1:]:     # bit-reproducible run to run
        assert np.array_equal(zp, runs[0][0]) and np.array_equal(rp, runs[0][1])
    rel = np.max(np.abs(runs[0][0] - zk) / np.abs(zk))
    print(f"10 outer iterations: kernels {tk:.3f} ms, persistent {tp:.3f} ms; "
          f"zeta rel diff {rel:.2e}")
    assert rel <= 1e-12


def test_persistent_needs_the_sell_layout(gpu):
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    with pytest.raises(gpu.NpbcgError) as e:
        gpu.Context(m, engine=2, spmv_variant=1)
    assert e.value.code == gpu.ERR_INVALID
    with pytest.raises(gpu.NpbcgError) as e:
        gpu.Context(m, engine=7)
    assert e.value.code == gpu.ERR_INVALID
    with pytest.raises(gpu.NpbcgError) as e:
        gpu.Context(m, dot_pass=3)
    assert e.value.code == gpu.ERR_INVALID
