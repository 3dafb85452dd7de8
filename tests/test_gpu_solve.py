"""Solve-to-tolerance on the GPU (FP64 CG and FP64-refined mixed precision).

Every solve is judged by its FP64 true relative residual ||b - A x|| / ||b||,
recomputed here on the host with the oracle's SpMV, and the solution is
compared with the oracle's own FP64 CG solve (oracle.cg_solve)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond,
                      m.seed)


def _true_relres(m, x, b):
    return float(np.linalg.norm(b - oracle.spmv(_ocsr(m), x)) / np.linalg.norm(b))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,nonzer,shift,seed", [(1400, 7, 10.0, 12345.0),
                                                  (20000, 11, 30.0, 271.0)])
def test_solve_small(gpu, mode, n, nonzer, shift, seed):
    m = gpu.makea(n, nonzer, shift, seed)
    b = oracle.make_rhs(n, seed)
    with gpu.Context(m, spmv_variant=5) as ctx:
        ctx.set_rhs(b)
        st = ctx.solve(mode, RTOL, 2000)
        x = ctx.solution()
    rel = _true_relres(m, x, b)
    ox, oit, orel = oracle.cg_solve(_ocsr(m), b, rtol=RTOL, maxit=5000)
    err = float(np.max(np.abs(x - ox)) / np.max(np.abs(ox)))
    print(f"mode {mode} n={n}: {st.as_dict()} host relres {rel:.2e}; oracle its {oit}; "
          f"x rel err {err:.2e}")
    assert st.converged == 1 and rel <= RTOL and st.relres <= RTOL
    assert abs(rel - st.relres) <= 1e-3 * RTOL + 0.05 * rel
    assert err <= 1e-8
    if mode == 1:
        assert st.sweeps >= 2 and st.fp32_spmvs == st.iterations


@pytest.mark.parametrize("mode", [0, 1])
def test_solve_cfg3(gpu, mode):
    # The solve workload: n=150000, nonzer=15, shift=110, seed 314, b = 2u-1.
    m = gpu.makea(150000, 15, 110.0, 314.0)
    b = oracle.make_rhs(m.n, 314.0)
    with gpu.Context(m) as ctx:
        ctx.set_rhs(b)
        st = ctx.solve(mode, RTOL, 200)
        x = ctx.solution()
    rel = _true_relres(m, x, b)
    print(f"cfg3 mode {mode}: {st.as_dict()} host relres {rel:.2e}")
    assert st.converged == 1 and rel <= RTOL


def test_solve_fp64_csr_variant_and_restart(gpu):
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    b = oracle.make_rhs(m.n, 12345.0)
    with gpu.Context(m, spmv_variant=1, solve_batch=3) as ctx:
        ctx.set_rhs(b)
        st = ctx.solve(0, RTOL, 2000)
        x = ctx.solution()
    assert st.converged == 1 and _true_relres(m, x, b) <= RTOL
    assert st.sweeps >= 2  # batches of 3 steps: several host checks


def test_solve_errors_and_zero_rhs(gpu):
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    with gpu.Context(m, spmv_variant=1) as ctx:
        with pytest.raises(gpu.NpbcgError) as e:
            ctx.solve(0, RTOL, 10)
        assert e.value.code == gpu.ERR_INVALID  # no right-hand side yet
        ctx.set_rhs(oracle.make_rhs(m.n, 1.0))
        with pytest.raises(gpu.NpbcgError) as e:
            ctx.solve(1, RTOL, 10)
        assert e.value.code == gpu.ERR_INVALID  # mixed needs the LDS-staged layout
        st = ctx.solve(0, RTOL, 2, check=False)
        assert st.converged == 0
        ctx.set_rhs(np.zeros(m.n))
        st = ctx.solve(0, RTOL, 10)
        assert st.converged == 1 and not np.any(ctx.solution())


def test_solve_leaves_power_state(gpu, golden):
    c = golden["S_npb"]
    m = gpu.makea(c["n"], c["nonzer"], c["shift"], c["seed"])
    with gpu.Context(m, spmv_variant=5) as ctx:
        z1, _ = ctx.power(3, c["shift"])
        ctx.set_rhs(oracle.make_rhs(m.n, 3.0))
        ctx.solve(1, RTOL, 500)
        z2, _ = ctx.power(3, c["shift"])   # continues from the x left by the first run
    ref = [float.fromhex(h) for h in c["zeta_hex"][:6]]
    assert np.all(np.abs(np.concatenate([z1, z2]) - ref) / np.abs(ref) <= 1e-12)


@pytest.mark.parametrize("mode", [0, 1])
def test_solve_jacobi_pcg(gpu, mode):
    m = gpu.makea(20000, 11, 30.0, 271.0)
    om = _ocsr(m)
    b = oracle.make_rhs(m.n, 271.0)
    dinv = 1.0 / oracle.diagonal(om)
    with gpu.Context(m, spmv_variant=5) as ctx:
        assert np.all(ctx.preconditioner() == 1.0)          # plain CG by default
        with pytest.raises(gpu.NpbcgError) as e:
            ctx.set_preconditioner(gpu.PRECOND_LEARNED)      # no learned diagonal yet
        assert e.value.code == gpu.ERR_INVALID
        ctx.set_preconditioner(gpu.PRECOND_JACOBI)
        minv = ctx.preconditioner()
        ctx.set_rhs(b)
        st = ctx.solve(mode, RTOL, 2000)
        x = ctx.solution()
        ctx.set_preconditioner(gpu.PRECOND_NONE)
        st0 = ctx.solve(mode, RTOL, 2000)
    # 1/a_jj is one correctly rounded division on both sides
    assert np.array_equal(minv, dinv)
    rel = _true_relres(m, x, b)
    ox, oit, orel = oracle.pcg_solve(om, b, dinv, rtol=RTOL, maxit=5000)
    err = float(np.max(np.abs(x - ox)) / np.max(np.abs(ox)))
    print(f"jacobi mode {mode}: {st.as_dict()} host relres {rel:.2e}; oracle PCG its {oit}; "
          f"plain {st0.iterations}; x rel err {err:.2e}")
    assert st.converged == 1 and rel <= RTOL and st.precond == gpu.PRECOND_JACOBI
    assert st.fallback == 0 and st0.precond == gpu.PRECOND_NONE
    assert err <= 1e-8
    if mode == 0:   # same recurrence as the oracle's PCG: step counts agree closely
        assert abs(st.iterations - oit) <= 2


@pytest.mark.parametrize("mode", [0, 1])
def test_conj_grad_after_solve_is_unchanged(gpu, golden, mode):
    # A solve leaves its stop flag behind on the device; the drop-in conj_grad
    # on the same context must not see it.
    c = golden["S_npb"]
    m = gpu.makea(c["n"], c["nonzer"], c["shift"], c["seed"])
    x = np.ones(m.n + 2)
    with gpu.Context(m, spmv_variant=5) as ctx:
        before = ctx.conj_grad(x)
        ctx.set_rhs(oracle.make_rhs(m.n, 3.0))
        ctx.solve(mode, RTOL, 500)
        after = ctx.conj_grad(x)
    for u, v in zip(before[:4], after[:4]):
        assert np.array_equal(u, v)
    assert before[4] == after[4]
