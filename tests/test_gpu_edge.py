"""Edge orders on the GPU: systems smaller than a wavefront, a 64-row slice
or a workgroup, and sizes either side of those boundaries, through the
drop-in conj_grad, the power loop, solves and batched solves, each against
the oracle on the same generated matrix (tolerances of the other GPU tests).

The reference's fixed 25-step conj_grad (cg.cu:653-801) divides 0 by 0 once
an order-1 or order-2 system has converged exactly; the oracle returns NaN
there.  So the fixed-step paths start at n = 3, and the solves, which stop
at their tolerance, start at n = 1."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

FIXED_STEP_SIZES = [3, 5, 17, 63, 64, 65, 129, 257]
SOLVE_SIZES = [1, 2, 3, 17, 64, 65, 257]
RTOL = 1e-10


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond,
                      m.seed)


def _mat(gpu, n):
    return gpu.makea(n, min(n, 5), 10.0, 1000.0 + n)


@pytest.mark.parametrize("variant", [0, 1, 3, 4, 5])
def test_small_orders_conj_grad_and_power(gpu, variant):
    for n in FIXED_STEP_SIZES:
        m = _mat(gpu, n)
        om = _ocsr(m)
        rng = np.random.default_rng(n)
        x = np.ones(n + 2)
        x[:n] += 0.25 * rng.standard_normal(n)
        x[n] = 3.0
        with gpu.Context(m, spmv_variant=variant) as ctx:
            z, p, q, r, rn = ctx.conj_grad(x)
        oz, op, oq, orr, orn = oracle.conj_grad(om, x)
        xmax = float(np.max(np.abs(x[:n])))
        for v, ref, lab in ((z, oz, "z"), (p, op, "p"), (q, oq, "q"), (r, orr, "r")):
            scale = max(1.0, xmax, float(np.max(np.abs(ref[:n]))))
            err = float(np.max(np.abs(v[:n] - ref[:n])))
            assert err <= 1e-12 * scale, f"n={n} variant {variant} {lab}: {err:.3e}"
        assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n]), f"n={n} rnorm"
        assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 3.0 and p[n] == 3.0
        with gpu.Context(m, spmv_variant=variant) as ctx:
            zeta, rnorm = ctx.power(4, m.shift)
        ozeta, ornorm, _ = oracle.power(om, 4)
        rel = np.abs(np.asarray(zeta) - ozeta) / np.abs(ozeta)
        assert np.all(rel <= 1e-12), f"n={n} variant {variant} zeta {rel.max():.3e}"
        assert np.all(np.abs(np.asarray(rnorm) - ornorm) <= 1e-12 * np.sqrt(n))


@pytest.mark.parametrize("mode,variant", [(0, 5), (1, 5), (0, 1)])
def test_small_orders_solve(gpu, mode, variant):
    for n in SOLVE_SIZES:
        m = _mat(gpu, n)
        om = _ocsr(m)
        b = oracle.make_rhs(n, 7.0 + n)
        with gpu.Context(m, spmv_variant=variant) as ctx:
            ctx.set_rhs(b)
            st = ctx.solve(mode, RTOL, 500)
            x = ctx.solution()
        rel = float(np.linalg.norm(b - oracle.spmv(om, x)) / np.linalg.norm(b))
        ox, _, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=1000)
        assert st.converged == 1 and rel <= RTOL, f"n={n} mode {mode}: {st.as_dict()} {rel:.2e}"
        assert np.max(np.abs(x - ox)) <= 1e-8 * np.max(np.abs(ox)), f"n={n} mode {mode}"


@pytest.mark.parametrize("mode,layout", [(1, 3), (0, 3), (1, 1)])
def test_small_systems_in_a_batch(gpu, mode, layout):
    mats = [_mat(gpu, n) for n in SOLVE_SIZES]
    rhs = [np.stack([oracle.make_rhs(m.n, 100.0 * s + k) for k in range(4)], axis=1)
           for s, m in enumerate(mats)]
    with gpu.Batch(mats, nrhs=4, mode=mode, layout=layout) as bt:
        bt.set_rhs(rhs)
        st, steps, rel = bt.solve(RTOL, 500)
        xs = bt.solution()
    assert st.layout == layout and st.converged == 4 * len(mats) and st.max_relres <= RTOL
    for s, m in enumerate(mats):
        om = _ocsr(m)
        for k in range(4):
            b = rhs[s][:, k]
            ox, _, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=1000)
            assert np.linalg.norm(b - oracle.spmv(om, xs[s][:, k])) <= RTOL * np.linalg.norm(b)
            assert np.max(np.abs(xs[s][:, k] - ox)) <= 1e-8 * np.max(np.abs(ox)), f"n={m.n}"
