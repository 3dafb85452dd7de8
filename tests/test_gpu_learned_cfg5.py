"""The learned preconditioner as shipped (npb-gpu_amd/npbcg/learned_precond.npz,
trained on the GPU on 2000 generated systems) on its evaluation system and on
held-out training-range systems, and the Jacobi-CG fallback.

Evaluation system: n = 150000, nonzer 15, shift 110, seed 314, b = 2u - 1 from
the LCG stream seeded 315 (oracle.make_rhs).  Every solve is judged by the
FP64 true relative residual recomputed on the host (<= 1e-10) and by its
distance to the oracle's FP64 CG solution (<= 1e-8 max-relative).  The
preconditioned CG recurrence is the reference's unpreconditioned one
(cg.cu:689-772) with z = M^-1 r; the reference has no preconditioner, so the
preconditioned solves are pinned by the solution and the residual, not by a
trajectory.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10
PARAMS_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "npb-gpu_amd", "npbcg", "learned_precond.npz")


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond, m.seed)


def _relres(om, x, b):
    return float(np.linalg.norm(b - oracle.spmv(om, x)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def cfg5(gpu):
    m = gpu.makea(150000, 15, 110.0, 314.0)
    om = _ocsr(m)
    b = oracle.make_rhs(m.n, 314.0)
    ox, oit, orel = oracle.cg_solve(om, b, rtol=1e-12, maxit=200)
    print(f"oracle CG: {oit} steps, relres {orel:.2e}", flush=True)
    return m, om, b, ox


def test_shipped_network_loads(gpu):
    assert os.path.exists(PARAMS_FILE), PARAMS_FILE
    with np.load(PARAMS_FILE, allow_pickle=False) as f:
        params, norm = f["params"], f["norm"]
    with gpu.Net(seed=1) as net:
        assert params.shape == (net.param_count,) and norm.shape == (2 * gpu.NET_FEATURES,)
        assert np.all(np.isfinite(params)) and np.all(np.isfinite(norm))
        net.load(PARAMS_FILE)
        p2, n2 = net.get_params()
    assert np.array_equal(p2, params) and np.array_equal(n2, norm)


@pytest.mark.parametrize("mode", [0, 1])
def test_cfg5_eval_system(gpu, cfg5, mode):
    m, om, b, ox = cfg5
    kinds = [("none", gpu.PRECOND_NONE), ("jacobi", gpu.PRECOND_JACOBI),
             ("neumann", gpu.PRECOND_NEUMANN), ("learned", gpu.PRECOND_LEARNED)]
    res = {}
    with gpu.Context(m) as ctx, gpu.Net(seed=1) as net:
        ctx.set_rhs(b)
        net.load(PARAMS_FILE)
        inf_ms = net.apply(ctx)
        for name, kind in kinds:
            ctx.set_preconditioner(kind)
            ctx.solve(mode, RTOL, 300)             # first solve builds and captures
            st = ctx.solve(mode, RTOL, 300)
            res[name] = (st, ctx.solution())
    for name, (st, x) in res.items():
        rel = _relres(om, x, b)
        err = float(np.max(np.abs(x - ox)) / np.max(np.abs(ox)))
        print(f"mode {mode} {name}: steps {st.iterations} sweeps {st.sweeps} "
              f"time {st.time_ms:.3f} ms host relres {rel:.2e} x err {err:.2e} "
              f"fallback {st.fallback}")
        assert st.converged == 1 and rel <= RTOL and st.fallback == 0
        assert err <= 1e-8
    print(f"inference {inf_ms:.3f} ms")
    it = {k: v[0].iterations for k, v in res.items()}
    # Measured gains at this conditioning (kappa about 1.36): the trained
    # network never needs more steps than plain CG, and the sparse operator
    # (Neumann or learned) never more than Jacobi-CG.
    assert it["learned"] <= it["none"]
    assert it["learned"] <= it["jacobi"] and it["neumann"] <= it["jacobi"]


def _held_out_systems():
    """Training-range systems the network has not seen: n in [10000, 75000],
    nonzer 7..15, shift (nonzer+1)^2 U[0.3, 0.5], drawn from numpy PCG64
    seeded 4242 (the training run draws its systems from splitmix64 of
    seed 20260101)."""
    rng = np.random.default_rng(4242)
    out = []
    for _ in range(4):
        n = int(rng.integers(10000, 75001))
        nz = int(rng.integers(7, 16))
        shift = (nz + 1) ** 2 * float(rng.uniform(0.3, 0.5))
        seed = float(int(rng.integers(1, 1 << 40)))
        out.append((n, nz, shift, seed))
    return out


def test_trained_network_beats_its_base_on_held_out_pairs(gpu):
    """Pair loss sum_k |M^-1 r_k - e_k|^2 / (P |e_k|^2) of the trained (g, h)
    against the network's base g = h = 0 (the Neumann form), on residual /
    correction pairs the solver logs for systems outside the training run."""
    ratios = []
    with gpu.Net(seed=1) as trained, gpu.Net(seed=1) as base:
        trained.load(PARAMS_FILE)
        for n, nz, shift, seed in _held_out_systems():
            m = gpu.makea(n, nz, shift, seed)
            with gpu.Context(m) as ctx:
                ctx.set_rhs(oracle.make_rhs(n, seed))
                ctx.log_pairs(8)
                lt, _, gh = trained.loss_grad(ctx, n)
                lb, _, gh0 = base.loss_grad(ctx, n)
            assert np.all(gh0 == 0.0) and np.all(np.isfinite(gh))
            ratios.append(lt / lb)
            print(f"n={n} nonzer={nz} shift={shift:.1f}: loss {lt:.4e} base {lb:.4e} "
                  f"ratio {lt / lb:.3f}")
    print(f"mean ratio {np.mean(ratios):.3f}")
    assert np.mean(ratios) < 0.95 and max(ratios) < 1.0


def _nan_net(gpu):
    net = gpu.Net(seed=1)
    p, nrm = net.get_params()
    p[:] = np.nan
    net.set_params(p, nrm)
    return net


@pytest.mark.parametrize("mode", [0, 1])
def test_fallback_to_jacobi_on_nan_operator(gpu, mode):
    """A learned operator that produces NaN cannot converge: the solve is redone
    with Jacobi-CG from x = 0 and must land on the oracle's solution."""
    m = gpu.makea(20000, 11, 30.0, 271.0)
    om = _ocsr(m)
    b = oracle.make_rhs(m.n, 271.0)
    ox, _, _ = oracle.cg_solve(om, b, rtol=1e-12, maxit=5000)
    with gpu.Context(m) as ctx:
        ctx.set_rhs(b)
        net = _nan_net(gpu)
        try:
            net.apply(ctx)
        finally:
            net.close()
        ctx.set_preconditioner(gpu.PRECOND_LEARNED)
        sts = []
        for _ in range(2):          # the second call replays captured sequences
            st = ctx.solve(mode, RTOL, 400)
            x = ctx.solution()
            sts.append(st)
            rel = _relres(om, x, b)
            err = float(np.max(np.abs(x - ox)) / np.max(np.abs(ox)))
            print(f"mode {mode}: {st.as_dict()} host relres {rel:.2e} x err {err:.2e}")
            assert st.fallback == 1 and st.precond == gpu.PRECOND_JACOBI
            assert st.converged == 1 and rel <= RTOL and err <= 1e-8
        # the context's choice stays LEARNED; a plain solve afterwards is unaffected
        ctx.set_preconditioner(gpu.PRECOND_NONE)
        st = ctx.solve(mode, RTOL, 400)
        assert st.fallback == 0 and st.precond == gpu.PRECOND_NONE and st.converged == 1
        assert _relres(om, ctx.solution(), b) <= RTOL


def test_fallback_when_maxit_cannot_be_met(gpu):
    m = gpu.makea(20000, 11, 30.0, 271.0)
    b = oracle.make_rhs(m.n, 271.0)
    with gpu.Context(m) as ctx, gpu.Net(seed=1) as net:
        ctx.set_rhs(b)
        net.load(PARAMS_FILE)
        net.apply(ctx)
        ctx.set_preconditioner(gpu.PRECOND_LEARNED)
        with pytest.raises(gpu.NpbcgError) as e:
            ctx.solve(0, RTOL, 2)
        assert e.value.code == gpu.ERR_NOCONV
        st = ctx.solve(0, RTOL, 2, check=False)
        print(st.as_dict())
        assert st.fallback == 1 and st.precond == gpu.PRECOND_JACOBI and st.converged == 0
        # with a budget it can meet, the learned solve converges without falling back
        st = ctx.solve(0, RTOL, 400)
        assert st.fallback == 0 and st.precond == gpu.PRECOND_LEARNED and st.converged == 1
