"""Learned sparse preconditioner on the GPU, against oracle.learned.

Graph and pair logging are integer/FP64 work checked exactly or to float32
rounding; the network's loss and gradient run BF16 MFMA GEMMs with FP32
accumulation and are compared with the oracle's BF16-emulating form within
the tolerances stated in each test."""
import os

import numpy as np
import pytest

import oracle
from oracle import learned as ol

pytestmark = pytest.mark.gpu

RTOL = 1e-10
PARAMS_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "npb-gpu_amd", "npbcg", "learned_precond.npz")


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond, m.seed)


def _setup(gpu, n=3000, nonzer=9, shift=25.0, seed=77.0, pairs=6):
    m = gpu.makea(n, nonzer, shift, seed)
    om = _ocsr(m)
    b = oracle.make_rhs(n, seed)
    return m, om, b


def test_graph_matches_oracle(gpu):
    m, om, _ = _setup(gpu)
    g = ol.graph(om)
    with gpu.Context(m, spmv_variant=1) as ctx:
        nbr, wgt, nval, feat = ctx.graph()
    assert np.array_equal(nbr, g.nbr)                        # selection: exact
    assert np.allclose(wgt, g.wgt, rtol=2e-7, atol=0)          # FP64 math, FP32 storage
    assert np.allclose(nval, g.nval, rtol=2e-7, atol=0)
    assert np.allclose(feat, g.feat, rtol=2e-6, atol=1e-7)


def test_log_pairs_match_oracle_cg(gpu):
    m, om, b = _setup(gpu)
    with gpu.Context(m, spmv_variant=5) as ctx:
        ctx.set_rhs(b)
        r, e = ctx.log_pairs(6)
        xs = ctx.solution()
    ox, _, _ = oracle.cg_solve(om, b, rtol=1e-13, maxit=4000)
    assert np.max(np.abs(xs - ox)) <= 1e-9 * np.max(np.abs(ox))
    R, E = ol.cg_pairs(om, b, ox, 6, lambda v: oracle.spmv(om, v))
    for k in range(6):   # float32 storage of FP64 CG vectors
        assert np.max(np.abs(r[k] - R[k])) <= 1e-6 * np.max(np.abs(R[k]))
        assert np.max(np.abs(e[k] - E[k])) <= 1e-5 * np.max(np.abs(E[k]))


def _random_params(seed=5):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(ol.N_PARAMS) * 0.15).astype(np.float32)
    p[ol.OFF_W0 + ol.F * ol.H:ol.OFF_B0] = 0.0
    return p


def test_loss_and_gradient_match_oracle(gpu):
    m, om, b = _setup(gpu)
    g = ol.graph(om)
    params = _random_params()
    norm = ol.feature_norm(g)
    with gpu.Context(m, spmv_variant=1) as ctx, gpu.Net(seed=3) as net:
        ctx.set_rhs(b)
        r, e = ctx.log_pairs(6)
        net.set_params(params, norm)
        loss, grad, gh = net.loss_grad(ctx, m.n)
        p2, n2 = net.get_params()
    assert np.array_equal(p2, params) and np.array_equal(n2, norm)
    R, E = r.astype(np.float64), e.astype(np.float64)
    want, wgrad, wgh = ol.loss_and_grad(params.astype(np.float64), norm, g, R, E, bf16_on=True)
    # BF16 activations: per-row outputs agree to a few BF16 ulps, sums closer
    assert np.max(np.abs(gh - wgh)) <= 2e-2, np.max(np.abs(gh - wgh))
    assert abs(loss - want) <= 2e-3 * want, (loss, want)
    cos = float(grad @ wgrad / (np.linalg.norm(grad) * np.linalg.norm(wgrad)))
    rel = float(np.linalg.norm(grad - wgrad) / np.linalg.norm(wgrad))
    print(f"loss {loss:.6e} oracle {want:.6e}; grad cos {cos:.6f} rel {rel:.3e}")
    assert cos >= 0.999 and rel <= 5e-2
    # the readout gradient (no MFMA below it) is tighter
    ro = slice(ol.OFF_R, ol.N_PARAMS)
    assert np.linalg.norm(grad[ro] - wgrad[ro]) <= 2e-2 * np.linalg.norm(wgrad[ro])


def test_preconditioner_operator_matches_dense(gpu):
    m, om, b = _setup(gpu, n=1500)
    g = ol.graph(om)
    g.nval = g.nval.astype(np.float32).astype(np.float64)   # the device keeps N in float32
    rng = np.random.default_rng(9)
    v = rng.standard_normal(m.n)
    with gpu.Context(m, spmv_variant=1) as ctx, gpu.Net(seed=11) as net:
        ctx.set_preconditioner(gpu.PRECOND_NEUMANN)
        z0 = ctx.apply_preconditioner(v)
        d0 = ctx.preconditioner()
        p = _random_params(6)
        net.set_params(p, ol.feature_norm(g))
        net.apply(ctx)
        ctx.set_preconditioner(gpu.PRECOND_LEARNED)
        z1 = ctx.apply_preconditioner(v)
        ctx.set_rhs(b)
        r, e = ctx.log_pairs(2)
        _, _, gh = net.loss_grad(ctx, m.n)
    assert np.array_equal(d0, 1.0 / g.adiag)
    M0 = ol.precond_dense(g, np.zeros(2 * m.n))
    assert np.allclose(z0, M0 @ v, rtol=1e-12, atol=1e-14 * np.max(np.abs(M0 @ v)))
    M1 = ol.precond_dense(g, gh.astype(np.float64))
    assert np.allclose(z1, M1 @ v, rtol=1e-12, atol=1e-14 * np.max(np.abs(M1 @ v)))


@pytest.mark.parametrize("mode", [0, 1])
def test_sparse_preconditioned_solves(gpu, mode):
    m = gpu.makea(20000, 11, 30.0, 271.0)
    om = _ocsr(m)
    b = oracle.make_rhs(m.n, 271.0)
    res = {}
    with gpu.Context(m, spmv_variant=5) as ctx, gpu.Net(seed=2) as net:
        ctx.set_rhs(b)
        for kind in (gpu.PRECOND_NONE, gpu.PRECOND_NEUMANN):
            ctx.set_preconditioner(kind)
            st = ctx.solve(mode, RTOL, 2000)
            res[kind] = (st, ctx.solution())
        p = _random_params(7) * 0.1
        net.set_params(p)
        net.apply(ctx)
        ctx.set_preconditioner(gpu.PRECOND_LEARNED)
        st = ctx.solve(mode, RTOL, 2000)
        res[gpu.PRECOND_LEARNED] = (st, ctx.solution())
    ox, _, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=5000)
    for kind, (st, x) in res.items():
        rel = float(np.linalg.norm(b - oracle.spmv(om, x)) / np.linalg.norm(b))
        print(f"mode {mode} precond {kind}: {st.as_dict()} host relres {rel:.2e}")
        assert st.converged == 1 and rel <= RTOL and st.precond == kind and st.fallback == 0
        assert np.max(np.abs(x - ox)) <= 1e-8 * np.max(np.abs(ox))
    assert res[gpu.PRECOND_NEUMANN][0].iterations <= res[gpu.PRECOND_NONE][0].iterations


def test_training_steps_reduce_the_pair_loss(gpu):
    m, om, b = _setup(gpu, n=6000, nonzer=11, shift=40.0, seed=91.0)
    with gpu.Context(m, spmv_variant=1) as ctx, gpu.Net(seed=4) as net:
        ctx.set_rhs(b)
        ctx.log_pairs(8)
        ctx.set_preconditioner(gpu.PRECOND_NEUMANN)
        first = net.step(ctx, 3e-3)          # untrained readout: the base loss
        losses = [net.step(ctx, 3e-3) for _ in range(60)]
    print(f"loss: base {first:.5e} after 60 steps {losses[-1]:.5e}")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.9 * first


def test_train_pipeline_small(gpu):
    o = gpu.default_train_opts()
    o.systems, o.n_min, o.n_max = 12, 2000, 5000
    o.nonzer_min, o.nonzer_max = 7, 11
    o.gen_threads = 4
    with gpu.Net(seed=8) as net, gpu.Net(seed=8) as net2:
        st, lh, bh = net.train(o)
        st2, lh2, bh2 = net2.train(o)    # same seed: same run
    print(st.as_dict())
    assert st.systems == 12 and st.steps == 12 * o.steps_per_system
    assert np.all(np.isfinite(lh)) and np.all(bh > 0)
    assert np.array_equal(lh, lh2) and np.array_equal(bh, bh2)
    assert lh[0] == bh[0]          # the first system sees the untrained (g = h = 0) network
