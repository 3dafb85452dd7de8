"""Pins of the CPU restatement of the learned preconditioner (oracle.learned).

The device kernels are checked against oracle.learned on the GPU; these tests
check oracle.learned itself against independent facts: torch's BF16
rounding, brute-force top-K selection, the dense form of M^-1, central finite
differences of the float64 network and loss, and CG identities of the pairs."""
import numpy as np
import pytest
import torch

import oracle
from oracle import learned as ol


def _mat(n=96, nonzer=4, shift=6.0, seed=271.0):
    return oracle.makea(n, nonzer, shift, seed)


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4096).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4096),
                        np.float32([0.0, -0.0, 1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -(1 + 2.0 ** -8)])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ol.bf16(x), ref)
    assert ol.bf16(np.float32([1 + 2.0 ** -8]))[0] == 1.0           # tie -> even
    assert ol.bf16(np.float32([1 + 3 * 2.0 ** -8]))[0] == 1 + 2.0 ** -6


def test_graph_top_k_brute_force_and_padding():
    m = _mat()
    g = ol.graph(m)
    d = m.dense()
    for j in range(m.n):
        cand = sorted(((-abs(d[j, k]), k) for k in range(m.n) if k != j and d[j, k] != 0.0))
        want = [k for _, k in cand[:ol.K]]
        got = [int(k) for k, w in zip(g.nbr[j], g.wgt[j]) if w != 0.0]
        assert got == want
        assert np.all(g.nbr[j, len(want):] == j) and np.all(g.wgt[j, len(want):] == 0.0)
        if want:
            assert abs(g.wgt[j].sum() - 1.0) < 1e-12
        for t, k in enumerate(want):
            assert g.nval[j, t] == d[j, k] / (d[j, j] * d[k, k])
    assert np.array_equal(g.adiag, np.diag(d))
    # features from their definitions on the dense row
    j = 7
    off = [k for k in range(m.n) if k != j and d[j, k] != 0]
    a = np.array([d[j, k] for k in off])
    dk = np.abs(np.array([d[k, k] for k in off]))
    dj = abs(d[j, j])
    want = [np.abs(a).sum() / dj, (a * a).sum() / dj ** 2, np.abs(a).max() / dj, a.sum() / dj,
            (a * a / (dj * dk)).sum(), (np.abs(a) / np.sqrt(dj * dk)).sum(),
            np.log2(len(off) + 1) / 8, np.log10(dj) / 3]
    assert np.allclose(g.feat[j], want, rtol=1e-13, atol=0)


def test_graph_ties_and_short_rows():
    # 5x5 symmetric: row 0 has ties |a| = 1 at columns 1, 2, 3; row 4 has one off-diagonal
    d = np.diag([4.0, 5.0, 6.0, 7.0, 8.0])
    for (i, k, v) in [(0, 1, 1.0), (0, 2, -1.0), (0, 3, 1.0), (1, 2, 0.5), (3, 4, 2.0)]:
        d[i, k] = d[k, i] = v
    rs = [0]
    ci, a = [], []
    for j in range(5):
        for k in range(5):
            if d[j, k] != 0:
                ci.append(k)
                a.append(d[j, k])
        rs.append(len(ci))
    m = oracle.CSR(5, np.array(rs, np.int32), np.array(ci, np.int32), np.array(a), 0, 0, 0, 0, 0)
    g = ol.graph(m)
    assert list(g.nbr[0]) == [1, 2, 3, 0]          # ties by column, then a self pad
    assert list(g.nbr[4]) == [3, 4, 4, 4] and list(g.wgt[4]) == [1.0, 0.0, 0.0, 0.0]


def test_sym_apply_and_dense_preconditioner():
    m = _mat()
    g = ol.graph(m)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((3, m.n))
    C = ol.sym_matrix(g)
    assert np.allclose(C, C.T)
    assert np.allclose(ol.sym_apply(g, v), v @ C.T, rtol=1e-13, atol=1e-15)
    gh = rng.uniform(-0.5, 0.5, 2 * m.n)
    M = ol.precond_dense(g, gh)
    assert np.allclose(M, M.T, rtol=0, atol=1e-15)
    # g = h = 0 is D^-1 - D^-1 (E + E^T) D^-1 with E the top-K off-diagonals,
    # built here from the dense matrix and a brute-force selection
    A = m.dense()
    E = np.zeros_like(A)
    for j in range(m.n):
        cand = sorted((-abs(A[j, k]), k) for k in range(m.n) if k != j and A[j, k] != 0.0)
        for _, k in cand[:ol.K]:
            E[j, k] = A[j, k]
    Di = np.diag(1.0 / np.diag(A))
    M0 = ol.precond_dense(g, np.zeros(2 * m.n))
    assert np.allclose(M0, Di - Di @ (E + E.T) @ Di, rtol=1e-13, atol=1e-17)
    # the scales enter as S M S on the diagonal part and Q C Q off it
    s = np.exp(gh[:m.n])
    q = np.exp(gh[:m.n] + gh[m.n:])
    assert np.allclose(M, np.diag(s * s / np.diag(A)) - np.diag(q) @ Di @ (E + E.T) @ Di @ np.diag(q),
                       rtol=1e-12, atol=1e-17)


def _pairs(m, P=4):
    b = oracle.make_rhs(m.n, 5.0)
    xs, _, _ = oracle.cg_solve(m, b, rtol=1e-14, maxit=2000)
    return ol.cg_pairs(m, b, xs, P, lambda v: oracle.spmv(m, v))


def test_cg_pairs_are_residual_correction_pairs():
    m = _mat()
    R, E = _pairs(m, 5)
    A = m.dense()
    # A e_k = r_k (e_k = x* - x_k, r_k = b - A x_k), and CG residuals are orthogonal
    for k in range(5):
        assert np.linalg.norm(A @ E[k] - R[k]) <= 1e-9 * np.linalg.norm(R[0])
    G = R @ R.T
    assert np.all(np.abs(G - np.diag(np.diag(G))) <= 1e-9 * np.max(np.diag(G)))


def test_pair_loss_dense_form_and_finite_differences():
    m = _mat()
    g = ol.graph(m)
    R, E = _pairs(m, 4)
    rng = np.random.default_rng(2)
    gh = rng.uniform(-0.3, 0.3, 2 * m.n)
    loss, dgh = ol.pair_loss(g, gh, R, E)
    M = ol.precond_dense(g, gh)
    want = sum(np.sum((M @ R[p] - E[p]) ** 2) / (4 * np.sum(E[p] ** 2)) for p in range(4))
    assert abs(loss - want) <= 1e-12 * want
    eps = 1e-6
    for i in rng.choice(2 * m.n, 24, replace=False):
        e = np.zeros(2 * m.n)
        e[i] = eps
        fd = (ol.pair_loss(g, gh + e, R, E)[0] - ol.pair_loss(g, gh - e, R, E)[0]) / (2 * eps)
        assert abs(fd - dgh[i]) <= 1e-6 * np.max(np.abs(dgh)) + 1e-7 * abs(fd)


def test_network_gradient_finite_differences_float64():
    m = _mat(n=80)
    g = ol.graph(m)
    R, E = _pairs(m, 3)
    rng = np.random.default_rng(3)
    params = rng.standard_normal(ol.N_PARAMS) * 0.2
    params[ol.OFF_W0 + ol.F * ol.H:ol.OFF_B0] = 0.0      # padded input rows
    norm = ol.feature_norm(g)
    loss, grad, _ = ol.loss_and_grad(params, norm, g, R, E, bf16_on=False)
    eps = 1e-6
    idx = np.concatenate([rng.choice(np.arange(ol.OFF_W0, ol.OFF_W0 + ol.F * ol.H), 6),
                          rng.choice(np.arange(ol.OFF_B0, ol.OFF_W1), 3),
                          rng.choice(np.arange(ol.OFF_W1, ol.OFF_R), 12),
                          np.arange(ol.OFF_R, ol.N_PARAMS)[[0, 5, ol.H, ol.H + 1, ol.H + 9,
                                                            2 * ol.H + 1]]])
    for i in idx:
        e = np.zeros(ol.N_PARAMS)
        e[i] = eps
        lp = ol.loss_and_grad(params + e, norm, g, R, E, bf16_on=False)[0]
        lm = ol.loss_and_grad(params - e, norm, g, R, E, bf16_on=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-5 * np.max(np.abs(grad)) + 1e-4 * abs(fd), (i, fd, grad[i])


def test_untrained_readout_gives_the_base_preconditioner():
    m = _mat()
    g = ol.graph(m)
    params = np.random.default_rng(4).standard_normal(ol.N_PARAMS) * 0.1
    params[ol.OFF_R:] = 0.0
    act = ol.forward(params, ol.feature_norm(g), g)
    assert not np.any(act["gh"])
