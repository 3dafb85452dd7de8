"""Pins the CPU oracle (oracle/npbcg_oracle.c) before anything is compared to it.

Checks, each chosen so a plausible slip in the restatement fails one of them:
  * randlc against the exact integer closed form a^k x0 mod 2^46;
  * the generator and the serial power loop against the reference's own serial
    code (golden fixtures written by tests/golden/make_golden.py): CSR digests,
    ratio bits and every per-iteration (zeta, rnorm), bit for bit;
  * NPB's published verification zetas (cg.cu:373-393, eps 1e-10);
  * dense linear algebra at tiny n: matvec, solve, and the inverse-power limit
    zeta -> shift + lambda_min|.|(A);
  * structural invariants of the assembled CSR.
"""
import hashlib

import numpy as np
import pytest

import oracle

A46 = 1220703125
M46 = 1 << 46


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("seed", [314159265, 12345, 271828, 314, 1, M46 - 1])
def test_randlc_matches_integer_closed_form(seed):
    x = float(seed)
    xi = seed
    for k in range(1, 3001):
        u, x = oracle.randlc(x)
        xi = (A46 * xi) % M46
        assert int(x) == xi, f"step {k}"
        assert u == xi / float(M46)
    # jump ahead: x_k = a^k x0 mod 2^46
    assert int(x) == (pow(A46, 3000, M46) * seed) % M46


def _check_case(m, case, niter_check=None):
    assert m.nnz == case["nnz"]
    assert _sha(m.rowstr) == case["sha256_rowstr"]
    assert _sha(m.colidx) == case["sha256_colidx"]
    assert _sha(m.a) == case["sha256_a"]
    assert float(m.ratio).hex() == case["ratio_hex"]
    k = case["niter"] if niter_check is None else min(niter_check, case["niter"])
    zeta, rnorm, x = oracle.power(m, k)
    ref_zeta = [float.fromhex(h) for h in case["zeta_hex"][:k]]
    ref_rnorm = [float.fromhex(h) for h in case["rnorm_hex"][:k]]
    assert [float(v) for v in zeta] == ref_zeta
    assert [float(v) for v in rnorm] == ref_rnorm
    return zeta, x


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "W_npb", "A_npb"])
def test_oracle_bit_exact_with_reference_small(golden, name):
    c = golden[name]
    m = oracle.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
    zeta, x = _check_case(m, c)
    assert hashlib.sha256(x[: c["n"]].tobytes()).hexdigest() == c["x_final_sha256"]


@pytest.mark.parametrize("name", ["B_npb", "cfg2_n75000_seed271828"])
def test_oracle_bit_exact_with_reference_class_b(golden, name):
    c = golden[name]
    m = oracle.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
    _check_case(m, c, niter_check=2)


@pytest.mark.slow
@pytest.mark.parametrize("name", ["C_npb", "cfg3_n150000_seed314"])
def test_oracle_bit_exact_with_reference_class_c(golden, name):
    c = golden[name]
    m = oracle.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
    _check_case(m, c, niter_check=1)


@pytest.mark.parametrize("cls", ["S", "W", "A"])
def test_oracle_npb_verification_zeta(cls):
    n, nonzer, niter, shift, zref = oracle.NPB_CLASSES[cls]
    m = oracle.makea(n, nonzer, shift)
    zeta, _, _ = oracle.power(m, niter)
    assert abs(zeta[-1] - zref) / zref <= oracle.NPB_EPSILON


@pytest.mark.parametrize("name,cls", [("B_npb", "B"), ("C_npb", "C")])
def test_reference_trajectory_reaches_npb_zeta(golden, name, cls):
    # The reference's own final zeta (fixture) passes NPB verification, which
    # ties the fixtures of the long classes to the published constants.
    zref = oracle.NPB_CLASSES[cls][4]
    z = float.fromhex(golden[name]["zeta_hex"][-1])
    assert abs(z - zref) / zref <= oracle.NPB_EPSILON


def _tiny(n=48, nonzer=4, shift=3.0, seed=31415.0):
    return oracle.makea(n, nonzer, shift, seed)


def test_spmv_matches_dense():
    m = _tiny()
    d = m.dense()
    rng = np.random.default_rng(1)
    v = rng.standard_normal(m.n)
    np.testing.assert_allclose(oracle.spmv(m, v), d @ v, rtol=1e-13, atol=1e-13)


def test_conj_grad_solves_tiny_system_and_postconditions():
    m = _tiny()
    n = m.n
    d = m.dense()
    x = np.ones(n + 2)
    x[n] = 7.0  # index n is copied through, never used
    z, p, q, r, rnorm = oracle.conj_grad(m, x, cgitmax=3 * n)
    np.testing.assert_allclose(z[:n], np.linalg.solve(d, x[:n]), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(r[:n], d @ z[:n], rtol=1e-12, atol=1e-12)
    assert abs(rnorm - np.linalg.norm(x[:n] - d @ z[:n])) < 1e-12
    assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 7.0 and p[n] == 7.0


def test_power_loop_converges_to_eigenvalue():
    m = _tiny(n=40, nonzer=3, shift=2.5, seed=4711.0)
    d = m.dense()
    lam = np.linalg.eigvalsh(0.5 * (d + d.T))
    lstar = lam[np.argmin(np.abs(lam))]
    zeta, _, _ = oracle.power(m, 400, cgitmax=120)
    assert abs(zeta[-1] - (m.shift + lstar)) < 1e-9 * abs(m.shift + lstar)


def test_cg_solve_reaches_true_residual():
    m = _tiny(n=64, nonzer=4, shift=6.0, seed=271.0)
    b = oracle.make_rhs(m.n, 271.0)
    x, it, rel = oracle.cg_solve(m, b, rtol=1e-12, maxit=1000)
    d = m.dense()
    assert rel <= 1e-12
    assert np.linalg.norm(b - d @ x) / np.linalg.norm(b) <= 1e-12
    assert 0 < it < 1000


def test_pcg_solve_reduces_to_cg_and_solves_with_diagonal():
    m = _tiny(n=64, nonzer=4, shift=6.0, seed=271.0)
    b = oracle.make_rhs(m.n, 271.0)
    d = m.dense()
    assert np.array_equal(oracle.diagonal(m), np.diag(d))
    x0, it0, rel0 = oracle.cg_solve(m, b, rtol=1e-12, maxit=1000)
    x1, it1, rel1 = oracle.pcg_solve(m, b, np.ones(m.n), rtol=1e-12, maxit=1000)
    # M = I is plain CG, step for step
    assert np.array_equal(x0, x1) and it0 == it1 and rel0 == rel1
    # Jacobi and a positive random diagonal both reach the true residual
    rng = np.random.default_rng(5)
    for minv in (1.0 / np.diag(d), np.exp(rng.uniform(-1, 1, m.n)) / np.diag(d)):
        x, it, rel = oracle.pcg_solve(m, b, minv, rtol=1e-12, maxit=1000)
        assert rel <= 1e-12 and 0 < it < 1000
        assert np.linalg.norm(b - d @ x) / np.linalg.norm(b) <= 1e-12
    # With M^-1 = A^-1 (the diagonal of a diagonal matrix) PCG takes one step.
    dm = oracle.CSR(m.n, np.arange(m.n + 1, dtype=np.int32), np.arange(m.n, dtype=np.int32),
                    np.diag(d).copy(), m.ratio, m.nonzer, m.shift, m.rcond, m.seed)
    x, it, rel = oracle.pcg_solve(dm, b, 1.0 / np.diag(d), rtol=1e-12, maxit=10)
    assert it == 1 and rel <= 1e-15


def test_make_rhs_definition():
    b = oracle.make_rhs(5, 314.0)
    t = 315.0
    for i in range(5):
        u, t = oracle.randlc(t)
        assert b[i] == 2.0 * u - 1.0


@pytest.mark.parametrize("n,nonzer,shift,seed", [(1400, 7, 10.0, 314159265.0),
                                                  (500, 5, 4.0, 99.0), (64, 2, 1.0, 7.0)])
def test_csr_invariants(n, nonzer, shift, seed):
    m = oracle.makea(n, nonzer, shift, seed)
    rs, ci, a = m.rowstr, m.colidx, m.a
    assert rs[0] == 0 and np.all(np.diff(rs) >= 1)
    for j in range(n):
        row = ci[rs[j]:rs[j + 1]]
        assert np.all(np.diff(row) > 0)
        assert j in row  # every vector carries its own index (vecset)
    d = m.dense()
    assert np.array_equal(d != 0, (d != 0).T)
    off = d - d.T
    assert np.max(np.abs(off)) <= 4 * np.finfo(float).eps * np.max(np.abs(d))


def test_makea_capacity_error():
    import ctypes

    n, nonzer = 200, 5
    rowstr = np.zeros(n + 1, dtype=np.int32)
    cap = 10
    colidx = np.zeros(cap, dtype=np.int32)
    a = np.zeros(cap)
    rc = oracle._lib().orc_makea(n, nonzer, 0.1, 3.0, 5.0, cap, rowstr, colidx, a,
                                 ctypes.byref(ctypes.c_double()))
    assert rc == -2
