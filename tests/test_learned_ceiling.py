"""What any preconditioner can gain on the evaluation family (nonzer 15,
shift 110): the premises behind the measured result that the learned and
Neumann forms do not shorten the n = 150000 solve.

On a dense order-2000 member of the family (seed 314) this computes the
condition number of A, of the Jacobi-preconditioned D^-1/2 A D^-1/2 and of
the sparse top-K form M^-1 = D^-1 - D^-1 (E + E^T) D^-1 (the network's base,
g = h = 0), the classical CG bound on the steps to a 1e-10 relative residual,
sqrt(kappa)-based, and the steps plain and preconditioned CG actually take.
The matrix is NPB's makea (cg.cu:1303-1368) at a smaller order; the solves
are the reference's CG recurrence (cg.cu:689-772) with z = M^-1 r.
"""
import numpy as np
import pytest

import oracle
from oracle import learned

pytestmark = pytest.mark.slow

N, NONZER, SHIFT, SEED = 2000, 15, 110.0, 314.0
RTOL = 1e-10


def _dense(m):
    A = np.zeros((m.n, m.n))
    for j in range(m.n):
        A[j, m.colidx[m.rowstr[j]:m.rowstr[j + 1]]] = m.a[m.rowstr[j]:m.rowstr[j + 1]]
    return A


def _kappa(A, Minv=None):
    if Minv is None:
        ev = np.linalg.eigvalsh(A)
    else:
        L = np.linalg.cholesky(Minv)            # Minv = L L^T, SPD
        ev = np.linalg.eigvalsh(L.T @ A @ L)
    assert ev[0] > 0
    return ev[-1] / ev[0]


def _cg_bound(kappa, rtol=RTOL):
    rho = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    return int(np.ceil(np.log(2 / rtol) / -np.log(rho)))


def _pcg_steps(A, b, Minv, rtol=RTOL, maxit=200):
    x = np.zeros_like(b)
    r = b.copy()
    z = Minv @ r
    p = z.copy()
    rz = r @ z
    for k in range(1, maxit + 1):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if np.linalg.norm(b - A @ x) <= rtol * np.linalg.norm(b):
            return k
        z = Minv @ r
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return maxit


def test_preconditioning_ceiling_on_the_evaluation_family():
    m = oracle.makea(N, NONZER, SHIFT, SEED)
    A = _dense(m)
    assert np.allclose(A, A.T, rtol=1e-14, atol=0.0)
    A = 0.5 * (A + A.T)
    b = oracle.make_rhs(N, SEED)
    d = np.diag(A).copy()
    g = learned.graph(m)
    M_neu = learned.precond_dense(g, np.zeros(2 * N))
    M_neu = 0.5 * (M_neu + M_neu.T)
    # at this shift the diagonal (rcond - shift + ...) is negative and A is
    # negative definite; CG and every M^-1 below act on -A alike, so the
    # spectra are taken of -A and -M^-1
    assert np.all(d < 0)
    A, M_neu = -A, -M_neu
    forms = {"none": np.eye(N), "jacobi": np.diag(1.0 / -d), "neumann": M_neu}
    kap = {k: _kappa(A, None if k == "none" else M) for k, M in forms.items()}
    steps = {k: _pcg_steps(A, b, M) for k, M in forms.items()}
    for k in forms:
        print(f"{k}: kappa {kap[k]:.4f}, CG bound {_cg_bound(kap[k])} steps, "
              f"CG takes {steps[k]} steps")
    # the family is well conditioned: plain CG needs few steps to 1e-10
    assert kap["none"] < 1.5 and steps["none"] <= 10
    # a diagonal does not help, and the sparse form lowers kappa only a little
    assert kap["jacobi"] >= 0.99 * kap["none"]
    assert kap["neumann"] < kap["none"] and kap["neumann"] > 0.9 * kap["none"]
    # so neither saves more than one of the handful of steps
    assert steps["none"] - min(steps["jacobi"], steps["neumann"]) <= 1
    assert _cg_bound(kap["none"]) - _cg_bound(kap["neumann"]) <= 1
