"""Contracts of the host-array entry points on the GPU.

* The drop-in conj_grad sees in-place edits of a[] / colidx[] (the full check,
  the default): the reference's serial conj_grad reads its arrays on every call
  (cg.cu:703-709), so a stale device copy must never be used silently.  The
  sampled check keeps the reference GPU path's contract instead (setup_gpu
  uploads once, cg.cu:1390-1603).
* The exact-signature shim re-creates its context after an in-place edit and
  returns results for the edited matrix.
* Step budgets learned by one solve do not slow a harder later one: after a
  restart the budget grows back and the next solve of that kind needs one
  sweep again.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _ocsr(m, a=None):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a if a is None else a, m.ratio, m.nonzer,
                      m.shift, m.rcond, m.seed)


def _close(v, ref, x):
    scale = max(1.0, float(np.max(np.abs(x))), float(np.max(np.abs(ref))))
    return float(np.max(np.abs(v - ref))) <= 1e-12 * scale


def _diag_positions(m):
    rows = np.repeat(np.arange(m.n), np.diff(m.rowstr))
    return np.nonzero(m.colidx == rows)[0]


@pytest.mark.parametrize("variant", [3, 5])
def test_dropin_detects_in_place_edit(gpu, variant):
    m = gpu.makea(20000, 11, 30.0, 271.0)
    x = np.ones(m.n + 2)
    with gpu.Context(m, spmv_variant=variant) as ctx:
        z, p, q, r, rn = ctx.conj_grad(x)
        oz, _, _, _, orn = oracle.conj_grad(_ocsr(m), x)
        assert _close(z[: m.n], oz[: m.n], x)
        # a symmetric in-place edit of the diagonal: same pointers, same pattern
        a_saved = m.a.copy()
        m.a[_diag_positions(m)] += 1.0
        try:
            with pytest.raises(gpu.NpbcgError) as e:
                ctx.conj_grad(x)
            assert e.value.code == gpu.ERR_MISMATCH
            # one entry of colidx moved within its row: also seen
        finally:
            m.a[:] = a_saved
        z2, _, _, _, _ = ctx.conj_grad(x)   # restored contents: accepted again
        assert np.array_equal(z2, z)
        k = int(m.rowstr[1]) - 1
        saved = int(m.colidx[k])
        m.colidx[k] = saved - 1 if saved - 1 > m.colidx[k - 1] else saved
        if m.colidx[k] != saved:
            with pytest.raises(gpu.NpbcgError):
                ctx.conj_grad(x)
        m.colidx[k] = saved


def test_sampled_check_keeps_the_upload_once_contract(gpu):
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    x = np.ones(m.n + 2)
    with gpu.Context(m, matrix_check=gpu.CHECK_SAMPLED) as ctx:
        z, _, _, _, _ = ctx.conj_grad(x)
        a_saved = m.a.copy()
        m.a[_diag_positions(m)] += 1.0
        try:
            z2, _, _, _, _ = ctx.conj_grad(x)   # same pointers: the resident copy is used
        finally:
            m.a[:] = a_saved
    assert np.array_equal(z, z2)


def test_compat_shim_follows_in_place_edits(gpu):
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    n = m.n
    L = gpu.LIB
    assert L.npbcg_compat_set_n(n) == 0
    a = m.a.copy()
    x = np.ones(n + 2)
    outs = [np.zeros(n + 2) for _ in range(4)]
    rn = ctypes.c_double(0.0)
    try:
        for edit in (0.0, 2.0):
            a[_diag_positions(m)] += edit
            L.npbcg_compat_conj_grad(m.colidx, m.rowstr, x, outs[0], a, outs[1], outs[2],
                                     outs[3], ctypes.byref(rn))
            assert np.isfinite(rn.value), gpu.last_error()
            oz, op, oq, orr, orn = oracle.conj_grad(_ocsr(m, a), x)
            for v, ref in zip(outs, (oz, op, oq, orr)):
                assert _close(v[:n], ref[:n], x)
            assert abs(rn.value - orn) <= 1e-12 * np.sqrt(n)
    finally:
        L.npbcg_compat_release()


@pytest.mark.parametrize("mode", [0, 1])
def test_budget_recovers_after_a_harder_solve(gpu, mode):
    # n = 1400, shift 10: kappa about 7, a random right-hand side needs ~30
    # CG steps; an eigenvector right-hand side converges in one.
    m = gpu.makea(1400, 7, 10.0, 12345.0)
    om = _ocsr(m)
    w, v = np.linalg.eigh(om.dense())
    easy = v[:, 700].copy()
    hard = oracle.make_rhs(m.n, 12345.0)
    batch = 60
    with gpu.Context(m, spmv_variant=5, solve_batch=batch) as ctx:
        seq = []
        for b in (easy, hard, hard, easy, hard):
            ctx.set_rhs(b)
            st = ctx.solve(mode, 1e-10, 3000)
            rel = float(np.linalg.norm(b - oracle.spmv(om, ctx.solution())) / np.linalg.norm(b))
            assert st.converged == 1 and rel <= 1e-10
            seq.append((st.iterations, st.sweeps))
    print(f"mode {mode}: (steps, sweeps) {seq}")
    if mode == 0:
        assert seq[0][1] == 1 and seq[1][1] >= 2      # the easy budget is too small
        assert seq[2][1] == 1                         # it has grown back
        assert seq[4][1] <= 2
    else:
        assert seq[2][1] <= seq[1][1] and seq[4][1] <= seq[1][1] + 1
