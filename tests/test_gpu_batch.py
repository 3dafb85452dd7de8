"""Batched solves on the GPU: every (system, right-hand side) solution is
checked against the oracle's FP64 CG solve of that system, and by its FP64
true residual recomputed on the host."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond, m.seed)


def _systems(gpu, count=5, seed=3):
    rng = np.random.default_rng(seed)
    params = [(int(rng.integers(1500, 5000)), int(rng.integers(7, 11)), 20.0, float(s + 1))
              for s in range(count)]
    return gpu.make_many(params, workers=4)


def _rhs(mats, nrhs):
    return [np.stack([oracle.make_rhs(m.n, 1000.0 * s + k) for k in range(nrhs)], axis=1)
            for s, m in enumerate(mats)]


@pytest.mark.parametrize("mode,nrhs,layout", [(1, 4, 2), (0, 4, 2), (1, 1, 2), (0, 2, 2),
                                              (1, 4, 1), (0, 1, 1), (1, 4, 3), (0, 4, 3),
                                              (1, 1, 3), (0, 1, 3), (1, 2, 3), (0, 2, 3)])
def test_batch_solves_match_oracle(gpu, mode, nrhs, layout):
    mats = _systems(gpu)
    rhs = _rhs(mats, nrhs)
    with gpu.Batch(mats, nrhs=nrhs, mode=mode, layout=layout) as bt:
        bt.set_rhs(rhs)
        st, steps, rel = bt.solve(RTOL, 3000)
        xs = bt.solution()
    print(f"mode {mode} nrhs {nrhs}: {st.as_dict()}; steps {steps.ravel().tolist()}")
    assert st.layout == layout
    assert st.converged == len(mats) * nrhs and st.max_relres <= RTOL
    for s, m in enumerate(mats):
        om = _ocsr(m)
        for k in range(nrhs):
            b = rhs[s][:, k]
            x = xs[s][:, k]
            host = np.linalg.norm(b - oracle.spmv(om, x)) / np.linalg.norm(b)
            assert host <= RTOL and abs(host - rel[s, k]) <= 1e-3 * RTOL + 0.05 * host
            ox, oit, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=5000)
            assert np.max(np.abs(x - ox)) <= 1e-8 * np.max(np.abs(ox))
            if mode == 0:   # FP64 batch steps follow the oracle's CG closely
                assert abs(int(steps[s, k]) - oit) <= 3


def test_batch_launch_shapes_agree(gpu):
    mats = _systems(gpu, 3, seed=9)
    rhs = _rhs(mats, 4)
    sols = []
    for lpr, rt, layout in [(4, 64, 1), (16, 256, 1), (64, 1024, 1), (8, 64, 2), (16, 512, 2),
                            (64, 4096, 2), (8, 512, 3), (16, 3000, 3), (16, 4096, 3)]:
        with gpu.Batch(mats, nrhs=4, mode=0, lanes_per_row=lpr, rows_per_tile=rt,
                       layout=layout) as bt:
            bt.set_rhs(rhs)
            st, _, _ = bt.solve(RTOL, 3000)
            assert st.layout == layout and st.converged == st.solves
            if layout == 3:  # one 2048-row sorting window per item, whatever is asked
                assert st.rows_per_tile == 2048
            sols.append(np.concatenate(bt.solution()))
    for x in sols[1:]:
        assert np.max(np.abs(x - sols[0])) <= 1e-9 * np.max(np.abs(sols[0]))


def test_batch_autotuner_persists(gpu, tmp_path):
    mats = _systems(gpu, 4, seed=5)
    cache = str(tmp_path / "tune.txt")
    with gpu.Batch(mats, nrhs=4) as bt:
        lpr, rt, ms = bt.tune(cache)
        bt.set_rhs(_rhs(mats, 4))
        st, _, _ = bt.solve(RTOL, 3000)
    assert lpr in (4, 8, 16, 32, 64) and 64 <= rt <= 4096 and ms > 0
    assert st.lanes_per_row == lpr and st.rows_per_tile == rt and st.converged == st.solves
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].split()[1:3] == [str(lpr), str(rt)]
    with gpu.Batch(mats, nrhs=4) as bt:     # same shape: the record is reused, not re-timed
        assert bt.tune(cache)[:2] == (lpr, rt)
    assert len(open(cache).read().splitlines()) == 1


def test_batch_reports_unconverged_and_errors(gpu):
    mats = _systems(gpu, 2, seed=4)
    with gpu.Batch(mats, nrhs=1, mode=0) as bt:
        with pytest.raises(gpu.NpbcgError) as e:
            bt.solve(RTOL, 10)
        assert e.value.code == gpu.ERR_INVALID        # no right-hand side yet
        bt.set_rhs(_rhs(mats, 1))
        st, steps, rel = bt.solve(RTOL, 3, check=False)
        assert st.converged < st.solves and np.all(steps <= 3)
    with pytest.raises(gpu.NpbcgError):
        gpu.Batch(mats, nrhs=3)


def test_batch_many_systems_xcd_placement(gpu):
    # 12 systems: tiles are placed per XCD (at least 8 systems), uneven sizes
    mats = _systems(gpu, 12, seed=21)
    rhs = _rhs(mats, 4)
    with gpu.Batch(mats, nrhs=4, mode=1) as bt:
        bt.set_rhs(rhs)
        st, steps, rel = bt.solve(RTOL, 3000)
        xs = bt.solution()
    assert st.layout == 3  # the default layout: SELL slices
    assert st.converged == 48 and st.max_relres <= RTOL
    for s, m in enumerate(mats):
        om = _ocsr(m)
        for k in range(4):
            b = rhs[s][:, k]
            assert np.linalg.norm(b - oracle.spmv(om, xs[s][:, k])) <= RTOL * np.linalg.norm(b)


def test_batch_sell_many_blocks_and_windows(gpu):
    # systems wide enough for several column blocks (8064 columns with four
    # FP32 right-hand sides, 3008 with four FP64 ones) and several 2048-row
    # sorting windows, checked against the oracle's CG
    params = [(9000, 9, 20.0, 7.0), (20000, 10, 20.0, 8.0), (33000, 8, 20.0, 9.0)]
    mats = gpu.make_many(params, workers=3)
    rhs = _rhs(mats, 4)
    for mode in (1, 0):
        with gpu.Batch(mats, nrhs=4, mode=mode, layout=3, rows_per_tile=4096) as bt:
            bt.set_rhs(rhs)
            st, steps, rel = bt.solve(RTOL, 3000)
            xs = bt.solution()
        assert st.layout == 3 and st.converged == 12 and st.max_relres <= RTOL
        for s, m in enumerate(mats):
            om = _ocsr(m)
            for k in (0, 3):
                b = rhs[s][:, k]
                ox, _, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=5000)
                assert np.max(np.abs(xs[s][:, k] - ox)) <= 1e-8 * np.max(np.abs(ox))


@pytest.mark.parametrize("mode,nrhs", [(1, 4), (0, 4), (0, 1), (1, 2)])
def test_batch_sell_window_edges(gpu, mode, nrhs):
    # SELL window items at the edges of their shapes: systems smaller than one
    # 64-row slice, exactly one 2048-row sorting window and one row more, one
    # column block (3008 columns for four FP64 right-hand sides) and one
    # column more, several blocks with a one-column last block, and more than
    # one FP32 block (8064 columns).  Every solution is checked against the
    # oracle's CG and the host residual.
    params = [(40, 5, 20.0, 11.0), (2048, 7, 20.0, 12.0), (2049, 8, 20.0, 13.0),
              (3008, 7, 20.0, 14.0), (3009, 9, 20.0, 15.0), (6017, 8, 20.0, 16.0),
              (8065, 7, 20.0, 17.0), (100, 6, 20.0, 18.0)]
    mats = gpu.make_many(params, workers=4)
    rhs = _rhs(mats, nrhs)
    with gpu.Batch(mats, nrhs=nrhs, mode=mode, layout=3) as bt:
        bt.set_rhs(rhs)
        st, steps, rel = bt.solve(RTOL, 3000)
        xs = bt.solution()
    print(f"mode {mode} nrhs {nrhs}: {st.as_dict()}; steps {steps.ravel().tolist()}")
    assert st.layout == 3 and st.converged == len(mats) * nrhs and st.max_relres <= RTOL
    for s, m in enumerate(mats):
        om = _ocsr(m)
        for k in range(nrhs):
            b = rhs[s][:, k]
            x = xs[s][:, k]
            host = np.linalg.norm(b - oracle.spmv(om, x)) / np.linalg.norm(b)
            assert host <= RTOL and abs(host - rel[s, k]) <= 1e-3 * RTOL + 0.05 * host
            ox, _, _ = oracle.cg_solve(om, b, rtol=RTOL, maxit=5000)
            assert np.max(np.abs(x - ox)) <= 1e-8 * np.max(np.abs(ox))


@pytest.mark.parametrize("mode", [1, 0])
def test_batch_full_workload(gpu, mode):
    # The full batched workload: 64 systems, n ~ U{2000..40000}, nonzer ~
    # U{7..13} (numpy PCG64 seeded s), shift 20, matrix seeds 1..64, four
    # right-hand sides uniform(-1, 1) each (PCG64 seeded 1000 + s), default
    # launch shape.  Every solve must meet 1e-10; the FP64 residual of every
    # eighth system is recomputed on the host.
    params = []
    for s in range(1, 65):
        rng = np.random.default_rng(s)
        n = int(rng.integers(2000, 40001))
        nz = int(rng.integers(7, 14))
        params.append((n, nz, 20.0, float(s)))
    mats = gpu.make_many(params)
    rhs = [np.random.default_rng(1000 + s).uniform(-1.0, 1.0, (m.n, 4))
           for s, m in enumerate(mats, start=1)]
    with gpu.Batch(mats, nrhs=4, mode=mode) as bt:
        bt.set_rhs(rhs)
        st, steps, rel = bt.solve(RTOL, 3000)
        xs = bt.solution()
    print(f"mode {mode}: {st.as_dict()}; max steps {int(steps.max())}")
    assert st.solves == 256 and st.converged == 256 and st.max_relres <= RTOL
    assert np.all(rel <= RTOL)
    for s in range(0, 64, 8):
        om = _ocsr(mats[s])
        for k in range(4):
            b = rhs[s][:, k]
            host = np.linalg.norm(b - oracle.spmv(om, xs[s][:, k])) / np.linalg.norm(b)
            assert host <= RTOL and abs(host - rel[s, k]) <= 1e-3 * RTOL + 0.05 * host
