"""General CSR inputs on the GPU: random symmetric, diagonally dominant
matrices (not NPB's generator) with short, medium and long rows, columns
sorted or shuffled within rows (shuffled rows make the library fall back from
the LDS-staged layout to the CSR row kernels), and rows with no off-diagonal
entry.  Every SpMV variant, the drop-in conj_grad and the power loop are
checked against the oracle's SpMV / conj_grad / power loop on the same arrays,
with the tolerances of the NPB tests (1e-12 relative); so are explicit
LDS-staged shapes (block widths 1000-19000, 64-19000 rows per range).  FP64
and mixed-precision solves reach a host-recomputed relative residual of
1e-10 and match the oracle's CG solution to 1e-8."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = [  # n, off-diagonal draws per row, shuffle columns, seed
    (300, 1, False, 1),
    (5000, 15, False, 2),
    (5000, 15, True, 3),
    (20000, 75, False, 4),
    (20000, 75, True, 5),
]


def _random_matrix(gpu, n, k, shuffle, seed):
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), k)
    cols = rng.integers(0, n, size=n * k)
    vals = rng.uniform(-1.0, 1.0, size=n * k)
    keep = rows != cols
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    off = sp.coo_matrix((np.concatenate([vals, vals]),
                         (np.concatenate([rows, cols]), np.concatenate([cols, rows]))),
                        shape=(n, n)).tocsr()
    off.sum_duplicates()
    d = np.asarray(abs(off).sum(axis=1)).ravel() + 1.0
    a_mat = (off + sp.diags(d)).tocsr()
    a_mat.sort_indices()
    rowstr = a_mat.indptr.astype(np.int32)
    colidx = a_mat.indices.astype(np.int32)
    a = a_mat.data.astype(np.float64)
    if shuffle:
        for j in range(n):
            s, e = rowstr[j], rowstr[j + 1]
            perm = s + rng.permutation(e - s)
            colidx[s:e], a[s:e] = colidx[perm].copy(), a[perm].copy()
    m = gpu.Matrix(n, rowstr, colidx, a)
    om = oracle.CSR(n, rowstr, colidx, a, 1.0, 0, 10.0, 0.1, float(seed))
    return m, om


def _close(v, ref, scale, label):
    err = float(np.max(np.abs(v - ref)))
    assert err <= 1e-12 * max(1.0, scale, float(np.max(np.abs(ref)))), f"{label}: {err:.3e}"


@pytest.mark.parametrize("n,k,shuffle,seed", CASES)
def test_general_csr_spmv_variants(gpu, n, k, shuffle, seed):
    m, om = _random_matrix(gpu, n, k, shuffle, seed)
    rng = np.random.default_rng(seed + 100)
    v = np.zeros(n + 1)
    v[:n] = rng.standard_normal(n)
    ref = oracle.spmv(om, v[:n])
    for variant in (0, 1, 2, 3, 4, 5):
        with gpu.Context(m, spmv_variant=variant) as ctx:
            got = ctx.describe()["spmv_variant"]
            assert got == variant or (variant == 0) or (variant == 5 and shuffle and got == 1)
            ctx.set_vector(gpu.VEC_P, v)
            ctx.spmv_bench(1)
            q = ctx.get_vector(gpu.VEC_Q)
        _close(q[:n], ref, float(np.max(np.abs(v))), f"n={n} k={k} shuffle={shuffle} v{variant}")


@pytest.mark.parametrize("n,k,shuffle,seed", CASES)
def test_general_csr_conj_grad_and_power(gpu, n, k, shuffle, seed):
    m, om = _random_matrix(gpu, n, k, shuffle, seed)
    rng = np.random.default_rng(seed + 200)
    x = np.ones(n + 2)
    x[:n] += 0.25 * rng.standard_normal(n)
    x[n] = 1.5
    oz, op, oq, orr, orn = oracle.conj_grad(om, x)
    ozeta, ornorm, _ = oracle.power(om, 3)
    xmax = float(np.max(np.abs(x[:n])))
    for variant, form in ((0, 0), (5, 1), (5, 2), (1, 0)):
        with gpu.Context(m, spmv_variant=variant, cg_form=form) as ctx:
            z, p, q, r, rn = ctx.conj_grad(x)
            ctx.fill_x(1.0)
            zeta, rnorm = ctx.power(3, 10.0)
        lab = f"n={n} k={k} shuffle={shuffle} v{variant} form{form}"
        for vec, ref, name in ((z, oz, "z"), (p, op, "p"), (q, oq, "q"), (r, orr, "r")):
            _close(vec[:n], ref[:n], xmax, f"{lab} {name}")
        assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n]), lab
        assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 1.5 and p[n] == 1.5
        rel = np.abs(np.asarray(zeta) - ozeta) / np.abs(ozeta)
        assert np.all(rel <= 1e-12), f"{lab} zeta {rel.max():.3e}"


# Explicit LDS-staged shapes (column block width, rows per range): many and
# few column blocks, one to hundreds of slices per workgroup, 512- and
# 1024-thread workgroups, and a last block narrower than the others.
SHAPES = [(1000, 128), (2048, 64), (4096, 512), (9000, 3000), (19000, 19000)]


@pytest.mark.parametrize("width,rows", SHAPES)
def test_general_csr_sell_shapes(gpu, width, rows):
    m, om = _random_matrix(gpu, 20000, 75, False, 6)
    rng = np.random.default_rng(6 + width)
    v = np.zeros(m.n + 1)
    v[:m.n] = rng.standard_normal(m.n)
    ref = oracle.spmv(om, v[:m.n])
    x = np.ones(m.n + 2)
    x[:m.n] += 0.25 * rng.standard_normal(m.n)
    x[m.n] = 0.5
    oz, op, oq, orr, orn = oracle.conj_grad(om, x)
    for form in (1, 2):
        with gpu.Context(m, spmv_variant=5, tile_width=width, tile_rows=rows,
                         cg_form=form) as ctx:
            d = ctx.describe()
            assert d["spmv_variant"] == 5 and d["block_width"] == width, d
            ctx.set_vector(gpu.VEC_P, v)
            ctx.spmv_bench(2)  # two products: both slice orders
            q = ctx.get_vector(gpu.VEC_Q)
            z, p, qq, r, rn = ctx.conj_grad(x)
        lab = f"W={width} R={rows} form{form} ({d['col_blocks']} blocks, {d['sell_threads']} threads)"
        _close(q[:m.n], ref, float(np.max(np.abs(v))), lab + " q=Ap")
        xmax = float(np.max(np.abs(x[:m.n])))
        for vec, rf, name in ((z, oz, "z"), (p, op, "p"), (qq, oq, "q"), (r, orr, "r")):
            _close(vec[:m.n], rf[:m.n], xmax, f"{lab} {name}")
        assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:m.n]), lab


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,k,shuffle,seed", [(5000, 15, False, 7), (20000, 75, False, 8),
                                              (20000, 75, True, 9)])
def test_general_csr_solves(gpu, mode, n, k, shuffle, seed):
    # FP64 and mixed-precision solves to 1e-10 on general symmetric matrices.
    # The mixed solve needs the LDS-staged layout: with shuffled columns (CSR
    # row kernels) it is refused with an error, and the FP64 solve runs.
    m, om = _random_matrix(gpu, n, k, shuffle, seed)
    b = np.random.default_rng(seed + 300).uniform(-1.0, 1.0, n)
    if mode == 1 and shuffle:
        with gpu.Context(m) as ctx:
            ctx.set_rhs(b)
            with pytest.raises(gpu.NpbcgError):
                ctx.solve(mode, 1e-10, 2000)
        return
    with gpu.Context(m, spmv_variant=0 if shuffle else 5) as ctx:
        ctx.set_rhs(b)
        st = ctx.solve(mode, 1e-10, 2000)
        xs = ctx.solution()
    rel = float(np.linalg.norm(b - oracle.spmv(om, xs)) / np.linalg.norm(b))
    ox, oit, orel = oracle.cg_solve(om, b, rtol=1e-10, maxit=5000)
    err = float(np.max(np.abs(xs - ox)) / np.max(np.abs(ox)))
    print(f"mode {mode} n={n} k={k} shuffle={shuffle}: {st.as_dict()} host relres {rel:.2e}, "
          f"oracle its {oit}, x rel err {err:.2e}")
    assert st.converged == 1 and rel <= 1e-10 and st.relres <= 1e-10
    assert err <= 1e-8
