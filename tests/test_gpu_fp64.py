"""GPU parity of the FP64 conj_grad path (the HIP library through its C ABI)
against the CPU oracle and against the reference's own trajectories.

Tolerances (FP64; the GPU sums in a different, fixed order):
  * zeta: relative 1e-12 per outer iteration (the target the path is held to);
  * vectors z, p, q, r: max abs difference <= 1e-12 * max(1, max|x|, max|v|);
  * rnorm: abs difference <= 1e-12 * ||x||.
NPB verification (relative 1e-10 to the published zeta) is checked as well.
"""
import hashlib

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ZETA_RTOL = 1e-12
VEC_TOL = 1e-12


def _ocsr(m):
    return oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond,
                      m.seed)


def _gen(lib, c):
    return lib.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])


def _vec_close(v, ref, xmax, label):
    scale = max(1.0, xmax, float(np.max(np.abs(ref))))
    err = float(np.max(np.abs(v - ref)))
    print(f"{label}: max abs err {err:.3e} (scale {scale:.3e})")
    assert err <= VEC_TOL * scale, label


def _zeta_close(z, ref, label):
    z, ref = np.asarray(z), np.asarray(ref)
    rel = np.abs(z - ref) / np.abs(ref)
    print(f"{label}: max rel zeta err {rel.max():.3e}")
    assert np.all(rel <= ZETA_RTOL), label


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "W_npb"])
def test_conj_grad_dropin_matches_oracle(gpu, golden, name):
    c = golden[name]
    m = _gen(gpu, c)
    n = m.n
    rng = np.random.default_rng(7)
    x = np.ones(n + 2)
    x[:n] += 0.25 * rng.standard_normal(n)
    x[n] = 3.0
    with gpu.Context(m) as ctx:
        z, p, q, r, rn = ctx.conj_grad(x)
    oz, op, oq, orr, orn = oracle.conj_grad(_ocsr(m), x)
    xmax = float(np.max(np.abs(x[:n])))
    for v, ref, lab in [(z, oz, "z"), (p, op, "p"), (q, oq, "q"), (r, orr, "r")]:
        _vec_close(v[:n], ref[:n], xmax, f"{name} {lab}")
    assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n])
    # index n: q = z = 0, r = p = x[n] (cg.cu:671-676)
    assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 3.0 and p[n] == 3.0


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "W_npb", "A_npb"])
def test_power_matches_oracle_and_reference(gpu, golden, name):
    c = golden[name]
    m = _gen(gpu, c)
    niter = c["niter"]
    with gpu.Context(m) as ctx:
        zeta, rnorm = ctx.power(niter, c["shift"])
        xg = ctx.get_vector(gpu.VEC_X)
    ozeta, ornorm, ox = oracle.power(_ocsr(m), niter)
    _zeta_close(zeta, ozeta, f"{name} vs oracle")
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"]], f"{name} vs reference")
    assert np.all(np.abs(rnorm - ornorm) <= 1e-12 * np.sqrt(m.n))
    _vec_close(xg[: m.n], ox[: m.n], 1.0, f"{name} final x")
    assert xg[m.n] == 1.0
    if name.endswith("_npb"):
        zref = oracle.NPB_CLASSES[c["npb_class"]][4]
        assert abs(zeta[-1] - zref) / zref <= oracle.NPB_EPSILON


@pytest.mark.parametrize("name", ["B_npb", "cfg2_n75000_seed271828", "C_npb",
                                  "cfg3_n150000_seed314"])
def test_power_large_matches_reference(gpu, golden, name):
    c = golden[name]
    m = _gen(gpu, c)
    with gpu.Context(m) as ctx:
        zeta, rnorm = ctx.power(c["niter"], c["shift"])
    ref = [float.fromhex(h) for h in c["zeta_hex"]]
    _zeta_close(zeta, ref, f"{name} vs reference")
    rref = np.array([float.fromhex(h) for h in c["rnorm_hex"]])
    print(f"{name}: max abs rnorm err {np.max(np.abs(rnorm - rref)):.3e}")
    assert np.all(np.abs(rnorm - rref) <= 1e-12 * np.sqrt(m.n))
    if name.endswith("_npb"):
        zref = oracle.NPB_CLASSES[c["npb_class"]][4]
        assert abs(zeta[-1] - zref) / zref <= oracle.NPB_EPSILON


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("name", ["cfg1_n1400_seed12345", "A_npb"])
def test_spmv_variants(gpu, golden, name, variant):
    c = golden[name]
    m = _gen(gpu, c)
    rng = np.random.default_rng(variant)
    v = np.zeros(m.n + 1)
    v[: m.n] = rng.standard_normal(m.n)
    with gpu.Context(m, spmv_variant=variant) as ctx:
        ctx.set_vector(gpu.VEC_P, v)
        ctx.spmv_bench(1)
        q = ctx.get_vector(gpu.VEC_Q)
        zeta, _ = ctx.power(3, c["shift"])
    _vec_close(q[: m.n], oracle.spmv(_ocsr(m), v[: m.n]), float(np.max(np.abs(v))),
               f"{name} spmv variant {variant}")
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"][:3]],
                f"{name} variant {variant}")


@pytest.mark.parametrize("tile_rows,tile_width", [(1, 2), (7, 130), (64, 511), (1400, 1400),
                                                  (333, 4096)])
def test_tiled_layout_shapes(gpu, golden, tile_rows, tile_width):
    # Many column blocks, an odd-width last block, partial last tiles, one
    # row per tile: every shape must give the oracle's answer.
    c = golden["cfg1_n1400_seed12345"]
    m = _gen(gpu, c)
    n = m.n
    x = np.ones(n + 2)
    x[:n] += np.linspace(-0.3, 0.3, n)
    with gpu.Context(m, spmv_variant=5, tile_rows=tile_rows, tile_width=tile_width,
                     fuse_p_update=tile_rows % 2) as ctx:
        z, p, q, r, rn = ctx.conj_grad(x)
        ctx.fill_x(1.0)  # the drop-in call left x on the device
        zeta, _ = ctx.power(4, c["shift"])
    oz, op, oq, orr, orn = oracle.conj_grad(_ocsr(m), x)
    for v, ref, lab in [(z, oz, "z"), (q, oq, "q"), (r, orr, "r")]:
        _vec_close(v[:n], ref[:n], 1.3, f"tiled {tile_rows}x{tile_width} {lab}")
    assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n])
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"][:4]], "tiled power")


@pytest.mark.parametrize("variant,fuse", [(1, 0), (5, 0), (5, 1)])
def test_class_b_variants_match_reference(gpu, golden, variant, fuse):
    c = golden["B_npb"]
    m = _gen(gpu, c)
    with gpu.Context(m, spmv_variant=variant, fuse_p_update=fuse) as ctx:
        zeta, _ = ctx.power(c["niter"], c["shift"])
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"]], f"B variant {variant}")


def test_power_is_deterministic_and_graph_matches_direct(gpu, golden):
    c = golden["A_npb"]
    m = _gen(gpu, c)
    with gpu.Context(m) as g, gpu.Context(m, use_graph=0) as d:
        z1, r1 = g.power(6, c["shift"])
        g.fill_x(1.0)
        z2, r2 = g.power(6, c["shift"])
        z3, r3 = d.power(6, c["shift"])
    assert z1.tobytes() == z2.tobytes() and r1.tobytes() == r2.tobytes()
    assert z1.tobytes() == z3.tobytes() and r1.tobytes() == r3.tobytes()


def test_device_steps_reproduce_power_loop(gpu, golden):
    # conj_grad_gpu + gpu_kernel_ten + gpu_kernel_eleven, as cg.cu:500-520 calls them
    c = golden["S_npb"]
    m = _gen(gpu, c)
    with gpu.Context(m) as ctx:
        zs = []
        for _ in range(4):
            ctx.conj_grad_device()
            t1, t2 = ctx.norms()
            zs.append(c["shift"] + 1.0 / t1)
            ctx.scale_x(1.0 / np.sqrt(t2))
    ref = [float.fromhex(h) for h in c["zeta_hex"][:4]]
    _zeta_close(zs, ref, "device steps")


def test_compat_shim_reference_signature(gpu, golden):
    import ctypes

    c = golden["cfg1_n1400_seed12345"]
    m = _gen(gpu, c)
    n = m.n
    x = np.ones(n + 1)
    z, p, q, r = (np.zeros(n + 1) for _ in range(4))
    rn = ctypes.c_double(0.0)
    assert gpu.LIB.npbcg_compat_set_n(n) == 0
    gpu.LIB.npbcg_compat_conj_grad(m.colidx, m.rowstr, x, z, m.a, p, q, r, ctypes.byref(rn))
    gpu.LIB.npbcg_compat_release()
    oz, _, _, _, orn = oracle.conj_grad(_ocsr(m), np.ones(n + 2))
    _vec_close(z[:n], oz[:n], 1.0, "compat z")
    assert abs(rn.value - orn) <= 1e-12 * np.sqrt(n)


def test_mismatched_arrays_are_rejected(gpu, golden):
    c = golden["cfg1_n1400_seed12345"]
    m = _gen(gpu, c)
    other = gpu.makea(m.n, 7, 10.0, 99.0)
    with gpu.Context(m) as ctx:
        x = np.ones(m.n + 1)
        with pytest.raises(gpu.NpbcgError) as e:
            ctx.conj_grad(x, colidx=other.colidx, rowstr=other.rowstr, a=other.a)
        assert e.value.code == gpu.ERR_MISMATCH
        a2 = m.a.copy()
        a2[::7] *= 2.0
        with pytest.raises(gpu.NpbcgError):
            ctx.conj_grad(x, a=a2)
        ctx.conj_grad(x, colidx=m.colidx.copy(), rowstr=m.rowstr.copy(), a=m.a.copy())


def test_vector_roundtrip_and_sampling(gpu, golden):
    c = golden["S_npb"]
    m = _gen(gpu, c)
    with gpu.Context(m) as ctx:
        v = np.arange(m.n + 1, dtype=np.float64)
        for w in range(5):
            ctx.set_vector(w, v + w)
            assert np.array_equal(ctx.get_vector(w), v + w)
        ctx.fill_x(1.0)
        ctx.set_spmv_sampling(1)
        zeta, _ = ctx.power(2, c["shift"])
        t = ctx.timing()
        assert t.spmv_samples == 25 and t.spmv_avg_ms > 0 and t.total_ms > 0
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"][:2]], "sampled run")


@pytest.mark.parametrize("size", ["cfg1", "n20000"])
def test_single_system_tuner_persists_and_matches_oracle(gpu, golden, tmp_path, size):
    if size == "cfg1":
        c = golden["cfg1_n1400_seed12345"]
        m, niter, shift = _gen(gpu, c), c["niter"], c["shift"]
    else:
        m, niter, shift = gpu.makea(20000, 11, 30.0, 271.0), 3, 30.0
    cache = str(tmp_path / "tune.txt")
    shape, ms = gpu.tune(m, cache)
    print(f"{size}: tuned {shape}, {ms * 1e3:.2f} us")
    assert shape["spmv_variant"] in (1, 2, 3, 4, 5) and ms > 0
    lines = open(cache).read().splitlines()
    assert len(lines) == 1
    assert lines[0].split()[1:4] == [str(shape[k]) for k in ("spmv_variant", "tile_width",
                                                              "tile_rows")]
    shape2, ms2 = gpu.tune(m, cache)  # the record is read back, not re-timed
    assert shape2 == shape and abs(ms2 - ms) <= 1e-5
    assert len(open(cache).read().splitlines()) == 1
    with gpu.Context(m, **shape) as ctx:
        assert ctx.describe()["spmv_variant"] == shape["spmv_variant"]
        zeta, rnorm = ctx.power(niter, shift)
    ozeta, _, _ = oracle.power(_ocsr(m), niter)
    _zeta_close(zeta, ozeta, f"{size} tuned shape vs oracle")


@pytest.mark.parametrize("cg_form,dot_pass", [(1, 0), (2, 1), (2, 2)])
@pytest.mark.parametrize("name", ["S_npb", "A_npb", "B_npb", "cfg3_n150000_seed314"])
def test_cg_forms_match_reference(gpu, golden, name, cg_form, dot_pass):
    # cg_form 1: the reference's step order; 2: the one-reduction form (s = A r,
    # p.q from r.s and the previous step's scalars), with r.s formed in the
    # SpMV (dot_pass 1) or in a combine pass of its own (dot_pass 2, the
    # default from 24 column blocks up).  All are held to the reference's own
    # zeta and rnorm trajectories.
    c = golden[name]
    m = _gen(gpu, c)
    with gpu.Context(m, spmv_variant=5, cg_form=cg_form, dot_pass=dot_pass) as ctx:
        d = ctx.describe()
        assert d["cg_form"] == cg_form
        assert d["dot_pass"] == (dot_pass if cg_form == 2 else 1)
        zeta, rnorm = ctx.power(c["niter"], c["shift"])
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"]],
                f"{name} form {cg_form} dot pass {dot_pass}")
    rref = np.array([float.fromhex(h) for h in c["rnorm_hex"]])
    assert np.all(np.abs(rnorm - rref) <= 1e-12 * np.sqrt(m.n))


def test_capped_range_shape_matches_oracle(gpu):
    # At n = 1,000,000 one row range's per-slice p.q slots no longer fit in
    # the LDS beside a full vector block, so the default shape narrows the
    # block and caps the range, then picks the range count whose last round
    # of workgroups fills the most CU slots (the class D path, here with a
    # sparse matrix the oracle runs in seconds).  The power loop on that
    # shape is held to the oracle's zeta and rnorm.  (With 13 entries per row
    # the automatic choice would take a CSR row kernel, so the layout is named.)
    m = gpu.makea(1000000, 3, 10.0, 314159265.0)
    niter = 2
    with gpu.Context(m, spmv_variant=5) as ctx:
        d = ctx.describe()
        zeta, rnorm = ctx.power(niter, m.shift)
    print(d)
    assert d["spmv_variant"] == 5
    nwg = d["col_blocks"] * d["row_ranges"]
    assert d["sell_workgroups"] == nwg and d["range_rows"] * d["row_ranges"] >= m.n
    # the capped ranges hold at most the slices the LDS leaves beside the block
    assert (d["block_width"] + (d["range_rows"] + 63) // 64) * 8 <= 160 * 1024
    assert d["sort_window"] == 8192
    ozeta, ornorm, _ = oracle.power(_ocsr(m), niter)
    _zeta_close(zeta, ozeta, "capped range shape")
    assert np.all(np.abs(rnorm - ornorm) <= 1e-12 * np.sqrt(m.n) * np.maximum(1.0, np.abs(ornorm)))


@pytest.mark.parametrize("window", [64, 2048, 8192])
def test_sort_windows_match_reference(gpu, golden, window):
    # Rows are sorted by their length in a column block within sort windows
    # before they are cut into 64-row slices.  The window changes which rows
    # share a slice (the padding, and the order of the per-slice p.q
    # partials), never the order of a row's own sum.  The n = 150000 matrix
    # has 4688-row ranges, so 64, 2048 and 8192 (the whole range) give three
    # layouts; each is held to the reference's zeta and rnorm trajectories,
    # and a mixed solve on it (FP32 row sums collected in LDS, ranges wider
    # than 2048 rows) to the FP64 true residual.
    c = golden["cfg3_n150000_seed314"]
    m = _gen(gpu, c)
    b = oracle.make_rhs(m.n, 314.0)
    with gpu.Context(m, spmv_variant=5, sort_window=window) as ctx:
        d = ctx.describe()
        zeta, rnorm = ctx.power(c["niter"], c["shift"])
        ctx.set_rhs(b)
        st = ctx.solve(1, 1e-10, 200)
        x = ctx.solution()
    assert d["sort_window"] == window
    _zeta_close(zeta, [float.fromhex(h) for h in c["zeta_hex"]], f"sort window {window}")
    rref = np.array([float.fromhex(h) for h in c["rnorm_hex"]])
    assert np.all(np.abs(rnorm - rref) <= 1e-12 * np.sqrt(m.n))
    rel = float(np.linalg.norm(b - oracle.spmv(_ocsr(m), x)) / np.linalg.norm(b))
    assert st.converged == 1 and rel <= 1e-10, (st.as_dict(), rel)


def test_sort_window_padding_and_default(gpu, golden):
    # Cutting a descending run into consecutive 64-row slices gives the least
    # padding of any cut into slices of 64, so padding cannot grow as
    # windows nest: 64 within 2048 within the whole 4688-row range.  The
    # default sorts the whole range (4736 rows, from 2048 up to 8192).
    c = golden["cfg3_n150000_seed314"]
    m = _gen(gpu, c)
    stored = {}
    for w in (64, 2048, 8192, 0):
        with gpu.Context(m, spmv_variant=5, sort_window=w) as ctx:
            d = ctx.describe()
        stored[w] = d["sell_stored"]
        if w == 0:
            assert d["sort_window"] == 4736 and d["range_rows"] == 4688
    print(stored)
    assert stored[64] >= stored[2048] >= stored[8192] >= m.nnz
    assert stored[2048] > stored[8192] and stored[0] == stored[8192]
    for bad in (-64, 100):
        with pytest.raises(gpu.NpbcgError) as e:
            gpu.Context(m, sort_window=bad)
        assert e.value.code == gpu.ERR_INVALID


@pytest.mark.parametrize("cg_form,dot_pass", [(1, 0), (2, 1), (2, 2)])
def test_cg_forms_dropin_vectors(gpu, golden, cg_form, dot_pass):
    c = golden["cfg1_n1400_seed12345"]
    m = _gen(gpu, c)
    n = m.n
    rng = np.random.default_rng(11)
    x = np.ones(n + 2)
    x[:n] += 0.25 * rng.standard_normal(n)
    x[n] = 2.0
    with gpu.Context(m, spmv_variant=5, cg_form=cg_form, dot_pass=dot_pass) as ctx:
        z, p, q, r, rn = ctx.conj_grad(x)
    oz, op, oq, orr, orn = oracle.conj_grad(_ocsr(m), x)
    xmax = float(np.max(np.abs(x[:n])))
    for v, ref, lab in [(z, oz, "z"), (p, op, "p"), (q, oq, "q"), (r, orr, "r")]:
        _vec_close(v[:n], ref[:n], xmax, f"form {cg_form} {lab}")
    assert abs(rn - orn) <= 1e-12 * np.linalg.norm(x[:n])
    assert z[n] == 0.0 and q[n] == 0.0 and r[n] == 2.0 and p[n] == 2.0
