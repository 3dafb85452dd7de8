"""Per-kernel profiling (npbcg_profile_kernels, the counterpart of the
reference's PROFILING build, cg.cu:565-589) on the GPU: the profiled run
computes the same zeta as npbcg_power, bit for bit, and as the reference's
trajectory (relative 1e-12); its table lists every kernel of the path with
the launch counts the step structure fixes; the kernel times add up to the
reported total."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ZETA_RTOL = 1e-12


def _gen(lib, c):
    return lib.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])


def _calls(kernels):
    return {k["name"]: k["calls"] for k in kernels}


@pytest.mark.parametrize("variant,cg_form", [(5, 2), (5, 1), (1, 0)])
def test_profile_matches_power_and_counts(gpu, golden, variant, cg_form):
    c = golden["S_npb"]
    m = _gen(gpu, c)
    niter, cgit = c["niter"], 25
    opts = {"spmv_variant": variant}
    if cg_form:
        opts["cg_form"] = cg_form
    with gpu.Context(m, **opts) as ctx:
        ctx.fill_x(1.0)
        zeta_p, _ = ctx.power(niter, c["shift"])
        ctx.fill_x(1.0)
        kernels, total_ms, zeta = ctx.profile_kernels(niter, c["shift"])
        t = ctx.timing()
    for k in kernels:
        print(f"{k['name']:45s} {k['threads']:5d} {k['calls']:5d} {k['ms']:9.4f} ms")
    print(f"total {total_ms:.4f} ms, launches {t.launches}")
    assert np.array_equal(zeta, zeta_p)
    ref = np.array([float.fromhex(h) for h in c["zeta_hex"]])
    assert np.all(np.abs(zeta - ref) / np.abs(ref) <= ZETA_RTOL)

    calls = _calls(kernels)
    assert calls["cg_init_kernel"] == niter
    assert calls["scale_x_kernel"] == 1
    assert calls["resid_finish_kernel"] == niter
    # one CG SpMV per inner step, whatever its form
    cg_spmv = sum(v for k, v in calls.items() if k.startswith("spmv_") and "A z" not in k)
    assert cg_spmv == niter * cgit
    assert sum(v for k, v in calls.items() if "A z" in k) == niter
    if variant == 5 and cg_form == 2:
        assert calls["spmv_sell_kernel (s = A r)"] == niter * cgit
        assert calls["update_cg_kernel"] == niter * cgit
        assert calls["update_p_kernel"] == niter
        assert calls["resid_combine_kernel"] == niter
    if variant == 1:
        assert calls["update_zr_kernel"] == niter * cgit
        assert calls["update_p_kernel"] == niter * cgit
        assert "resid_combine_kernel" not in calls
    assert sum(calls.values()) + 1 == t.launches
    assert all(k["ms"] >= 0.0 for k in kernels)
    assert abs(sum(k["ms"] for k in kernels) - total_ms) <= 1e-9 * max(total_ms, 1.0)
    assert abs(t.total_ms - total_ms) <= 1e-9 * max(total_ms, 1.0)


def test_profile_kernel_limit_is_an_error(gpu, golden):
    c = golden["S_npb"]
    m = _gen(gpu, c)
    with gpu.Context(m, spmv_variant=5) as ctx:
        ctx.fill_x(1.0)
        with pytest.raises(gpu.NpbcgError):
            ctx.profile_kernels(1, c["shift"], max_kinds=2)
