"""npbcg_conj_grad_device_noalloc (no context, no device allocation: its
partial sums and scalars live in the library's fixed device array) against
the existing entry point it is a variant of, npbcg_conj_grad_device on a
context created with spmv_variant = 1 (the same kernels in the same order,
cg.cu:803-846 computed on the device).  Two calls in a row on the same GPU,
the second with a smaller matrix, each compared with the context's z, p, q,
r and rnorm: bit-identical is expected, and the test bound is ||d|| <=
1e-12 * scale as for every FP64 vector of the drop-in.  The inputs and
outputs are device arrays the test allocates through the HIP runtime, as a
caller that cannot allocate inside its own loop would own them."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        lib = ctypes.CDLL("libamdhip64.so")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _HIP = lib
    return _HIP


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        hip = _hip()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert _hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _start_vector(n, seed):
    rng = np.random.default_rng(seed)
    x = np.ones(n + 2)
    x[:n] += 0.25 * rng.standard_normal(n)
    x[n] = 3.0
    x[n + 1] = 0.0
    return x


def _noalloc(gpu, m, x):
    n = m.n
    bufs = {"rowstr": DevArray(m.rowstr), "colidx": DevArray(m.colidx), "a": DevArray(m.a),
            "x": DevArray(x)}
    for k in ("z", "p", "q", "r"):
        bufs[k] = DevArray(np.full(n + 2, np.nan))   # every output entry must be written
    bufs["rnorm"] = DevArray(np.full(1, np.nan))
    try:
        rc = gpu.LIB.npbcg_conj_grad_device_noalloc(
            n, bufs["rowstr"].ptr, bufs["colidx"].ptr, bufs["a"].ptr, bufs["x"].ptr,
            bufs["z"].ptr, bufs["p"].ptr, bufs["q"].ptr, bufs["r"].ptr, 25, bufs["rnorm"].ptr,
            None)
        assert rc == 0, gpu.last_error()
        assert _hip().hipDeviceSynchronize() == 0
        out = {k: bufs[k].read()[: n + 1] for k in ("z", "p", "q", "r")}
        out["rnorm"] = float(bufs["rnorm"].read()[0])
        return out
    finally:
        for b in bufs.values():
            b.free()


def _existing(gpu, m, x):
    n = m.n
    with gpu.Context(m, spmv_variant=1) as ctx:
        assert ctx.describe()["spmv_variant"] == 1
        ctx.set_vector(gpu.VEC_X, x)
        rn = ctx.conj_grad_device()
        out = {k: ctx.get_vector(w) for k, w in (("z", gpu.VEC_Z), ("p", gpu.VEC_P),
                                                 ("q", gpu.VEC_Q), ("r", gpu.VEC_R))}
    out["rnorm"] = rn
    assert all(v.shape == (n + 1,) for k, v in out.items() if k != "rnorm")
    return out


@pytest.mark.parametrize("cases", [[("cfg2_n75000_seed271828", 5), ("S_npb", 6)]])
def test_noalloc_matches_conj_grad_device(gpu, golden, cases):
    for name, seed in cases:          # second call: a smaller matrix on the same GPU
        c = golden[name]
        m = gpu.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
        x = _start_vector(m.n, seed)
        print(f"noalloc vs conj_grad_device: {name} n={m.n} nnz={m.nnz}", flush=True)
        got = _noalloc(gpu, m, x)
        ref = _existing(gpu, m, x)
        scale = max(1.0, float(np.max(np.abs(x[: m.n]))))
        for k in ("z", "p", "q", "r"):
            assert np.all(np.isfinite(got[k])), k
            err = float(np.max(np.abs(got[k] - ref[k])))
            bound = 1e-12 * max(scale, float(np.max(np.abs(ref[k]))))
            print(f"  {k}: max |d| {err:.3e} (bound {bound:.1e}), identical "
                  f"{np.array_equal(got[k], ref[k])}")
            assert err <= bound, k
        assert abs(got["rnorm"] - ref["rnorm"]) <= 1e-12 * max(1.0, abs(ref["rnorm"]))
        # index n as conj_grad leaves it (cg.cu:671-676)
        n = m.n
        assert got["z"][n] == 0.0 and got["q"][n] == 0.0
        assert got["r"][n] == 3.0 and got["p"][n] == 3.0
        if m.n <= 14000:  # and the oracle's serial conj_grad
            om = oracle.CSR(m.n, m.rowstr, m.colidx, m.a, m.ratio, m.nonzer, m.shift, m.rcond,
                            m.seed)
            oz, op, oq, orr, orn = oracle.conj_grad(om, x)
            for k, v in (("z", oz), ("p", op), ("q", oq), ("r", orr)):
                err = float(np.max(np.abs(got[k][:n] - v[:n])))
                assert err <= 1e-12 * max(scale, float(np.max(np.abs(v[:n])))), k
            assert abs(got["rnorm"] - orn) <= 1e-12 * np.linalg.norm(x[:n])


def test_noalloc_rejects_bad_arguments(gpu):
    one = DevArray(np.zeros(4))
    try:
        p = one.ptr
        rc = gpu.LIB.npbcg_conj_grad_device_noalloc(0, p, p, p, p, p, p, p, p, 25, p, None)
        assert rc == gpu.ERR_INVALID
        rc = gpu.LIB.npbcg_conj_grad_device_noalloc(3, p, p, p, p, None, p, p, p, 25, p, None)
        assert rc == gpu.ERR_INVALID
        rc = gpu.LIB.npbcg_conj_grad_device_noalloc(3, p, p, p, p, p, p, p, p, 0, p, None)
        assert rc == gpu.ERR_INVALID
    finally:
        one.free()
