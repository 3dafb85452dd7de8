"""CPU oracle for the NPB CG ``conj_grad`` path — TEST INFRASTRUCTURE ONLY.

ctypes front end for ``liboracle.so`` (the plain-C restatement in
``npbcg_oracle.c``) and, where it was built, for ``_ref/libnpbref_<CLASS>.so``
(the reference's own serial code compiled by ``make -C oracle ref``).

Only ``tests/``, the smoke check and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker.  The product library never
routes through it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB: Optional[ctypes.CDLL] = None

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_DP = ctypes.POINTER(ctypes.c_double)

# NPB CG problem classes: setparams.cpp:762-775 (na, nonzer, niter, shift)
# and the verification zetas of cg.cu:373-393 (valid for seed 314159265).
NPB_SEED = 314159265.0
NPB_RCOND = 0.1
NPB_CLASSES = {
    "S": (1400, 7, 15, 10.0, 8.5971775078648),
    "W": (7000, 8, 15, 12.0, 10.362595087124),
    "A": (14000, 11, 15, 20.0, 17.130235054029),
    "B": (75000, 13, 75, 60.0, 22.712745482631),
    "C": (150000, 15, 75, 110.0, 28.973605592845),
    "D": (1500000, 21, 100, 500.0, 52.514532105794),
    "E": (9000000, 26, 100, 1500.0, 77.522164599383),
}
NPB_EPSILON = 1.0e-10  # cg.cu:534


def lib_path() -> str:
    return os.path.join(_HERE, "liboracle.so")


def _lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        lib = ctypes.CDLL(path)
        lib.orc_randlc.restype = ctypes.c_double
        lib.orc_randlc.argtypes = [_DP, ctypes.c_double]
        lib.orc_makea.restype = ctypes.c_int
        lib.orc_makea.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                  _I32P, _I32P, _F64P, _DP]
        lib.orc_conj_grad.restype = None
        lib.orc_conj_grad.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P,
                                      _F64P, _F64P, _F64P, _F64P, _F64P, _F64P, _DP]
        lib.orc_power.restype = None
        lib.orc_power.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, _I32P, _I32P, _F64P, _F64P,
                                  _F64P, _F64P, _F64P, _F64P, _F64P, _F64P]
        lib.orc_spmv.restype = None
        lib.orc_spmv.argtypes = [ctypes.c_int, _I32P, _I32P, _F64P, _F64P, _F64P]
        lib.orc_cg_solve.restype = ctypes.c_int
        lib.orc_cg_solve.argtypes = [ctypes.c_int, _I32P, _I32P, _F64P, _F64P,
                                     _F64P, ctypes.c_double, ctypes.c_int, _DP, _F64P]
        lib.orc_pcg_solve.restype = ctypes.c_int
        lib.orc_pcg_solve.argtypes = [ctypes.c_int, _I32P, _I32P, _F64P, _F64P, _F64P,
                                      _F64P, ctypes.c_double, ctypes.c_int, _DP, _F64P]
        lib.orc_make_rhs.restype = None
        lib.orc_make_rhs.argtypes = [ctypes.c_int, ctypes.c_double, _F64P]
        _LIB = lib
    return _LIB


def randlc(x: float, a: float = 1220703125.0) -> tuple[float, float]:
    """One LCG step; returns (u, new_x)."""
    xv = ctypes.c_double(x)
    u = _lib().orc_randlc(ctypes.byref(xv), a)
    return u, xv.value


class CSR:
    """A generated NPB matrix, 0-based CSR with int32 indices."""

    def __init__(self, n, rowstr, colidx, a, ratio, nonzer, shift, rcond, seed):
        self.n = n
        self.rowstr = rowstr
        self.colidx = colidx
        self.a = a
        self.ratio = ratio
        self.nonzer = nonzer
        self.shift = shift
        self.rcond = rcond
        self.seed = seed

    @property
    def nnz(self) -> int:
        return int(self.rowstr[self.n])

    def dense(self) -> np.ndarray:
        m = np.zeros((self.n, self.n))
        for j in range(self.n):
            s, e = self.rowstr[j], self.rowstr[j + 1]
            m[j, self.colidx[s:e]] = self.a[s:e]
        return m


def makea(n: int, nonzer: int, shift: float, seed: float = NPB_SEED,
          rcond: float = NPB_RCOND) -> CSR:
    """The reference generator (cg.cu:406-408 + makea/sparse)."""
    cap = n * (nonzer + 1) * (nonzer + 1)
    rowstr = np.zeros(n + 1, dtype=np.int32)
    colidx = np.zeros(cap, dtype=np.int32)
    a = np.zeros(cap, dtype=np.float64)
    ratio = ctypes.c_double(0.0)
    nnz = _lib().orc_makea(n, nonzer, rcond, shift, float(seed), cap, rowstr,
                           colidx, a, ctypes.byref(ratio))
    if nnz < 0:
        raise RuntimeError(f"orc_makea failed with code {nnz}")
    return CSR(n, rowstr, colidx[:nnz].copy(), a[:nnz].copy(), ratio.value,
               nonzer, shift, rcond, seed)


def conj_grad(m: CSR, x: np.ndarray, cgitmax: int = 25):
    """Serial conj_grad; x has >= n+1 entries.  Returns (z, p, q, r, rnorm)."""
    n = m.n
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.shape[0] >= n + 1
    z, p, q, r = (np.zeros_like(x) for _ in range(4))
    rn = ctypes.c_double(0.0)
    _lib().orc_conj_grad(n, cgitmax, m.colidx, m.rowstr, x, z, m.a, p, q, r,
                         ctypes.byref(rn))
    return z, p, q, r, rn.value


def power(m: CSR, niter: int, cgitmax: int = 25, x0: Optional[np.ndarray] = None):
    """Serial inverse-power loop from x = 1 (or x0).  Returns (zeta[], rnorm[], x)."""
    n = m.n
    x = np.ones(n + 2) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    z, p, q, r = (np.zeros(n + 2) for _ in range(4))
    zeta = np.zeros(niter)
    rnorm = np.zeros(niter)
    _lib().orc_power(n, niter, cgitmax, m.shift, m.colidx, m.rowstr, m.a, x, z,
                     p, q, r, zeta, rnorm)
    return zeta, rnorm, x


def spmv(m: CSR, x: np.ndarray) -> np.ndarray:
    y = np.zeros(m.n)
    _lib().orc_spmv(m.n, m.rowstr, m.colidx, m.a,
                    np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def make_rhs(n: int, seed: float) -> np.ndarray:
    b = np.zeros(n)
    _lib().orc_make_rhs(n, float(seed), b)
    return b


def cg_solve(m: CSR, b: np.ndarray, rtol: float = 1e-10, maxit: int = 10000):
    """Plain FP64 CG to a true relative residual.  Returns (x, iters, relres)."""
    n = m.n
    x = np.zeros(n)
    work = np.zeros(4 * n)
    rr = ctypes.c_double(0.0)
    it = _lib().orc_cg_solve(n, m.rowstr, m.colidx, m.a,
                             np.ascontiguousarray(b, dtype=np.float64), x, rtol,
                             maxit, ctypes.byref(rr), work)
    return x, it, rr.value


def pcg_solve(m: CSR, b: np.ndarray, minv: np.ndarray, rtol: float = 1e-10,
              maxit: int = 10000):
    """FP64 CG preconditioned by the diagonal minv.  Returns (x, iters, relres)."""
    n = m.n
    x = np.zeros(n)
    work = np.zeros(5 * n)
    rr = ctypes.c_double(0.0)
    it = _lib().orc_pcg_solve(n, m.rowstr, m.colidx, m.a,
                              np.ascontiguousarray(minv, dtype=np.float64),
                              np.ascontiguousarray(b, dtype=np.float64), x, rtol,
                              maxit, ctypes.byref(rr), work)
    return x, it, rr.value


def diagonal(m: CSR) -> np.ndarray:
    """a_jj of every row (0 where the row stores no diagonal entry)."""
    d = np.zeros(m.n)
    rows = np.repeat(np.arange(m.n), np.diff(m.rowstr[: m.n + 1]))
    on = m.colidx[: m.rowstr[m.n]] == rows
    d[rows[on]] = m.a[: m.rowstr[m.n]][on]
    return d


# ---------------------------------------------------------------------------
# The reference's own serial code (oracle/_ref, built from the reference tree)
# ---------------------------------------------------------------------------

def ref_lib_path(cls: str) -> str:
    return os.path.join(_HERE, "_ref", f"libnpbref_{cls}.so")


def ref_available(cls: str) -> bool:
    return os.path.exists(ref_lib_path(cls))


class RefClass:
    """One class-specific build of the reference serial CG (sizes are macros)."""

    def __init__(self, cls: str):
        path = ref_lib_path(cls)
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path} (make -C oracle ref)")
        lib = ctypes.CDLL(path)
        lib.ref_params.restype = None
        lib.ref_params.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3 + [_DP, _DP]
        lib.ref_makea.restype = ctypes.c_long
        lib.ref_makea.argtypes = [ctypes.c_double, _I32P, _I32P, _F64P, ctypes.c_long]
        lib.ref_conj_grad.restype = None
        lib.ref_conj_grad.argtypes = [_F64P] * 5 + [_DP]
        lib.ref_power.restype = None
        lib.ref_power.argtypes = [ctypes.c_int, _F64P, _F64P, _F64P]
        self.lib = lib
        na, nz, ni = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        sh, rc = ctypes.c_double(), ctypes.c_double()
        lib.ref_params(ctypes.byref(na), ctypes.byref(nz), ctypes.byref(ni),
                       ctypes.byref(sh), ctypes.byref(rc))
        self.n, self.nonzer, self.niter = na.value, nz.value, ni.value
        self.shift, self.rcond = sh.value, rc.value

    def makea(self, seed: float = NPB_SEED) -> CSR:
        n = self.n
        cap = n * (self.nonzer + 1) ** 2
        rowstr = np.zeros(n + 1, dtype=np.int32)
        colidx = np.zeros(cap, dtype=np.int32)
        a = np.zeros(cap)
        nnz = self.lib.ref_makea(float(seed), rowstr, colidx, a, cap)
        if nnz < 0:
            raise RuntimeError("ref_makea capacity")
        ratio = float(np.power(self.rcond, 1.0 / n))
        return CSR(n, rowstr, colidx[:nnz].copy(), a[:nnz].copy(), ratio,
                   self.nonzer, self.shift, self.rcond, seed)

    def conj_grad(self, x: np.ndarray):
        """Reference conj_grad on the matrix of the last makea call."""
        n = self.n
        xx = np.ascontiguousarray(x[: n + 1], dtype=np.float64).copy()
        z, p, q, r = (np.zeros(n + 1) for _ in range(4))
        rn = ctypes.c_double(0.0)
        self.lib.ref_conj_grad(xx, z, p, q, r, ctypes.byref(rn))
        return z, p, q, r, rn.value

    def power(self, niter: int):
        """Reference serial power loop from x = 1 on the last makea matrix."""
        zeta = np.zeros(niter)
        rnorm = np.zeros(niter)
        x = np.zeros(self.n)
        self.lib.ref_power(niter, zeta, rnorm, x)
        return zeta, rnorm, x
