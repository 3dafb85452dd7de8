"""Python bindings (ctypes) for libnpbcg, the gfx950 NPB CG conj_grad library.

The compute path is the HIP library; this module only marshals arrays across
the C ABI declared in ``include/npbcg.h``.  Importing it raises if the built
library is absent — there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnpbcg.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "npbcg.h")

OK = 0
ERR_INVALID = -1
ERR_NOMEM = -2
ERR_DEVICE = -3
ERR_NODEVICE = -4
ERR_CAPACITY = -5
ERR_INTERNAL = -6
ERR_MISMATCH = -7
ERR_NOCONV = -8

ENGINE_AUTO, ENGINE_KERNELS, ENGINE_PERSISTENT = 0, 1, 2
CHECK_FULL, CHECK_SAMPLED = 0, 1
VEC_X, VEC_Z, VEC_P, VEC_Q, VEC_R = 0, 1, 2, 3, 4
SPMV_AUTO, SPMV_SEG64, SPMV_SEG32, SPMV_SEG16, SPMV_SEG64_U2, SPMV_TILED = 0, 1, 2, 3, 4, 5

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_DP = ctypes.POINTER(ctypes.c_double)


class NpbcgError(RuntimeError):
    def __init__(self, code: int, where: str, text: str):
        super().__init__(f"{where}: status {code}: {text}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("cgitmax", ctypes.c_int),
                ("engine", ctypes.c_int), ("spmv_variant", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("tile_rows", ctypes.c_int),
                ("tile_width", ctypes.c_int), ("fuse_p_update", ctypes.c_int),
                ("solve_batch", ctypes.c_int), ("cg_form", ctypes.c_int),
                ("matrix_check", ctypes.c_int), ("dot_pass", ctypes.c_int),
                ("sort_window", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


SOLVE_FP64, SOLVE_MIXED = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_NEUMANN = 0, 1, 2, 3
MAX_PAIRS = 8
NET_FEATURES = 8


class BatchOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("nrhs", ctypes.c_int), ("mode", ctypes.c_int),
                ("lanes_per_row", ctypes.c_int), ("rows_per_tile", ctypes.c_int),
                ("inner_steps", ctypes.c_int), ("layout", ctypes.c_int),
                ("reserved", ctypes.c_int * 9)]


class BatchStats(ctypes.Structure):
    _fields_ = [("solves", ctypes.c_int), ("converged", ctypes.c_int), ("sweeps", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("lanes_per_row", ctypes.c_int),
                ("rows_per_tile", ctypes.c_int), ("layout", ctypes.c_int),
                ("reserved_i", ctypes.c_int),
                ("time_ms", ctypes.c_double), ("max_relres", ctypes.c_double),
                ("reserved_d", ctypes.c_double * 4)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class TrainOpts(ctypes.Structure):
    _fields_ = [("systems", ctypes.c_int), ("n_min", ctypes.c_int), ("n_max", ctypes.c_int),
                ("nonzer_min", ctypes.c_int), ("nonzer_max", ctypes.c_int),
                ("shift_min", ctypes.c_double), ("shift_max", ctypes.c_double),
                ("rcond", ctypes.c_double), ("seed", ctypes.c_ulonglong),
                ("pairs", ctypes.c_int), ("steps_per_system", ctypes.c_int),
                ("lr", ctypes.c_float), ("gen_threads", ctypes.c_int),
                ("first_system", ctypes.c_int), ("reserved", ctypes.c_int * 5)]


class TrainStats(ctypes.Structure):
    _fields_ = [("systems", ctypes.c_int), ("steps", ctypes.c_int),
                ("skipped", ctypes.c_int), ("reserved_i", ctypes.c_int),
                ("loss_first", ctypes.c_double), ("loss_last", ctypes.c_double),
                ("base_first", ctypes.c_double), ("base_last", ctypes.c_double),
                ("wait_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("reserved", ctypes.c_double * 3)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class SolveStats(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("sweeps", ctypes.c_int), ("fp64_spmvs", ctypes.c_int),
                ("fp32_spmvs", ctypes.c_int), ("precond", ctypes.c_int),
                ("fallback", ctypes.c_int), ("reserved_i", ctypes.c_int),
                ("relres", ctypes.c_double), ("time_ms", ctypes.c_double),
                ("reserved_d", ctypes.c_double * 4)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k in ("converged", "iterations", "sweeps",
                                               "fp64_spmvs", "fp32_spmvs", "precond",
                                               "fallback", "relres", "time_ms")}


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("spmv_avg_ms", ctypes.c_double),
                ("spmv_samples", ctypes.c_int), ("launches", ctypes.c_int),
                ("nnz", ctypes.c_longlong), ("n", ctypes.c_int), ("engine", ctypes.c_int)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("threads", ctypes.c_int), ("calls", ctypes.c_int),
                ("ms", ctypes.c_double)]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libnpbcg.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "npbcg_default_opts": (None, [ctypes.POINTER(Opts)]),
        "npbcg_abi_version": (ctypes.c_int, []),
        "npbcg_last_error": (ctypes.c_char_p, []),
        "npbcg_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "npbcg_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
        "npbcg_makea": (ctypes.c_long, [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                        _I32P, _I32P, _F64P, _DP]),
        "npbcg_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, _I32P, _I32P,
                                        _F64P, ctypes.POINTER(Opts)]),
        "npbcg_destroy": (None, [vp]),
        "npbcg_conj_grad": (ctypes.c_int, [vp, _I32P, _I32P, _F64P, _F64P, _F64P, _F64P,
                                           _F64P, _F64P, _DP]),
        "npbcg_conj_grad_device": (ctypes.c_int, [vp, _DP]),
        "npbcg_norms": (ctypes.c_int, [vp, _DP, _DP]),
        "npbcg_scale_x": (ctypes.c_int, [vp, ctypes.c_double]),
        "npbcg_power": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, _F64P, _F64P]),
        "npbcg_set_vector": (ctypes.c_int, [vp, ctypes.c_int, _F64P]),
        "npbcg_get_vector": (ctypes.c_int, [vp, ctypes.c_int, _F64P]),
        "npbcg_fill_x": (ctypes.c_int, [vp, ctypes.c_double]),
        "npbcg_set_spmv_sampling": (ctypes.c_int, [vp, ctypes.c_int]),
        "npbcg_get_timing": (ctypes.c_int, [vp, ctypes.POINTER(Timing)]),
        "npbcg_profile_kernels": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, _F64P,
                                                 ctypes.POINTER(KernelTime), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int), _DP]),
        "npbcg_spmv_bench": (ctypes.c_int, [vp, ctypes.c_int, _DP]),
        "npbcg_spmv_kernel_bench": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _DP]),
        "npbcg_tune": (ctypes.c_int, [ctypes.c_int, _I32P, _I32P, _F64P, ctypes.c_void_p,
                                      ctypes.c_char_p, _DP]),
        "npbcg_describe": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
        "npbcg_copy_bench": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_int, _DP]),
        "npbcg_stream_bench": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _DP]),
        "npbcg_set_rhs": (ctypes.c_int, [vp, _F64P]),
        "npbcg_solve": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(SolveStats)]),
        "npbcg_get_solution": (ctypes.c_int, [vp, _F64P]),
        "npbcg_set_preconditioner": (ctypes.c_int, [vp, ctypes.c_int]),
        "npbcg_get_preconditioner": (ctypes.c_int, [vp, _F64P]),
        "npbcg_apply_preconditioner": (ctypes.c_int, [vp, _F64P, _F64P]),
        "npbcg_log_pairs": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
        "npbcg_set_pairs": (ctypes.c_int, [vp, ctypes.c_int, _F32P, _F32P]),
        "npbcg_graph": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "npbcg_net_param_count": (ctypes.c_int, []),
        "npbcg_net_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int,
                                            ctypes.c_ulonglong]),
        "npbcg_net_destroy": (None, [vp]),
        "npbcg_net_get_params": (ctypes.c_int, [vp, vp, vp]),
        "npbcg_net_set_params": (ctypes.c_int, [vp, vp, vp]),
        "npbcg_net_default_train_opts": (None, [ctypes.POINTER(TrainOpts)]),
        "npbcg_net_train": (ctypes.c_int, [vp, ctypes.POINTER(TrainOpts),
                                           ctypes.POINTER(TrainStats), vp, vp]),
        "npbcg_net_loss_grad": (ctypes.c_int, [vp, vp, _DP, vp, vp]),
        "npbcg_net_step": (ctypes.c_int, [vp, vp, ctypes.c_float, _DP]),
        "npbcg_net_apply": (ctypes.c_int, [vp, vp, _DP]),
        "npbcg_set_generator_threads": (None, [ctypes.c_int]),
        "npbcg_batch_default_opts": (None, [ctypes.POINTER(BatchOpts)]),
        "npbcg_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, vp, vp, vp, vp,
                                              ctypes.POINTER(BatchOpts)]),
        "npbcg_batch_destroy": (None, [vp]),
        "npbcg_batch_set_rhs": (ctypes.c_int, [vp, _F64P]),
        "npbcg_batch_solve": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(BatchStats), vp, vp]),
        "npbcg_batch_get_solution": (ctypes.c_int, [vp, _F64P]),
        "npbcg_batch_spmm_bench": (ctypes.c_int, [vp, ctypes.c_int, _DP]),
        "npbcg_batch_tune": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int), _DP]),
        "npbcg_compat_set_n": (ctypes.c_int, [ctypes.c_int]),
        "npbcg_compat_conj_grad": (None, [_I32P, _I32P, _F64P, _F64P, _F64P, _F64P,
                                          _F64P, _F64P, _DP]),
        "npbcg_compat_release": (None, []),
        "npbcg_conj_grad_device_noalloc": (ctypes.c_int, [ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                                          vp, vp, ctypes.c_int, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def header_functions() -> list[str]:
    """Names of the functions include/npbcg.h declares."""
    with open(HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(npbcg_[a-z0-9_]+)\s*\(", text)))


def last_error() -> str:
    return (LIB.npbcg_last_error() or b"").decode()


def _check(code: int, where: str) -> int:
    if code < 0:
        raise NpbcgError(code, where, last_error() or LIB.npbcg_status_string(code).decode())
    return code


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(LIB.npbcg_device_count(ctypes.byref(c)), "npbcg_device_count")
    return c.value


def default_opts(**kw) -> Opts:
    o = Opts()
    LIB.npbcg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Matrix:
    """An NPB CG matrix in 0-based CSR (int32 indices, float64 values)."""

    def __init__(self, n, rowstr, colidx, a, ratio=None, nonzer=None, shift=None,
                 rcond=None, seed=None):
        self.n, self.rowstr, self.colidx, self.a = n, rowstr, colidx, a
        self.ratio, self.nonzer, self.shift, self.rcond, self.seed = (
            ratio, nonzer, shift, rcond, seed)

    @property
    def nnz(self) -> int:
        return int(self.rowstr[self.n])


def makea(n: int, nonzer: int, shift: float, seed: float = 314159265.0,
          rcond: float = 0.1, nz_cap: Optional[int] = None) -> Matrix:
    """The NPB CG generator (bit-exact with the reference's makea)."""
    cap = n * (nonzer + 1) * (nonzer + 1) if nz_cap is None else nz_cap
    rowstr = np.zeros(n + 1, dtype=np.int32)
    colidx = np.zeros(max(cap, 1), dtype=np.int32)
    a = np.zeros(max(cap, 1), dtype=np.float64)
    ratio = ctypes.c_double(0.0)
    nnz = _check(LIB.npbcg_makea(n, nonzer, rcond, shift, float(seed), cap, rowstr,
                                 colidx, a, ctypes.byref(ratio)), "npbcg_makea")
    return Matrix(n, rowstr, colidx[:nnz].copy(), a[:nnz].copy(), ratio.value, nonzer,
                  shift, rcond, seed)


def tune(m: Matrix, cache_path: Optional[str] = None, **opts):
    """Single-system autotuner (npbcg_tune): returns ({spmv_variant,
    tile_width, tile_rows}, SpMV ms) for Context(m, **opts, **shape)."""
    o = default_opts(**opts)
    ms = ctypes.c_double(0.0)
    _check(LIB.npbcg_tune(m.n, m.rowstr, m.colidx, m.a, ctypes.byref(o),
                          None if cache_path is None else cache_path.encode(), ctypes.byref(ms)),
           "npbcg_tune")
    return ({"spmv_variant": o.spmv_variant, "tile_width": o.tile_width,
             "tile_rows": o.tile_rows}, ms.value)


class Context:
    """A device-resident problem (npbcg_create .. npbcg_destroy)."""

    def __init__(self, m: Matrix, **opts):
        self.m = m
        self.n = m.n
        self._h = ctypes.c_void_p()
        o = default_opts(**opts)
        _check(LIB.npbcg_create(ctypes.byref(self._h), m.n, m.rowstr, m.colidx, m.a,
                                ctypes.byref(o)), "npbcg_create")

    def close(self):
        if self._h:
            LIB.npbcg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def conj_grad(self, x: np.ndarray, colidx=None, rowstr=None, a=None):
        """Drop-in conj_grad on host arrays; returns (z, p, q, r, rnorm)."""
        n = self.n
        x = np.ascontiguousarray(x[: n + 1], dtype=np.float64)
        z, p, q, r = (np.zeros(n + 1) for _ in range(4))
        rn = ctypes.c_double(0.0)
        _check(LIB.npbcg_conj_grad(self._h,
                                   self.m.colidx if colidx is None else colidx,
                                   self.m.rowstr if rowstr is None else rowstr,
                                   x, z, self.m.a if a is None else a, p, q, r,
                                   ctypes.byref(rn)), "npbcg_conj_grad")
        return z, p, q, r, rn.value

    def conj_grad_device(self) -> float:
        rn = ctypes.c_double(0.0)
        _check(LIB.npbcg_conj_grad_device(self._h, ctypes.byref(rn)), "npbcg_conj_grad_device")
        return rn.value

    def norms(self):
        t1, t2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(LIB.npbcg_norms(self._h, ctypes.byref(t1), ctypes.byref(t2)), "npbcg_norms")
        return t1.value, t2.value

    def scale_x(self, s: float):
        _check(LIB.npbcg_scale_x(self._h, s), "npbcg_scale_x")

    def power(self, niter: int, shift: float):
        zeta = np.zeros(max(niter, 1))
        rnorm = np.zeros(max(niter, 1))
        _check(LIB.npbcg_power(self._h, niter, shift, zeta, rnorm), "npbcg_power")
        return zeta[:niter], rnorm[:niter]

    def profile_kernels(self, niter: int, shift: float, max_kinds: int = 32):
        """Per-kernel device times of niter outer iterations (npbcg_profile_kernels).

        Returns (kernels, total_ms, zeta): kernels is a list of dicts
        {name, threads, calls, ms} in order of first launch."""
        zeta = np.zeros(niter)
        kt = (KernelTime * max_kinds)()
        nk = ctypes.c_int(0)
        total = ctypes.c_double(0.0)
        _check(LIB.npbcg_profile_kernels(self._h, niter, shift, zeta, kt, max_kinds,
                                         ctypes.byref(nk), ctypes.byref(total)),
               "npbcg_profile_kernels")
        kernels = [{"name": kt[i].name.decode(), "threads": kt[i].threads,
                    "calls": kt[i].calls, "ms": kt[i].ms} for i in range(nk.value)]
        return kernels, total.value, zeta

    def set_vector(self, which: int, v: np.ndarray):
        v = np.ascontiguousarray(v[: self.n + 1], dtype=np.float64)
        _check(LIB.npbcg_set_vector(self._h, which, v), "npbcg_set_vector")

    def get_vector(self, which: int) -> np.ndarray:
        out = np.zeros(self.n + 1)
        _check(LIB.npbcg_get_vector(self._h, which, out), "npbcg_get_vector")
        return out

    def fill_x(self, value: float = 1.0):
        _check(LIB.npbcg_fill_x(self._h, value), "npbcg_fill_x")

    def set_spmv_sampling(self, every: int):
        _check(LIB.npbcg_set_spmv_sampling(self._h, every), "npbcg_set_spmv_sampling")

    def timing(self) -> Timing:
        t = Timing()
        _check(LIB.npbcg_get_timing(self._h, ctypes.byref(t)), "npbcg_get_timing")
        return t

    def describe(self) -> dict:
        import json

        buf = ctypes.create_string_buffer(1024)
        _check(LIB.npbcg_describe(self._h, buf, 1024), "npbcg_describe")
        return json.loads(buf.value.decode())

    def set_rhs(self, b: np.ndarray):
        _check(LIB.npbcg_set_rhs(self._h, np.ascontiguousarray(b[: self.n], dtype=np.float64)),
               "npbcg_set_rhs")

    def solve(self, mode: int = SOLVE_FP64, rtol: float = 1e-10, maxit: int = 1000,
              check: bool = True):
        """Solves A x = b (resident b); returns SolveStats."""
        st = SolveStats()
        rc = LIB.npbcg_solve(self._h, mode, rtol, maxit, ctypes.byref(st))
        if check:
            _check(rc, "npbcg_solve")
        return st

    def solution(self) -> np.ndarray:
        x = np.zeros(self.n)
        _check(LIB.npbcg_get_solution(self._h, x), "npbcg_get_solution")
        return x

    def set_preconditioner(self, kind: int):
        _check(LIB.npbcg_set_preconditioner(self._h, kind), "npbcg_set_preconditioner")

    def preconditioner(self) -> np.ndarray:
        """The diagonal of the active M^-1 (ones for plain CG)."""
        m = np.zeros(self.n)
        _check(LIB.npbcg_get_preconditioner(self._h, m), "npbcg_get_preconditioner")
        return m

    def apply_preconditioner(self, v: np.ndarray) -> np.ndarray:
        z = np.zeros(self.n)
        _check(LIB.npbcg_apply_preconditioner(
            self._h, np.ascontiguousarray(v, dtype=np.float64), z),
            "npbcg_apply_preconditioner")
        return z

    def log_pairs(self, pairs: int = MAX_PAIRS):
        """CG residual/correction pairs of the resident right-hand side."""
        r = np.zeros((pairs, self.n), np.float32)
        e = np.zeros((pairs, self.n), np.float32)
        _check(LIB.npbcg_log_pairs(self._h, pairs, r.ctypes.data, e.ctypes.data),
               "npbcg_log_pairs")
        return r, e

    def set_pairs(self, r: np.ndarray, e: np.ndarray):
        r = np.ascontiguousarray(r, dtype=np.float32)
        e = np.ascontiguousarray(e, dtype=np.float32)
        _check(LIB.npbcg_set_pairs(self._h, r.shape[0], r, e), "npbcg_set_pairs")

    def graph(self):
        """(nbr, wgt, nval, feat) of the top-K row graph."""
        k = 4
        nbr = np.zeros((self.n, k), np.int32)
        wgt = np.zeros((self.n, k), np.float32)
        nval = np.zeros((self.n, k), np.float32)
        feat = np.zeros((self.n, NET_FEATURES), np.float32)
        kk = _check(LIB.npbcg_graph(self._h, nbr.ctypes.data, wgt.ctypes.data,
                                    nval.ctypes.data, feat.ctypes.data), "npbcg_graph")
        assert kk == k
        return nbr, wgt, nval, feat

    def copy_bench(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of a device copy of nbytes (read nbytes + write nbytes)."""
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_copy_bench(self._h, nbytes, reps, ctypes.byref(ms)),
               "npbcg_copy_bench")
        return ms.value

    def stream_bench(self, mode: int, nbytes: int, reps: int = 5) -> float:
        """Mean ms of a device copy (mode 0) or read-only stream (mode 1) of
        nbytes (npbcg_stream_bench)."""
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_stream_bench(self._h, mode, nbytes, reps, ctypes.byref(ms)),
               "npbcg_stream_bench")
        return ms.value

    def spmv_bench(self, reps: int = 20) -> float:
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_spmv_bench(self._h, reps, ctypes.byref(ms)), "npbcg_spmv_bench")
        return ms.value

    def spmv_kernel_bench(self, precision: int = 0, reps: int = 20) -> float:
        """Mean event-timed duration (ms) of the CG SpMV kernel alone,
        launched as successive CG steps launch it (npbcg_spmv_kernel_bench);
        precision 0 = FP64, 1 = the mixed solve's FP32 kernel."""
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_spmv_kernel_bench(self._h, precision, reps, ctypes.byref(ms)),
               "npbcg_spmv_kernel_bench")
        return ms.value


def default_train_opts() -> TrainOpts:
    o = TrainOpts()
    LIB.npbcg_net_default_train_opts(ctypes.byref(o))
    return o


class Net:
    """The learned-preconditioner network (npbcg_net_*)."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._h = ctypes.c_void_p()
        _check(LIB.npbcg_net_create(ctypes.byref(self._h), device, seed), "npbcg_net_create")
        self.param_count = LIB.npbcg_net_param_count()

    def close(self):
        if self._h:
            LIB.npbcg_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.param_count, np.float32)
        nrm = np.zeros(2 * NET_FEATURES, np.float32)
        _check(LIB.npbcg_net_get_params(self._h, p.ctypes.data, nrm.ctypes.data),
               "npbcg_net_get_params")
        return p, nrm

    def set_params(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        nrm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        _check(LIB.npbcg_net_set_params(self._h, None if p is None else p.ctypes.data,
                                        None if nrm is None else nrm.ctypes.data),
               "npbcg_net_set_params")

    def save(self, path: str):
        p, nrm = self.get_params()
        np.savez(path, params=p, norm=nrm)

    def load(self, path: str):
        with np.load(path, allow_pickle=False) as f:
            self.set_params(f["params"], f["norm"])

    def train(self, opts: Optional[TrainOpts] = None, history: bool = True):
        o = opts if opts is not None else default_train_opts()
        st = TrainStats()
        lh = np.zeros(o.systems) if history else None
        bh = np.zeros(o.systems) if history else None
        _check(LIB.npbcg_net_train(self._h, ctypes.byref(o), ctypes.byref(st),
                                   None if lh is None else lh.ctypes.data,
                                   None if bh is None else bh.ctypes.data), "npbcg_net_train")
        return st, lh, bh

    def loss_grad(self, ctx: "Context", n: int):
        loss = ctypes.c_double(0.0)
        grad = np.zeros(self.param_count, np.float32)
        gh = np.zeros(2 * n, np.float32)
        _check(LIB.npbcg_net_loss_grad(self._h, ctx._h, ctypes.byref(loss), grad.ctypes.data,
                                       gh.ctypes.data), "npbcg_net_loss_grad")
        return loss.value, grad, gh

    def step(self, ctx: "Context", lr: float = 1e-3) -> float:
        loss = ctypes.c_double(0.0)
        _check(LIB.npbcg_net_step(self._h, ctx._h, lr, ctypes.byref(loss)), "npbcg_net_step")
        return loss.value

    def apply(self, ctx: "Context") -> float:
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_net_apply(self._h, ctx._h, ctypes.byref(ms)), "npbcg_net_apply")
        return ms.value


def set_generator_threads(threads: int):
    """Generator threads for npbcg_makea calls made from the calling thread."""
    LIB.npbcg_set_generator_threads(threads)


def make_many(params, workers: int = 0):
    """makea for a list of (n, nonzer, shift, seed[, rcond]) tuples, one system
    per host thread (each single-threaded)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(p):
        set_generator_threads(1)
        return makea(*p)

    workers = workers or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, params))


class Batch:
    """Independent systems solved together (npbcg_batch_*)."""

    def __init__(self, mats, nrhs: int = 1, mode: int = SOLVE_MIXED, device: int = 0,
                 lanes_per_row: int = 0, rows_per_tile: int = 0, inner_steps: int = 0,
                 layout: int = 0):
        self.mats = list(mats)
        self.nrhs = nrhs
        self.n = np.array([m.n for m in self.mats], np.int32)
        o = BatchOpts()
        LIB.npbcg_batch_default_opts(ctypes.byref(o))
        o.device, o.nrhs, o.mode = device, nrhs, mode
        o.lanes_per_row, o.rows_per_tile, o.inner_steps = lanes_per_row, rows_per_tile, inner_steps
        o.layout = layout
        k = len(self.mats)
        rs = (ctypes.c_void_p * k)(*[m.rowstr.ctypes.data for m in self.mats])
        ci = (ctypes.c_void_p * k)(*[m.colidx.ctypes.data for m in self.mats])
        av = (ctypes.c_void_p * k)(*[m.a.ctypes.data for m in self.mats])
        self._h = ctypes.c_void_p()
        _check(LIB.npbcg_batch_create(ctypes.byref(self._h), k, self.n.ctypes.data,
                                      ctypes.cast(rs, ctypes.c_void_p),
                                      ctypes.cast(ci, ctypes.c_void_p),
                                      ctypes.cast(av, ctypes.c_void_p), ctypes.byref(o)),
               "npbcg_batch_create")
        self.rows = int(self.n.sum())

    def close(self):
        if self._h:
            LIB.npbcg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, rhs):
        """rhs: list (one per system) of [n_s, nrhs] arrays."""
        b = np.concatenate([np.asarray(r, np.float64).reshape(-1, self.nrhs) for r in rhs])
        _check(LIB.npbcg_batch_set_rhs(self._h, np.ascontiguousarray(b).reshape(-1)),
               "npbcg_batch_set_rhs")

    def tune(self, cache_path=None):
        lpr, rt = ctypes.c_int(0), ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_batch_tune(self._h, None if cache_path is None else cache_path.encode(),
                                    ctypes.byref(lpr), ctypes.byref(rt), ctypes.byref(ms)),
               "npbcg_batch_tune")
        return lpr.value, rt.value, ms.value

    def solve(self, rtol: float = 1e-10, maxit: int = 2000, check: bool = True):
        st = BatchStats()
        ns = len(self.mats) * self.nrhs
        steps = np.zeros(ns, np.int32)
        rel = np.zeros(ns)
        rc = LIB.npbcg_batch_solve(self._h, rtol, maxit, ctypes.byref(st), steps.ctypes.data,
                                   rel.ctypes.data)
        if check:
            _check(rc, "npbcg_batch_solve")
        return st, steps.reshape(-1, self.nrhs), rel.reshape(-1, self.nrhs)

    def spmm_bench(self, reps: int = 20) -> float:
        ms = ctypes.c_double(0.0)
        _check(LIB.npbcg_batch_spmm_bench(self._h, reps, ctypes.byref(ms)), "npbcg_batch_spmm_bench")
        return ms.value

    def solution(self):
        x = np.zeros(self.rows * self.nrhs)
        _check(LIB.npbcg_batch_get_solution(self._h, x), "npbcg_batch_get_solution")
        x = x.reshape(-1, self.nrhs)
        out, off = [], 0
        for n in self.n:
            out.append(x[off:off + n])
            off += n
        return out
