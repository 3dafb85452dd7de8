"""The C-ABI library on a host without a GPU: it loads, exports every function
include/npbcg.h declares, and its error paths return status codes (the
reference exits the process instead).  No kernel is launched here."""
import ctypes
import subprocess

import numpy as np
import pytest


def test_exports_every_declared_function(npbcg_lib):
    declared = npbcg_lib.header_functions()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", npbcg_lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    for f in declared:
        getattr(npbcg_lib.LIB, f)


def test_abi_version_and_status_strings(npbcg_lib):
    assert npbcg_lib.LIB.npbcg_abi_version() == 3
    for code in range(0, -9, -1):
        s = npbcg_lib.LIB.npbcg_status_string(code).decode()
        assert s and s != "unknown status"
    assert npbcg_lib.LIB.npbcg_status_string(-99).decode() == "unknown status"


def test_default_opts(npbcg_lib):
    o = npbcg_lib.default_opts()
    assert o.cgitmax == 25 and o.use_graph == 1 and o.engine == 0


def test_create_rejects_malformed_csr(npbcg_lib):
    n = 4
    h = ctypes.c_void_p()
    o = npbcg_lib.default_opts()
    good_rs = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    ci = np.array([0, 1, 2, 3], dtype=np.int32)
    a = np.ones(4)
    for rs, cc in [(np.array([1, 1, 2, 3, 4], dtype=np.int32), ci),
                   (np.array([0, 2, 1, 3, 4], dtype=np.int32), ci),
                   (good_rs, np.array([0, 1, 9, 3], dtype=np.int32)),
                   (good_rs, np.array([0, -1, 2, 3], dtype=np.int32))]:
        rc = npbcg_lib.LIB.npbcg_create(ctypes.byref(h), n, rs, cc, a, ctypes.byref(o))
        assert rc == npbcg_lib.ERR_INVALID
        assert npbcg_lib.last_error()
    rc = npbcg_lib.LIB.npbcg_create(ctypes.byref(h), 0, good_rs, ci, a, ctypes.byref(o))
    assert rc == npbcg_lib.ERR_INVALID


def test_create_without_gpu_reports_nodevice(npbcg_lib):
    if npbcg_lib.device_count() > 0:
        pytest.skip("a GPU is visible; the no-device path is exercised on GPU-less hosts")
    m = npbcg_lib.makea(100, 3, 2.0, 5.0)
    with pytest.raises(npbcg_lib.NpbcgError) as e:
        npbcg_lib.Context(m)
    assert e.value.code == npbcg_lib.ERR_NODEVICE


def test_compat_shim_without_n_returns_nan(npbcg_lib):
    n = 4
    rs = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    ci = np.array([0, 1, 2, 3], dtype=np.int32)
    a = np.ones(4)
    vecs = [np.zeros(n + 1) for _ in range(5)]
    rn = ctypes.c_double(0.0)
    npbcg_lib.LIB.npbcg_compat_conj_grad(ci, rs, vecs[0], vecs[1], a, vecs[2], vecs[3],
                                         vecs[4], ctypes.byref(rn))
    assert np.isnan(rn.value)
    assert npbcg_lib.LIB.npbcg_compat_set_n(0) == npbcg_lib.ERR_INVALID


def test_null_context_calls_are_errors(npbcg_lib):
    L = npbcg_lib.LIB
    d = ctypes.c_double()
    assert L.npbcg_conj_grad_device(None, ctypes.byref(d)) == npbcg_lib.ERR_INVALID
    assert L.npbcg_norms(None, ctypes.byref(d), ctypes.byref(d)) == npbcg_lib.ERR_INVALID
    assert L.npbcg_scale_x(None, 1.0) == npbcg_lib.ERR_INVALID
    assert L.npbcg_fill_x(None, 1.0) == npbcg_lib.ERR_INVALID
    kt = (npbcg_lib.KernelTime * 4)()
    nk = ctypes.c_int(0)
    assert L.npbcg_profile_kernels(None, 1, 10.0, np.zeros(1), kt, 4, ctypes.byref(nk),
                                   ctypes.byref(d)) == npbcg_lib.ERR_INVALID
    L.npbcg_destroy(None)


def test_python_structs_match_the_header(npbcg_lib, tmp_path):
    # Every struct the binding passes by pointer has the header's size and
    # field offsets (compiled with the host C compiler against include/npbcg.h).
    import os
    pairs = [("npbcg_opts", npbcg_lib.Opts), ("npbcg_timing", npbcg_lib.Timing),
             ("npbcg_kernel_time", npbcg_lib.KernelTime),
             ("npbcg_solve_stats", npbcg_lib.SolveStats),
             ("npbcg_train_opts", npbcg_lib.TrainOpts), ("npbcg_train_stats", npbcg_lib.TrainStats),
             ("npbcg_batch_opts", npbcg_lib.BatchOpts), ("npbcg_batch_stats", npbcg_lib.BatchStats)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "npbcg.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line.strip():
            name, field, val = line.split()
            got[(name, field)] = int(val)
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)

