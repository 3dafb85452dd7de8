"""Host checks of the column-blocked SELL layout builder (CPU, no GPU).

tests/c/sell_layout_check.cpp is compiled with g++ against the product's
npb-gpu_amd/csrc/tiled_layout.cpp and run on generated ragged CSR matrices
(empty rows, n not a multiple of 64, block widths and row ranges that do not
divide n).  For each case it checks that every row has one slot per column
block, that the slices hold each row's block segment in CSR order with
(0, 0) padding, that the wavefront placement only permutes a workgroup's
slices and keeps each wavefront's slice count under the kernels' round order,
and that y = A x taken through the layout (block sums in entry order, blocks
in order) equals the same sums from the CSR bit for bit.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "npb-gpu_amd", "csrc")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("sell") / "sell_layout_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", CSRC, "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "sell_layout_check.cpp"),
                    os.path.join(CSRC, "tiled_layout.cpp"), "-o", str(exe)], check=True)
    return exe


def _run(checker, n, lmax, W, R, window, waves, seed):
    out = subprocess.run([str(checker), str(n), str(lmax), str(W), str(R), str(window),
                          str(waves), str(seed)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    assert f[-1] == "ok"
    return {"stored": int(f[7]), "nnz": int(f[9]), "worst": int(f[13]), "worst_sorted": int(f[15])}


@pytest.mark.parametrize("n,lmax,W,R", [(1, 3, 2, 1), (63, 4, 10, 7), (500, 5, 600, 100),
                                        (3000, 40, 700, 1000), (4999, 90, 1250, 4999)])
@pytest.mark.parametrize("window", [64, 2048, 8192])
@pytest.mark.parametrize("waves", [0, 8, 16])
def test_layout_invariants(checker, n, lmax, W, R, window, waves):
    _run(checker, n, lmax, W, R, window, waves, seed=n * 31 + window + waves)


def test_wider_windows_pad_less(checker):
    # Nested windows (64 within 2048 within 8192, all inside one 20000-row
    # range): the descending cut of a window is the least padding of any cut
    # into 64-row slices, so stored entries cannot grow with the window.
    st = [_run(checker, 20000, 30, 5000, 20000, w, 16, seed=7)["stored"]
          for w in (64, 2048, 8192)]
    nnz = _run(checker, 20000, 30, 5000, 20000, 64, 0, seed=7)["nnz"]
    assert st[0] >= st[1] >= st[2] >= nnz, st
    assert st[0] > st[2]


def test_placement_evens_wavefront_widths(checker):
    # One fixed matrix (ranges of 5000 rows, 79 slices for 8 wavefronts): the
    # busiest wavefront's total width under the round order drops with the
    # placement (sum over workgroups; deterministic for this seed).
    r = _run(checker, 5000, 60, 1250, 5000, 4096, 8, seed=3)
    assert r["worst"] < r["worst_sorted"], r
