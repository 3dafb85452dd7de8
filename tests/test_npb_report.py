"""The NPB-style driver: class table, operation count, argument rules (CPU),
and a class S run with its verification line (GPU)."""
import os
import re
import subprocess
import sys

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "npb-gpu_amd"))


def _mod():
    from npbcg import npb_cg
    return npb_cg


def test_class_table_matches_oracle(npbcg_lib):
    m = _mod()
    assert m.CLASSES == oracle.NPB_CLASSES
    assert m.NPB_SEED == oracle.NPB_SEED and m.EPSILON == oracle.NPB_EPSILON


def test_mops_is_the_reference_operation_count(npbcg_lib):
    m = _mod()
    # class B: 2*75*75000*(3 + 182 + 25*(5 + 182) + 3) operations in 1 s
    assert m.mops(75000, 13, 75, 1.0) == 2 * 75 * 75000 * (3 + 182 + 25 * 187 + 3) / 1e6
    assert m.mops(75000, 13, 75, 0.0) == 0.0


def test_arguments(npbcg_lib):
    m = _mod()
    with pytest.raises(SystemExit):
        m.main(["--class", "S", "--n", "100"])
    with pytest.raises(SystemExit):
        m.main(["--n", "100"])


def test_kernel_table_layout(npbcg_lib):
    # the columns of the reference's PROFILING table (cg.cu:566-568), plus launches
    m = _mod()
    text = m.kernel_table([{"name": "spmv_sell_kernel (s = A r)", "threads": 1024,
                            "calls": 375, "ms": 3.0},
                           {"name": "update_cg_kernel", "threads": 256, "calls": 375,
                            "ms": 1.0}], 5.0)
    head, row1, row2 = text.splitlines()
    assert head.split("\t")[:4] == ["GPU Kernel", " " * 8 + "Threads Per Block",
                                     " " * 10 + "Time in Seconds", " " * 7 + "Time in Percentage"]
    f = row1.split("\t")
    assert f[0].strip() == "spmv_sell_kernel (s = A r)" and len(f[0]) == 29
    assert int(f[1]) == 1024 and float(f[2]) == 0.003 and f[3].strip() == "60.00%"
    assert int(f[4]) == 375
    assert row2.split("\t")[3].strip() == "20.00%"


@pytest.mark.gpu
def test_class_s_report_verifies(gpu):
    out = subprocess.run([sys.executable, os.path.join(ROOT, "npb-gpu_amd", "npbcg", "npb_cg.py"),
                          "--class", "S", "--json", "--profile"], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    lines = out.stdout.splitlines()
    assert " VERIFICATION SUCCESSFUL" in lines
    assert " Verification    =               SUCCESSFUL" in lines
    it_lines = [l for l in lines if re.fullmatch(r" {4,}\d+ +\S+e[-+]\d+ *\S+e[-+]\d+", l)]
    assert len(it_lines) == 15 and it_lines[-1].split()[0] == "15"
    assert any("zeta identical to the timed run" in l for l in lines)
    assert any(l.strip().startswith("spmv_") for l in lines)
