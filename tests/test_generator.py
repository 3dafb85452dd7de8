"""The product's matrix generator (npbcg_makea, host code in libnpbcg.so).

It must reproduce the reference's assembled CSR bit for bit: against the
oracle on a grid of small problems (edge cases included) and against the
reference's own digests (golden fixtures) at every configured size.
"""
import hashlib

import numpy as np
import pytest

import oracle


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _same(m, o):
    return (np.array_equal(m.rowstr, o.rowstr) and np.array_equal(m.colidx, o.colidx)
            and np.array_equal(m.a.view(np.int64), o.a.view(np.int64))
            and float(m.ratio).hex() == float(o.ratio).hex())


@pytest.mark.parametrize("n,nonzer,shift,seed", [
    (1, 0, 10.0, 3.0),          # one row: only the forced diagonal
    (1, 1, 10.0, 7.0),          # nonzer = n: the one position drawn
    (2, 1, 1.0, 5.0),
    (3, 3, 10.0, 1003.0),       # nonzer = n: every position drawn
    (17, 16, 2.0, 11.0),        # nonzer = n - 1
    (65, 5, 10.0, 1065.0),
    (100, 0, 1.5, 12.0),        # diagonal-only matrix
    (1000, 3, 4.0, 271828.0),   # even seed
    (1400, 7, 10.0, 12345.0),
    (2000, 13, 20.0, 1.0),
    (4097, 9, 8.0, 314159265.0),  # n just above a power of two (nn1 = 8192)
    (7000, 8, 12.0, 314159265.0),
])
def test_generator_matches_oracle(npbcg_lib, n, nonzer, shift, seed):
    m = npbcg_lib.makea(n, nonzer, shift, seed)
    o = oracle.makea(n, nonzer, shift, seed)
    assert _same(m, o)


@pytest.mark.parametrize("name", ["S_npb", "cfg1_n1400_seed12345", "W_npb", "A_npb",
                                  "B_npb", "cfg2_n75000_seed271828", "C_npb",
                                  "cfg3_n150000_seed314"])
def test_generator_matches_reference_digests(npbcg_lib, golden, name):
    c = golden[name]
    m = npbcg_lib.makea(c["n"], c["nonzer"], c["shift"], c["seed"], c["rcond"])
    assert m.nnz == c["nnz"]
    assert _sha(m.rowstr) == c["sha256_rowstr"]
    assert _sha(m.colidx) == c["sha256_colidx"]
    assert _sha(m.a) == c["sha256_a"]
    assert float(m.ratio).hex() == c["ratio_hex"]
    for j, row in c["rows_head"].items():
        j = int(j)
        s = int(m.rowstr[j])
        assert int(m.rowstr[j + 1]) - s == row["len"]
        assert [int(v) for v in m.colidx[s:s + 8]] == row["colidx"]
        assert [float(v).hex() for v in m.a[s:s + 8]] == row["a"]


def test_generator_thread_count_does_not_change_bits(npbcg_lib, monkeypatch):
    ref = npbcg_lib.makea(9000, 11, 20.0, 777.0)
    for t in ("1", "3", "7"):
        monkeypatch.setenv("NPBCG_GEN_THREADS", t)
        m = npbcg_lib.makea(9000, 11, 20.0, 777.0)
        assert _same(m, ref)


def test_generator_errors(npbcg_lib):
    with pytest.raises(npbcg_lib.NpbcgError) as e:
        npbcg_lib.makea(200, 5, 3.0, 5.0, nz_cap=10)
    assert e.value.code == npbcg_lib.ERR_CAPACITY
    for bad in [dict(n=0, nonzer=1), dict(n=10, nonzer=11), dict(n=10, nonzer=-1)]:
        with pytest.raises(npbcg_lib.NpbcgError) as e:
            npbcg_lib.makea(bad["n"], bad["nonzer"], 1.0, 5.0)
        assert e.value.code == npbcg_lib.ERR_INVALID
    for seed in (0.0, 2.0 ** 46, 1.5):
        with pytest.raises(npbcg_lib.NpbcgError) as e:
            npbcg_lib.makea(10, 2, 1.0, seed)
        assert e.value.code == npbcg_lib.ERR_INVALID
