"""The host side of the drop-in surface through the C ABI, no GPU needed:
the partition rule against the oracle's (cc:139-158), the info strings of
test17:128-152 / test18:342-355, print_matlab_script (cc:2493-2534) read back
into a dense matrix, debug-mode NaN rejection (cc:805, 2538-2659), the
documented block-size limit, dup and t_solve."""
import re

import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import AlglinError, BorderedCR


@pytest.mark.parametrize("N", [1, 2, 3, 5, 6, 7, 16, 33, 100])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 7, 16])
def test_partition_rule_matches_oracle(oracle_mod, N, P):
    o = oracle_mod.OracleBorderedCR(np.float64, P, P)
    o.allocate(N, 2, 0, 0, 0, 0)
    g = BorderedCR()
    g.allocate(N, 2, 0, 0, 0, 0)
    g.set_num_parallel_block(P)
    assert g.used_parallel_block() == o.used_parallel_block()
    assert f"nblk:{o.used_parallel_block()}" in g.info()


def test_info_strings():
    g = BorderedCR()
    g.allocate(16, 2, 0, 0, 0, 0)
    assert "nblk:1" in g.info()
    g.set_num_parallel_block(4)
    assert "nblk:4" in g.info()
    g.set_num_parallel_block(2)
    assert "nblk:2" in g.info()
    g.select("LU", "LU", "PINV")
    assert "last   = LastBlock LU" in g.info() and "last2  = LastBlock PINV" in g.info()
    h = BorderedCR()
    h.allocate(6, 2, 1, 2, 3, 2)
    h.select("QRP", "PINV", "NONE")
    h.set_num_parallel_block(4)
    assert "nblk:3" in h.info() and "CyclicReduction+QRP" in h.info() and "LastBlock PINV" in h.info()
    for bad in (("BAD", "LU", "NONE"), ("LU", "BAD", "NONE"), ("LU", "LU", "BAD")):
        with pytest.raises(AlglinError, match="unknown/unsupported"):
            h.select(*bad)


def test_matlab_script_reads_back_as_the_matrix():
    s = systems.stable_system()
    g = BorderedCR()
    g.load_system(s)
    text = g.matlab_script()
    vals = {}
    for name in ("I", "J", "V"):
        m = re.search(name + r" = \[ (.*?) \];", text, re.S)
        vals[name] = np.array([float(t) for t in re.split(r",\s*(?:\.\.\.)?\s*", m.group(1).strip()) if t])
    nr = int(re.search(r"nr  = (\d+);", text).group(1))
    nc = int(re.search(r"nc  = (\d+);", text).group(1))
    nnz = int(re.search(r"nnz = (\d+);", text).group(1))
    assert "MAT = sparse( I, J, V, nr, nc );" in text
    A = np.zeros((nr, nc))
    np.add.at(A, (vals["I"].astype(int) - 1, vals["J"].astype(int) - 1), vals["V"])
    assert nnz == len(vals["V"]) and np.array_equal(A, systems.dense(s))
    # 20 entries per line
    first = text.split("\n")[0]
    assert first.startswith("I = [ 1") and first.count(",") == 20


def test_debug_mode_rejects_nan_on_the_host():
    s = systems.gaussian_system(10, 4, 1, 1, 1, 1, seed=2)
    s["D"][3, 1, 2] = np.nan
    g = BorderedCR()
    g.load_system(s)
    g.set_debug(True)
    assert not g.factorize()
    assert "check_matrix" in g.last_error() and "matrix D" in g.last_error()
    with pytest.raises(AlglinError, match="check_matrix"):
        g.check_matrix()


def test_block_size_limit_is_reported():
    g = BorderedCR()
    g.allocate(4, 65, 1, 1, 1, 1)       # the host surface takes it
    x = np.ones(g.ncols())
    assert g.Mv(x).shape == (g.nrows(),)
    assert not g.factorize()
    assert "n <= 64 and nr <= 16" in g.last_error()
    h = BorderedCR()
    h.allocate(4, 8, 1, 17, 17, 1)
    assert not h.factorize() and "nr <= 16" in h.last_error()


def test_dup_and_t_solve():
    s = systems.gaussian_system(12, 3, 1, 1, 2, 2, seed=4)
    a = BorderedCR()
    a.load_system(s)
    a.set_num_parallel_block(3)
    b = BorderedCR()
    b.dup(a)
    pa, pb = a.packed(), b.packed()
    assert all(np.array_equal(pa[k], pb[k]) for k in pa)
    assert b.used_parallel_block() == a.used_parallel_block() == 3
    with pytest.raises(AlglinError, match="t_solve"):
        b.t_solve(np.zeros(b.nrows()))
    assert not b.solve(np.ones(b.nrows()))[0] and "not factorized" in b.last_error()
