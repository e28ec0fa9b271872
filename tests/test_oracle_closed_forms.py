"""The oracle against the reference tests' closed forms.

* test18-BorderedCR-coverage.cc:920-1059 -- the stable 6-block system,
  x_i = 0.25 + 0.5 i, every internal {LU,QR,QRP} x last
  {LU,LUPQ,QR,QRP,SVD,LSS,LSY,PINV} x threading mode, 1 RHS and 3 RHS,
  tolerance 5e-9 on x and on the residual.
* test18:1061-1098 -- singular last block: LU fails without a fallback,
  PINV fallback recovers x = (2, -1, 0).
* test9-Cinterface.c -- the 7x7 (x = 1..7) and bordered 8x8 systems.
* test7-BorderedCR.cc -- N=10000, n=8, mt19937(2) fill, x_i = 1 + i%100,
  max error < 1e-8, residual < 1e-6.
* test18:342-346 / test17:128-139 -- the partition shows up in info().
"""
import numpy as np
import pytest

from alglin_amd import systems

INTERNALS = ["LU", "QR", "QRP"]
LASTS = ["LU", "LUPQ", "QR", "QRP", "SVD", "LSS", "LSY", "PINV"]


def make(oracle_mod, s, internal="LU", last="LU", last2="NONE", nthreads=1, npb=0, dtype=np.float64):
    o = oracle_mod.OracleBorderedCR(dtype, nthreads, npb)
    o.allocate(s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"])
    o.load(s)
    o.select(internal, last, last2)
    return o


@pytest.mark.parametrize("internal", INTERNALS)
@pytest.mark.parametrize("last", LASTS)
@pytest.mark.parametrize("fth,sth", [(False, False), (True, False), (False, True), (True, True)])
def test_stable_system_all_combinations(oracle_mod, internal, last, fth, sth):
    s = systems.stable_system()
    o = make(oracle_mod, s, internal, last, nthreads=4)
    o.set_num_parallel_block(4)
    o.set_use_thread(fth, sth)
    nrows, ncols = systems.dims(s)
    xref = 0.25 + 0.5 * np.arange(ncols)
    b = systems.mv(s, xref)
    assert o.factorize(), o.last_error()
    ok, x = o.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 5e-9
    assert np.abs(systems.mv(s, x) - b).max() < 5e-9
    nrhs = 3
    X = np.stack([(1 + c) * xref - 0.1 * c for c in range(nrhs)], axis=1)
    Bm = systems.mv(s, X)
    ok, Xs = o.solve_n(Bm)
    assert ok
    assert np.abs(Xs - X).max() < 5e-9
    assert np.abs(systems.mv(s, Xs) - Bm).max() < 5e-9


def test_partition_info(oracle_mod):
    s = systems.stable_system()
    o = make(oracle_mod, s, nthreads=4)
    o.set_num_parallel_block(4)
    # 6 blocks, 4 requested ranges -> 3 used (every range keeps 2 blocks), cc:141-158
    assert "nblk:3" in o.info()
    assert o.used_parallel_block() == 3
    assert "LastBlock LU" in o.info()


def test_singular_last_block_fallback(oracle_mod):
    s = systems.singular_fallback()
    o = make(oracle_mod, s, "LU", "LU", "NONE")
    assert not o.factorize()
    assert "load_and_factorize_last" in o.last_error()
    o = make(oracle_mod, s, "LU", "LU", "PINV")
    assert o.factorize()
    xref = np.array([2.0, -1.0, 0.0])
    b = systems.mv(s, xref)
    ok, x = o.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 5e-9


@pytest.mark.parametrize("bordered", [False, True])
def test_cinterface_systems(oracle_mod, bordered):
    s = systems.cinterface_system(bordered)
    nrows, ncols = systems.dims(s)
    A = systems.dense(s)
    if bordered:
        b = np.array([32, 81, 66, 165, 121, -52, 237, 4], float)
        xref = np.array([1, 2, 3, 4, 5, 6, 7, 0], float)
        # the printed 8x8 of test9:106-116
        assert A[7].tolist() == [1, -1, 1, -1, 1, -1, 1, -1]
    else:
        b = np.array([32, 81, 66, 165, 121, -52, 237], float)
        xref = np.arange(1, 8, dtype=float)
        assert A[4].tolist() == [0, 1, 0, 0, 20, 2, 1]
    assert np.allclose(A @ xref, b)
    o = make(oracle_mod, s)
    assert o.factorize()
    ok, x = o.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 1e-12


@pytest.mark.parametrize("nthreads", [1, 4])
def test_test7_large_system(oracle_mod, nthreads):
    s = systems.test7_system()
    nrows, ncols = systems.dims(s)
    xref = 1.0 + (np.arange(ncols) % 100)
    b = systems.mv(s, xref)
    o = make(oracle_mod, s, nthreads=nthreads)
    assert o.factorize()
    ok, x = o.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 1e-8
    assert np.abs(x - xref).sum() / ncols < 1e-8
    r = systems.mv(s, x) - b
    assert np.linalg.norm(r) < 1e-6
    ok, X = o.solve_n(np.stack([b, b], 1))
    assert ok and np.abs(X[:, 1] - xref).max() < 1e-8
