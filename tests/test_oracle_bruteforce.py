"""The oracle against brute force: dense assembly + numpy.linalg.solve.

Covers what the closed forms leave open: random shapes including N = 1, 2, 3
(no or one elimination level), odd N (unpaired rows, cc:1133 jpp clamp),
empty borders (qr/qx/nr/nx = 0), every internal elimination and a threaded
partition.  Also pins the sparse pattern (cc:2178-2413) against an
independent numpy assembly.
"""
import numpy as np
import pytest

from alglin_amd import systems

SHAPES = [
    (1, 3, 1, 1, 1, 1), (2, 2, 1, 1, 1, 1), (3, 4, 2, 2, 2, 2), (5, 3, 1, 1, 2, 2),
    (7, 2, 0, 0, 0, 0), (8, 4, 2, 2, 0, 0), (11, 5, 1, 3, 3, 1), (16, 1, 1, 1, 1, 1),
    (17, 6, 3, 3, 2, 2), (31, 4, 4, 4, 4, 4), (64, 3, 1, 1, 1, 1), (100, 2, 2, 2, 1, 1),
]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("internal", ["LU", "QR", "QRP"])
@pytest.mark.parametrize("nthreads", [1, 3])
def test_against_dense_solve(oracle_mod, shape, internal, nthreads):
    N, n, qr, qx, nr, nx = shape
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N * 131 + n)
    A = systems.dense(s)
    nrows, ncols = A.shape
    rng = np.random.default_rng(5)
    b = rng.standard_normal(nrows)
    xd = np.linalg.solve(A, b)
    o = oracle_mod.OracleBorderedCR(np.float64, nthreads)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    o.select(internal, "LU")
    assert np.array_equal(o.dense(), A)
    assert o.factorize(), o.last_error()
    ok, x = o.solve(b)
    assert ok
    cond = np.linalg.cond(A)
    assert np.abs(x - xd).max() <= 1e-13 * cond * max(1.0, np.abs(xd).max())
    Bm = rng.standard_normal((nrows, 3))
    ok, X = o.solve_n(Bm)
    assert ok
    assert np.abs(X - np.linalg.solve(A, Bm)).max() <= 1e-13 * cond * max(1.0, np.abs(X).max())


def test_mv_matches_dense(oracle_mod):
    s = systems.gaussian_system(9, 3, 2, 1, 2, 3, seed=3)
    A = systems.dense(s)
    o = oracle_mod.OracleBorderedCR()
    o.allocate(9, 3, 2, 1, 2, 3)
    o.load(s)
    x = np.random.default_rng(1).standard_normal(A.shape[1])
    assert np.allclose(o.Mv(x), A @ x, rtol=0, atol=1e-13)
    assert np.allclose(systems.mv(s, x), A @ x, rtol=0, atol=1e-13)
    nnz = 9 * 3 * (2 * 3 + 3 + 2) + 2 * (3 + 1 + 3) + (3 + 2) * (2 * 3 + 1 + 3)
    I, J, V = o.sparse()
    assert len(I) == nnz


def test_float_path(oracle_mod):
    s = systems.gaussian_system(40, 4, 2, 2, 2, 2, shift=4.0, seed=9)
    A = systems.dense(s)
    xref = np.linspace(1, 2, A.shape[1])
    b = A @ xref
    o = oracle_mod.OracleBorderedCR(np.float32)
    o.allocate(40, 4, 2, 2, 2, 2)
    o.load(s)
    assert o.factorize()
    ok, x = o.solve(b.astype(np.float32))
    assert ok
    # forward error of a backward-stable float solve: O(eps_f32 * cond(A))
    bound = 10 * np.finfo(np.float32).eps * np.linalg.cond(A) * np.abs(xref).max()
    assert np.abs(x - xref).max() < bound
