"""Last-block factorizations on the GPU (BABD_BorderedCR.cc:1281-1479 with
select_last LU / LUPQ / QR / QRP / SVD / LSS / LSY / PINV and the second
solver of cc:1347-1369): closed forms of test18 and agreement with the
oracle's dense classes on the same inputs (FP64, relative 1e-10 on x; 1e-9
for the SVD family, whose Jacobi sweeps are ordered differently), including
a rectangular last block (Nr > Nc) and a 152 x 152 block factored in HBM
instead of LDS."""
import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import BorderedCR

pytestmark = pytest.mark.gpu

LASTS = ["LU", "LUPQ", "QR", "QRP"]


def gpu(s, internal, last):
    g = BorderedCR(np.float64)
    g.load_system(s)
    g.select(internal, last, "NONE")
    return g


@pytest.mark.parametrize("internal", ["LU", "QR", "QRP"])
@pytest.mark.parametrize("last", LASTS)
def test_stable_system_every_last_block(internal, last):
    """test18-BorderedCR-coverage.cc:954-1059: x_i = 0.25 + 0.5 i, tol 5e-9."""
    print(f"case internal={internal} last={last}", flush=True)
    s = systems.stable_system()
    nrows, ncols = systems.dims(s)
    xref = 0.25 + 0.5 * np.arange(ncols)
    g = gpu(s, internal, last)
    assert g.factorize(), g.last_error()
    ok, x = g.solve(systems.mv(s, xref))
    assert ok
    assert np.abs(x - xref).max() < 5e-9
    X = np.stack([(1 + c) * xref - 0.1 * c for c in range(3)], 1)
    ok, Xs = g.solve_n(systems.mv(s, X))
    assert ok and np.abs(Xs - X).max() < 5e-9
    assert f"LastBlock {last}" in g.info()


SHAPES = [
    (9, 6, 3, 3, 2, 2),      # square last block, LDS
    (17, 8, 2, 2, 3, 3),
    (12, 5, 3, 2, 2, 2),     # Nr = 2n+5 > Nc = 2n+4: QR / QRP only
    (6, 64, 16, 16, 8, 8),   # Nc = 152: factored in HBM
]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("last", LASTS)
def test_last_block_matches_oracle(oracle_mod, shape, last):
    N, n, qr, qx, nr, nx = shape
    Nr, Nc = 2 * n + qr + nr, 2 * n + qx + nx
    if last in ("LU", "LUPQ") and Nr != Nc:
        pytest.skip("LU-type last blocks need Nr == Nc")
    print(f"case shape={shape} last={last}", flush=True)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N * 7 + n)
    nrows, ncols = systems.dims(s)
    L = max(nrows, ncols)
    B = np.zeros((L, 2))
    B[:nrows] = np.random.default_rng(3).standard_normal((nrows, 2))
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    o.select("LU", last)
    assert o.factorize(), o.last_error()
    ok, Xo = o.solve_n(B[:nrows])
    assert ok
    g = gpu(s, "LU", last)
    assert g.factorize(), g.last_error()
    ok, Xg = g.solve_n(B[:nrows])
    assert ok, g.last_error()
    err = np.abs(Xg[:ncols] - Xo[:ncols]).max() / np.abs(Xo[:ncols]).max()
    assert err < 1e-10, err


@pytest.mark.parametrize("last", ["SVD", "LSS", "LSY", "PINV"])
@pytest.mark.parametrize("shape", [(9, 6, 3, 3, 2, 2), (12, 5, 3, 2, 2, 2), (6, 64, 16, 16, 8, 8)])
def test_svd_family_matches_oracle(oracle_mod, shape, last):
    """SVD / LSS / LSY / PINV last blocks: the minimum-norm least-squares
    solution through a one-sided Jacobi SVD on the GPU, against the oracle's."""
    N, n, qr, qx, nr, nx = shape
    print(f"case shape={shape} last={last}", flush=True)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N * 5 + n)
    nrows, ncols = systems.dims(s)
    B = np.random.default_rng(4).standard_normal((nrows, 2))
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    o.select("LU", last)
    assert o.factorize(), o.last_error()
    ok, Xo = o.solve_n(B)
    assert ok
    g = gpu(s, "LU", last)
    assert g.factorize(), g.last_error()
    ok, Xg = g.solve_n(B)
    assert ok, g.last_error()
    err = np.abs(Xg[:ncols] - Xo[:ncols]).max() / np.abs(Xo[:ncols]).max()
    assert err < 1e-9, err


@pytest.mark.parametrize("last2", ["SVD", "LSS", "LSY", "PINV"])
def test_second_solver_on_singular_last_block(last2):
    """test18-BorderedCR-coverage.cc:1061-1098: an exactly singular last block
    with LU first and a singular-value second solver gives x = (2, -1, 0)."""
    s = systems.singular_fallback()
    nrows, ncols = systems.dims(s)
    xref = np.array([2.0, -1.0, 0.0])
    b = systems.mv(s, xref)
    g = gpu(s, "LU", "LU")
    g.select("LU", "LU", last2)
    assert g.factorize(), g.last_error()
    ok, x = g.solve(b)
    assert ok
    assert np.abs(x[:ncols] - xref).max() < 5e-9, x


def test_second_solver_idle_on_regular_block(oracle_mod):
    """A regular last block with a second solver selected: the first solver's
    result, untouched."""
    s = systems.gaussian_system(9, 6, 3, 3, 2, 2, shift=3.0, seed=3)
    nrows, _ = systems.dims(s)
    b = np.random.default_rng(1).standard_normal(nrows)
    g1 = gpu(s, "LU", "LU")
    assert g1.factorize()
    ok, x1 = g1.solve(b)
    g2 = gpu(s, "LU", "LU")
    g2.select("LU", "LU", "SVD")
    assert g2.factorize()
    ok, x2 = g2.solve(b)
    assert np.array_equal(x1, x2)


def test_singular_last_block_fails_like_the_reference():
    """An exactly singular last block: LU / LUPQ / QR / QRP fail with the
    reference's message (no second solver selected, cc:1347-1374)."""
    s = systems.gaussian_system(5, 3, 1, 1, 1, 1, shift=3.0, seed=11)
    s["H"][:] = 0.0          # boundary rows all zero -> last block singular
    for last in LASTS:
        g = gpu(s, "LU", last)
        assert not g.factorize(), last
        assert "load_and_factorize_last" in g.last_error()
