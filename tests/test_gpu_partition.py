"""The partitioned elimination order on the GPU (num_parallel_block = P):
P block ranges reduced by cyclic reduction on their own, then the chain of
their end rows (src/BABD_BorderedCR.cc:139-158, 1055-1171, 1175-1266, and
the reduced forward/backward sweeps 1723-1795, 1953-2075).  The oracle runs
the reference's ordering with the same P; LU pivots must agree exactly, FP64
factors and solutions to 1e-11 (same operations in the same order)."""
import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import BorderedCR

pytestmark = pytest.mark.gpu

SHAPES = [(16, 4, 2, 2, 2, 2), (37, 8, 2, 2, 1, 1), (100, 3, 1, 1, 1, 1), (45, 16, 4, 4, 3, 3),
          (19, 32, 8, 8, 4, 4), (9, 40, 5, 5, 3, 3), (12, 64, 16, 16, 8, 8), (7, 5, 0, 0, 0, 0)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("internal", ["LU", "QR", "QRP"])
@pytest.mark.parametrize("npb", [2, 3, 7])
def test_partition_matches_oracle(oracle_mod, shape, internal, npb):
    N, n, qr, qx, nr, nx = shape
    print(f"case shape={shape} internal={internal} npb={npb}", flush=True)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N * 17 + n + npb)
    nrows, ncols = systems.dims(s)
    B = np.random.default_rng(11).standard_normal((nrows, 2))
    o = oracle_mod.OracleBorderedCR(np.float64, npb, npb)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    o.select(internal, "LU")
    assert o.factorize(), o.last_error()
    ok, Xo = o.solve_n(B)
    assert ok
    g = BorderedCR(np.float64)
    g.load_system(s)
    g.select(internal, "LU", "NONE")
    g.set_num_parallel_block(npb)
    assert g.used_parallel_block() == o.used_parallel_block()
    assert f"nblk:{o.used_parallel_block()}" in g.info()
    assert g.factorize(), g.last_error()
    ok, Xg = g.solve_n(B)
    assert ok, g.last_error()
    scale = max(1.0, float(np.abs(Xo).max()))
    assert np.abs(Xg - Xo).max() / scale < 1e-11
    ok1, xg = g.solve(B[:, 1])
    assert ok1 and np.abs(xg - Xo[:, 1]).max() / scale < 1e-11
    r = systems.mv(s, Xg) - B
    assert np.linalg.norm(r) / np.linalg.norm(B) < 1e-12
    if internal == "LU" and N > 1:
        f, oi = g.factors(), o.internals()
        np.testing.assert_array_equal(f["perm"][n:], oi["perm"][n:])
        ts = 2 * n * n + n
        got = f["T"][ts:].reshape(-1, ts)[:, : 2 * n * n]
        ref = oi["T"][ts:].reshape(-1, ts)[:, : 2 * n * n]
        assert np.abs(got - ref).max() / max(1.0, np.abs(ref).max()) < 1e-11
        for key in ("D", "E"):
            assert np.abs(f[key][n * n:] - oi[key][n * n:]).max() / max(1.0, np.abs(oi[key]).max()) < 1e-11, key


def test_partition_float32_and_mixed(oracle_mod):
    """FP32 factors in the partitioned order agree with the FP32 oracle at
    FP32 accuracy; the mixed solver refines them to 1e-12."""
    N, n, qr, qx, nr, nx = 200, 16, 4, 4, 2, 2
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=4.0, seed=5)
    nrows, _ = systems.dims(s)
    b = np.random.default_rng(2).standard_normal(nrows)
    o = oracle_mod.OracleBorderedCR(np.float32, 4, 4)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    assert o.factorize()
    ok, xo = o.solve(b.astype(np.float32))
    g = BorderedCR(np.float32)
    g.load_system(s)
    g.set_num_parallel_block(4)
    assert g.factorize()
    ok, xg = g.solve(b.astype(np.float32))
    assert ok
    assert np.abs(xg - xo).max() / np.abs(xo).max() < 1e-4
    m = BorderedCR(np.float64)
    m.allocate(N, n, qr, qx, nr, nx)
    m.select_precision("MIXED", 1e-12, 30)
    m.load_system(s)
    m.set_num_parallel_block(4)
    assert m.factorize()
    ok, x = m.solve(b)
    assert ok and m.refine_report()["status"] == 0
    assert np.linalg.norm(systems.mv(s, x) - b) / np.linalg.norm(b) <= 1e-12
