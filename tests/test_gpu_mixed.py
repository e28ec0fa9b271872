"""Mixed precision on the GPU: FP32 cyclic-reduction factors of the FP64
operator + FP64 iterative refinement (no counterpart in the reference; the
FP64 oracle is the judge).  Tolerance: relative residual <= 1e-12 as
requested (or twice the FP64 oracle's own residual where that is larger),
forward error against the FP64 oracle's solution within cond(A) times the
two residuals (forward error <= condition number x backward error)."""
import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import BorderedCR

pytestmark = pytest.mark.gpu


def mixed_solver(s, tol=1e-12, max_iter=30):
    g = BorderedCR(np.float64)
    g.allocate(s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"])
    g.select_precision("MIXED", tol, max_iter)
    g.load_system(s)
    g.select("LU", "LU", "NONE")
    return g


@pytest.mark.parametrize("N,n,qr,nr", [(200, 8, 2, 2), (300, 16, 4, 2), (257, 32, 8, 4), (33, 40, 4, 4),
                                        (1, 6, 2, 2), (2, 4, 1, 1), (20, 64, 16, 8)])
def test_mixed_reaches_fp64_residual(oracle_mod, N, n, qr, nr):
    print(f"case N={N} n={n} qr={qr} nr={nr}", flush=True)
    s = systems.conditioned_system(N, n, qr, qr, nr, nr, kmin=1e2, kmax=1e4, seed=N + n)
    nrows, ncols = systems.dims(s)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(nrows)
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(N, n, qr, qr, nr, nr)
    o.load(s)
    assert o.factorize()
    ok, xo = o.solve(b)
    # the residual the FP64 direct solve itself attains on this system
    floor = np.linalg.norm(systems.mv(s, xo) - b) / np.linalg.norm(b)
    tol = max(1e-12, 2.0 * floor)
    g = mixed_solver(s, tol=tol)
    assert g.factorize(), g.last_error()
    ok, x = g.solve(b)
    rep = g.refine_report()
    assert ok, (g.last_error(), rep, floor)
    assert rep["status"] == 0 and rep["rel_residual"] <= tol
    assert rep["first_residual"] > rep["rel_residual"]   # refinement did the work
    r = systems.mv(s, x) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 2 * tol
    A = systems.dense(s) if nrows <= 3000 else None
    cond = np.linalg.cond(A) if A is not None else 1e6
    bound = max(1e-13, 2.0 * cond * (rep["rel_residual"] + floor))
    assert np.abs(x - xo).max() / np.abs(xo).max() <= bound


def test_mixed_multi_rhs_and_device_path(oracle_mod):
    from alglin_amd.bcr import DeviceArray, packed_blocks
    s = systems.conditioned_system(500, 16, 4, 4, 2, 2, kmin=1e2, kmax=1e3, seed=9)
    nrows, ncols = systems.dims(s)
    B = np.random.default_rng(4).standard_normal((nrows, 4))
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(500, 16, 4, 4, 2, 2)
    o.load(s)
    assert o.factorize()
    ok, Xo = o.solve_n(B)
    floor = max(np.linalg.norm(systems.mv(s, Xo[:, c]) - B[:, c]) / np.linalg.norm(B[:, c]) for c in range(4))
    tol = max(1e-12, 2.0 * floor)
    g = mixed_solver(s, tol=tol)
    assert g.factorize()
    ok, X = g.solve_n(B)
    assert ok, (g.last_error(), g.refine_report(), floor)
    for c in range(4):
        assert np.linalg.norm(systems.mv(s, X[:, c]) - B[:, c]) / np.linalg.norm(B[:, c]) <= 2 * tol
    # device-resident operator
    pk = packed_blocks(s)
    dev = {k: DeviceArray.from_host(v) for k, v in pk.items()}
    h = BorderedCR(np.float64)
    h.allocate(500, 16, 4, 4, 2, 2)
    h.select_precision("MIXED", tol)
    h.bind_device(dev)
    assert h.factorize_device(sync=True), h.last_error()
    xd = DeviceArray.from_host(B[:, 0].copy())
    h.solve_device(xd, 1, nrows)
    x = xd.to_host()
    assert np.linalg.norm(systems.mv(s, x) - B[:, 0]) / np.linalg.norm(B[:, 0]) <= 2 * tol
    # a copy queued on the solver's stream is ordered with its solve
    bd = DeviceArray.from_host(B[:, 1].copy())
    for _ in range(2):
        xd.copy_from(bd, stream=h.stream())
        h.solve_device(xd, 1, nrows)
        x = xd.to_host()
        assert np.linalg.norm(systems.mv(s, x) - B[:, 1]) / np.linalg.norm(B[:, 1]) <= 2 * tol
    assert h.refine_report()["iterations"] >= 1


def test_mixed_reports_stall_on_fp32_singular_blocks():
    """Blocks with cond ~1e10: FP32 factors carry no digits, refinement
    cannot converge; without the fallback the solve must say so instead of
    returning garbage."""
    s = systems.conditioned_system(64, 8, 2, 2, 2, 2, kmin=1e10, kmax=1e11, seed=3)
    nrows, _ = systems.dims(s)
    g = mixed_solver(s, tol=1e-12, max_iter=10)
    g.set_fallback(False)
    g.factorize()
    ok, x = g.solve(np.ones(nrows))
    rep = g.refine_report()
    if ok:   # converged after all: then the residual must really be small
        assert rep["rel_residual"] <= 1e-12
    else:
        assert rep["status"] in (1, 2)
        assert "refinement" in g.last_error()


def test_mixed_falls_back_to_fp64_qrp(oracle_mod):
    """Same ill-conditioned blocks with the fallback on: refinement stalls,
    the FP64 QRP cyclic reduction solves the system (status 4) and the result
    agrees with the FP64 QRP oracle to the accuracy the FP64 solve has."""
    N, n, q, r = 64, 8, 2, 2
    s = systems.conditioned_system(N, n, q, q, r, r, kmin=1e10, kmax=1e11, seed=3)
    nrows, _ = systems.dims(s)
    b = np.random.default_rng(5).standard_normal((nrows, 2))
    g = mixed_solver(s, tol=1e-12, max_iter=10)
    assert g.factorize()
    ok, X = g.solve_n(b)
    rep = g.refine_report()
    assert ok, (g.last_error(), rep)
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(N, n, q, q, r, r)
    o.load(s)
    o.select("QRP", "LU")
    assert o.factorize()
    ok, Xo = o.solve_n(b)
    floor = max(np.linalg.norm(systems.mv(s, Xo[:, c]) - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(2))
    if rep["status"] == 4:
        for c in range(2):
            res = np.linalg.norm(systems.mv(s, X[:, c]) - b[:, c]) / np.linalg.norm(b[:, c])
            assert res <= max(1e-12, 10 * floor), (res, floor)
        assert abs(rep["rel_residual"] - max(
            np.linalg.norm(systems.mv(s, X[:, c]) - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(2))) \
            <= 1e-3 * rep["rel_residual"] + 1e-18
    else:   # refinement converged after all
        assert rep["status"] == 0 and rep["rel_residual"] <= 1e-12
