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
    if ok:   # converged after all, or at the FP64 floor: then it must say so
        assert rep["status"] in (0, 5)
        if rep["status"] == 0:
            assert rep["rel_residual"] <= 1e-12
        else:
            assert 0 <= rep["backward_error"] <= 64 * 1.1102230246251565e-16
    else:
        assert rep["status"] in (1, 2)
        assert "refinement" in g.last_error()


@pytest.mark.parametrize("N,n,q,r,kmax", [(64, 8, 2, 2, 1e4), (20, 64, 16, 8, 1e6), (33, 40, 8, 4, 1e5)])
def test_mixed_falls_back_to_fp64_qrp(oracle_mod, N, n, q, r, kmax):
    """max_iter = 0 ends the refinement after the first FP32 solve, so every
    column takes the FP64 QRP path (the chunked QR pair kernel, T and a
    column chunk in LDS, also at n = 64): the result is that path's, status
    4, agreeing with the FP64 QRP oracle and at its residual."""
    print(f"case N={N} n={n} q={q} r={r} kmax={kmax:g}", flush=True)
    s = systems.conditioned_system(N, n, q, q, r, r, kmin=1e2, kmax=kmax, seed=3 + n)
    nrows, _ = systems.dims(s)
    b = np.random.default_rng(5).standard_normal((nrows, 2))
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(N, n, q, q, r, r)
    o.load(s)
    o.select("QRP", "LU")
    assert o.factorize()
    ok, Xo = o.solve_n(b)
    floor = max(np.linalg.norm(systems.mv(s, Xo[:, c]) - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(2))
    tol = max(1e-12, 4 * floor)
    g = mixed_solver(s, tol=tol, max_iter=0)
    assert g.factorize()
    ok, X = g.solve_n(b)
    rep = g.refine_report()
    assert ok, (g.last_error(), rep)
    assert rep["status"] == 4, rep
    res = [np.linalg.norm(systems.mv(s, X[:, c]) - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(2)]
    assert max(res) <= tol, (res, floor)
    # a residual at this level is mostly the rounding of its own evaluation:
    # two FP64 evaluations of b - A x (the device's, numpy's) differ by up to
    # a few u |A| |x| + |b| in norm, relative to ||b||
    A = systems.dense(s) if nrows <= 3000 else None
    if A is not None:
        noise = max(8 * np.finfo(float).eps * np.linalg.norm(np.abs(A) @ np.abs(X[:A.shape[1], c]) + np.abs(b[:, c]))
                    / np.linalg.norm(b[:, c]) for c in range(2))
    else:
        noise = 1e-13
    assert abs(rep["rel_residual"] - max(res)) <= 0.1 * max(res) + noise, (rep["rel_residual"], res, noise)
    cond = np.linalg.cond(A) if A is not None else 1e8
    assert np.abs(X - Xo).max() / np.abs(Xo).max() <= max(1e-12, 4 * cond * tol)


def test_mixed_without_fallback_reports_max_iter():
    s = systems.conditioned_system(64, 8, 2, 2, 2, 2, kmin=1e2, kmax=1e4, seed=4)
    nrows, _ = systems.dims(s)
    g = mixed_solver(s, tol=1e-12, max_iter=0)
    g.set_fallback(False)
    assert g.factorize()
    ok, _ = g.solve(np.random.default_rng(1).standard_normal(nrows))
    assert not ok and "refinement" in g.last_error()
    assert g.refine_report()["status"] == 2


@pytest.mark.parametrize("N", [600, 20000])
def test_mixed_kappa_1e6_four_rhs(oracle_mod, N):
    """The mixed-precision workload at n = 64, qr = qx = 16, nr = nx = 8 with
    kappa(D_k) in [1e2, 1e6] and 4 random right-hand sides (+ b = A 1).
    Random right-hand sides have no FP64 solution with a 1e-12 residual
    (plain FP64 stays near 1e-9 at N = 20000): each column either reaches
    1e-12 (status 0) or stops at the FP64 floor (status 5: componentwise
    backward error <= 64 u), and its residual is never worse than plain
    FP64's.  At N = 600 the FP64 oracle's solution is the reference for the
    forward error; at N = 20000 (full size) the properties are size-free."""
    from alglin_amd.bcr import DeviceArray, packed_blocks
    n, q, r = 64, 16, 8
    print(f"case N={N}", flush=True)
    s = systems.conditioned_system(N, n, q, q, r, r, kmin=1e2, kmax=1e6, seed=2024)
    nrows, ncols = systems.dims(s)
    L = max(nrows, ncols)
    B = np.zeros((nrows, 5))
    B[:, :4] = np.random.default_rng(7).standard_normal((nrows, 4))
    B[:, 4] = systems.mv(s, np.ones(ncols))
    dev = {k: DeviceArray.from_host(v) for k, v in packed_blocks(s).items()}

    def run(mode):
        g = BorderedCR(np.float64)
        g.allocate(N, n, q, q, r, r)
        g.select("LU", "LU", "NONE")
        if mode == "mixed":
            g.select_precision("MIXED", 1e-12, 30)
            g.set_fallback(True)
        g.bind_device(dev)
        Xd = DeviceArray.from_host(np.asfortranarray(np.vstack([B, np.zeros((L - nrows, 5))])).reshape(-1, order="F"))
        assert g.factorize_device(sync=True), g.last_error()
        g.solve_device(Xd, 5, L)
        X = Xd.to_host().reshape(L, 5, order="F")[:ncols]
        return g, X

    g64, X64 = run("fp64")
    gm, Xm = run("mixed")
    rep = gm.refine_report()
    assert rep["status"] in (0, 5), rep
    if rep["status"] == 5:
        assert 0 <= rep["backward_error"] <= 64 * 1.1102230246251565e-16
    for c in range(5):
        r64 = np.linalg.norm(systems.mv(s, X64[:, c]) - B[:, c]) / np.linalg.norm(B[:, c])
        rm = np.linalg.norm(systems.mv(s, Xm[:, c]) - B[:, c]) / np.linalg.norm(B[:, c])
        print(f"col {c}: plain FP64 {r64:.2e}  mixed {rm:.2e}", flush=True)
        assert rm <= max(1e-12, 2 * r64), (c, rm, r64)
    # b = A 1: the forward error is the conditioning's
    assert np.abs(Xm[:, 4] - 1).max() <= 1e-5
    if N <= 1000:
        o = oracle_mod.OracleBorderedCR(np.float64)
        o.allocate(N, n, q, q, r, r)
        o.load(s)
        assert o.factorize()
        ok, Xo = o.solve_n(B)
        assert ok
        # both are backward stable: they differ by at most cond x residuals
        for c in range(5):
            d = np.abs(Xm[:, c] - Xo[:, c]).max() / np.abs(Xo[:, c]).max()
            assert d <= 1e-4, (c, d)


def test_mixed_residual_device_after_host_load():
    """MIXED + host blocks: only the mixed engine holds the operator; the
    device residual runs there."""
    from alglin_amd.bcr import DeviceArray
    s = systems.conditioned_system(100, 8, 2, 2, 2, 2, kmin=1e2, kmax=1e3, seed=8)
    nrows, ncols = systems.dims(s)
    g = mixed_solver(s)
    assert g.factorize()
    x = np.random.default_rng(3).standard_normal(ncols)
    b = np.random.default_rng(4).standard_normal(nrows)
    xd, bd, rd = DeviceArray.from_host(x), DeviceArray.from_host(b), DeviceArray(nrows)
    g.residual_device(xd, bd, rd)
    assert np.abs(rd.to_host() - (b - systems.mv(s, x))).max() <= 1e-12 * np.abs(b).max() * 10


def test_solve_device_rejects_bad_calls():
    from alglin_amd.bcr import AlglinError, DeviceArray
    s = systems.gaussian_system(20, 4, 1, 1, 1, 1, seed=2)
    nrows, ncols = systems.dims(s)
    L = max(nrows, ncols)
    g = BorderedCR(np.float64)
    g.load_system(s)
    X = DeviceArray(2 * L)
    with pytest.raises(AlglinError, match="not factorized"):
        g.solve_device(X, 1, L)
    assert g.factorize()
    with pytest.raises(AlglinError, match="ldRhs"):
        g.solve_device(X, 2, L - 1)
    with pytest.raises(AlglinError, match="nrhs"):
        g.solve_device(X, -1, L)
