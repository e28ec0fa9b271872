"""GPU parity of the MI355X BorderedCR against the oracle and the reference's
closed forms.  Every call goes through the C ABI (libalglin_amd.so).

Tolerances: float64 results agree with the float64 oracle (same elimination
order and pivots) to 1e-11 relative; float32 to 50*eps32*cond(A).  Integer
pivot arrays of the LU path are compared exactly.
"""
import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import BorderedCR, lib

pytestmark = pytest.mark.gpu

SHAPES = [
    (1, 3, 1, 1, 1, 1), (2, 2, 1, 1, 1, 1), (3, 4, 2, 2, 2, 2), (5, 3, 1, 1, 2, 2),
    (7, 2, 0, 0, 0, 0), (8, 4, 2, 2, 0, 0), (11, 5, 1, 3, 3, 1), (16, 1, 1, 1, 1, 1),
    (17, 6, 3, 3, 2, 2), (31, 4, 4, 4, 4, 4), (64, 3, 1, 1, 1, 1), (100, 2, 2, 2, 1, 1),
    (45, 8, 4, 4, 2, 2), (33, 16, 4, 4, 3, 3), (19, 32, 8, 8, 4, 4), (9, 40, 5, 5, 3, 3),
    # last block 116 x 116 (staged in LDS) and 152 x 152 (read from global memory);
    # n = 64 FP64 panels spill to the global scratch of the generic kernel
    (5, 48, 12, 12, 8, 8), (6, 64, 16, 16, 8, 8),
]


def oracle_solve(oracle_mod, s, internal, dtype, B):
    o = oracle_mod.OracleBorderedCR(dtype)
    o.allocate(s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"])
    o.load(s)
    o.select(internal, "LU")
    assert o.factorize(), o.last_error()
    ok, X = o.solve_n(B.astype(dtype))
    assert ok
    return o, X


def gpu_solver(s, internal="LU", dtype=np.float64):
    g = BorderedCR(dtype)
    g.load_system(s)
    g.select(internal, "LU", "NONE")
    return g


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("internal", ["LU", "QR", "QRP"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matches_oracle(oracle_mod, shape, internal, dtype):
    N, n, qr, qx, nr, nx = shape
    print(f"case shape={shape} internal={internal} dtype={np.dtype(dtype).name}", flush=True)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N * 131 + n)
    A = systems.dense(s)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((A.shape[0], 3))
    o, Xo = oracle_solve(oracle_mod, s, internal, dtype, B)
    g = gpu_solver(s, internal, dtype)
    assert g.factorize(), g.last_error()
    ok, Xg = g.solve_n(B.astype(dtype))
    assert ok, g.last_error()
    ok1, xg = g.solve(B[:, 0].astype(dtype))
    assert ok1
    scale = max(1.0, float(np.abs(Xo).max()))
    if dtype == np.float64:
        tol = 1e-11
    else:
        tol = 50 * np.finfo(np.float32).eps * np.linalg.cond(A)
    assert np.abs(Xg - Xo).max() / scale < tol
    assert np.abs(xg - Xo[:, 0]).max() / scale < tol
    if dtype == np.float64:
        Xd = np.linalg.solve(A, B)
        assert np.abs(Xg - Xd).max() / max(1.0, np.abs(Xd).max()) < 1e-13 * np.linalg.cond(A)
    # a second factorization on the same handle reproduces the first bit for
    # bit (no pending border contribution survives between factorizations)
    assert g.factorize(), g.last_error()
    ok2, Xg2 = g.solve_n(B.astype(dtype))
    assert ok2 and np.array_equal(Xg, Xg2)
    f = g.factors()
    if internal == "LU" and N > 1:
        oi = o.internals()
        # LU pivot rows are integers: exact (a tie would need equal |values|)
        np.testing.assert_array_equal(f["perm"][n:], oi["perm"][n:])
        if dtype == np.float64:
            for key in ("T", "D", "E"):
                ref = oi[key][(2 * n * n + n if key == "T" else n * n):]
                got = f[key][(2 * n * n + n if key == "T" else n * n):]
                if key == "T":   # the tau slots are unused by LU
                    ts = 2 * n * n + n
                    ref = ref.reshape(-1, ts)[:, : 2 * n * n]
                    got = got.reshape(-1, ts)[:, : 2 * n * n]
                assert np.abs(got - ref).max() / max(1.0, np.abs(ref).max()) < 1e-11, key


@pytest.mark.parametrize("internal", ["LU", "QR", "QRP"])
def test_stable_system_closed_form(internal):
    """test18-BorderedCR-coverage.cc:920-1059 (x_i = 0.25 + 0.5 i, tol 5e-9)."""
    s = systems.stable_system()
    nrows, ncols = systems.dims(s)
    xref = 0.25 + 0.5 * np.arange(ncols)
    b = systems.mv(s, xref)
    g = gpu_solver(s, internal)
    assert g.factorize(), g.last_error()
    ok, x = g.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 5e-9
    assert np.abs(systems.mv(s, x) - b).max() < 5e-9
    X = np.stack([(1 + c) * xref - 0.1 * c for c in range(3)], 1)
    ok, Xs = g.solve_n(systems.mv(s, X))
    assert ok and np.abs(Xs - X).max() < 5e-9
    assert "LastBlock LU" in g.info()
    # host-side operator and COO export agree with the dense assembly
    A = systems.dense(s)
    assert np.allclose(g.Mv(xref), A @ xref, atol=1e-13)
    I, J, V = g.sparse()
    Ad = np.zeros_like(A)
    np.add.at(Ad, (I, J), V)
    assert np.array_equal(Ad, A)


def test_c_interface_test9():
    """src_tests/test9-Cinterface.c through the BABD_* entry points."""
    L = lib()
    import ctypes

    def p(a):
        return ctypes.c_void_p(a.ctypes.data)

    DE = np.array([10, 2, 2, 40, 2, 1, 3, -1, 10, 2, 2, 40, 2, 1, 3, -1], float)
    H0 = np.array([0, 1, 1, 1, 0, 2], float)
    HN = np.array([20, 0, 2, 2, -10, 2], float)
    Hq = np.array([1, 1, 30], float)
    rhs = np.array([32, 81, 66, 165, 121, -52, 237], float)
    assert L.BABD_factorize(1, 0, 0, 2, 2, 1, 1, p(DE), 2, p(H0), 3, p(HN), 3, p(Hq), 3) == 0, L.BABD_get_last_error()
    assert L.BABD_solve(1, p(rhs)) == 0
    assert np.abs(rhs - np.arange(1, 8)).max() < 1e-12
    B = np.ones(7)
    C = np.array([1, -1, 1, -1, 1, -1, 1], float)
    D = np.array([-1.0])
    rhs1 = np.array([32, 81, 66, 165, 121, -52, 237, 4], float)
    assert L.BABD_factorize_bordered(1, 0, 0, 2, 2, 1, 1, 1, 1, p(DE), 2, p(H0), 3, p(HN), 3, p(Hq), 3,
                                     p(B), 7, p(C), 1, p(D), 1) == 0, L.BABD_get_last_error()
    assert L.BABD_solve(1, p(rhs1)) == 0
    assert np.abs(rhs1 - np.array([1, 2, 3, 4, 5, 6, 7, 0])).max() < 1e-12
    assert L.BABD_free(1) == 0


def test_test7_system():
    """test7-BorderedCR.cc: N=10000, n=8, mt19937(2) fill; err < 1e-8, residual < 1e-6."""
    s = systems.test7_system()
    nrows, ncols = systems.dims(s)
    xref = 1.0 + (np.arange(ncols) % 100)
    b = systems.mv(s, xref)
    g = gpu_solver(s)
    assert g.factorize(), g.last_error()
    assert g.fast_path()
    ok, x = g.solve(b)
    assert ok
    assert np.abs(x - xref).max() < 1e-8
    assert np.linalg.norm(systems.mv(s, x) - b) < 1e-6
    ok, X = g.solve_n(np.stack([b, b], 1))
    assert ok and np.abs(X - xref[:, None]).max() < 1e-8


def test_large_n32_matches_oracle(oracle_mod):
    """Block size and borders of the large single-system workload (n=32,
    qr=qx=8, nr=nx=4) at N=4096, 1 and 16 right-hand sides."""
    s = systems.gaussian_system(4096, 32, 8, 8, 4, 4, shift=3.0, seed=1)
    nrows, _ = systems.dims(s)
    B = np.random.default_rng(3).standard_normal((nrows, 16))
    _, Xo = oracle_solve(oracle_mod, s, "LU", np.float64, B)
    g = gpu_solver(s)
    assert g.factorize(), g.last_error()
    assert g.fast_path()
    ok, X = g.solve_n(B)
    assert ok
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-11
    r = systems.mv(s, X) - B
    assert np.linalg.norm(r) / np.linalg.norm(B) < 1e-13


@pytest.mark.parametrize("n,nr,nx", [(40, 3, 2), (48, 8, 16), (33, 2, 9), (20, 5, 9), (6, 1, 1), (8, 0, 0),
                                     (56, 8, 3), (48, 16, 2), (60, 8, 2), (64, 2, 2), (44, 0, 2)])
def test_device_residual_shapes(n, nr, nx):
    """The GPU residual r = b - A x (BABD_BorderedCR.cc:2111-2166): block rows
    with the border partials C_k x_k formed in the same column loop (on
    lanes n..n+nr-1 of their own when n + nr <= 64 and one block row fills a
    wave: (56, 8) and (48, 16) fill the wave exactly, (60, 8) and (64, 2) do
    not fit, nr = 0 has no border), the first B columns and b loaded ahead,
    several block rows per wave for small n; against the host product of the
    same blocks."""
    from alglin_amd.bcr import DeviceArray, packed_blocks
    N = 37
    qr = max(min(4, n), nx - nr)
    qx = qr + nr - nx   # a square last block (the LU last-block solver)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=4.0, seed=n + nr)
    dev = {k: DeviceArray.from_host(v) for k, v in packed_blocks(s).items()}
    g = BorderedCR()
    g.allocate(N, n, qr, qx, nr, nx)
    g.bind_device(dev)
    assert g.factorize_device(sync=True), g.last_error()
    nrows, ncols = systems.dims(s)
    rng = np.random.default_rng(n)
    x, b = rng.standard_normal(ncols), rng.standard_normal(nrows)
    X, bd, r = DeviceArray.from_host(x), DeviceArray.from_host(b), DeviceArray(nrows)
    g.residual_device(X, bd, r)
    ref = b - systems.mv(s, x)
    assert np.abs(r.to_host() - ref).max() <= 1e-12 * (1.0 + np.abs(ref).max())


@pytest.mark.parametrize("n,nx", [(32, 4), (24, 8), (12, 2), (6, 1)])
def test_full_size_residual(n, nx):
    """Size-independent property at the large single-system size: the FP64
    relative residual of factorize + solve stays at rounding level.  At this
    size many pairs of a level run concurrently, which exercises the
    ordering between the reads and writes of neighbouring pairs."""
    N = 100000 if n >= 24 else 200000
    s = systems.gaussian_system(N, n, 8, 8, nx, nx, shift=3.0, seed=n)
    nrows, ncols = systems.dims(s)
    b = systems.mv(s, np.ones(ncols))
    g = gpu_solver(s)
    assert g.factorize(), g.last_error()
    ok, x = g.solve(b)
    assert ok
    assert np.linalg.norm(systems.mv(s, x) - b) / np.linalg.norm(b) < 1e-13
    assert np.abs(x - 1.0).max() < 1e-10


def test_zero_pivot_reports_failure():
    s = systems.gaussian_system(4, 2, 1, 1, 1, 1, seed=2)
    s["E"][0][:] = 0.0
    s["D"][1][:] = 0.0    # T = [E_0; D_1] = 0 -> getrf info > 0 (cc:856-863)
    g = gpu_solver(s)
    assert not g.factorize()
    assert "pivot" in g.last_error()


def test_device_resident_path():
    """bind_device / factorize_device / solve_device / residual_device on
    device arrays allocated through the library itself."""
    from alglin_amd.bcr import DeviceArray, packed_blocks
    s = systems.gaussian_system(300, 8, 2, 2, 2, 2, shift=4.0, seed=5)
    pk = packed_blocks(s)
    dev = {k: DeviceArray.from_host(v) for k, v in pk.items()}
    g = BorderedCR()
    g.allocate(300, 8, 2, 2, 2, 2)
    g.bind_device(dev)
    assert g.factorize_device(sync=True), g.last_error()
    nrows, ncols = systems.dims(s)
    xref = np.linspace(-1, 1, ncols)
    b = systems.mv(s, xref)
    X = DeviceArray.from_host(b)
    g.solve_device(X, 1, nrows)
    x = X.to_host()
    assert np.abs(x - xref).max() < 1e-12
    bd = DeviceArray.from_host(b)
    r = DeviceArray(nrows)
    g.residual_device(X, bd, r)
    rr = b - systems.mv(s, x)
    assert np.abs(r.to_host() - rr).max() < 1e-12
    g.set_timing(True)
    assert g.factorize_device(sync=True)
    g.solve_device(X, 1, nrows)
    t = g.timing()
    assert t["pair_launches"] == 9 and t["pair_ms"] > 0 and t["forward_ms"] > 0


@pytest.mark.parametrize("shape", [(19, 32, 8, 8, 4, 4), (40, 24, 3, 3, 8, 8), (33, 17, 2, 2, 1, 1),
                                   (21, 5, 2, 2, 2, 2), (17, 40, 5, 5, 3, 3), (9, 64, 16, 16, 8, 8)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pair_kernel_variants_agree(oracle_mod, shape, dtype):
    """Every LU pair kernel (default, register-tiled, column-split waves,
    matrix-core right-hand update) gives the oracle's pivots and the
    same solution; the register-tiled and column-split kernels do the same
    arithmetic in the same order, so their factor panels agree bit for bit.
    The matrix-core kernel (FP32 factors n <= 64, FP64 factors n <= 48)
    factors T with the same arithmetic (T bit for bit at the first level,
    where its inputs are the same) and forms the right-hand blocks as GEMMs
    (D, E to the factor precision's rounding)."""
    N, n, qr, qx, nr, nx = shape
    print(f"case shape={shape} dtype={np.dtype(dtype).name}", flush=True)
    s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=3.0, seed=N + n)
    nrows, ncols = systems.dims(s)
    b = np.random.default_rng(0).standard_normal(nrows).astype(dtype)
    o = oracle_mod.OracleBorderedCR(dtype)
    o.allocate(N, n, qr, qx, nr, nx)
    o.load(s)
    o.select("LU", "LU")
    assert o.factorize()
    ok, xo = o.solve(b)
    tol = 1e-11 if dtype == np.float64 else 2e-4
    panels = np.arange(1, N)                              # blocks j that are a pair's pivot block
    fac = {}
    for variant in (0, 3, 4, 5):
        g = gpu_solver(s, "LU", dtype)
        g.set_pair_kernel(variant)
        assert g.factorize(), g.last_error()
        ok, x = g.solve(b)
        assert ok
        err = np.abs(x[:ncols] - xo[:ncols]).max() / max(1.0, np.abs(xo[:ncols]).max())
        assert err < tol, (variant, err)
        f = g.factors()
        perm = f["perm"].reshape(N, n)[panels]
        assert np.array_equal(perm, o.internals()["perm"].reshape(N, n)[panels]), variant
        fac[variant] = dict(T=f["T"].reshape(N, -1)[panels, :2 * n * n], D=f["D"].reshape(N, -1)[panels],
                            E=f["E"].reshape(N, -1)[panels])
    for k in ("T", "D", "E"):
        assert np.array_equal(fac[3][k], fac[4][k]), k
    mfma_shape = nr <= 8 and nx <= 16 and ((n <= 32 and n + nr <= 48) or (n <= 48 and n + nr <= 64)
                                           or (n <= 64 and n + nr <= 80))
    if dtype == np.float64:
        mfma_shape = mfma_shape and n <= 48 and n + nr <= 64
    if mfma_shape:
        first = np.arange(1, N, 2) - 1                    # level-0 pairs among `panels`
        assert np.array_equal(fac[5]["T"][first], fac[4]["T"][first])
        rel = 1e-4 if dtype == np.float32 else 1e-12
        for k in ("D", "E"):
            ref = fac[4][k]
            assert np.abs(fac[5][k] - ref).max() <= rel * max(1.0, np.abs(ref).max()), k
