"""Batched uneven systems on the GPU: every system of the batch must equal the
single-system algorithm (the oracle's BorderedCR at num_parallel_block = 1,
same pair method, LU last block) on the same inputs, FP64 within 1e-10
relative on x; mixed precision must reach the requested residual per
system; graph replays must equal the eager launches bit for bit."""
import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import AlglinError, Batch, DeviceArray

pytestmark = pytest.mark.gpu

METHODS = ["LU", "QR", "QRP"]


def mixed_dims():
    """Shapes covering every kernel class: n in 1..48, N from 1 (no
    reduction level) to 300, borders 0..3, all three pair methods."""
    rows = []
    ns = [1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 17, 24, 32, 33, 40, 48]
    Ns = [1, 2, 3, 5, 16, 37, 100, 300]
    k = 0
    for n in ns:
        for N in Ns[(k % 3)::3]:
            q = n // 4
            r = k % 4
            rows.append((N, n, q, q, r, r, k % 3))
            k += 1
    return np.asarray(rows, np.int32)


def random_rhs(bt, seed):
    L = bt.size("x")
    return np.random.default_rng(seed).standard_normal(L)


def oracle_x(oracle_mod, sysd, method, b):
    o = oracle_mod.OracleBorderedCR(np.float64)
    o.allocate(sysd["N"], sysd["n"], sysd["qr"], sysd["qx"], sysd["nr"], sysd["nx"])
    o.load(sysd)
    o.select(METHODS[method], "LU")
    assert o.factorize(), o.last_error()
    ok, x = o.solve(b)
    assert ok
    return x


@pytest.mark.parametrize("kernel", ["tuned", "mfma"])
def test_batch_fp64_matches_oracle_per_system(oracle_mod, kernel):
    """FP64 batch against the oracle per system, with the tuned class kernels
    and with the matrix-core pair forced on the 16 < n <= 32 and 32 < n <= 48
    LU classes (FP64 MFMA for the right-hand blocks)."""
    dims = mixed_dims()
    print(f"case batch fp64 nsys={len(dims)} kernel={kernel}", flush=True)
    bt = Batch(mixed=False)
    bt.plan(dims)
    if kernel == "mfma":
        bt.set_class_kernel(3, 0, fp32_factors=False)
        bt.set_class_kernel(4, 0, fp32_factors=False)
    bt.generate(seed=7, shift=4.0)
    b = random_rhs(bt, 1)
    B = DeviceArray.from_host(b)
    X = DeviceArray(bt.size("x"))
    # no refinement sweep: the first FP64 solve itself must match the oracle
    # (refinement would hide wrong factors that still precondition well)
    rc = bt.solve(B, X, tol=1e-12, max_iter=0, graph=False)
    rep = bt.report()
    assert rc == 0, rep
    assert rep["sweeps"].max() == 0
    x = X.to_host()
    for s, row in enumerate(dims):
        sysd = bt.system(s)
        nrows, ncols = systems.dims(sysd)
        o = bt.offset("x", s)
        bs = b[o:o + nrows]
        xo = oracle_x(oracle_mod, sysd, int(row[6]), bs)
        xs = x[o:o + ncols]
        err = np.abs(xs - xo[:ncols]).max() / max(1.0, np.abs(xo[:ncols]).max())
        assert err < 1e-10, (s, tuple(row), err)
        res = np.linalg.norm(systems.mv(sysd, xs) - bs) / np.linalg.norm(bs)
        # the GPU's own FP64 residual agrees with numpy's (both are FP64
        # evaluations of b - A x, which cancels: agreement to ~1e-14 absolute)
        assert res <= 1e-12 and abs(res - rep["rel_residual"][s]) <= 0.1 * res + 1e-14, (s, res)


@pytest.mark.parametrize("rows", [
    [(5, 4, 1, 1, 1, 1, 0), (37, 9, 2, 2, 2, 2, 0), (16, 16, 4, 4, 3, 3, 0)],      # last blocks Nc <= 64
    [(9, 64, 16, 16, 8, 8, 0), (20, 61, 15, 15, 3, 3, 0), (3, 8, 2, 2, 1, 1, 0)],  # Nc up to 152
], ids=["nc64", "nc152"])
def test_batch_last_block_solve_widths(oracle_mod, rows):
    """The one-wavefront last-block solve holds one, two or four rows per
    lane by the batch's largest last block: FP64 batches whose last blocks
    need one and four rows per lane against the oracle per system."""
    dims = np.asarray(rows, np.int32)
    print(f"case batch fp64 last-block widths Ncmax={max(2 * r[1] + r[2] + r[4] for r in rows)}", flush=True)
    bt = Batch(mixed=False)
    bt.plan(dims)
    bt.generate(seed=11, shift=4.0)
    b = random_rhs(bt, 3)
    B = DeviceArray.from_host(b)
    X = DeviceArray(bt.size("x"))
    rc = bt.solve(B, X, tol=1e-12, max_iter=0, graph=False)
    assert rc == 0, bt.report()
    x = X.to_host()
    for s, row in enumerate(dims):
        sysd = bt.system(s)
        nrows, ncols = systems.dims(sysd)
        o = bt.offset("x", s)
        xo = oracle_x(oracle_mod, sysd, int(row[6]), b[o:o + nrows])
        err = np.abs(x[o:o + ncols] - xo[:ncols]).max() / max(1.0, np.abs(xo[:ncols]).max())
        assert err < 1e-10, (s, tuple(row), err)


@pytest.mark.parametrize("variant", [0, 4, 5])
def test_batch_mixed_reaches_tolerance_and_graph_is_exact(variant):
    """Mixed batch to 1e-12 with each choice of kernels for the 16 < n <= 48
    LU classes (default, the rank-1 kernels, the matrix-core pair)."""
    dims = mixed_dims()
    dims[:, 6] = 0                                  # LU everywhere
    print(f"case batch mixed nsys={len(dims)} pair kernel {variant}", flush=True)
    bt = Batch(mixed=True)
    bt.plan(dims)
    bt.set_pair_kernel(variant)
    bt.generate(seed=11, shift=4.0)
    b = random_rhs(bt, 2)
    B = DeviceArray.from_host(b)
    X0 = DeviceArray(bt.size("x"))
    assert bt.solve(B, X0, tol=1e-12, max_iter=30, graph=False) == 0, bt.report()
    rep0 = bt.report()
    x0 = X0.to_host()
    assert rep0["max_rel_residual"] <= 1e-12
    assert rep0["sweeps"].max() >= 1                 # FP32 factors alone do not get there
    for s in range(0, len(dims), 5):
        sysd = bt.system(s)
        nrows, ncols = systems.dims(sysd)
        o = bt.offset("x", s)
        res = np.linalg.norm(systems.mv(sysd, x0[o:o + ncols]) - b[o:o + nrows]) / np.linalg.norm(b[o:o + nrows])
        assert res <= 2e-12, (s, res)
    # graph replays: every replay, with eager work between them, gives the
    # same bits (the eager path narrows its sweeps to the unconverged systems,
    # the graph keeps the whole batch, so the two agree to tolerance only)
    X1 = DeviceArray(bt.size("x"))
    first = None
    for rep in range(3):
        assert bt.solve(B, X1, tol=1e-12, max_iter=30, graph=True) == 0
        xg = X1.to_host()
        assert bt.report()["max_rel_residual"] <= 1e-12
        if first is None:
            first = xg
        assert np.array_equal(xg, first), rep
        assert bt.solve(B, X0, tol=1e-12, max_iter=30, graph=False) == 0
        assert np.array_equal(X0.to_host(), x0)
    assert np.abs(first - x0).max() <= 1e-9 * max(1.0, np.abs(x0).max())


def test_batch_rejects_rectangular_last_block():
    bt = Batch(mixed=False)
    with pytest.raises(AlglinError):
        bt.plan([(10, 4, 2, 1, 1, 1, 0)])


def test_time_factorize_covers_every_launch():
    """alglin_batch_time_factorize: one row per (level, kernel class) launch
    of one factorization, each with a positive duration and the algorithmic
    bytes and flops of the pairs it eliminates; the per-launch bytes add up
    to the whole batch's pair bytes (FP32 factors, FP64 level-0 reads)."""
    dims = np.asarray([(N, n, n // 4, n // 4, 2, 2, 0) for N, n in
                       [(300, 8), (300, 20), (200, 40), (37, 4), (100, 33)]], np.int32)
    bt = Batch(mixed=True)
    bt.plan(dims)
    bt.generate(seed=5, shift=4.0)
    rows = bt.time_factorize()
    levels = int(np.ceil(np.log2(dims[:, 0].max())))
    assert rows.shape[1] == 6 and len(rows) >= levels
    assert set(np.unique(rows[:, 0]).astype(int)) == set(range(levels))
    assert np.all(rows[:, 3] > 0) and np.all(rows[:, 4] > 0) and np.all(rows[:, 5] > 0)
    want = 0.0
    for N, n, qr, qx, nr, nx, _ in dims:
        L = 1
        while L < N:
            p = (N + L - 1) // (2 * L)
            rd = 4 * n * n + 2 * n * nx + 2 * nr * n
            wr = 6 * n * n + 2 * n * nx + 3 * nr * n + nr * nx
            want += p * (rd * (8 if L == 1 else 4) + wr * 4 + 4 * n)
            L *= 2
    assert abs(rows[:, 4].sum() - want) <= 1e-9 * want


def test_selected_fallback_stages_reach_tolerance():
    """solve_selected with the FP64 fallback on systems with nearly singular
    pairs: every system ends at the tolerance; the ones FP32 LU + refinement
    cannot bring there go to plain FP64 LU (stage 1), and only what that
    misses to FP64 QRP (stage 2)."""
    from alglin_amd import selector as S
    bt, B, sing = S.make_batch(512, seed=91, Nmin=64, Nmax=512)
    X = DeviceArray(bt.size("x"))
    rc = bt.solve_selected(B, X, tol=1e-12, selector=None, fallback=True)
    rep = bt.report()
    sel = bt.selection()
    print(f"stages {np.bincount(sel['fallback_stage'], minlength=3)} failed {rep['failed']}", flush=True)
    assert rc == 0 and rep["failed"] == 0
    assert np.all(rep["rel_residual"] <= 1e-12)
    st = sel["fallback_stage"]
    assert set(np.unique(st)) <= {0, 1, 2}
    assert (st >= 1).sum() > 0
    # the systems that need the fallback are mostly the nearly singular ones
    assert (st[sing >= 1] >= 1).sum() > (st[sing < 1] >= 1).sum()


@pytest.mark.parametrize("streams,fork_level", [(1, 0), (2, 1)])
def test_batch_streams_give_the_same_bits(streams, fork_level):
    """The stream layout (one stream, or the n > 32 class on a second stream
    from a given level) changes only the order in which independent systems'
    launches run: the mixed solution is bit-identical to the default's."""
    dims = mixed_dims()
    print(f"case batch streams={streams} fork_level={fork_level}", flush=True)
    bt = Batch(mixed=True)
    bt.plan(dims)
    bt.generate(seed=9, shift=4.0)
    B = DeviceArray.from_host(random_rhs(bt, 2))
    X0, X1 = DeviceArray(bt.size("x")), DeviceArray(bt.size("x"))
    assert bt.solve(B, X0, tol=1e-12, max_iter=30, graph=True) == 0, bt.report()
    bt.set_streams(streams, fork_level)
    assert bt.solve(B, X1, tol=1e-12, max_iter=30, graph=True) == 0, bt.report()
    assert np.array_equal(X0.to_host(), X1.to_host())
