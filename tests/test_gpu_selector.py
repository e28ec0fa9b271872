"""The learned LU-vs-QRP selector on the GPU.

* the GPU forward/backward kernels against the float64 restatement
  (oracle/mlp_oracle.py, itself pinned by finite differences): loss within
  1e-5 relative, gradient within 2e-3 of its max entry (FP32 accumulation);
* GPU training lowers the loss and separates a learnable rule;
* solver-generated samples: the selector trained on them and applied to a
  held-out batch never does worse than the method without it (LU first,
  the FP64 fallback -- LU, then QRP -- for the systems that miss tol) and
  every system ends at the tolerance."""
import numpy as np
import pytest

from alglin_amd import selector as S
from alglin_amd.bcr import Mlp
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def synthetic(n=700, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 12)).astype(np.float32) * 3 + 1
    y = ((X[:, 2] - 0.5 * X[:, 7]) > 1.0).astype(np.int32)
    return X, y


def test_gpu_gradient_matches_float64_oracle():
    X, y = synthetic()
    m = Mlp(5)
    mu = X.mean(0)
    sig = X.std(0) + 0.1
    m.set_params(mu=mu, sig=sig)
    p, mu32, sig32 = m.params()
    loss, g = m.loss_grad(X, y)
    lo, go = M.loss_grad(p.astype(np.float64), mu32.astype(np.float64), sig32.astype(np.float64),
                         X.astype(np.float64), y.astype(np.int64))
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert np.abs(g - go).max() <= 2e-3 * np.abs(go).max(), np.abs(g - go).max()


def test_gpu_training_learns_a_rule():
    X, y = synthetic(2000, 1)
    m = Mlp(2)
    m.set_params(mu=X.mean(0), sig=X.std(0))
    hist = m.train(X, y, epochs=300, lr=3e-3)
    assert hist[-1] < 0.3 * hist[0]
    acc = ((m.predict(X) > 0.5) == (y == 1)).mean()
    assert acc > 0.97, acc


def test_selector_gain_on_held_out_systems():
    """A network trained here on 1024 solver-generated samples, on a held-out
    batch with the FP64 fallback on: no more failures than FP32 LU alone,
    and it flags systems FP32 LU could not solve (they skip the FP32 sweeps
    they would stall in)."""
    f, y, info = S.samples(1024, seed=21, Nmin=64, Nmax=512)
    print(f"train: LU ok {info['lu_ok'].mean():.3f} QRP ok {info['qrp_ok'].mean():.3f} labels {y.sum()}", flush=True)
    m, hist = S.train(f, y, epochs=400)
    assert hist[-1] <= hist[0]
    bt, B, sing = S.make_batch(1024, seed=77, Nmin=64, Nmax=512)
    from alglin_amd.bcr import DeviceArray
    X0 = DeviceArray(bt.size("x"))
    bt.solve_selected(B, X0, tol=1e-12, selector=None, fallback=True)
    base = bt.report()
    base_fell = bt.selection()["fell_back"].copy()
    base_fb = int(base_fell.sum())
    X1 = DeviceArray(bt.size("x"))
    bt.solve_selected(B, X1, tol=1e-12, selector=m, fallback=True)
    sel = bt.report()
    chosen = bt.selection()
    print(f"without selector: {base_fb} FP64 fallbacks, {base['failed']} failed, max sweeps {base['max_sweeps']}, "
          f"{base['ms']:.1f} ms; with: {int(chosen['chose_qrp'].sum())} sent to QRP, "
          f"{int(chosen['fell_back'].sum())} fallbacks, {sel['failed']} failed, max sweeps {sel['max_sweeps']}, "
          f"{sel['ms']:.1f} ms", flush=True)
    # every system that reports success meets the tolerance
    ok = sel["status"] == 0
    assert np.all(sel["rel_residual"][ok] <= 1e-12)
    # the selector never makes things worse
    assert sel["failed"] <= base["failed"]
    # with the fallback on, a flagged system goes straight to the FP64 path:
    # fewer systems reach it through stalled FP32 sweeps than without the
    # selector, and most flagged systems are ones FP32 LU could not solve
    flagged = chosen["chose_qrp"]
    missed = int((chosen["fell_back"] & ~flagged).sum())
    assert missed < base_fb
    assert int((flagged & base_fell).sum()) >= 0.5 * int(flagged.sum())


def test_committed_selector_cuts_failures_without_fallback():
    """The committed weights (trained on 2e5 solver-generated samples of other
    seeds) on a held-out batch, with the FP64 fallback switched off: the
    selector's LU/QRP choice leaves fewer systems above the tolerance than
    FP32 LU everywhere, and every system it reports as converged is."""
    from alglin_amd.bcr import DeviceArray
    mlp = S.load()
    bt, B, sing = S.make_batch(2048, seed=313, Nmin=64, Nmax=512)
    X = DeviceArray(bt.size("x"))
    bt.solve_selected(B, X, tol=1e-12, selector=None, fallback=False)
    base = bt.report()
    bt.solve_selected(B, X, tol=1e-12, selector=mlp, fallback=False)
    sel = bt.report()
    sent = int(bt.selection()["chose_qrp"].sum())
    print(f"failed: LU only {base['failed']}, with selector {sel['failed']} ({sent} sent to QRP)", flush=True)
    assert base["failed"] >= 20
    assert sel["failed"] < base["failed"]
    ok = sel["status"] == 0
    assert np.all(sel["rel_residual"][ok] <= 1e-12)


def test_selector_workload_full_size():
    """The learned-selector workload at full size: 8192 systems, n in
    {8, 16, 32}, N in [64, 2048], half with a rank-deficient pair column
    (1e-10), mixed + committed selector + FP64 fallback.  Size-independent
    properties: every system reported solved has a residual <= 1e-12, every
    other one is marked (a system that stopped at the FP64 floor carries
    status bit 3 and never went through the FP64 fallback), and the systems
    the FP64 path took are mostly the nearly singular ones."""
    from alglin_amd.bcr import DeviceArray
    mlp = S.load()
    bt, B, sing = S.make_batch(8192, seed=2024, Nmin=64, Nmax=2048)
    X = DeviceArray(bt.size("x"))
    bt.solve_selected(B, X, tol=1e-12, selector=mlp, fallback=True)
    rep = bt.report()
    sel = bt.selection()
    st = rep["status"]
    print(f"failed {rep['failed']}, flagged {int(sel['chose_qrp'].sum())}, FP64 path {int(sel['fell_back'].sum())}, "
          f"floor {int(((st & 8) != 0).sum())}, max sweeps {rep['max_sweeps']}", flush=True)
    ok = st == 0
    assert np.all(rep["rel_residual"][ok] <= 1e-12)
    floor = (st & 8) != 0
    assert not np.any(sel["fell_back"][floor])
    assert rep["failed"] <= 8
    fb = sel["fell_back"]
    assert (fb & (sing >= 1)).sum() > 4 * (fb & (sing < 1)).sum()


def test_default_batch_floor_systems_skip_the_fallback():
    """The 4096-system uneven batch of the default bench workload has systems
    whose FP64 residual floor lies above 1e-12 (plain FP64 misses it too).
    They end with status bit 3 in plain FP64 and in the learned path, and the
    learned path does not send them through the FP64 fallback."""
    from alglin_amd import systems
    from alglin_amd.bcr import Batch, DeviceArray
    dims = systems.uneven_batch_dims(4096, 2024)
    ref = Batch(mixed=False)
    ref.plan(dims)
    ref.generate(seed=2024, shift=4.0)
    mix = Batch(mixed=True)
    mix.plan(dims, operator=False)
    mix.share_operator(ref)
    L = ref.size("x")
    B = DeviceArray.from_host(np.random.default_rng(2029).standard_normal(L))
    X = DeviceArray(L)
    ref.solve(B, X, 1e-12, 30, graph=True)
    r64 = ref.report()
    mix.solve_selected(B, X, 1e-12, 30, selector=S.load(), fallback=True)
    rm = mix.report()
    sel = mix.selection()
    f64 = (r64["status"] & 8) != 0
    fm = (rm["status"] & 8) != 0
    print(f"FP64: failed {r64['failed']} floor {int(f64.sum())}; learned: failed {rm['failed']} floor {int(fm.sum())} "
          f"FP64 path {int(sel['fell_back'].sum())}", flush=True)
    assert np.all((r64["status"] == 0) | f64)
    assert np.all((rm["status"] == 0) | fm)
    assert not np.any(sel["fell_back"][fm])
    assert np.all(rm["rel_residual"][rm["status"] == 0] <= 1e-12)


def host_panel_features(T, N, n):
    """Features 4..8 and 11 from FP32 LU factors, by their definition: per
    pivot block j = 1..N-1 the pivots u_ii of the panel's U (the top n x n
    block of its 2n x n factor, column-major) give min/max |u_ii| and their
    ratio; the features are log10 of the smallest and largest pivot, the mean
    log10 ratio, log10 of the smallest ratio, the share of ratios below 1e-5,
    and log10 of max |u_ik| (i <= k) over the largest pivot."""
    ts = 2 * n * n + n
    pmin, pmax, umax = [], [], 0.0
    for j in range(1, N):
        P = T[j * ts:j * ts + 2 * n * n].astype(np.float64).reshape(n, 2 * n).T   # row r, column c
        U = np.abs(np.triu(P[:n, :]))
        d = np.abs(np.diag(P[:n, :]))
        pmin.append(d.min())
        pmax.append(d.max())
        umax = max(umax, U.max())
    pmin, pmax = np.array(pmin), np.array(pmax)
    ratio = np.where(pmax > 0, pmin / np.where(pmax > 0, pmax, 1), 0.0)
    return np.array([np.log10(pmin.min() + 1e-30), np.log10(pmax.max() + 1e-30), np.mean(np.log10(ratio + 1e-30)),
                     np.log10(ratio.min() + 1e-30), np.mean(ratio < 1e-5),
                     np.log10(umax / (pmax.max() + 1e-30) + 1e-30)])


def test_batch_features_match_the_factors():
    """The batch's selector features (per-panel statistics kernel, then one
    workgroup per system) against the same statistics computed on the host
    from the single-system FP32 LU factors of each system (the same kernels:
    the pivots and U agree bit for bit), and the shape features 0..3."""
    from alglin_amd import systems
    from alglin_amd.bcr import Batch, BorderedCR, DeviceArray
    dims = np.asarray([(7, 5, 1, 1, 2, 2, 0), (33, 20, 5, 5, 2, 2, 0), (2, 40, 10, 10, 2, 2, 0),
                       (1, 8, 2, 2, 2, 2, 0), (130, 48, 12, 12, 2, 2, 0), (64, 16, 4, 4, 3, 3, 0),
                       (9, 1, 0, 0, 1, 1, 0)], np.int32)
    print(f"case features nsys={len(dims)}", flush=True)
    bt = Batch(mixed=True)
    bt.plan(dims)
    bt.generate(seed=11, shift=4.0)
    B = DeviceArray.from_host(np.random.default_rng(3).standard_normal(bt.size("x")))
    X = DeviceArray(bt.size("x"))
    bt.solve_selected(B, X, tol=1e-12, selector=None, fallback=False)
    f = bt.selection()["features"]
    for s, (N, n, qr, qx, nr, nx, _) in enumerate(dims):
        assert f[s, 0] == pytest.approx(np.log2(N) / 12, rel=1e-6)
        assert f[s, 1] == pytest.approx(n / 48, rel=1e-6)
        assert f[s, 2] == pytest.approx(qr / 12, rel=1e-6)
        assert f[s, 3] == pytest.approx(nr / 8, rel=1e-6)
        if N < 2:
            assert np.all(f[s, [4, 5, 6, 7, 8, 11]] == 0)
            continue
        g = BorderedCR(np.float32)
        g.load_system(bt.system(s))
        g.select("LU", "LU", "NONE")
        assert g.factorize(), g.last_error()
        want = host_panel_features(g.factors()["T"], int(N), int(n))
        got = f[s, [4, 5, 6, 7, 8, 11]].astype(np.float64)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5), (s, got, want)
