"""Synthetic bordered BABD systems in the BorderedCR block layout.

A system is a dict of numpy arrays holding the blocks as mathematical
matrices (row index first):

    D (N, n, n)   E (N, n, n)   B (N, n, nx)   C (N+1, nr, n)
    H (n+qr, Nc)  = [H0 | HN | Hq | Hp]   Cq (nr, qx)   F (nr, nx)

plus the integer dims N, n, qr, qx, nr, nx.  The global matrix is the one of
src/Alglin/BABD_BorderedCR.hxx:72-102: block row k holds D_k at block column
k, E_k at block column k+1 and B_k in the xb columns; then the n+qr boundary
rows H; then the nr border rows [C_0 ... C_N | Cq | F].

Generators:
  * stable_system           -- the closed-form 6-block system of
                               src_tests/test18-BorderedCR-coverage.cc:163-230
  * singular_fallback       -- test18:238-254 (exactly singular last block)
  * cinterface_system       -- the 7x7 / 8x8 systems of src_tests/test9-Cinterface.c
  * test7_system            -- src_tests/test7-BorderedCR.cc:39-100 (mt19937 seed 2)
  * gaussian_system         -- the seeded Gaussian workloads (D = shift*I + G/sqrt(n))
  * conditioned_system      -- D_k = Q diag(s) Q^T with log-uniform s
"""
from __future__ import annotations

import numpy as np


def empty_system(N, n, qr, qx, nr, nx, dtype=np.float64):
    Nc = 2 * n + qx + nx
    return dict(
        N=N, n=n, qr=qr, qx=qx, nr=nr, nx=nx,
        D=np.zeros((N, n, n), dtype), E=np.zeros((N, n, n), dtype),
        B=np.zeros((N, n, nx), dtype), C=np.zeros((N + 1, nr, n), dtype),
        H=np.zeros((n + qr, Nc), dtype), Cq=np.zeros((nr, qx), dtype),
        F=np.zeros((nr, nx), dtype),
    )


def dims(sys):
    N, n, qr, qx, nr, nx = (sys[k] for k in ("N", "n", "qr", "qx", "nr", "nx"))
    nrows = n * (N + 1) + qr + nr
    ncols = n * (N + 1) + qx + nx
    return nrows, ncols


def stable_system():
    """test18:163-230 with N=6, n=2, qr=1, qx=2, nr=3, nx=2 (test18:43-48)."""
    N, n, qr, qx, nr, nx = 6, 2, 1, 2, 3, 2
    s = empty_system(N, n, qr, qx, nr, nx)
    m = n + qr
    for k in range(N):
        s["D"][k] = [[5.0 + 0.2 * k, -0.08], [0.15, 4.7 + 0.2 * k]]
        s["E"][k] = [[-0.35, 0.06], [0.04, -0.28]]
        for j in range(nx):
            for i in range(n):
                s["B"][k, i, j] = 0.04 * (1 + k + i + 2 * j)
        for j in range(n):
            for i in range(nr):
                s["C"][k, i, j] = 0.03 * (1 + 2 * k + i + j)
    for j in range(n):
        for i in range(nr):
            s["C"][N, i, j] = 0.05 * (1 + i + j)
    for j in range(qx):
        for i in range(nr):
            s["Cq"][i, j] = 0.07 * (1 + i + 2 * j)
    for j in range(nx):
        for i in range(nr):
            s["F"][i, j] = 0.02 * (1 + i + j)
    s["F"][0, 0] += 3.2
    s["F"][1, 1] += 2.8
    H = s["H"]
    for j in range(n):
        for i in range(m):
            H[i, j] = 0.02 * (1 + i + j)
            H[i, n + j] = 0.03 * (1 + i + j)
    H[0, n + 0] += 5.1
    H[1, n + 1] += 4.9
    H[2, n + 0] += 0.2
    H[2, n + 1] -= 0.25
    for j in range(qx):
        for i in range(m):
            H[i, 2 * n + j] = 0.04 * (1 + i + j)
    for j in range(nx):
        for i in range(m):
            H[i, 2 * n + qx + j] = 0.05 * (1 + i + j)
    return s


def singular_fallback():
    """test18:238-254: N=1, n=1, qr=qx=0, nr=nx=1; the last block is singular."""
    s = empty_system(1, 1, 0, 0, 1, 1)
    s["D"][0, 0, 0] = 1
    s["H"][0, :] = [0, 1, 0]
    s["C"][0, 0, 0] = 1
    s["C"][1, 0, 0] = 1
    s["F"][0, 0] = 0
    return s


def cinterface_system(bordered: bool):
    """test9-Cinterface.c: nblock=2, n=2, qr=qx=1 (+ nr=nx=1 when bordered)."""
    nr = nx = 1 if bordered else 0
    s = empty_system(2, 2, 1, 1, nr, nx)
    for k in range(2):
        s["D"][k] = [[10, 2], [2, 40]]
        s["E"][k] = [[2, 3], [1, -1]]
    H0 = np.array([[0, 1], [1, 0], [1, 2]], float)
    HN = np.array([[20, 2], [0, -10], [2, 2]], float)
    Hq = np.array([[1], [1], [30]], float)
    s["H"][:, 0:2] = H0
    s["H"][:, 2:4] = HN
    s["H"][:, 4:5] = Hq
    if bordered:
        s["B"][:, :, 0] = 1.0        # full right column of ones (ldB = 7)
        s["H"][:, 5] = 1.0
        crow = np.array([1, -1, 1, -1, 1, -1, 1], float)
        for k in range(3):
            s["C"][k, 0, :] = crow[2 * k: 2 * k + 2]
        s["Cq"][0, 0] = crow[6]
        s["F"][0, 0] = -1.0
    return s


def test7_system(N=10000, n=8, qr=4, qx=4, nr=1, nx=1):
    """test7-BorderedCR.cc:39-100: std::mt19937(2) fill, diagonal dominance 1.01*n.

    numpy's legacy RandomState(seed) shares std::mt19937's seeding and output
    stream, so the blocks are the reference test's own values.
    """
    rs = np.random.RandomState(2)
    mx = 4294967295.0

    def rand(lo, hi, count):
        u = rs.randint(0, 2 ** 32, size=count, dtype=np.uint64).astype(np.float64) / mx
        return lo + (hi - lo) * u

    s = empty_system(N, n, qr, qx, nr, nx)
    diag = 1.01 * n
    Nc = 2 * n + qx + nx
    H = rand(-1, 0, (n + qr) * Nc).reshape(n + qr, Nc)
    for i in range(n + qr):
        H[i, i + n] += diag
    s["H"][:] = H
    for k in range(N):
        v = rand(-1, 0, 2 * n * n).reshape(n, n, 2)   # (i, j) loops, D then E
        s["D"][k] = v[:, :, 0]
        s["E"][k] = v[:, :, 1]
        s["D"][k][np.arange(n), np.arange(n)] += diag
        for i in range(n):
            s["B"][k, i, :] = rand(-0.1, 0.1, nx)
            s["C"][k, :, i] = rand(-0.1, 0.1, nr)
    for i in range(nr):
        s["F"][i, :] = rand(-0.1, 0.1, nx)
        s["Cq"][i, :] = rand(-0.1, 0.1, qx)
        s["F"][i, i] += diag
    s["C"][N, :, :] = 1.0
    return s


def gaussian_system(N, n, qr, qx, nr, nx, shift=4.0, seed=0, dtype=np.float64):
    """D_k = shift*I + G/sqrt(n), E_k = G/sqrt(n); borders and corners ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    s = empty_system(N, n, qr, qx, nr, nx, dtype)
    sc = 1.0 / np.sqrt(n)
    s["D"][:] = rng.standard_normal((N, n, n)) * sc
    s["D"][:, np.arange(n), np.arange(n)] += shift
    s["E"][:] = rng.standard_normal((N, n, n)) * sc
    s["B"][:] = rng.standard_normal((N, n, nx))
    s["C"][:] = rng.standard_normal((N + 1, nr, n))
    s["H"][:] = rng.standard_normal(s["H"].shape)
    s["Cq"][:] = rng.standard_normal((nr, qx))
    s["F"][:] = rng.standard_normal((nr, nx))
    return s


def conditioned_system(N, n, qr, qx, nr, nx, kmin=1e2, kmax=1e6, seed=0):
    """D_k = Q diag(s) Q^T with s log-uniform so cond(D_k) in [kmin, kmax]."""
    rng = np.random.default_rng(seed)
    s = gaussian_system(N, n, qr, qx, nr, nx, shift=0.0, seed=seed + 1)
    kappa = np.exp(rng.uniform(np.log(kmin), np.log(kmax), N))
    for k in range(N):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        sv = np.exp(np.linspace(0.0, -np.log(kappa[k]), n))
        rng.shuffle(sv)
        s["D"][k] = (Q * sv) @ Q.T
    return s


def dense(sys) -> np.ndarray:
    """Assemble the full matrix (brute-force reference for small systems)."""
    N, n, qr, qx, nr, nx = (sys[k] for k in ("N", "n", "qr", "qx", "nr", "nx"))
    nrows, ncols = dims(sys)
    A = np.zeros((nrows, ncols), np.result_type(sys["D"], np.float64))
    cb = (N + 1) * n + qx  # first xb column
    for k in range(N):
        r = slice(k * n, (k + 1) * n)
        A[r, k * n:(k + 1) * n] = sys["D"][k]
        A[r, (k + 1) * n:(k + 2) * n] = sys["E"][k]
        A[r, cb:cb + nx] = sys["B"][k]
    r = slice(N * n, N * n + n + qr)
    H = sys["H"]
    A[r, 0:n] = H[:, 0:n]
    A[r, N * n:(N + 1) * n] = H[:, n:2 * n]
    A[r, (N + 1) * n:(N + 1) * n + qx] = H[:, 2 * n:2 * n + qx]
    A[r, cb:cb + nx] = H[:, 2 * n + qx:]
    r0 = (N + 1) * n + qr
    r = slice(r0, r0 + nr)
    for k in range(N + 1):
        A[r, k * n:(k + 1) * n] = sys["C"][k]
    A[r, (N + 1) * n:(N + 1) * n + qx] = sys["Cq"]
    A[r, cb:cb + nx] = sys["F"]
    return A


def mv(sys, x: np.ndarray) -> np.ndarray:
    """y = A x in float64 without assembling A (vectorised block products)."""
    N, n, qr, qx, nr, nx = (sys[k] for k in ("N", "n", "qr", "qx", "nr", "nx"))
    nrows, ncols = dims(sys)
    x = np.asarray(x, np.float64)
    squeeze = x.ndim == 1
    X = x.reshape(ncols, -1)
    Y = np.zeros((nrows, X.shape[1]))
    xb = X[(N + 1) * n + qx:]
    xblk = X[: (N + 1) * n].reshape(N + 1, n, -1)
    D = sys["D"].astype(np.float64)
    E = sys["E"].astype(np.float64)
    Bm = sys["B"].astype(np.float64)
    Yb = np.matmul(D, xblk[:-1]) + np.matmul(E, xblk[1:]) + np.matmul(Bm, xb[None, :, :])
    Y[: N * n] = Yb.reshape(N * n, -1)
    H = sys["H"].astype(np.float64)
    xq = X[(N + 1) * n:(N + 1) * n + qx]
    Y[N * n:N * n + n + qr] = (H[:, :n] @ xblk[0] + H[:, n:2 * n] @ xblk[N]
                               + H[:, 2 * n:2 * n + qx] @ xq + H[:, 2 * n + qx:] @ xb)
    if nr > 0:
        C = sys["C"].astype(np.float64)
        Y[(N + 1) * n + qr:] = (np.einsum("kij,kjr->ir", C, xblk) + sys["Cq"] @ xq + sys["F"] @ xb)
    return Y[:, 0] if squeeze else Y


def uneven_batch_dims(nsys, seed=0, nmin=4, nmax=48, Nmin=16, Nmax=4096, method=0):
    """Shapes of a batch of uneven systems: n uniform in {nmin..nmax}, N
    log-uniform in [Nmin, Nmax] (rounded), qr = qx = n // 4, nr = nx = 2.
    Rows (N, n, qr, qx, nr, nx, method)."""
    rng = np.random.default_rng(seed)
    n = rng.integers(nmin, nmax + 1, nsys)
    N = np.rint(np.exp(rng.uniform(np.log(Nmin), np.log(Nmax), nsys))).astype(np.int64)
    q = n // 4
    return np.stack([N, n, q, q, np.full(nsys, 2), np.full(nsys, 2), np.full(nsys, method)], 1).astype(np.int32)

