"""Writes the golden fixtures under tests/golden/ (run once; the outputs are committed).

It is a second, independent restatement of alglin::BorderedCR's pure cyclic
reduction (src/BABD_BorderedCR.cc:1055-1171 with one block range,
cc:832-834) in which every dense step is the LAPACK routine the reference
calls, taken from scipy.linalg.lapack (OpenBLAS/reference LAPACK, the
family the reference links by default, Rakefile:26-33):

    buildT  (cc:845-898):  ?getrf / ?geqrf / ?geqp3 + permutation_to_exchange
    applyT  (cc:917-982):  ?laswp + ?trsm + ?gemm, or ?ormqr('L','T')
    C rows  (cc:1107-1154): QRP column exchanges, ?trsm('R','U'), ?gemm
    last    (cc:1281-1374): assembly of the Nr x Nc block, ?getrf / ?getrs

The fixtures hold, per case: the packed input blocks, the factor arrays after
factorize() (T with ipiv/tau, Perm, D, E, B, C, F, the last block) and the
solution of b = A*x_ref.  The oracle's own dense kernels are checked against
them in tests/test_oracle_golden.py.

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np
from scipy.linalg import blas, lapack

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from alglin_amd import systems  # noqa: E402


def lp(dtype, name):
    return getattr(lapack, ("d" if dtype == np.float64 else "s") + name)


def bl(dtype, name):
    return getattr(blas, ("d" if dtype == np.float64 else "s") + name)


class LapackCR:
    """Pure CR on one range, dense steps through scipy LAPACK/BLAS."""

    def __init__(self, s, internal, dtype):
        self.s = s
        self.internal = internal
        self.dt = dtype
        N, n, nr, nx = s["N"], s["n"], s["nr"], s["nx"]
        f = lambda a: np.array(a, dtype=dtype, order="F")  # noqa: E731
        # block lists of Fortran-ordered matrices (math orientation)
        self.D = [f(s["D"][k]) for k in range(N)]
        self.E = [f(s["E"][k]) for k in range(N)]
        self.B = [f(s["B"][k]) for k in range(N)]
        self.C = [f(s["C"][k]) for k in range(N + 1)]
        self.F = f(s["F"])
        self.T = [np.zeros((2 * n, n), dtype, order="F") for _ in range(N)]
        self.tau = [np.zeros(n, dtype) for _ in range(N)]
        self.perm = [np.zeros(n, np.int32) for _ in range(N)]

    def build_T(self, j, T):
        n = self.s["n"]
        if self.internal == "LU":
            lu, piv, info = lp(self.dt, "getrf")(T)
            assert info == 0
            self.T[j] = np.asfortranarray(lu)
            self.perm[j] = (piv + 1).astype(np.int32)  # scipy returns 0-based pivots
        elif self.internal == "QR":
            qr, tau, work, info = lp(self.dt, "geqrf")(T)
            assert info == 0
            self.T[j] = np.asfortranarray(qr)
            self.tau[j] = tau
        else:
            qr, jpvt, tau, work, info = lp(self.dt, "geqp3")(T)
            assert info == 0
            self.T[j] = np.asfortranarray(qr)
            self.tau[j] = tau
            P = list(jpvt)          # 1-based permutation (hxx:265-275)
            S = np.zeros(n, np.int32)
            for i in range(n):
                jj = i
                while jj < n and P[jj] != i + 1:
                    jj += 1
                P[jj], P[i] = P[i], P[jj]
                S[i] = jj
            self.perm[j] = S

    def apply_T(self, j, top, bottom):
        n = self.s["n"]
        W = np.asfortranarray(np.vstack([top, bottom]).astype(self.dt))
        T = self.T[j]
        if self.internal == "LU":
            for i in range(n):                          # ?laswp rows 1..n
                p = self.perm[j][i] - 1
                if p != i:
                    W[[i, p], :] = W[[p, i], :]
            W[:n] = bl(self.dt, "trsm")(1.0, T[:n, :n], W[:n], side=0, lower=1, trans_a=0, diag=1)
            W[n:] = bl(self.dt, "gemm")(-1.0, T[n:, :n], W[:n], 1.0, W[n:])
        else:
            cq, work, info = lp(self.dt, "ormqr")("L", "T", T, self.tau[j], W, max(1, 64 * W.shape[1]))
            assert info == 0
            W = np.asfortranarray(cq)
        return W[n:].copy(order="F"), W[:n].copy(order="F")

    def factorize(self):
        s = self.s
        N, n, nr, nx = s["N"], s["n"], s["nr"], s["nx"]
        k = 1
        while k < N:
            for j in range(k, N, 2 * k):
                jp, jpp = j - k, min(j + k, N)
                T = np.asfortranarray(np.vstack([self.E[jp], self.D[j]]))
                self.build_T(j, T)
                zero = np.zeros((n, n), self.dt, order="F")
                self.D[jp], self.D[j] = self.apply_T(j, self.D[jp], zero)
                self.E[jp], self.E[j] = self.apply_T(j, zero, self.E[j])
                if nx > 0:
                    self.B[jp], self.B[j] = self.apply_T(j, self.B[jp], self.B[j])
                if nr > 0:
                    Cj = self.C[j]
                    if self.internal == "QRP":
                        S = self.perm[j]
                        for i in range(n - 1, -1, -1):
                            if S[i] > i:
                                Cj[:, [i, S[i]]] = Cj[:, [S[i], i]]
                    R = np.triu(self.T[j][:n, :n])
                    Cj = bl(self.dt, "trsm")(1.0, R, Cj, side=1, lower=0, trans_a=0, diag=0)
                    self.C[j] = np.asfortranarray(Cj)
                    self.C[jp] = self.C[jp] - Cj @ self.D[j]
                    self.C[jpp] = self.C[jpp] - Cj @ self.E[j]
                    if nx > 0:
                        self.F = self.F - Cj @ self.B[j]
            k *= 2
        # last block (cc:1304-1316)
        qr, qx = s["qr"], s["qx"]
        Nr, Nc = 2 * n + nr + qr, 2 * n + nx + qx
        Hm = np.zeros((Nr, Nc), self.dt)
        Hm[:n, :n] = self.D[0]
        Hm[:n, n:2 * n] = self.E[0]
        Hm[:n, 2 * n + qx:] = self.B[0]
        Hm[n:2 * n + qr, :] = s["H"]
        o = 2 * n + qr
        Hm[o:, :n] = self.C[0]
        Hm[o:, n:2 * n] = self.C[N]
        Hm[o:, 2 * n:2 * n + qx] = s["Cq"]
        Hm[o:, 2 * n + qx:] = self.F
        self.Hmat = Hm
        self.klast = k

    def solve(self, b):
        """Forward / last (LU) / backward, cc:1489-1542 with one range."""
        s = self.s
        N, n, nr, nx, qr, qx = s["N"], s["n"], s["nr"], s["nx"], s["qr"], s["qx"]
        x = np.array(b, self.dt)
        xbr = (N + 1) * n + qr
        work = np.zeros(nr, self.dt)
        k = 1
        while k < N:
            for j in range(k, N, 2 * k):
                jp = j - k
                top, bot = self.apply_T(j, x[jp * n:(jp + 1) * n, None], x[j * n:(j + 1) * n, None])
                x[jp * n:(jp + 1) * n] = top[:, 0]
                x[j * n:(j + 1) * n] = bot[:, 0]
                if nr > 0:
                    work -= self.C[j] @ x[j * n:(j + 1) * n]
            k *= 2
        x[xbr:xbr + nr] += work
        X0 = (N - 1) * n
        x[[*range(n)] + [*range(X0, X0 + n)]] = x[[*range(X0, X0 + n)] + [*range(n)]]
        lu, piv, info = lp(self.dt, "getrf")(self.Hmat)
        assert info == 0
        sol, info = lp(self.dt, "getrs")(lu, piv, x[X0:X0 + self.Hmat.shape[0]])
        x[X0:X0 + self.Hmat.shape[1]] = sol
        x[[*range(n)] + [*range(X0, X0 + n)]] = x[[*range(X0, X0 + n)] + [*range(n)]]
        xn = x[(N + 1) * n + qx:(N + 1) * n + qx + nx]
        k = self.klast
        while True:
            k //= 2
            if k == 0:
                break
            for j in range(k, N, 2 * k):
                jp, jpp = j - k, min(j + k, N)
                xj = x[j * n:(j + 1) * n] - self.D[j] @ x[jp * n:(jp + 1) * n] - self.E[j] @ x[jpp * n:(jpp + 1) * n]
                if nx > 0:
                    xj = xj - self.B[j] @ xn
                R = np.triu(self.T[j][:n, :n])
                xj = bl(self.dt, "trsv")(R, xj, lower=0, trans=0, diag=0)
                if self.internal == "QRP":
                    S = self.perm[j]
                    for i in range(n):
                        if S[i] > i:
                            xj[[i, S[i]]] = xj[[S[i], i]]
                x[j * n:(j + 1) * n] = xj
        return x[: (N + 1) * n + qx + nx]

    def packed_internals(self):
        """Factor arrays in the reference's internal layout (hxx:148-228)."""
        s = self.s
        N, n = s["N"], s["n"]
        Tsize = 2 * n * n + n
        Tp = np.zeros(N * Tsize, self.dt)
        for j in range(N):
            Tp[j * Tsize:j * Tsize + 2 * n * n] = self.T[j].ravel(order="F")
            Tp[j * Tsize + 2 * n * n:(j + 1) * Tsize] = self.tau[j]
        return dict(
            T=Tp,
            perm=np.concatenate(self.perm).astype(np.int32),
            D=np.concatenate([a.ravel(order="F") for a in self.D]),
            E=np.concatenate([a.ravel(order="F") for a in self.E]),
            B=np.concatenate([a.ravel(order="F") for a in self.B]) if s["nx"] else np.zeros(0, self.dt),
            C=np.concatenate([a.ravel(order="F") for a in self.C]) if s["nr"] else np.zeros(0, self.dt),
            F=self.F.ravel(order="F"),
            Hmat=self.Hmat.ravel(order="F"),
        )


CASES = [
    # name, (N, n, qr, qx, nr, nx), seed, shift
    ("n4_N16", (16, 4, 1, 1, 2, 2), 11, 4.0),
    ("n3_N13", (13, 3, 2, 2, 1, 1), 12, 3.0),
    ("n5_N7_noborder", (7, 5, 2, 2, 0, 0), 13, 4.0),
    ("n2_N9_wide", (9, 2, 0, 1, 2, 1), 14, 4.0),
    ("n8_N33", (33, 8, 4, 4, 2, 2), 15, 3.0),
]


def main():
    out = {}
    for name, dims_, seed, shift in CASES:
        N, n, qr, qx, nr, nx = dims_
        s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=shift, seed=seed)
        nrows, ncols = systems.dims(s)
        assert nrows == ncols
        xref = np.linspace(-1.0, 2.0, ncols)
        b = systems.mv(s, xref)
        for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
            for internal in ("LU", "QR", "QRP"):
                cr = LapackCR(s, internal, dt)
                cr.factorize()
                x = cr.solve(b.astype(dt))
                key = f"{name}.{tag}.{internal}"
                for k, v in cr.packed_internals().items():
                    out[f"{key}.{k}"] = v
                out[f"{key}.x"] = x
        for k in ("D", "E", "B", "C", "H", "Cq", "F"):
            out[f"{name}.in.{k}"] = s[k]
        out[f"{name}.dims"] = np.array(dims_, np.int32)
        out[f"{name}.b"] = b
        out[f"{name}.xref"] = xref
    path = os.path.join(HERE, "lapack_cr_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
