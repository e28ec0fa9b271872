"""ORACLE — test infrastructure only (ctypes view of oracle/build/liboracle_bcr.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It drives the CPU restatement of alglin::BorderedCR
(src/BABD_BorderedCR.cc) with the reference's method names so that the
parity tests read like the reference's own tests (src_tests/test18-*.cc).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle_bcr.so")
_lib = None

INTERNAL = {"LU": 0, "QR": 1, "QRP": 2}
LAST = {"LU": 0, "LUPQ": 1, "QR": 2, "QRP": 3, "SVD": 4, "LSS": 5, "LSY": 6, "PINV": 7}
LAST2 = {"NONE": 0, "SVD": 1, "LSS": 2, "LSY": 3, "PINV": 4}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_allocate.argtypes = [vp] + [i32] * 6
        L.orc_set_num_parallel_block.argtypes = [vp, i32]
        L.orc_set_threads.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_select.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_load.argtypes = [vp] * 8
        L.orc_get.argtypes = [vp] * 9
        L.orc_factorize.argtypes = [vp]
        L.orc_factorize.restype = ctypes.c_int
        L.orc_solve.argtypes = [vp, vp]
        L.orc_solve.restype = ctypes.c_int
        L.orc_solve_n.argtypes = [vp, i32, vp, i32]
        L.orc_solve_n.restype = ctypes.c_int
        L.orc_mv.argtypes = [vp, vp, vp]
        L.orc_sparse_nnz.argtypes = [vp]
        L.orc_sparse_nnz.restype = i32
        L.orc_sparse.argtypes = [vp, vp, vp, vp, i32]
        L.orc_dims.argtypes = [vp, i32]
        L.orc_dims.restype = i32
        L.orc_info.argtypes = [vp]
        L.orc_info.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleBorderedCR:
    """CPU BorderedCR<double|float> restatement (oracle)."""

    def __init__(self, dtype=np.float64, nthreads: int = 1, num_parallel_block: int = 0):
        self.dtype = np.dtype(dtype)
        self._h = lib().orc_create(0 if self.dtype == np.float64 else 1, nthreads, num_parallel_block)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_destroy(self._h)
            self._h = None

    # --- setup (cc:42-170, hxx:407-540) --------------------------------------
    def allocate(self, N, n, qr, qx, nr, nx):
        self.N, self.n, self.qr, self.qx, self.nr, self.nx = N, n, qr, qx, nr, nx
        lib().orc_allocate(self._h, N, n, qr, qx, nr, nx)

    def set_num_parallel_block(self, npb):
        lib().orc_set_num_parallel_block(self._h, npb)

    def set_use_thread(self, factorize: bool, solve: bool):
        lib().orc_set_threads(self._h, int(factorize), int(solve))

    def select(self, internal="LU", last="LU", last2="NONE"):
        lib().orc_select(self._h, INTERNAL[internal], LAST[last], LAST2[last2])

    def load(self, blocks: dict):
        """blocks: dict of packed arrays D (N,n,n) E B C H Cq F, column-major blocks."""
        def col(a):
            return None if a is None else np.asfortranarray(a, dtype=self.dtype)
        arrs = {}
        for key in ("D", "E", "B", "C", "H", "Cq", "F"):
            a = blocks.get(key)
            if a is None:
                arrs[key] = None
            elif key in ("D", "E", "B", "C"):
                # (count, rows, cols) -> each block column-major, blocks contiguous
                arrs[key] = np.ascontiguousarray(np.transpose(np.asarray(a, self.dtype), (0, 2, 1)))
            else:
                arrs[key] = np.ascontiguousarray(np.asarray(a, self.dtype).T)
        lib().orc_load(self._h, *[_ptr(arrs[k]) for k in ("D", "E", "B", "C", "H", "Cq", "F")])
        self._keep = arrs

    # --- dims ---------------------------------------------------------------
    def nrows(self):
        return lib().orc_dims(self._h, 0)

    def ncols(self):
        return lib().orc_dims(self._h, 1)

    def used_parallel_block(self):
        return lib().orc_dims(self._h, 4)

    def info(self) -> str:
        return lib().orc_info(self._h).decode()

    def last_error(self) -> str:
        return lib().orc_last_error(self._h).decode()

    # --- compute ------------------------------------------------------------
    def factorize(self) -> bool:
        return lib().orc_factorize(self._h) == 0

    def solve(self, b: np.ndarray):
        """Returns (ok, x) for one rhs of length nrows (x has ncols entries)."""
        L = max(self.nrows(), self.ncols())
        x = np.zeros(L, self.dtype)
        x[: len(b)] = b
        ok = lib().orc_solve(self._h, _ptr(x)) == 0
        return ok, x[: self.ncols()]

    def solve_n(self, Bm: np.ndarray):
        """Bm: (nrows, nrhs). Returns (ok, X (ncols, nrhs))."""
        L = max(self.nrows(), self.ncols())
        nrhs = Bm.shape[1]
        X = np.zeros((L, nrhs), self.dtype, order="F")
        X[: Bm.shape[0], :] = Bm
        ok = lib().orc_solve_n(self._h, nrhs, _ptr(X), L) == 0
        return ok, np.array(X[: self.ncols(), :])

    def Mv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, self.dtype)
        y = np.zeros(self.nrows(), self.dtype)
        lib().orc_mv(self._h, _ptr(x), _ptr(y))
        return y

    def sparse(self, offs=0):
        nnz = lib().orc_sparse_nnz(self._h)
        I = np.zeros(nnz, np.int32)
        J = np.zeros(nnz, np.int32)
        V = np.zeros(nnz, self.dtype)
        lib().orc_sparse(self._h, _ptr(I), _ptr(J), _ptr(V), offs)
        return I, J, V

    def dense(self) -> np.ndarray:
        I, J, V = self.sparse(0)
        A = np.zeros((self.nrows(), self.ncols()), self.dtype)
        np.add.at(A, (I, J), V)
        return A

    def internals(self):
        """Factor arrays after factorize (reference internal layout)."""
        N, n, nr, nx = self.N, self.n, self.nr, self.nx
        Nr = lib().orc_dims(self._h, 2)
        Nc = lib().orc_dims(self._h, 3)
        out = dict(
            D=np.zeros(N * n * n, self.dtype), E=np.zeros(N * n * n, self.dtype),
            B=np.zeros(N * n * nx, self.dtype), C=np.zeros((N + 1) * nr * n, self.dtype),
            F=np.zeros(nr * nx, self.dtype), T=np.zeros(N * (2 * n * n + n), self.dtype),
            perm=np.zeros(N * n, np.int32), Hmat=np.zeros(Nr * Nc, self.dtype),
        )
        lib().orc_get(self._h, *[_ptr(out[k]) for k in ("D", "E", "B", "C", "F", "T", "perm", "Hmat")])
        return out
