"""Python view of the alglin_amd C ABI (include/alglin_amd.h) via ctypes.

`BorderedCR` mirrors alglin::BorderedCR<T> (src/Alglin/BABD_BorderedCR.hxx)
method by method; compute runs in libalglin_amd.so on the GPU.  There is no
CPU fallback: a missing library raises, and calls on a machine without a
GPU fail inside the library.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalglin_amd.so")
# per-class kernel choices of the autotuner (alglin_amd/autotune.py)
TUNING_PATH = os.path.join(_HERE, "tuning.json")
_lib = None


class AlglinError(RuntimeError):
    pass


def lib():
    """Load libalglin_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AlglinError(f"{LIB_PATH} is missing: run `python -m alglin_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, ci = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int
        cs = ctypes.c_char_p
        sig = {
            "alglin_bcr_create": (vp, [ci]),
            "alglin_bcr_destroy": (None, [vp]),
            "alglin_bcr_allocate": (ci, [vp] + [i32] * 6),
            "alglin_bcr_select": (ci, [vp, cs, cs, cs]),
            "alglin_bcr_dims": (ci, [vp, vp]),
            "alglin_bcr_load": (ci, [vp, cs, i32, vp, i32]),
            "alglin_bcr_load_bottom": (ci, [vp, vp, i32, vp, i32, vp, i32, vp, i32]),
            "alglin_bcr_load_bottom2": (ci, [vp, vp, i32, vp, i32, vp, i32, vp, i32]),
            "alglin_bcr_load_packed": (ci, [vp] * 8),
            "alglin_bcr_factorize": (ci, [vp]),
            "alglin_bcr_solve": (ci, [vp, vp]),
            "alglin_bcr_solve_n": (ci, [vp, i32, vp, i32]),
            "alglin_bcr_mv": (ci, [vp, vp, vp]),
            "alglin_bcr_sparse_nnz": (i32, [vp]),
            "alglin_bcr_sparse_pattern": (ci, [vp, vp, vp, i32]),
            "alglin_bcr_sparse_values": (ci, [vp, vp]),
            "alglin_bcr_sparse_load": (ci, [vp, vp, vp, i32, vp, i32, i32, ci]),
            "alglin_bcr_check_matrix": (ci, [vp]),
            "alglin_bcr_last_error": (cs, [vp]),
            "alglin_bcr_info": (cs, [vp]),
            "alglin_bcr_set_num_parallel_block": (ci, [vp, i32]),
            "alglin_bcr_set_debug": (ci, [vp, ci]),
            "alglin_bcr_used_parallel_block": (i32, [vp]),
            "alglin_bcr_dup": (ci, [vp, vp]),
            "alglin_bcr_t_solve": (ci, [vp, vp]),
            "alglin_bcr_matlab_script": (i64, [vp, ctypes.c_char_p, i64]),
            "alglin_bcr_get_packed": (ci, [vp] * 8),
            "alglin_bcr_bind_device": (ci, [vp] * 8),
            "alglin_bcr_factorize_device": (ci, [vp, ci]),
            "alglin_bcr_solve_device": (ci, [vp, vp, i32, i64, ci]),
            "alglin_bcr_residual_device": (ci, [vp, vp, vp, vp, ci]),
            "alglin_bcr_sync": (ci, [vp]),
            "alglin_bcr_stream": (vp, [vp]),
            "alglin_bcr_get_factors": (ci, [vp] * 9),
            "alglin_bcr_fast_path": (ci, [vp]),
            "alglin_bcr_set_timing": (ci, [vp, ci]),
            "alglin_bcr_set_precision": (ci, [vp, cs, ctypes.c_double, i32]),
            "alglin_bcr_refine_report": (ci, [vp, vp]),
            "alglin_bcr_set_fallback": (ci, [vp, ctypes.c_int]),
            "alglin_bcr_set_pair_kernel": (ci, [vp, ctypes.c_int]),
            "alglin_bcr_timing": (ci, [vp, vp, ci]),
            "alglin_bcr_factor_solve_noalloc": (ci, [vp] + [vp] * 8 + [i32, i64, vp, ci, ctypes.c_char_p, i32]),
            "alglin_dev_malloc": (vp, [i64]),
            "alglin_dev_free": (ci, [vp]),
            "alglin_memcpy_htod": (ci, [vp, vp, i64]),
            "alglin_memcpy_dtoh": (ci, [vp, vp, i64]),
            "alglin_memcpy_dtod": (ci, [vp, vp, i64]),
            "alglin_memcpy_dtod_async": (ci, [vp, vp, i64, vp]),
            "alglin_batch_time_factorize": (ci, [vp, vp, i32]),
            "alglin_batch_set_pair_kernel": (ci, [vp, ctypes.c_int]),
            "alglin_batch_set_streams": (ci, [vp, ctypes.c_int, ctypes.c_int]),
            "alglin_batch_set_class_kernel": (ci, [vp, ci, ci, ci, ci]),
            "alglin_batch_create": (vp, [ctypes.c_int]),
            "alglin_batch_destroy": (None, [vp]),
            "alglin_batch_plan": (ci, [vp, i32, vp]),
            "alglin_batch_size": (i64, [vp, i32]),
            "alglin_batch_offset": (i64, [vp, i32, i32]),
            "alglin_batch_bind_device": (ci, [vp] * 8),
            "alglin_batch_generate": (ci, [vp, ctypes.c_uint64, ctypes.c_double]),
            "alglin_batch_solve": (ci, [vp, vp, vp, ctypes.c_double, i32, ctypes.c_int]),
            "alglin_batch_report": (ci, [vp, vp, vp, vp, vp]),
            "alglin_batch_last_error": (cs, [vp]),
            "alglin_batch_stream": (vp, [vp]),
            "alglin_batch_generate_singular": (ci, [vp, ctypes.c_uint64, ctypes.c_double, i32, vp, ctypes.c_double]),
            "alglin_batch_set_methods": (ci, [vp, i32, vp]),
            "alglin_batch_solve_selected": (ci, [vp, vp, vp, ctypes.c_double, i32, vp, ctypes.c_float,
                                                 ctypes.c_int]),
            "alglin_batch_selection": (ci, [vp, vp, vp, vp]),
            "alglin_mlp_num_params": (i32, []),
            "alglin_mlp_create": (vp, [ctypes.c_uint64]),
            "alglin_mlp_destroy": (None, [vp]),
            "alglin_mlp_last_error": (cs, [vp]),
            "alglin_mlp_get_params": (ci, [vp, vp, vp, vp]),
            "alglin_mlp_set_params": (ci, [vp, vp, vp, vp]),
            "alglin_mlp_load": (ci, [vp, cs]),
            "alglin_mlp_save": (ci, [vp, cs]),
            "alglin_mlp_loss_grad": (ci, [vp, vp, vp, i32, vp, vp]),
            "alglin_mlp_train": (ci, [vp, vp, vp, i32, i32, ctypes.c_float, vp]),
            "alglin_mlp_predict": (ci, [vp, vp, i32, vp]),
            "alglin_device_synchronize": (ci, []),
            "alglin_device_info": (ci, [ctypes.c_char_p, i32, vp, vp]),
            "BABD_factorize": (ci, [i32] * 7 + [vp, i32] * 4),
            "BABD_factorize_bordered": (ci, [i32] * 9 + [vp, i32] * 7),
            "BABD_solve": (ci, [i32, vp]),
            "BABD_get_last_error": (cs, []),
            "BABD_free": (ci, [i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class BorderedCR:
    """GPU BorderedCR<double|float> with the reference's method names."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self._h = lib().alglin_bcr_create(0 if self.dtype == np.float64 else 1)
        if not self._h:
            raise AlglinError("alglin_bcr_create failed (no HIP device?)")

    def close(self):
        if getattr(self, "_h", None):
            lib().alglin_bcr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise AlglinError(f"{what}: {self.last_error()}")

    # --- setup ----------------------------------------------------------------
    def allocate(self, N, n, qr, qx, nr, nx):
        self._check(lib().alglin_bcr_allocate(self._h, N, n, qr, qx, nr, nx), "allocate")

    def select(self, internal="LU", last="LU", last2="NONE"):
        self._check(lib().alglin_bcr_select(self._h, internal.encode(), last.encode(), last2.encode()), "select")

    def dims(self):
        out = np.zeros(10, np.int32)
        self._check(lib().alglin_bcr_dims(self._h, _p(out)), "dims")
        keys = ("N", "n", "qr", "qx", "nr", "nx", "nrows", "ncols", "Nr", "Nc")
        return dict(zip(keys, (int(v) for v in out)))

    def nrows(self):
        return self.dims()["nrows"]

    def ncols(self):
        return self.dims()["ncols"]

    def last_error(self) -> str:
        return lib().alglin_bcr_last_error(self._h).decode()

    def info(self) -> str:
        return lib().alglin_bcr_info(self._h).decode()

    def load(self, what: str, nbl: int, A: np.ndarray):
        """Reference loaders: what in B,C,D,E,DE,DEB,F,Cq,CqF,H,add_B,add_C,add_C2,add_C2F,add_F."""
        A = np.asfortranarray(A, dtype=self.dtype)
        self._check(lib().alglin_bcr_load(self._h, what.encode(), nbl, _p(A), max(1, A.shape[0])), f"load_{what}")

    def load_bottom(self, H0, HN, Hq, Hp):
        arr = [np.asfortranarray(a, dtype=self.dtype) for a in (H0, HN, Hq, Hp)]
        args = []
        for a in arr:
            args += [_p(a), max(1, a.shape[0])]
        self._check(lib().alglin_bcr_load_bottom(self._h, *args), "load_bottom")

    def load_bottom2(self, C0, CN, Cq, F):
        arr = [np.asfortranarray(a, dtype=self.dtype) for a in (C0, CN, Cq, F)]
        args = []
        for a in arr:
            args += [_p(a), max(1, a.shape[0])]
        self._check(lib().alglin_bcr_load_bottom2(self._h, *args), "load_bottom2")

    def load_system(self, s: dict):
        """All blocks of a system dict (alglin_amd.systems layout)."""
        self.allocate(s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"])
        packed = packed_blocks(s, self.dtype)
        self._keep = packed
        self._check(lib().alglin_bcr_load_packed(self._h, *[_p(packed[k]) for k in
                                                           ("D", "E", "B", "C", "H", "Cq", "F")]), "load_packed")

    # --- compute --------------------------------------------------------------
    def factorize(self) -> bool:
        return lib().alglin_bcr_factorize(self._h) == 0

    def solve(self, b: np.ndarray):
        """One rhs (nrows) -> (ok, x[ncols]); in place on a max(nrows,ncols) buffer."""
        d = self.dims()
        L = max(d["nrows"], d["ncols"])
        x = np.zeros(L, self.dtype)
        x[: len(b)] = b
        ok = lib().alglin_bcr_solve(self._h, _p(x)) == 0
        return ok, x[: d["ncols"]]

    def solve_n(self, Bm: np.ndarray):
        d = self.dims()
        L = max(d["nrows"], d["ncols"])
        nrhs = Bm.shape[1]
        X = np.zeros((L, nrhs), self.dtype, order="F")
        X[: Bm.shape[0]] = Bm
        ok = lib().alglin_bcr_solve_n(self._h, nrhs, _p(X), L) == 0
        return ok, np.array(X[: d["ncols"]])

    def Mv(self, x):
        x = np.ascontiguousarray(x, self.dtype)
        y = np.zeros(self.nrows(), self.dtype)
        self._check(lib().alglin_bcr_mv(self._h, _p(x), _p(y)), "Mv")
        return y

    def sparse(self, offs=0):
        nnz = lib().alglin_bcr_sparse_nnz(self._h)
        I = np.zeros(nnz, np.int32)
        J = np.zeros(nnz, np.int32)
        V = np.zeros(nnz, self.dtype)
        self._check(lib().alglin_bcr_sparse_pattern(self._h, _p(I), _p(J), offs), "sparse_pattern")
        self._check(lib().alglin_bcr_sparse_values(self._h, _p(V)), "sparse_values")
        return I, J, V

    def sparse_load(self, V, I, J, r_offs=0, c_offs=0, do_error=True):
        V = np.ascontiguousarray(V, self.dtype)
        I = np.ascontiguousarray(I, np.int32)
        J = np.ascontiguousarray(J, np.int32)
        self._check(lib().alglin_bcr_sparse_load(self._h, _p(V), _p(I), r_offs, _p(J), c_offs, len(V),
                                                int(do_error)), "sparse_load")

    def check_matrix(self):
        self._check(lib().alglin_bcr_check_matrix(self._h), "check_matrix")

    def set_debug(self, on: bool):
        self._check(lib().alglin_bcr_set_debug(self._h, int(on)), "set_debug")

    def set_num_parallel_block(self, npb: int):
        """Elimination order of the reference with npb threads (cc:139-158)."""
        self._check(lib().alglin_bcr_set_num_parallel_block(self._h, npb), "set_num_parallel_block")

    def used_parallel_block(self) -> int:
        return int(lib().alglin_bcr_used_parallel_block(self._h))

    def dup(self, other: "BorderedCR"):
        self._check(lib().alglin_bcr_dup(self._h, other._h), "dup")

    def t_solve(self, b):
        x = np.array(b, self.dtype)
        self._check(lib().alglin_bcr_t_solve(self._h, _p(x)), "t_solve")
        return x

    def matlab_script(self) -> str:
        n = lib().alglin_bcr_matlab_script(self._h, None, 0)
        if n < 0:
            raise AlglinError(f"matlab_script: {self.last_error()}")
        buf = ctypes.create_string_buffer(int(n))
        lib().alglin_bcr_matlab_script(self._h, buf, n)
        return buf.value.decode()

    def packed(self) -> dict:
        """Host copies of the loaded blocks (packed layout)."""
        d = self.dims()
        N, n, qr, qx, nr, nx, Nc = d["N"], d["n"], d["qr"], d["qx"], d["nr"], d["nx"], d["Nc"]
        out = dict(D=np.zeros(N * n * n, self.dtype), E=np.zeros(N * n * n, self.dtype),
                   B=np.zeros(max(1, N * n * nx), self.dtype), C=np.zeros(max(1, (N + 1) * nr * n), self.dtype),
                   H=np.zeros((n + qr) * Nc, self.dtype), Cq=np.zeros(max(1, nr * qx), self.dtype),
                   F=np.zeros(max(1, nr * nx), self.dtype))
        self._check(lib().alglin_bcr_get_packed(self._h, *[_p(out[k]) for k in ("D", "E", "B", "C", "H", "Cq", "F")]),
                    "get_packed")
        return out

    def factors(self):
        d = self.dims()
        N, n, nr, nx = d["N"], d["n"], d["nr"], d["nx"]
        out = dict(
            T=np.zeros(N * (2 * n * n + n), self.dtype), perm=np.zeros(N * n, np.int32),
            D=np.zeros(N * n * n, self.dtype), E=np.zeros(N * n * n, self.dtype),
            B=np.zeros(max(1, N * n * nx), self.dtype), C=np.zeros((N + 1) * nr * n, self.dtype),
            last=np.zeros(d["Nr"] * d["Nc"]), pend=np.zeros(1),
        )
        self._check(lib().alglin_bcr_get_factors(self._h, *[_p(out[k]) for k in
                                                           ("T", "perm", "D", "E", "B", "C", "last", "pend")]),
                    "get_factors")
        return out

    def fast_path(self) -> bool:
        return lib().alglin_bcr_fast_path(self._h) == 1

    # --- device-resident (DeviceArray operands) --------------------------------
    def bind_device(self, t: dict):
        """t: dict of DeviceArray in packed layout (see packed_blocks / DeviceArray.from_host)."""
        self._bound = t
        ptr = [ctypes.c_void_p(t[k].data_ptr()) for k in ("D", "E", "B", "C", "H", "Cq", "F")]
        self._check(lib().alglin_bcr_bind_device(self._h, *ptr), "bind_device")

    def select_precision(self, mode="FP64", tol=1e-12, max_iter=30):
        """'FP64' (reference arithmetic) or 'MIXED' (FP32 factors + FP64 refinement)."""
        self._check(lib().alglin_bcr_set_precision(self._h, mode.encode(), tol, max_iter), "select_precision")

    def set_pair_kernel(self, variant: int):
        """LU pair kernel: 0 default, 3 register tiles for every n,
        4 column-split waves for every n <= 64, 5 the matrix-core pair (FP32 factors
        n <= 64, FP64 factors n <= 48; nr <= 8, nx <= 16).  The default takes the
        matrix-core pair for 16 < n where it applies."""
        self._check(lib().alglin_bcr_set_pair_kernel(self._h, int(variant)), "set_pair_kernel")

    def set_fallback(self, on: bool):
        self._check(lib().alglin_bcr_set_fallback(self._h, int(on)), "set_fallback")

    def refine_report(self) -> dict:
        """status: 0 converged, 1 stalled, 2 max_iter, 3 fallback factorization
        failed, 4 FP64 QRP fallback reached tol, 5 FP64 residual floor (the
        componentwise backward error is at roundoff), 6 fallback missed tol."""
        out = np.zeros(5)
        self._check(lib().alglin_bcr_refine_report(self._h, _p(out)), "refine_report")
        return dict(iterations=int(out[0]), rel_residual=float(out[1]), first_residual=float(out[2]),
                    status=int(out[3]), backward_error=float(out[4]))

    def set_timing(self, on=True):
        self._check(lib().alglin_bcr_set_timing(self._h, int(on)), "set_timing")

    def timing(self, reset=True) -> dict:
        out = np.zeros(8)
        self._check(lib().alglin_bcr_timing(self._h, _p(out), int(reset)), "timing")
        keys = ("pair_ms", "forward_ms", "backward_ms", "last_ms", "reduce_ms", "residual_ms",
                "pair_launches", "sweep_launches")
        return dict(zip(keys, (float(v) for v in out)))

    def factorize_device(self, sync=True) -> bool:
        return lib().alglin_bcr_factorize_device(self._h, int(sync)) == 0

    def solve_device(self, X, nrhs: int, ld: int, sync=True):
        self._check(lib().alglin_bcr_solve_device(self._h, ctypes.c_void_p(X.data_ptr()), nrhs, ld, int(sync)),
                    "solve_device")

    def residual_device(self, x, b, r, sync=True):
        self._check(lib().alglin_bcr_residual_device(self._h, ctypes.c_void_p(x.data_ptr()),
                                                    ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                                    int(sync)), "residual_device")

    def sync(self):
        self._check(lib().alglin_bcr_sync(self._h), "sync")

    def stream(self) -> int:
        return lib().alglin_bcr_stream(self._h)


class DeviceArray:
    """A flat device allocation made through the library's own HIP runtime."""

    def __init__(self, count: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(1, self.count * self.dtype.itemsize)
        self.ptr = lib().alglin_dev_malloc(self.nbytes)
        if not self.ptr:
            raise AlglinError(f"device allocation of {self.nbytes} bytes failed")

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None):
        a = np.ascontiguousarray(a, dtype or a.dtype)
        d = cls(a.size, a.dtype)
        if a.size and lib().alglin_memcpy_htod(ctypes.c_void_p(d.ptr), _p(a), a.nbytes) != 0:
            raise AlglinError("H2D copy failed")
        return d

    def data_ptr(self) -> int:
        return self.ptr

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.count and lib().alglin_memcpy_dtoh(_p(out), ctypes.c_void_p(self.ptr), out.nbytes) != 0:
            raise AlglinError("D2H copy failed")
        return out

    def copy_from(self, other: "DeviceArray", stream: int | None = None):
        """Device copy; with `stream` (a solver's stream()) it is queued there
        and ordered with that solver's work, else it completes before return."""
        nb = min(self.nbytes, other.nbytes)
        if stream is None:
            rc = lib().alglin_memcpy_dtod(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), nb)
        else:
            rc = lib().alglin_memcpy_dtod_async(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), nb,
                                                ctypes.c_void_p(stream))
        if rc != 0:
            raise AlglinError("D2D copy failed")

    def free(self):
        if self.ptr:
            lib().alglin_dev_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    if lib().alglin_device_synchronize() != 0:
        raise AlglinError("hipDeviceSynchronize failed")


def device_info() -> dict:
    name = ctypes.create_string_buffer(128)
    cus = ctypes.c_int32(0)
    hbm = ctypes.c_int64(0)
    if lib().alglin_device_info(name, 128, ctypes.byref(cus), ctypes.byref(hbm)) != 0:
        raise AlglinError("no HIP device")
    return dict(arch=name.value.decode(), cus=cus.value, hbm_bytes=hbm.value)


def packed_blocks(s: dict, dtype=np.float64) -> dict:
    """System dict -> arrays in the reference's internal block layout (hxx:972-998)."""
    dt = np.dtype(dtype)

    def blocks(a):  # (K, rows, cols) -> K column-major blocks, contiguous
        return np.ascontiguousarray(np.transpose(np.asarray(a, dt), (0, 2, 1))).reshape(-1)

    def mat(a):
        return np.ascontiguousarray(np.asarray(a, dt).T).reshape(-1)

    return dict(D=blocks(s["D"]), E=blocks(s["E"]), B=blocks(s["B"]), C=blocks(s["C"]),
                H=mat(s["H"]), Cq=mat(s["Cq"]), F=mat(s["F"]))


def system_from_packed(N, n, qr, qx, nr, nx, D, E, B, C, H, Cq, F) -> dict:
    """Inverse of packed_blocks: reference-layout arrays of one system -> system dict."""
    Nc = 2 * n + qx + nx

    def blocks(a, K, rows, cols):
        return np.asarray(a).reshape(K, cols, rows).transpose(0, 2, 1).copy()

    def mat(a, rows, cols):
        return np.asarray(a).reshape(cols, rows).T.copy()

    return dict(N=N, n=n, qr=qr, qx=qx, nr=nr, nx=nx,
                D=blocks(D, N, n, n), E=blocks(E, N, n, n), B=blocks(B, N, n, nx), C=blocks(C, N + 1, nr, n),
                H=mat(H, n + qr, Nc), Cq=mat(Cq, nr, qx), F=mat(F, nr, nx))


class Batch:
    """Many independent systems solved together (alglin_batch_* in alglin_amd.h).

    dims: rows (N, n, qr, qx, nr, nx, method) with method 0 LU, 1 QR, 2 QRP.
    The operator lives in packed device arrays owned by this object
    (kinds DE, B, C, H, Cq, F); right-hand sides and solutions are packed
    vectors of kind x."""

    KINDS = ("DE", "B", "C", "H", "Cq", "F", "x")

    def __init__(self, mixed: bool = False):
        self._h = lib().alglin_batch_create(int(mixed))
        if not self._h:
            raise AlglinError("alglin_batch_create failed")
        self.mixed = mixed
        self.ops = None
        self._shared = False

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().alglin_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise AlglinError(f"{what}: {lib().alglin_batch_last_error(self._h).decode()}")

    def last_error(self) -> str:
        return lib().alglin_batch_last_error(self._h).decode()

    def plan(self, dims, operator: bool = True, tuning: bool = True):
        """Lays the systems out; with operator=False no operator storage is
        allocated (share_operator must follow before any solve).  tuning:
        apply the autotuner's per-class kernel choices (alglin_amd/tuning.json)
        when the file exists."""
        d = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 7))
        self.dims = d
        self._check(lib().alglin_batch_plan(self._h, len(d), _p(d)), "plan")
        if tuning:
            self.load_tuning()
        # operator storage (device, packed)
        if self.ops is not None and not getattr(self, "_shared", False):
            for a in self.ops.values():
                a.free()
        self.ops = None
        self._shared = False
        if not operator:
            return
        sz = {k: max(1, int(lib().alglin_batch_size(self._h, i))) for i, k in enumerate(self.KINDS)}
        self.ops = dict(D=DeviceArray(sz["DE"]), E=DeviceArray(sz["DE"]), B=DeviceArray(sz["B"]),
                        C=DeviceArray(sz["C"]), H=DeviceArray(sz["H"]), Cq=DeviceArray(sz["Cq"]),
                        F=DeviceArray(sz["F"]))
        ptr = [ctypes.c_void_p(self.ops[k].data_ptr()) for k in ("D", "E", "B", "C", "H", "Cq", "F")]
        self._check(lib().alglin_batch_bind_device(self._h, *ptr), "bind_device")

    def share_operator(self, other: "Batch"):
        """Use another batch's device operator (same plan) instead of our own copy."""
        if not np.array_equal(other.dims[:, :6], self.dims[:, :6]):
            raise AlglinError("share_operator: the two batches have different shapes")
        if self.ops is not None and not self._shared:
            for a in self.ops.values():
                a.free()
        self.ops = other.ops      # the arrays stay alive as long as either batch holds them
        self._shared = True
        ptr = [ctypes.c_void_p(self.ops[k].data_ptr()) for k in ("D", "E", "B", "C", "H", "Cq", "F")]
        self._check(lib().alglin_batch_bind_device(self._h, *ptr), "bind_device")

    def size(self, kind: str) -> int:
        return int(lib().alglin_batch_size(self._h, self.KINDS.index(kind)))

    def offset(self, kind: str, s: int) -> int:
        return int(lib().alglin_batch_offset(self._h, self.KINDS.index(kind), s))

    def generate(self, seed: int, shift: float, sing_k=None, eps: float = 0.0):
        """Seeded Gaussian systems on the GPU; sing_k[s] >= 1 makes the pair
        column [E_{k-1}; D_k] of system s rank deficient up to eps."""
        if sing_k is None:
            self._check(lib().alglin_batch_generate(self._h, ctypes.c_uint64(seed), shift), "generate")
        else:
            k = np.ascontiguousarray(sing_k, np.int32)
            self._check(lib().alglin_batch_generate_singular(self._h, ctypes.c_uint64(seed), shift, len(k), _p(k),
                                                             eps), "generate")

    def set_methods(self, methods):
        m = np.ascontiguousarray(methods, np.int32)
        self.dims[:, 6] = m
        self._check(lib().alglin_batch_set_methods(self._h, len(m), _p(m)), "set_methods")

    def solve_selected(self, B, X, tol=1e-12, max_iter=30, selector=None, threshold=0.5, fallback=True) -> int:
        """Mixed solve with per-system LU/QRP choice (see alglin_batch_solve_selected)."""
        rc = lib().alglin_batch_solve_selected(self._h, ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(X.data_ptr()),
                                               tol, max_iter, selector._h if selector is not None else None,
                                               threshold, int(fallback))
        if rc == 1:
            raise AlglinError(f"batch solve_selected: {self.last_error()}")
        return rc

    def selection(self) -> dict:
        ns = len(self.dims)
        f = np.zeros((ns, 12), np.float32)
        q = np.zeros(ns, np.int32)
        b = np.zeros(ns, np.int32)
        self._check(lib().alglin_batch_selection(self._h, _p(f), _p(q), _p(b)), "selection")
        return dict(features=f, chose_qrp=q.astype(bool), fell_back=b.astype(bool), fallback_stage=b.copy())

    def load_system(self, s: int, sysd: dict):
        """Upload one system's blocks (system dict) into its packed slices."""
        pk = packed_blocks(sysd)
        for key, kind in (("D", "DE"), ("E", "DE"), ("B", "B"), ("C", "C"), ("H", "H"), ("Cq", "Cq"), ("F", "F")):
            a = pk[key]
            if a.size:
                o = self.offset(kind, s) * 8
                if lib().alglin_memcpy_htod(ctypes.c_void_p(self.ops[key].data_ptr() + o), _p(a), a.nbytes) != 0:
                    raise AlglinError("H2D copy failed")

    def system(self, s: int) -> dict:
        """Download system s from the packed device arrays (system dict)."""
        N, n, qr, qx, nr, nx, _ = (int(v) for v in self.dims[s])
        Nc = 2 * n + qx + nx
        cnt = dict(D=N * n * n, E=N * n * n, B=N * n * nx, C=(N + 1) * nr * n, H=(n + qr) * Nc, Cq=nr * qx,
                   F=nr * nx)
        kinds = dict(D="DE", E="DE", B="B", C="C", H="H", Cq="Cq", F="F")
        out = {}
        for key, c in cnt.items():
            a = np.empty(c)
            if c and lib().alglin_memcpy_dtoh(_p(a), ctypes.c_void_p(self.ops[key].data_ptr() +
                                                                     8 * self.offset(kinds[key], s)), a.nbytes) != 0:
                raise AlglinError("D2H copy failed")
            out[key] = a
        return system_from_packed(N, n, qr, qx, nr, nx, **out)

    def solve(self, B: "DeviceArray", X: "DeviceArray", tol=1e-12, max_iter=30, graph=True) -> int:
        """X = A^{-1} B per system; 0 all systems met tol, 2 some did not."""
        rc = lib().alglin_batch_solve(self._h, ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(X.data_ptr()), tol,
                                      max_iter, int(graph))
        if rc == 1:
            raise AlglinError(f"batch solve: {self.last_error()}")
        return rc

    def report(self) -> dict:
        ns = len(self.dims)
        st = np.zeros(ns, np.int32)
        rel = np.zeros(ns)
        sw = np.zeros(ns, np.int32)
        summ = np.zeros(4)
        self._check(lib().alglin_batch_report(self._h, _p(st), _p(rel), _p(sw), _p(summ)), "report")
        return dict(status=st, rel_residual=rel, sweeps=sw, max_sweeps=int(summ[0]), max_rel_residual=float(summ[1]),
                    failed=int(summ[2]), ms=float(summ[3]))

    def stream(self) -> int:
        return lib().alglin_batch_stream(self._h)

    def set_pair_kernel(self, variant: int):
        """Kernels of the 16 < n <= 32 and 32 < n <= 64 LU pair classes: 0 default
        (the matrix-core pair where it applies: n <= 48, nr <= 8, nx <= 16;
        else the rank-1 kernels), 4 the rank-1 kernels (two-wave split,
        column-split), 5 the matrix-core pair."""
        self._check(lib().alglin_batch_set_pair_kernel(self._h, int(variant)), "set_pair_kernel")

    def set_streams(self, streams: int, fork_level: int = 0):
        """1: every launch on the batch's stream; 2 (default): the n > 32 LU
        class on a second stream beside the other classes, from reduction
        level fork_level on."""
        self._check(lib().alglin_batch_set_streams(self._h, int(streams), int(fork_level)), "set_streams")

    def set_class_kernel(self, cls: int, kernel: int, waves: int = 0, fp32_factors=None):
        """Kernel of LU pair class cls (3: 16 < n <= 32, 4: 32 < n <= 64):
        -1 default, 0 matrix-core pair (FP32 or FP64 factors, n <= 48), 1
        rank-1 (column-split with `waves` waves per pair for class 4), 2
        register-tiled (class 4)."""
        fp32 = self.mixed if fp32_factors is None else bool(fp32_factors)
        self._check(lib().alglin_batch_set_class_kernel(self._h, int(fp32), cls, kernel, waves), "set_class_kernel")

    def load_tuning(self, path=None):
        """Apply the per-class kernel choices the autotuner stored (JSON)."""
        import json
        path = path or TUNING_PATH
        if not os.path.exists(path):
            return False
        t = json.load(open(path))
        for key, prec in (("fp32_factors", True), ("fp64_factors", False)):
            for cls, c in t.get(key, {}).items():
                self.set_class_kernel(int(cls), int(c["kernel"]), int(c.get("waves", 0)), fp32_factors=prec)
        return True

    def time_factorize(self) -> np.ndarray:
        """One eager factorization with every pair-elimination launch timed
        (HIP events): rows of (level, class, workgroups, ms, bytes, flops)."""
        cap = 4096
        rows = np.zeros((cap, 6))
        n = lib().alglin_batch_time_factorize(self._h, _p(rows), cap)
        if n < 0:
            raise AlglinError(f"time_factorize: {self.last_error()}")
        return rows[: min(n, cap)].copy()


class Mlp:
    """The GPU-trained LU-vs-QRP selector (alglin_mlp_* in alglin_amd.h)."""

    def __init__(self, seed: int = 0):
        self._h = lib().alglin_mlp_create(ctypes.c_uint64(seed))
        if not self._h:
            raise AlglinError("alglin_mlp_create failed")
        self.nparams = int(lib().alglin_mlp_num_params())

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().alglin_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise AlglinError(f"{what}: {lib().alglin_mlp_last_error(self._h).decode()}")

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        mu = np.zeros(12, np.float32)
        sig = np.zeros(12, np.float32)
        self._check(lib().alglin_mlp_get_params(self._h, _p(p), _p(mu), _p(sig)), "get_params")
        return p, mu, sig

    def set_params(self, p=None, mu=None, sig=None):
        pp = None if p is None else np.ascontiguousarray(p, np.float32)
        m = None if mu is None else np.ascontiguousarray(mu, np.float32)
        sg = None if sig is None else np.ascontiguousarray(sig, np.float32)
        self._check(lib().alglin_mlp_set_params(self._h, None if pp is None else _p(pp), None if m is None else _p(m),
                                                None if sg is None else _p(sg)), "set_params")

    def load(self, path):
        self._check(lib().alglin_mlp_load(self._h, str(path).encode()), "load")

    def save(self, path):
        self._check(lib().alglin_mlp_save(self._h, str(path).encode()), "save")

    @staticmethod
    def _dev(X, y=None):
        Xd = DeviceArray.from_host(np.ascontiguousarray(X, np.float32))
        yd = None if y is None else DeviceArray.from_host(np.ascontiguousarray(y, np.int32))
        return Xd, yd

    def loss_grad(self, X, y):
        Xd, yd = self._dev(X, y)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.nparams, np.float32)
        self._check(lib().alglin_mlp_loss_grad(self._h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(yd.data_ptr()),
                                               len(y), _p(loss), _p(g)), "loss_grad")
        return float(loss[0]), g

    def train(self, X, y, epochs=300, lr=3e-3):
        Xd, yd = self._dev(X, y)
        hist = np.zeros(epochs, np.float32)
        self._check(lib().alglin_mlp_train(self._h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(yd.data_ptr()),
                                           len(y), epochs, lr, _p(hist)), "train")
        return hist

    def predict(self, X):
        Xd, _ = self._dev(X)
        out = DeviceArray(len(X), np.float32)
        self._check(lib().alglin_mlp_predict(self._h, ctypes.c_void_p(Xd.data_ptr()), len(X),
                                             ctypes.c_void_p(out.data_ptr())), "predict")
        return out.to_host()

