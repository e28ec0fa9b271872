"""alglin_bcr_factor_solve_noalloc against the entry points it mirrors
(alglin_bcr_bind_device + alglin_bcr_factorize_device +
alglin_bcr_solve_device): the same kernels on the same data, so the
solutions agree bit for bit.  Two calls in a row, the second on a smaller
system (the fixed scratch array keeps the first call's contents), then a
system whose working set exceeds the array is refused with a message."""
import ctypes

import numpy as np
import pytest

from alglin_amd import systems
from alglin_amd.bcr import BorderedCR, DeviceArray, device_synchronize, lib, packed_blocks

pytestmark = pytest.mark.gpu


def _existing(s, dev, B):
    nrows, ncols = systems.dims(s)
    L = max(nrows, ncols)
    g = BorderedCR(np.float64)
    g.allocate(s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"])
    g.select("LU", "LU", "NONE")
    g.bind_device(dev)
    X = DeviceArray.from_host(np.asfortranarray(B).reshape(-1, order="F"))
    assert g.factorize_device(sync=True), g.last_error()
    g.solve_device(X, B.shape[1], L)
    out = X.to_host().reshape(L, B.shape[1], order="F")
    g.close()
    return out


def _noalloc(s, dev, B, sync=1):
    nrows, ncols = systems.dims(s)
    L = max(nrows, ncols)
    dims = np.array([s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"]], np.int32)
    X = DeviceArray.from_host(np.asfortranarray(B).reshape(-1, order="F"))
    err = ctypes.create_string_buffer(256)
    ptr = [ctypes.c_void_p(dev[k].data_ptr()) for k in ("D", "E", "B", "C", "H", "Cq", "F")]
    rc = lib().alglin_bcr_factor_solve_noalloc(ctypes.c_void_p(dims.ctypes.data), *ptr,
                                               ctypes.c_void_p(X.data_ptr()), B.shape[1], L, None, sync, err, 256)
    device_synchronize()
    return rc, err.value.decode(), X.to_host().reshape(L, B.shape[1], order="F")


@pytest.mark.parametrize("first,second", [((700, 8, 2, 2, 2, 2, 3), (90, 4, 1, 1, 1, 1, 2)),
                                          ((400, 32, 8, 8, 4, 4, 2), (33, 17, 3, 3, 2, 2, 1))])
def test_noalloc_matches_existing_entry_point(first, second):
    for k, (N, n, qr, qx, nr, nx, nrhs) in enumerate((first, second)):
        print(f"case call={k} N={N} n={n} nrhs={nrhs}", flush=True)
        s = systems.gaussian_system(N, n, qr, qx, nr, nx, shift=4.0, seed=17 + k)
        nrows, ncols = systems.dims(s)
        L = max(nrows, ncols)
        B = np.zeros((L, nrhs))
        B[:nrows] = np.random.default_rng(3 + k).standard_normal((nrows, nrhs))
        dev = {key: DeviceArray.from_host(v) for key, v in packed_blocks(s).items()}
        ref = _existing(s, dev, B)
        rc, msg, got = _noalloc(s, dev, B)
        assert rc == 0, msg
        assert np.array_equal(got[:ncols], ref[:ncols])
        r = systems.mv(s, got[:ncols]) - B[:nrows]
        assert np.linalg.norm(r) / np.linalg.norm(B[:nrows]) < 1e-12


def test_noalloc_refuses_oversized_system():
    # N = 20000, n = 64 needs far more than the 64 MiB array: refused before
    # any device work (the operator pointers are never touched)
    s = dict(N=20000, n=64, qr=16, qx=16, nr=8, nx=8)
    dims = np.array([s["N"], s["n"], s["qr"], s["qx"], s["nr"], s["nx"]], np.int32)
    err = ctypes.create_string_buffer(256)
    X = DeviceArray(16)
    nul = [ctypes.c_void_p(0)] * 7
    rc = lib().alglin_bcr_factor_solve_noalloc(ctypes.c_void_p(dims.ctypes.data), *nul, ctypes.c_void_p(X.data_ptr()),
                                               1, 0, None, 1, err, 256)
    assert rc == 3
    assert "exceeds the fixed scratch array" in err.value.decode()
