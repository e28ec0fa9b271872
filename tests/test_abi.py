"""CPU checks of the drop-in boundary: libalglin_amd.so loads and exports
every entry point include/alglin_amd.h declares; the Python view refuses to
run without the native library (there is no CPU fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alglin_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:BABD|alglin)_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_c_interface():
    syms = declared_symbols()
    for name in ("BABD_factorize", "BABD_factorize_bordered", "BABD_solve", "BABD_get_last_error"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from alglin_amd import bcr
    if not os.path.exists(bcr.LIB_PATH):
        from alglin_amd import build
        build.build()
    out = subprocess.run(["nm", "-D", "--defined-only", bcr.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = bcr.lib()  # dlopen + ctypes signatures resolve
    for s in declared_symbols():
        assert hasattr(lib, s)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from alglin_amd import bcr
    monkeypatch.setattr(bcr, "_lib", None)
    monkeypatch.setattr(bcr, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(bcr.AlglinError):
        bcr.lib()
