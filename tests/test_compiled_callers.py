"""Compiled callers of the drop-in boundary.

* tests/cxx/surface_test.cc: a C++17 program built with g++ against
  include/alglin_amd/BorderedCR.hh and BABD.hh and linked to
  libalglin_amd.so.  Its host part (loaders and accessors against tagged
  values, views, partition and info strings, sparse export/import, Mv,
  check_matrix, t_solve, dup, the MATLAB script, the BABD wrapper's
  selection) needs no GPU; its gpu part solves the reference's closed-form
  test systems.
* tests/cxx/cinterface_test.c: a C99 program built with gcc against
  include/alglin_amd.h, shaped like src_tests/test9-Cinterface.c (gpu).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "alglin_amd")
CXX = os.path.join(ROOT, "tests", "cxx")


def _lib():
    from alglin_amd import bcr
    if not os.path.exists(bcr.LIB_PATH):
        from alglin_amd import build
        build.build()
    return bcr.LIB_PATH


def _link_flags():
    return [f"-L{LIBDIR}", "-lalglin_amd", f"-Wl,-rpath,{LIBDIR}", "-Wl,--allow-shlib-undefined",
            "-Wl,-rpath-link,/opt/rocm/lib"]


@pytest.fixture(scope="module")
def surface_exe(tmp_path_factory):
    _lib()
    exe = str(tmp_path_factory.mktemp("cxx") / "surface_test")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{INC}",
           os.path.join(CXX, "surface_test.cc"), "-o", exe] + _link_flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.fixture(scope="module")
def cinterface_exe(tmp_path_factory):
    _lib()
    exe = str(tmp_path_factory.mktemp("c") / "cinterface_test")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{INC}",
           os.path.join(CXX, "cinterface_test.c"), "-o", exe, "-lm"] + _link_flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cxx_surface_host(surface_exe):
    r = subprocess.run([surface_exe, "host"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_c_interface_compiles(cinterface_exe):
    assert os.path.exists(cinterface_exe)


@pytest.mark.gpu
def test_cxx_surface_gpu(surface_exe):
    r = subprocess.run([surface_exe, "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_c_interface_test9_compiled(cinterface_exe):
    r = subprocess.run([cinterface_exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
