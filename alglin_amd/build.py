"""Builds alglin_amd/libalglin_amd.so (HIP kernels + C++ host + C ABI) for gfx950.

Usage:  python -m alglin_amd.build [--force] [-j JOBS]

Every translation unit under alglin_amd/csrc is compiled by hipcc into
alglin_amd/_build/ (skipped when the object is newer than its sources), then
linked into alglin_amd/libalglin_amd.so in-tree, so the built library travels
with the source tree.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libalglin_amd.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ALGLIN_AMD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU ops are disabled on purpose (hipcc may otherwise emit
    # v_pk_fma_f32 & co from plain float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the register-resident kernels need their row loops fully unrolled
    "-mllvm", "-pragma-unroll-threshold=1000000",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-result", "-Wno-pass-failed",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cc")))


def headers():
    return (glob.glob(os.path.join(CSRC, "*.hh")) + glob.glob(os.path.join(CSRC, "*.inc"))
            + glob.glob(os.path.join(ROOT, "include", "*.h")))


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    deps = [src] + headers()
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + COMMON + lang + ["-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise RuntimeError("link of libalglin_amd.so failed")
        os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.force, a.j))
