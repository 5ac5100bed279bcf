"""Builds elastic_amd/libelastic_hip.so (gfx950) from elastic_amd/csrc, and
the Fortran binding (elastic_amd/fortran) when flang is present.

Every source is compiled by hipcc for gfx950 only; objects go to
elastic_amd/_build/ and the shared library is written in-tree so that it
travels with the repository.  Packed-FP32 instructions are disabled on the
device side.  The Fortran interface module and the solv replacement are
compiled into elastic_amd/_build/fortran (elastic_hip.mod, libelastic_solv.a)
for a Fortran program to link with -lelastic_hip.
Usage: python -m elastic_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libelastic_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FOBJ = os.path.join(OBJ, "fortran")

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(HERE, "..", "include")]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose: bool = False):
    """Compiles the ISO_C_BINDING module and the solv replacement; returns
    (module dir, static library path), or None when flang is absent."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(FOBJ, exist_ok=True)
    fdir = os.path.join(HERE, "fortran")
    objs = []
    for src in ("elastic_hip.f90", "solv_gpu.f90"):
        obj = os.path.join(FOBJ, src.replace(".f90", ".o"))
        cmd = [FLANG, "-c", "-O2", "-fPIC", "-module-dir", FOBJ, "-I", FOBJ,
               os.path.join(fdir, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
        objs.append(obj)
    lib = os.path.join(FOBJ, "libelastic_solv.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.run(["ar", "rcs", lib] + objs, check=True)
    if verbose:
        print("built", lib)
    return FOBJ, lib


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
    sys.exit(0)
