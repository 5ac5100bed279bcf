"""Builds tests/fortran/_build/drive_solv: the Fortran test program of the
solv drop-in, linked against the binding's static library and
libelastic_hip.so (found through an $ORIGIN-relative rpath, so the binary
works wherever the repository is copied)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "drive_solv")


def build(fortran_build):
    from elastic_amd import build as B

    moddir, lib = fortran_build
    os.makedirs(OUT, exist_ok=True)
    cmd = [B.FLANG, "-O2", "-I", moddir, os.path.join(HERE, "drive_solv.f90"), lib,
           "-L" + B.HERE, "-lelastic_hip", "-Wl,-rpath,$ORIGIN/../../../elastic_amd", "-o", EXE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return EXE
