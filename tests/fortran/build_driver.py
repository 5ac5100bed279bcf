"""Builds tests/fortran/_build/drive_solv and drive_batch: the Fortran test
programs of the solv drop-in and of the batch entry points, linked against
the binding's static library and libelastic_hip.so (found through an $ORIGIN-relative rpath, so the binary
works wherever the repository is copied)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "drive_solv")
EXE_BATCH = os.path.join(OUT, "drive_batch")


def build(fortran_build):
    """Builds drive_solv (the solv drop-in) and drive_batch (batch and tuning
    entry points through the ISO_C_BINDING module)."""
    from elastic_amd import build as B

    moddir, lib = fortran_build
    os.makedirs(OUT, exist_ok=True)
    for src, exe in (("drive_solv.f90", EXE), ("drive_batch.f90", EXE_BATCH)):
        cmd = [B.FLANG, "-O2", "-I", moddir, os.path.join(HERE, src), lib,
               "-L" + B.HERE, "-lelastic_hip", "-Wl,-rpath,$ORIGIN/../../../elastic_amd", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return EXE
