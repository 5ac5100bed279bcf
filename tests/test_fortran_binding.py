"""The Fortran side of the boundary: the ISO_C_BINDING module and the solv
replacement (elastic_amd/fortran) compile with flang and link against
libelastic_hip.so, and a Fortran program that calls solv exactly as
Main.f:820-826 does gets the reference's m and x back.

Bars: m = 1/ad bit for bit (Solv.f:65-69); x within 1e-10 relative L2 of the
reference's PCG solution (FP64, only the summation order differs); a zero
diagonal stops the program through stop_mef (Iters.f:337-342).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from elastic_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "fortran", "_build", "drive_solv")
DRIVER_BATCH = os.path.join(HERE, "fortran", "_build", "drive_batch")
HEADER = os.path.join(HERE, "..", "include", "elastic_hip.h")
MODULE = os.path.join(B.HERE, "fortran", "elastic_hip.f90")
have_flang = os.path.exists(B.FLANG)


@pytest.mark.skipif(not have_flang, reason="flang not installed")
def test_binding_compiles_and_links(tmp_path):
    objs = []
    for src in ("elastic_hip.f90", "solv_gpu.f90"):
        obj = str(tmp_path / src.replace(".f90", ".o"))
        subprocess.run([B.FLANG, "-c", "-O2", "-module-dir", str(tmp_path), "-I", str(tmp_path),
                        os.path.join(B.HERE, "fortran", src), "-o", obj], check=True)
        objs.append(obj)
    exe = str(tmp_path / "drive")
    subprocess.run([B.FLANG, "-O2", "-I", str(tmp_path), os.path.join(HERE, "fortran", "drive_solv.f90")]
                   + objs + ["-L" + B.HERE, "-lelastic_hip", "-o", exe], check=True)
    assert os.path.exists(exe)


def write_system(path, ia, ja, ad, al, b, tol, maxit):
    with open(path, "wb") as f:
        np.array([ia.shape[0] - 1, ja.shape[0]], np.int32).tofile(f)
        ia.astype(np.int32).tofile(f)
        ja.astype(np.int32).tofile(f)
        ad.astype(np.float64).tofile(f)
        al.astype(np.float64).tofile(f)
        b.astype(np.float64).tofile(f)
        np.array([tol], np.float64).tofile(f)
        np.array([maxit], np.int32).tofile(f)


def run_driver(tmp_path, g, ad=None):
    ad = g["ad"] if ad is None else ad
    fin, fout = str(tmp_path / "sys.bin"), str(tmp_path / "out.bin")
    write_system(fin, g["ia"], g["ja"], ad, g["al"], g["b"], float(g["tol"]), 10000)
    print("fortran drive_solv: neq=%d" % ad.shape[0], flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    out = np.fromfile(fout, np.float64) if os.path.exists(fout) else None
    return r, out


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
@pytest.mark.parametrize("name", ["hex8_5j", "quad4_8", "tet4_3j"])
def test_fortran_solv_dropin(tmp_path, golden, name):
    g = golden(name)
    r, out = run_driver(tmp_path, g)
    assert r.returncode == 0, r.stdout + r.stderr
    n = g["ad"].shape[0]
    m, x = out[:n], out[n:]
    assert np.array_equal(m, 1.0 / g["ad"])
    assert np.linalg.norm(x - g["x"]) <= 1e-10 * np.linalg.norm(g["x"])


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_solv_zero_diagonal_stops(tmp_path, golden):
    g = golden("hex8_5j")
    ad = g["ad"].copy()
    ad[4] = 0.0
    r, _ = run_driver(tmp_path, g, ad=ad)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "status" in r.stdout


def test_every_entry_point_has_a_fortran_interface():
    """Each function that include/elastic_hip.h declares is bound by name in
    the ISO_C_BINDING module, so a Fortran caller reaches all of them."""
    import re

    hdr = open(HEADER).read()
    declared = set(re.findall(r"^(?:int|int32_t|const char\*)\s+(elastic_\w+)\(", hdr, re.M))
    bound = set(re.findall(r'bind\(C, name="(elastic_\w+)"\)', open(MODULE).read()))
    assert len(declared) >= 40
    assert declared - bound == set()
    assert bound - declared == set()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER_BATCH), reason="Fortran driver not built")
@pytest.mark.parametrize("name", ["hex8_5j", "quad4_8"])
def test_fortran_batch_and_options(tmp_path, golden, name):
    """A Fortran program drives two contexts through one batch (FP64 and
    MIXED) and a tuning knob through elastic_set_option/elastic_get_option:
    x within 1e-10 (FP64) and 1e-8 (MIXED) of the reference's PCG solution."""
    g = golden(name)
    fin, fout = str(tmp_path / "sys.bin"), str(tmp_path / "out.bin")
    write_system(fin, g["ia"], g["ja"], g["ad"], g["al"], g["b"], float(g["tol"]), 10000)
    print("fortran drive_batch:", name, flush=True)
    r = subprocess.run([DRIVER_BATCH, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, np.float64)
    n = g["ad"].shape[0]
    x64, x32, knob, it64, it32 = out[:n], out[n:2 * n], out[2 * n], out[2 * n + 1], out[2 * n + 2]
    nx = np.linalg.norm(g["x"])
    assert np.linalg.norm(x64 - g["x"]) <= 1e-10 * nx
    assert np.linalg.norm(x32 - g["x"]) <= 1e-8 * nx
    assert knob == 8
    assert abs(it64 - int(g["pcg_iters"])) <= 2 and it32 > 0
