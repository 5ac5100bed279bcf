"""Result files against the reference's own writer.

tests/golden/*.npz hold, for each small case, the nodal stresses and the
displacements the reference computed (tform_mec, update_mec) and the three
files its write_mesh_res_mec (Write_vtk.f:742) wrote from them: legacy ASCII,
legacy big-endian binary and XML ASCII.  elastic_amd.vtk must write the same
bytes from the same arrays.
"""
import os

import numpy as np
import pytest

from elastic_amd import vtk
from tests.cases import CASES

GOLD = os.path.join(os.path.dirname(__file__), "golden")
STRESS_CASES = ["hex8_4", "hex8_5j", "tet4_3j", "quad4_8", "tria3_8", "quad4_6ps", "tria3_6ps"]


@pytest.mark.parametrize("name", STRESS_CASES)
@pytest.mark.parametrize("fmt", [("vtk_ascii", True, False), ("vtk_binary", True, True),
                                 ("vtu_ascii", False, False)])
def test_result_file_bytes(tmp_path, name, fmt):
    key, legacy, binary = fmt
    m = CASES[name][0]()
    g = np.load(os.path.join(GOLD, name + ".npz"))
    ntn = 6 if m.ndm == 3 else 4
    u = g["u"].reshape(m.nnode, m.ndf)
    tx = g["tx"].reshape(m.nnode, ntn)
    path = vtk.write_result(str(tmp_path / "res"), m.ix, m.x, u, tx, m.nen, m.ndm, m.ndf, 0,
                            legacy, binary)
    assert os.path.basename(path) == "res_istep_0." + ("vtk" if legacy else "vtu")
    assert open(path, "rb").read() == bytes(g[key])


def test_fortran_e_edit():
    # Ew.d / Ew.dEe as the reference's files show them
    assert vtk.f_e(0.0, 15, 5) == "    0.00000E+00"
    assert vtk.f_e(1.0, 15, 5) == "    0.10000E+01"
    assert vtk.f_e(-0.0673849, 15, 5, 3) == "  -0.67385E-001"
    assert vtk.f_e(9.999996, 15, 5) == "    0.10000E+02"   # carry into the exponent
    assert vtk.f_e(1e-120, 15, 5) == "    0.10000-119"   # Ew.d beyond exponent 99 drops the E
    assert vtk.f_i(12345678901, 10) == "*" * 10


def test_result_name():
    assert vtk.result_name("out/case1", 3, True) == "out/case1_istep_3.vtk"
    assert vtk.result_name("deck.dat", 0, False) == "deck_istep_0.vtu"
