"""Nodal stresses on the GPU (elastic_stress: tform_mec with the elements'
isw = 3) and a deck run end to end (deck -> GPU set-up, assembly, solve,
stresses -> result file).

* From the reference's u, the GPU stresses equal the reference's tx bit for
  bit (tests/golden, written by the reference's own tform_mec).
* From the GPU's own solution (resident x, update_mec on the device), they
  equal the oracle's tform of the same u bit for bit, at hex8 20^3.
* A deck written for the reference runs through elastic_amd.run and leaves
  the reference's result file: same text, numbers within the 5 digits the
  format prints (the solution differs from the reference's in its last bits).
"""
import os
import re

import numpy as np
import pytest

from elastic_amd import deck
from elastic_amd.run import run
from elastic_amd.system import System
from oracle import binding as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
STRESS_CASES = ["hex8_4", "hex8_5j", "tet4_3j", "quad4_8", "tria3_8", "quad4_6ps", "tria3_6ps"]


@pytest.mark.parametrize("name", STRESS_CASES)
def test_gpu_stress_matches_reference(name):
    m = CASES[name][0]()
    g = np.load(os.path.join(GOLD, name + ".npz"))
    ntn = 6 if m.ndm == 3 else 4
    s = System(m)
    tx = s.stress(u=g["u"].reshape(m.nnode, m.ndf))
    assert np.array_equal(tx, g["tx"].reshape(m.nnode, ntn))


def test_gpu_stress_from_resident_solution_hex8_20():
    m = CASES["hex8_20"][0]()
    s = System(m)
    x, st = s.solve(tol=1e-10)
    assert st["status"] == 0
    tx = s.stress(from_x=True)
    assert np.array_equal(tx, O.tform(m, s.displacements(x)))
    # and twice in a row (buffers reused)
    assert np.array_equal(s.stress(from_x=True), tx)


def test_gpu_stress_initial_stress():
    m = CASES["tet4_3j"][0]()
    g = np.load(os.path.join(GOLD, "tet4_3j.npz"))
    u = g["u"].reshape(m.nnode, m.ndf)
    tx0 = np.random.default_rng(9).normal(size=(m.nnode, 6))
    s = System(m)
    assert np.array_equal(s.stress(u=u, tx0=tx0), O.tform(m, u, tx0))


_NUM = re.compile(r"[-+]?\d\.\d+E[-+]?\d+|[-+]?\d\.\d+[-+]\d{3}")


def _split(text):
    """Lines with their E-format numbers taken out, and the numbers."""
    skel, nums = [], []
    for line in text.splitlines():
        nums += [float(t.replace("E", "e")) for t in _NUM.findall(line)]
        skel.append(_NUM.sub("#", line))
    return skel, np.array(nums)


@pytest.mark.parametrize("name,legacy", [("quad4_8", True), ("hex8_5j", False), ("tria3_6ps", True)])
def test_deck_run_end_to_end(tmp_path, name, legacy):
    m = CASES[name][0]()
    g = np.load(os.path.join(GOLD, name + ".npz"))
    dpath = str(tmp_path / (name + ".dat"))
    deck.write_deck(m, dpath)
    r = run(dpath, str(tmp_path / "res"), tol=float(g["tol"]), legacy=legacy)
    ntn = 6 if m.ndm == 3 else 4
    u_ref = g["u"].reshape(m.nnode, m.ndf)
    assert np.linalg.norm(r["u"] - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    tx_ref = g["tx"].reshape(m.nnode, ntn)
    assert np.linalg.norm(r["tx"] - tx_ref) <= 1e-7 * np.linalg.norm(tx_ref)
    got = open(r["path"]).read()
    ref = bytes(g["vtk_ascii" if legacy else "vtu_ascii"]).decode()
    sg, ng = _split(got)
    sr, nr = _split(ref)
    assert sg == sr                       # same structure, ids, connectivity, headers
    assert ng.shape == nr.shape
    scale = np.abs(nr).max()
    assert np.allclose(ng, nr, rtol=2e-5, atol=1e-5 * scale)
