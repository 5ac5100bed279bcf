"""The oracle's nodal stresses (tform_mec with the elements' isw = 3) against
the reference's own output and against the patch test.

* tests/golden/*.npz hold the reference's tx for u of its own solve: the
  restatement must reproduce it bit for bit.
* Patch test (independent of the reference): a linear displacement field is
  reproduced exactly by every element here, so every nodal stress must equal
  D eps of that field, on jittered meshes too.  This catches a wrong
  constitutive term, a swapped shear component or a sign that the reference
  vectors alone might leave open.
"""
import os

import numpy as np
import pytest

from elastic_amd import mesh as M
from oracle import binding as O
from tests.cases import CASES

GOLD = os.path.join(os.path.dirname(__file__), "golden")
STRESS_CASES = ["hex8_4", "hex8_5j", "tet4_3j", "quad4_8", "tria3_8", "quad4_6ps", "tria3_6ps"]


@pytest.mark.parametrize("name", STRESS_CASES)
def test_tform_matches_reference(name):
    m = CASES[name][0]()
    g = np.load(os.path.join(GOLD, name + ".npz"))
    ntn = 6 if m.ndm == 3 else 4
    tx = O.tform(m, g["u"].reshape(m.nnode, m.ndf))
    assert np.array_equal(tx, g["tx"].reshape(m.nnode, ntn))
    # update_mec of the reference's x gives the reference's u
    assert np.array_equal(O.System(m).displacements(g["x"]), g["u"].reshape(m.nnode, m.ndf))


def _d3(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam, mu


GRAD3 = np.array([[1.0e-3, 2.0e-4, -3.0e-4], [5.0e-4, -7.0e-4, 1.0e-4], [-2.0e-4, 4.0e-4, 9.0e-4]])
GRAD2 = np.array([[1.0e-3, 6.0e-4], [-2.0e-4, -8.0e-4]])


@pytest.mark.parametrize("kind", ["hex8", "tet4"])
def test_patch_3d(kind):
    m = (M.hex8_cube(4, jitter=0.2, seed=11) if kind == "hex8"
         else M.tet4_cube(3, jitter=0.2, seed=12, logE=(0.0, 0.0)))
    E, nu = m.e[:, 0], m.e[:, 1]
    u = m.x @ GRAD3.T + np.array([1e-3, -2e-3, 5e-4])   # u_i = G_ij x_j + rigid shift
    tx = O.tform(m, u)
    eps = np.array([GRAD3[0, 0], GRAD3[1, 1], GRAD3[2, 2], GRAD3[0, 1] + GRAD3[1, 0],
                    GRAD3[1, 2] + GRAD3[2, 1], GRAD3[0, 2] + GRAD3[2, 0]])
    assert np.ptp(E) == 0 and np.ptp(nu) == 0
    lam, mu = _d3(E[0], nu[0])
    tr = eps[:3].sum()
    want = np.concatenate([lam * tr + 2 * mu * eps[:3], mu * eps[3:]])
    assert np.allclose(tx, want[None, :], rtol=1e-10, atol=1e-16)


@pytest.mark.parametrize("tria", [False, True])
@pytest.mark.parametrize("plane_stress", [False, True])
def test_patch_2d(tria, plane_stress):
    build = M.tria3_square if tria else M.quad4_square
    m = build(6, E=1.7, nu=0.28, plane_stress=plane_stress, thickness=0.5)
    rng = np.random.default_rng(5)
    inner = (m.x[:, 0] > 1e-9) & (m.x[:, 0] < 1 - 1e-9) & (m.x[:, 1] > 1e-9) & (m.x[:, 1] < 1 - 1e-9)
    m.x[inner] += rng.uniform(-0.03, 0.03, (inner.sum(), 2))     # distorted but valid
    u = m.x @ GRAD2.T + np.array([2e-3, -1e-3])
    tx = O.tform(m, u)
    E, nu = 1.7, 0.28
    ex, ey, gxy = GRAD2[0, 0], GRAD2[1, 1], GRAD2[0, 1] + GRAD2[1, 0]
    if plane_stress:
        c = E / (1 - nu * nu)
        sxx, syy, szz = c * (ex + nu * ey), c * (nu * ex + ey), 0.0
    else:
        lam, mu = _d3(E, nu)
        sxx, syy = lam * (ex + ey) + 2 * mu * ex, lam * (ex + ey) + 2 * mu * ey
        szz = nu * (sxx + syy)
    sxy = E / (2 * (1 + nu)) * gxy
    assert np.allclose(tx, np.array([sxx, syy, szz, sxy])[None, :], rtol=1e-10, atol=1e-16)


def test_tform_initial_stress_and_count():
    m = CASES["quad4_8"][0]()
    g = np.load(os.path.join(GOLD, "quad4_8.npz"))
    u = g["u"].reshape(m.nnode, m.ndf)
    tx0 = np.random.default_rng(3).normal(size=(m.nnode, 4))
    a, b = O.tform(m, u), O.tform(m, u, tx0)
    assert np.allclose(b - a, tx0, rtol=0, atol=1e-12)
