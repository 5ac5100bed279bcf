"""Edge cases of the solve path and of the stress pass.

* b = 0: every method returns x = 0 at once (the reference's pcg stops at
  once too: sqrt|b.Mb| = 0 makes conv = 0 and its first test passes).
* One hex8 element with a clamped face (12 equations): the solution equals a
  dense direct solve of the same K (numpy), not only the reference's PCG.
* Index arrays outside the mesh are refused on the host (elastic_set_mesh).
* A hex8 whose Jacobian is positive at its Gauss points but not at one of its
  nodes assembles and solves, and its nodal stresses are refused with
  ELASTIC_NEG_JACOBIAN, as tform_mec's jacob3d_m stops on it; the oracle
  finds the same element.
"""


import numpy as np
import pytest

from elastic_amd import _lib
from elastic_amd import mesh as M
from elastic_amd.solver import BICGSTAB, FP64, GMRES, MIXED, PCG, ElasticError
from elastic_amd.system import System
from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver,precision", [(PCG, FP64), (PCG, MIXED), (BICGSTAB, FP64),
                                              (GMRES, FP64), (BICGSTAB, MIXED), (GMRES, MIXED)])
def test_zero_rhs(solver, precision):
    s = System(M.hex8_cube(3))
    x, st = s.ctx.solve(np.zeros(s.neq), solver=solver, precision=precision, tol=1e-10)
    assert st["status"] == 0
    assert np.all(x == 0.0)
    assert st["iters"] == 0


def _dense(s):
    K = np.zeros((s.neq, s.neq))
    ad, al = s.ctx.get_values()
    K[np.arange(s.neq), np.arange(s.neq)] = ad
    for i in range(s.neq):
        for k in range(s.ia[i] - 1, s.ia[i + 1] - 1):
            j = s.ja[k] - 1
            K[i, j] = K[j, i] = al[k]
    return K


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_single_element_against_dense_solve(precision):
    s = System(M.hex8_cube(1, E=2.0, nu=0.25))
    assert s.neq == 12
    K = _dense(s)
    xd = np.linalg.solve(K, s.b)
    x, st = s.ctx.solve(s.b, precision=precision, tol=1e-13)
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)


def test_set_mesh_refuses_out_of_range_nodes():
    m = M.hex8_cube(2)
    s = System(m, assemble=False)
    ix = np.ascontiguousarray(m.ix, np.int32).copy()
    ix[3, 2] = m.nnode + 1
    lt, ln, lv = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(3)
    rc = s.ctx._L.elastic_set_mesh(
        s.ctx._h, m.nnode, m.numel, m.nen, m.ndf, m.ndm, m.numat, ix,
        np.ascontiguousarray(m.x, np.float64), s.idn, np.ascontiguousarray(m.ie, np.int32),
        np.ascontiguousarray(m.e, np.float64), np.ascontiguousarray(m.eload, np.int32), 0, lt, ln, lv)
    assert rc == _lib.BAD_ARG


def _bad_node_hex():
    m = M.hex8_cube(1)
    n0 = m.ix[0, 0] - 1                       # local node 1, a free (z = 1) corner
    m.x[n0] = m.x[n0] + 0.45 * (np.array([0.5, 0.5, 0.5]) - m.x[n0]) * 2.0
    return m


def test_stress_refuses_non_positive_nodal_jacobian():
    m = _bad_node_hex()
    with pytest.raises(ValueError):
        O.tform(m, np.zeros((m.nnode, m.ndf)))       # the oracle stops on element 1
    s = System(m)                                    # Gauss points are fine: K assembles
    x, st = s.solve(tol=1e-10)
    assert st["status"] == 0
    with pytest.raises(ElasticError) as e:
        s.stress(from_x=True)
    assert e.value.code == _lib.NEG_JACOBIAN

