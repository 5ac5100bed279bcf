"""GPU assembly (pform_mec isw 2 and 4) against the reference.

The GPU adds element contributions colour by colour, as the reference's
OpenMP element loop does (Pform.f:108-184), and computes every element matrix
with the reference's operations in the reference's order.  So K and b are
compared bit for bit with the reference's coloured assembly (fixtures *_omp)
and to 1e-14 relative with its serial assembly (same sums, other order).
"""
import numpy as np
import pytest

from tests.cases import CASES, SMALL

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", SMALL)
def test_assembly_matches_reference(golden, name):
    from elastic_amd.system import System

    g = golden(name)
    print("case", name, flush=True)
    s = System(CASES[name][0]())
    ad, al = s.values()
    assert np.array_equal(s.ia, g["ia"]) and np.array_equal(s.ja, g["ja"])
    if "al_omp" in g:
        assert np.array_equal(ad, g["ad_omp"])
        assert np.array_equal(al, g["al_omp"])
        assert np.array_equal(s.b, g["b_omp"])
    assert rel(ad, g["ad"]) <= 1e-14 and rel(al, g["al"]) <= 1e-14 and rel(s.b, g["b"]) <= 1e-14
    s.ctx.close()


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8", "tria3_8"])
def test_assembly_residual_and_solve(golden, name):
    """b0 - K u at the reference's solution (second nonlinear pass) and the
    full mesh -> GPU assembly -> GPU PCG path."""
    from elastic_amd.system import System

    g = golden(name)
    s = System(CASES[name][0]())
    r = s.residual(g["x"])
    assert rel(r, g["resid"]) <= 1e-6 or np.linalg.norm(r) <= 1e-9 * np.linalg.norm(s.b)
    x, st = s.solve(tol=float(g["tol"]))
    assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
    assert rel(x, g["x"]) <= 1e-10
    s.ctx.close()


def test_negative_jacobian_is_an_error():
    from elastic_amd import mesh as M
    from elastic_amd.solver import ElasticError
    from elastic_amd.system import System

    m = M.hex8_cube(3)
    m.ix[5, [0, 1]] = m.ix[5, [1, 0]]  # swap two nodes: inverted element 6
    with pytest.raises(ElasticError) as e:
        System(m)
    assert e.value.code == 3


def test_workload1_gpu_assembly_digest(golden):
    """Workload 1 assembled on the GPU: serial-order digests of the reference
    are matched to 1e-14 (norms) and the GPU solve reproduces its x."""
    from elastic_amd.system import System

    g = golden("hex8_20")
    s = System(CASES["hex8_20"][0]())
    ad, al = s.values()
    for k, v in (("ad", ad), ("al", al), ("b", s.b)):
        n = g["norm_" + k]
        assert abs(np.linalg.norm(v) - n[0]) <= 1e-13 * n[0]
    x, st = s.solve(tol=1e-10)
    assert rel(x, g["x"]) <= 1e-10
    s.ctx.close()
