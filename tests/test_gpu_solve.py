"""GPU parity: symmetric CSRC SpMV and the Krylov solves against the
reference's own outputs (golden fixtures: the reference's assembled system,
its product y = A v and its PCG / GMRES / BiCGSTAB solutions).

Bars (written here, checked per case):
* FP64 SpMV: relative L2 error <= 1e-13 (only the summation order differs).
* FP32 SpMV (mixed path): relative L2 error <= 2e-6.
* FP64 PCG: iterations within +-2 of the reference, relative L2 error of x
  <= 1e-10.
* Mixed precision (FP32 inner, FP64 refinement): FP64 residual meets the
  reference's stopping rule and relative L2 error of x <= 1e-8.
"""
import numpy as np
import pytest

from tests.cases import SMALL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from elastic_amd.solver import Context

    c = Context(0)
    yield c
    c.close()


def load_system(ctx, g):
    ctx.set_structure(g["ia"], g["ja"])
    ctx.set_values(g["ad"], g["al"])


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", SMALL)
def test_spmv_fp64(ctx, golden, name):
    g = golden(name)
    print("case", name, flush=True)
    load_system(ctx, g)
    y = ctx.spmv(g["spmv_v"])
    assert rel(y, g["spmv_y"]) <= 1e-13


@pytest.mark.parametrize("name", SMALL)
def test_spmv_fp32(ctx, golden, name):
    from elastic_amd.solver import MIXED

    g = golden(name)
    load_system(ctx, g)
    y = ctx.spmv(g["spmv_v"], precision=MIXED)
    assert rel(y, g["spmv_y"]) <= 2e-6


@pytest.mark.parametrize("tile", [128, 300, 2048])
def test_spmv_tile_sizes(golden, tile):
    """Tiles of a few rows (many external columns) to one large tile."""
    from elastic_amd.solver import Context

    g = golden("hex8_5j")
    with Context(0, tile_nnz=tile) as c:
        load_system(c, g)
        assert rel(c.spmv(g["spmv_v"]), g["spmv_y"]) <= 1e-13


@pytest.mark.parametrize("name", SMALL)
def test_pcg_fp64(ctx, golden, name):
    g = golden(name)
    print("case", name, flush=True)
    load_system(ctx, g)
    x, st = ctx.solve(g["b"], tol=float(g["tol"]))
    assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
    assert rel(x, g["x"]) <= 1e-10
    assert st["norm_m_r"] < 3.16 * st["conv"]


@pytest.mark.parametrize("name", SMALL)
def test_pcg_mixed(ctx, golden, name):
    from elastic_amd.solver import MIXED

    g = golden(name)
    load_system(ctx, g)
    x, st = ctx.solve(g["b"], precision=MIXED, tol=float(g["tol"]))
    assert st["norm_m_r"] < st["conv"]
    assert rel(x, g["x"]) <= 1e-8
    assert st["outer"] >= 1


def test_pcg_deterministic(ctx, golden):
    g = golden("hex8_5j")
    load_system(ctx, g)
    x1, s1 = ctx.solve(g["b"], tol=1e-10)
    x2, s2 = ctx.solve(g["b"], tol=1e-10)
    assert np.array_equal(x1, x2) and s1["iters"] == s2["iters"]


def test_workload1_hex8_20(ctx, golden):
    """Workload 1 (hex8 20^3, 26,460 equations): reference system assembled by
    the oracle (pinned bit-exact to the reference's digests), solved on the
    GPU; the solution is compared with the reference's."""
    from oracle import binding as O
    from tests.cases import CASES
    from elastic_amd.solver import MIXED

    g = golden("hex8_20")
    s = O.System(CASES["hex8_20"][0]())
    ctx.set_structure(s.ia, s.ja)
    ctx.set_values(s.ad, s.al[: s.nad])
    x, st = ctx.solve(s.b, tol=1e-10)
    assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
    assert rel(x, g["x"]) <= 1e-10
    xm, sm = ctx.solve(s.b, tol=1e-10, precision=MIXED)
    assert rel(xm, g["x"]) <= 1e-8


def test_zero_diagonal_is_an_error(ctx, golden):
    from elastic_amd.solver import ElasticError

    g = golden("quad4_8")
    ad = g["ad"].copy()
    ad[5] = 0.0
    ctx.set_structure(g["ia"], g["ja"])
    ctx.set_values(ad, g["al"])
    with pytest.raises(ElasticError) as e:
        ctx.solve(g["b"])
    assert e.value.code == 1


def test_maxit_is_reported(ctx, golden):
    from elastic_amd.solver import ElasticError

    g = golden("quad4_8")
    load_system(ctx, g)
    with pytest.raises(ElasticError) as e:
        ctx.solve(g["b"], maxit=3)
    assert e.value.code == 2
    x, st = ctx.solve(g["b"], maxit=3, allow_maxit=True)
    assert st["iters"] == 3
