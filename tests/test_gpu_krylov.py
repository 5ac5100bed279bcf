"""GPU BiCGSTAB (Iters.f:1079-1349) and GMRES(50) (Iters.f:528-808) against
the reference's own solutions of the fixture systems (the harness calls
pbicgstab and gmres with their own signatures; their solver-call sites are
broken in the reference).

Bars (written here, checked per case):
* FP64: iterations within +-max(3, 5%) of the reference's (only the
  summation order of the dot products differs; BiCGSTAB's residual is not
  monotone, so rounding moves its exit point more than PCG's), x within 1e-8
  relative L2 of the reference's x, and the reference's stopping rule met by
  the explicit residual up to the 3.16 factor of its final check.
* Mixed precision (FP32-value operator, FP64 true residual): x within 1e-8
  relative L2 of the reference's x.
* 3x3 block Jacobi: x within 1e-8 of the reference's PCG solution.
"""
import numpy as np
import pytest

from tests.cases import CASES, SMALL

GMRES_CASES = [n for n in SMALL if CASES[n][2] & 2]      # fixtures with the reference's gmres
BICGSTAB_CASES = [n for n in SMALL if CASES[n][2] & 4]   # ... and pbicgstab solutions

pytestmark = pytest.mark.gpu

BICGSTAB, GMRES = 4, 2
MIXED = 32
BDIAG3 = 6


@pytest.fixture(scope="module")
def ctx():
    from elastic_amd.solver import Context

    c = Context(0)
    yield c
    c.close()


def load(ctx, g):
    ctx.set_structure(g["ia"], g["ja"])
    ctx.set_values(g["ad"], g["al"])


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", BICGSTAB_CASES)
def test_bicgstab_fp64(ctx, golden, name):
    g = golden(name)
    load(ctx, g)
    tol = float(g["tol"])
    print("bicgstab fp64", name, flush=True)
    x, st = ctx.solve(g["b"], solver=BICGSTAB, tol=tol)
    ref = int(g["bicgstab_iters"])
    assert abs(st["iters"] - ref) <= max(3, 0.05 * ref)
    assert rel(x, g["x_bicgstab"]) <= 1e-8
    assert st["norm_r"] < 3.16 * st["conv"]


@pytest.mark.parametrize("name", GMRES_CASES)
def test_gmres_fp64(ctx, golden, name):
    g = golden(name)
    load(ctx, g)
    tol = float(g["tol"])
    print("gmres fp64", name, flush=True)
    x, st = ctx.solve(g["b"], solver=GMRES, tol=tol, ngram=50)
    ref = int(g["gmres_iters"])
    assert abs(st["iters"] - ref) <= max(3, 0.05 * ref)
    assert rel(x, g["x_gmres"]) <= 1e-8


@pytest.mark.parametrize("solver", [BICGSTAB, GMRES])
@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8"])
def test_krylov_mixed(ctx, golden, name, solver):
    g = golden(name)
    load(ctx, g)
    print("krylov mixed", name, solver, flush=True)
    x, st = ctx.solve(g["b"], solver=solver, precision=MIXED, tol=float(g["tol"]))
    ref = g["x_bicgstab"] if solver == BICGSTAB else g["x_gmres"]
    assert rel(x, ref) <= 1e-8
    if solver == BICGSTAB:
        assert st["norm_r"] < st["conv"]


@pytest.mark.parametrize("solver", [BICGSTAB, GMRES])
def test_krylov_block_jacobi(ctx, golden, solver):
    g = golden("hex8_5j")
    load(ctx, g)
    x, st = ctx.solve(g["b"], solver=solver, precond=BDIAG3, tol=1e-12)
    assert rel(x, g["x"]) <= 1e-8


@pytest.mark.parametrize("solver", [BICGSTAB, GMRES])
def test_krylov_deterministic(ctx, golden, solver):
    g = golden("tet4_3j")
    load(ctx, g)
    x1, s1 = ctx.solve(g["b"], solver=solver, tol=1e-10)
    x2, s2 = ctx.solve(g["b"], solver=solver, tol=1e-10)
    assert np.array_equal(x1, x2) and s1["iters"] == s2["iters"]


def test_gmres_maxit_is_reported(ctx, golden):
    from elastic_amd._lib import MAXIT as MAXIT_CODE

    g = golden("hex8_5j")
    load(ctx, g)
    x, st = ctx.solve(g["b"], solver=GMRES, tol=1e-12, maxit=1, ngram=5, allow_maxit=True)
    assert st["status"] == MAXIT_CODE and st["iters"] == 5 and st["outer"] == 1


def test_mixed_gmres_converging_in_the_last_allowed_cycle(ctx, golden):
    """Mixed GMRES tests its FP64 residual at the top of the next cycle; when
    the cycle limit (maxit) ends the loop right after the cycle that
    converged, the final FP64 test still sees it: with maxit equal to the
    cycles a free run needs, the solve reports success and the same x."""
    g = golden("hex8_5j")
    load(ctx, g)
    x0, s0 = ctx.solve(g["b"], solver=GMRES, precision=MIXED, tol=1e-10, ngram=10, maxit=1000)
    assert s0["status"] == 0 and s0["outer"] >= 2
    x1, s1 = ctx.solve(g["b"], solver=GMRES, precision=MIXED, tol=1e-10, ngram=10,
                       maxit=s0["outer"])
    print("gmres cycles", s0["outer"], s1["outer"], s1["status"], flush=True)
    assert s1["status"] == 0 and s1["outer"] == s0["outer"]
    assert np.array_equal(x0, x1)
    assert rel(x1, g["x"]) <= 1e-8
