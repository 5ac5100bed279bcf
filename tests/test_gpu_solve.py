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


@pytest.mark.parametrize("precision", [64, 32])
def test_block_jacobi_pcg(ctx, golden, precision):
    """3x3 block-Jacobi (Precond.f:163-540) as the PCG preconditioner: the
    solution agrees with the reference's Jacobi-PCG solution."""
    from elastic_amd.solver import BDIAG3

    g = golden("hex8_5j")
    load_system(ctx, g)
    x, st = ctx.solve(g["b"], precond=BDIAG3, precision=precision, tol=1e-12)
    assert rel(x, g["x"]) <= 1e-8
    xj, sj = ctx.solve(g["b"], precision=precision, tol=1e-12)
    assert st["iters"] < sj["iters"]


def test_mixed_reliable_updates_keep_iterations(ctx, golden):
    """FP32 iteration with FP64 residual replacement: iteration count close to
    the FP64 PCG's (no restarts) and several replacements were made."""
    from elastic_amd.solver import MIXED
    from oracle import binding as O
    from tests.cases import CASES

    s = O.System(CASES["hex8_20"][0]())
    ctx.set_structure(s.ia, s.ja)
    ctx.set_values(s.ad, s.al[: s.nad])
    x64, s64 = ctx.solve(s.b, tol=1e-10)
    xm, sm = ctx.solve(s.b, tol=1e-10, precision=MIXED)
    assert sm["iters"] <= 1.3 * s64["iters"] + 10
    assert sm["outer"] >= 3
    assert sm["norm_m_r"] < sm["conv"]


@pytest.mark.parametrize("layout", ["scalar", "blocked"])
@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8", "tria3_8"])
def test_layouts_agree(golden, name, layout, monkeypatch):
    """Scalar CSRC tiles and node-blocked tiles give the same product and PCG."""
    from elastic_amd.solver import MIXED, Context

    if layout == "scalar":
        monkeypatch.setenv("ELASTIC_LAYOUT", "scalar")
    g = golden(name)
    with Context(0) as c:
        load_system(c, g)
        nb = c.layout_info()["nb"]
        assert nb == (1 if layout == "scalar" else (3 if name[:3] in ("hex", "tet") else 2))
        assert rel(c.spmv(g["spmv_v"]), g["spmv_y"]) <= 1e-13
        assert rel(c.spmv(g["spmv_v"], precision=MIXED), g["spmv_y"]) <= 2e-6
        x, st = c.solve(g["b"], tol=float(g["tol"]))
        assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
        assert rel(x, g["x"]) <= 1e-10


def test_partially_restrained_nodes_use_scalar_tiles():
    """Rollers (only uz fixed on z=0) leave nodes with 2 free dofs: the node
    blocks do not apply and the scalar tiles are used; results match the oracle."""
    from elastic_amd import mesh as M
    from elastic_amd.solver import Context
    from oracle import binding as O

    m = M.hex8_cube(4, jitter=0.1, seed=3)
    z0 = m.x[:, 2] == 0.0
    m.id[z0] = [0, 0, 1]
    m.id[(m.x[:, 0] == 0.0) & z0] = [1, 1, 1]
    s = O.System(m)
    with Context(0) as c:
        c.set_structure(s.ia, s.ja)
        c.set_values(s.ad, s.al[: s.nad])
        assert c.layout_info()["nb"] == 1
        x, st = c.solve(s.b, tol=1e-10)
    xo, so = O.pcg(s.ia, s.ja, s.ad, s.al[: s.nad], s.b, tol=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 2 and rel(x, xo) <= 1e-10


@pytest.mark.parametrize("name", ["hex8_5j", "quad4_8"])
def test_autotune_keeps_results(golden, name):
    """The autotuner rebuilds the layout with the fastest tile size and lanes
    per tile (timed inside a short PCG); the product and the solves stay
    within the same bars."""
    from elastic_amd.solver import FP64, MIXED, Context

    g = golden(name)
    with Context(0) as c:
        load_system(c, g)
        best, ms = c.autotune(precision=MIXED, reps=16)
        assert best in (256, 512, 768, 1024) and ms > 0
        assert best == c.get_option("tile_blocks")
        assert c.get_option("lanes32") in (256, 512)
        assert rel(c.spmv(g["spmv_v"]), g["spmv_y"]) <= 1e-13
        x, st = c.solve(g["b"], tol=float(g["tol"]))
        assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
        assert rel(x, g["x"]) <= 1e-10
        # lanes only: the tile size stays, the FP64 product still agrees
        best64, _ = c.autotune(precision=FP64, reps=16, tiles=False)
        assert best64 == best and c.get_option("lanes64") in (256, 512)
        assert rel(c.spmv(g["spmv_v"]), g["spmv_y"]) <= 1e-13
        xm, sm = c.solve(g["b"], precision=MIXED, tol=float(g["tol"]))
        assert rel(xm, g["x"]) <= 1e-8


@pytest.mark.parametrize("name", ["hex8_5j", "quad4_8", "tria3_8", "tet4_3j"])
def test_fp32_accumulation_option(golden, name):
    """The FP32 operator with FP32 staging and sums ("accum32"): the product
    within 2e-5 relative L2 of the reference's FP64 product (FP32 sums over
    up to 81 terms per row), and the mixed PCG still within 1e-8 of the
    reference's x (the FP64 true-residual replacement corrects it)."""
    from elastic_amd.solver import MIXED, Context

    g = golden(name)
    with Context(0) as c:
        c.set_option("accum32", 1)
        load_system(c, g)
        assert rel(c.spmv(g["spmv_v"], precision=MIXED), g["spmv_y"]) <= 2e-5
        x, st = c.solve(g["b"], precision=MIXED, tol=float(g["tol"]))
        assert st["norm_m_r"] < st["conv"]
        assert rel(x, g["x"]) <= 1e-8


def test_product_timing_option(golden):
    """The measurement option samples every S-th PCG product with HIP events:
    the count equals the sampled iterations that ran, the average is a
    positive duration, reset clears it, and the solution is unchanged."""
    from elastic_amd.solver import Context

    g = golden("hex8_5j")
    with Context(0) as c:
        load_system(c, g)
        x0, s0 = c.solve(g["b"], tol=1e-10)
        c.set_option("time_products", 4)
        c.product_timing(reset=True)
        x1, s1 = c.solve(g["b"], tol=1e-10)
        ms, n = c.product_timing(reset=True)
        assert np.array_equal(x0, x1) and s0["iters"] == s1["iters"]
        assert n == (s1["iters"] + 3) // 4 and ms > 0.0
        assert c.product_timing() == (0.0, 0)
        c.set_option("time_products", 0)


@pytest.mark.parametrize("accum32", [0, 1])
def test_mixed_hands_over_to_fp64(monkeypatch, accum32):
    """The progress guard's hand-over, forced from the host on hex8 12^3: with
    the replacement threshold at 2 every second iteration replaces the
    residual, which then falls by less than the guard's 10% per replacement,
    so after three replacements the solve continues with the FP64 operator
    (with FP32 sums the iteration's FP32 p and r are copied to their FP64
    vectors first), restarted from the true residual, as FP64 PCG with its
    stopping rule.  It still converges to the tight FP64 solution within
    1e-8, and its true residual meets the reference's final check
    (3.16 tol ||b||_M, Iters_omp.f:254)."""
    from elastic_amd import mesh as M
    from elastic_amd.solver import FP64, MIXED
    from elastic_amd.system import System

    s = System(M.hex8_cube(12, jitter=0.15, seed=M.BASE_SEED + 2))
    c = s.ctx
    xref, _ = c.solve(s.b, precision=FP64, tol=1e-12)
    monkeypatch.setenv("ELASTIC_RELIABLE_DELTA", "2.0")
    c.set_option("accum32", accum32)
    x, st = c.solve(s.b, precision=MIXED, tol=1e-10)
    assert 0 <= st["op64_from"] <= 20
    assert st["norm_m_r"] < 3.16 * st["conv"]
    assert rel(x, xref) <= 1e-8
    c.close()


@pytest.mark.parametrize("accum32", [0, 1])
def test_rows_not_a_multiple_of_four(accum32):
    """hex8 6^3 has 882 equations (882 = 4 * 220 + 2): the p updates handle
    four rows per lane and the last two rows in a tail loop, in FP64 and with
    the mixed iteration's FP32 vectors.  FP64 PCG matches the oracle's
    iteration count within 2 and its x within 1e-10; mixed PCG lands within
    1e-8 of the FP64 x."""
    from elastic_amd import mesh as M
    from elastic_amd.solver import FP64, MIXED
    from elastic_amd.system import System
    from oracle import binding as O

    m = M.hex8_cube(6, jitter=0.15, seed=M.BASE_SEED + 2)
    s = System(m)
    assert s.neq == 882
    ref = O.System(m)
    xo, so = O.pcg(ref.ia, ref.ja, ref.ad, ref.al, ref.b, tol=1e-10, maxit=5000)
    c = s.ctx
    c.set_option("accum32", accum32)
    x64, st64 = c.solve(s.b, precision=FP64, tol=1e-10)
    assert abs(st64["iters"] - so["iters"]) <= 2 and rel(x64, xo) <= 1e-10
    xm, stm = c.solve(s.b, precision=MIXED, tol=1e-10)
    assert stm["norm_m_r"] < stm["conv"] and rel(xm, x64) <= 1e-8
    c.close()


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j"])
@pytest.mark.parametrize("order", [0, 1])
def test_block_jacobi_fp32_vectors(golden, name, order):
    """The mixed PCG with FP32 sums keeps its vectors in FP32 with the 3x3
    node-block inverse too (applied per node in the update, the p update and
    the replacement pass, Precond.f:419-540): x within 1e-8 of the
    reference's, in both node orders, and the iteration count within 15% of
    the mixed block-Jacobi PCG that keeps FP64 vectors (FP64 sums).  The FP32
    values themselves cost iterations against FP64 PCG on these small meshes
    (tet4_3j: 168 against 121), in either form."""
    from elastic_amd.solver import BDIAG3, Context

    g = golden(name)
    with Context(0) as c:
        c.set_option("node_order", order)
        load_system(c, g)
        c.set_option("accum32", 0)
        x64, s64 = c.solve(g["b"], precond=BDIAG3, precision=32, tol=1e-12)
        c.set_option("accum32", 1)
        x, st = c.solve(g["b"], precond=BDIAG3, precision=32, tol=1e-12)
        print("bdiag fp32 vectors", name, order, st["iters"], s64["iters"], flush=True)
        assert st["status"] == 0 and rel(x, g["x"]) <= 1e-8
        assert st["iters"] <= 1.15 * s64["iters"] + 3


def test_block_jacobi_fp32_vectors_large():
    """hex8 40^3 (201,720 equations: FP32 increments of x, gated replacement
    passes, compact node order): the mixed block-Jacobi PCG with FP32 vectors
    lands within 1e-8 of the FP64 block-Jacobi solution at 1e-12, and alone
    and in a second run bit-identical."""
    from elastic_amd import mesh as M
    from elastic_amd.solver import BDIAG3, FP64, MIXED
    from elastic_amd.system import System

    s = System(M.hex8_cube(40, jitter=0.15, seed=M.BASE_SEED + 2))
    c = s.ctx
    c.upload_rhs(s.b)
    c.solve_resident(precond=BDIAG3, precision=FP64, tol=1e-12, maxit=100000)
    xref = c.download_x()
    st = c.solve_resident(precond=BDIAG3, precision=MIXED, tol=1e-10, maxit=100000)
    x = c.download_x()
    st2 = c.solve_resident(precond=BDIAG3, precision=MIXED, tol=1e-10, maxit=100000)
    x2 = c.download_x()
    print("hex8_40 bdiag fp32 vectors", st["iters"], st["outer"], rel(x, xref), flush=True)
    assert st["status"] == 0 and rel(x, xref) <= 1e-8
    assert np.array_equal(x, x2) and st["iters"] == st2["iters"]
    c.close()
