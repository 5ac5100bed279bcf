"""The compact node order of the node-blocked layout (3-dof systems; the
default from 32,768 nodes, forced here on the small fixtures with the option
"node_order" = 1).  Inside a solve every vector is in the layout's order; b
and x cross between orders at the solve's ends, so every result the caller
sees is in the CSRC order and is compared with the reference's own outputs.

Bars (written here, checked per case):
* FP64 SpMV: relative L2 error <= 1e-13 of the reference's y; FP32 <= 2e-6.
* FP64 PCG: iterations within +-2 of the reference's, x within 1e-10.
* Mixed PCG: x within 1e-8 of the reference's x.
* 3x3 block-Jacobi PCG: x within 1e-10; the block inverse returned by
  elastic_get_bdiag is still the reference's, bit for bit.
* BiCGSTAB / GMRES(50), FP64: x within 1e-8 of the reference's x (their
  iteration counts move with the order of the sums, so they are not pinned
  here; the CSRC-order tests pin them).
* Batch: x within 1e-10 (FP64) of the reference's and bit-identical alone
  and in a mix of compact and CSRC-order instances.
* The two orders on one system: x within 1e-10, stresses from the resident x
  equal to stresses from the downloaded x.
"""
import numpy as np
import pytest

from tests.test_gpu_batch import run_batch

pytestmark = pytest.mark.gpu

FIX3 = ["hex8_4", "hex8_5j", "tet4_3j"]
FP64, MIXED = 64, 32
PCG, GMRES, BICGSTAB = 1, 2, 4
BDIAG3 = 6


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def compact_ctx(g, rhs=False):
    from elastic_amd.solver import Context

    c = Context(0)
    c.set_option("node_order", 1)
    c.set_structure(g["ia"], g["ja"])
    c.set_values(g["ad"], g["al"])
    assert c.get_option("node_order") == 1
    if rhs:
        c.upload_rhs(g["b"])
    return c


@pytest.mark.parametrize("name", FIX3)
def test_compact_spmv(golden, name):
    g = golden(name)
    print("compact spmv", name, flush=True)
    c = compact_ctx(g)
    assert rel(c.spmv(g["spmv_v"]), g["spmv_y"]) <= 1e-13
    assert rel(c.spmv(g["spmv_v"], precision=MIXED), g["spmv_y"]) <= 2e-6
    c.close()


@pytest.mark.parametrize("name", FIX3)
def test_compact_pcg(golden, name):
    g = golden(name)
    tol = float(g["tol"])
    print("compact pcg", name, flush=True)
    c = compact_ctx(g)
    x, st = c.solve(g["b"], tol=tol)
    assert abs(st["iters"] - int(g["pcg_iters"])) <= 2
    assert rel(x, g["x"]) <= 1e-10
    xm, sm = c.solve(g["b"], tol=min(tol, 1e-11), precision=MIXED)
    assert rel(xm, g["x"]) <= 1e-8
    xb, sb = c.solve(g["b"], tol=1e-12, precond=BDIAG3)
    assert rel(xb, g["x"]) <= 1e-10
    assert np.array_equal(c.get_bdiag(), g["bdiag"])
    c.close()


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j"])
@pytest.mark.parametrize("solver", [BICGSTAB, GMRES])
def test_compact_krylov(golden, name, solver):
    g = golden(name)
    print("compact krylov", name, solver, flush=True)
    c = compact_ctx(g)
    x, st = c.solve(g["b"], solver=solver, tol=float(g["tol"]), ngram=50)
    assert rel(x, g["x_bicgstab" if solver == BICGSTAB else "x_gmres"]) <= 1e-8
    c.close()


def test_compact_batch(golden):
    from tests.test_gpu_batch import make_ctx

    print("compact batch", flush=True)
    gs = [golden(n) for n in ["hex8_5j", "tet4_3j", "quad4_8"]]
    comp = [compact_ctx(gs[0], rhs=True), compact_ctx(gs[1], rhs=True)]
    ctxs = comp + [make_ctx(gs[2])]
    for precision in (FP64, MIXED):
        alone = {i: run_batch([c], precision, 1e-10)[0] for i, c in enumerate(ctxs)}
        mix = run_batch(ctxs, precision, 1e-10, slots=2, stagger=1)
        for i in range(len(ctxs)):
            assert alone[i][1]["iters"] == mix[i][1]["iters"]
            assert np.array_equal(alone[i][0], mix[i][0])
            assert rel(mix[i][0], gs[i]["x"]) <= (1e-8 if precision == MIXED or i == 2 else 1e-10)
    for c in ctxs:
        c.close()


def test_orders_agree_and_stresses_from_resident_x():
    from elastic_amd import mesh as M
    from elastic_amd.solver import Context
    from elastic_amd.system import System

    print("orders agree: hex8 14^3", flush=True)
    m = M.hex8_cube(14, jitter=0.15, seed=11)
    v = None
    out = {}
    for order in (0, 1):
        c = Context(0)
        c.set_option("node_order", order)
        s = System(m, ctx=c)
        assert c.get_option("node_order") == order
        x, st = s.solve(tol=1e-12)
        if v is None:  # a generic vector (A x = b cancels to rounding in most rows)
            v = np.random.default_rng(3).standard_normal(s.neq)
        y = c.spmv(v)
        tx_res = s.stress(from_x=True)  # the resident x, back in the CSRC order
        tx_host = s.stress(u=s.displacements(x))
        assert np.array_equal(tx_res, tx_host)
        out[order] = (x, y, st["iters"])
        c.close()
    assert rel(out[1][0], out[0][0]) <= 1e-10
    assert rel(out[1][1], out[0][1]) <= 1e-13
    assert abs(out[1][2] - out[0][2]) <= 2
