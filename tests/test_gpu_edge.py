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



def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nring", [150, 2500])
def test_high_valence_fan(nring):
    """A hub node joined to `nring` ring nodes (Solv.f:13-15 accepts any
    CSRC system).  At 150 the hub's node row has 146 lower blocks, more than
    a node-blocked tile row holds: the library falls back to scalar row
    tiles.  At 2500 its CSRC row has 4,993 lower entries, more than a row
    tile holds: the library takes the gather-only product, whose FP64 result
    is the reference's matvec_csrcsym1 bit for bit.  Either way the solve
    matches the oracle's PCG: at 150, iterations within 2, x within 1e-10
    (FP64) and 1e-8 (mixed); at 2500 (condition number 5.5e9, where the
    oracle's own x is 3.5e-9 from a dense solve and its OpenMP form takes
    one iteration fewer) iterations within 1% and x within 1e-7."""
    m = M.tria3_fan(nring)
    print("fan", nring, flush=True)
    s = System(m)
    so = O.System(m)
    assert np.array_equal(so.ia, s.ia) and np.array_equal(so.ja, s.ja)
    info = s.ctx.layout_info()
    assert info["nb"] == (1 if nring == 150 else 0)
    if nring == 2500:  # the product of the GPU's own K, bit for bit
        ad, al = s.values()
        v = 1.0 + (np.arange(1, s.neq + 1) % 7) * 0.125
        assert np.array_equal(s.ctx.spmv(v), O.matvec(s.ia, s.ja, ad, al, v))
    xo, sto = O.pcg(so.ia, so.ja, so.ad, so.al, so.b, tol=1e-10, maxit=50000)
    assert sto["status"] == 0
    x, st = s.solve(tol=1e-10, maxit=50000)
    print("fan", nring, "iters", st["iters"], "oracle", sto["iters"], flush=True)
    small = nring == 150
    assert abs(st["iters"] - sto["iters"]) <= (2 if small else 0.01 * sto["iters"])
    assert rel(x, xo) <= (1e-10 if small else 1e-7)
    xm, sm = s.solve(tol=1e-10, precision=MIXED, maxit=50000)
    assert rel(xm, xo) <= (1e-8 if small else 1e-7)
    s.ctx.close()


@pytest.mark.parametrize("nring", [40, 60])
def test_fan_row_in_node_blocked_tiles(nring):
    """A hub row of 36 (nring 40) or 56 (nring 60) lower node blocks still
    fits a node-blocked tile row (at most 64), so the blocked product serves
    it: its row scan sums a run that spans several 16-lane DPP rows and
    whatever 64-block chunk boundary falls inside it.  The FP64 product of the
    GPU's own K against the reference's matvec_csrcsym1 (Matvec.f:528-621)
    within 1e-13; the solve as the oracle's PCG (iterations within 2, x within
    1e-10; mixed x within 1e-8)."""
    m = M.tria3_fan(nring)
    print("fan", nring, flush=True)
    s = System(m)
    so = O.System(m)
    assert np.array_equal(so.ia, s.ia) and np.array_equal(so.ja, s.ja)
    assert s.ctx.layout_info()["nb"] == 2
    ad, al = s.values()
    v = 1.0 + (np.arange(1, s.neq + 1) % 7) * 0.125
    assert rel(s.ctx.spmv(v), O.matvec(s.ia, s.ja, ad, al, v)) <= 1e-13
    xo, sto = O.pcg(so.ia, so.ja, so.ad, so.al, so.b, tol=1e-10, maxit=50000)
    assert sto["status"] == 0
    x, st = s.solve(tol=1e-10, maxit=50000)
    assert abs(st["iters"] - sto["iters"]) <= 2
    assert rel(x, xo) <= 1e-10
    xm, sm = s.solve(tol=1e-10, precision=MIXED, maxit=50000)
    assert rel(xm, xo) <= 1e-8
    s.ctx.close()


def _flipped(m):
    """The tria3 square with every other quad split along its other diagonal
    (same nodes and edge count, so the same neq and nad; another pattern)."""
    f = M.tria3_square(8)
    ix = f.ix.copy()
    a, b = m.ix[0::2, :3], m.ix[1::2, :3]
    p1, p2, p3, p4 = a[:, 0], a[:, 1], a[:, 2], b[:, 2]
    sel = (np.arange(a.shape[0]) % 2 == 0)[:, None]
    ix[0::2, :3] = np.where(sel, np.stack([p1, p2, p4], 1), a)
    ix[1::2, :3] = np.where(sel, np.stack([p2, p3, p4], 1), b)
    f.ix = ix
    return f


def test_solv_dropin_rebuilds_on_new_pattern_at_same_addresses():
    """elastic_solv_ (Solv.f:13-15) called twice with the same arrays,
    refilled in place with another system of the same neq and nad: the
    second solve is of the second system (the structure cache compares the
    pattern, not the addresses)."""
    import ctypes as C

    ma = M.tria3_square(8)
    mb = _flipped(ma)
    for m in (ma, mb):  # nodal force only, same right-hand side pattern
        m.eload[:] = 0
        m.loads = []
        m.f[40] = (0.3, -1.0)
    sa, sb = O.System(ma), O.System(mb)
    assert sa.neq == sb.neq and sa.nad == sb.nad and not np.array_equal(sa.ja, sb.ja)
    L = _lib.lib()
    fn = L.elastic_solv_
    fn.restype = C.c_int32
    n, nad = sa.neq, sa.nad
    bufs = dict(ia=np.zeros(n + 1, np.int32), ja=np.zeros(nad, np.int32), ad=np.zeros(n),
                al=np.zeros(nad), m=np.zeros(n), b=np.zeros(n), x=np.zeros(n))
    i32 = lambda v: C.byref(C.c_int32(v))  # noqa: E731
    i64 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    for s in (sa, sb):
        bufs["ia"][:] = s.ia
        bufs["ja"][:] = s.ja
        bufs["ad"][:] = s.ad
        bufs["al"][:] = s.al[:nad]
        bufs["b"][:] = s.b
        tol = C.c_double(1e-11)
        rc = fn(i32(n), i32(nad), p(bufs["ia"]), p(bufs["ja"]), p(bufs["ad"]), p(bufs["al"]),
                p(bufs["al"]), p(bufs["m"]), p(bufs["b"]), p(bufs["x"]), C.byref(tol), i32(10000),
                i32(50), i32(0), i32(1), i32(0), i32(0), i32(0), i32(0), i32(n), i64(0), i64(0),
                i64(0), i64(0))
        assert rc == 0
        xo, _ = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=1e-11)
        assert rel(bufs["x"], xo) <= 1e-10
        assert np.array_equal(bufs["m"], 1.0 / s.ad)


def test_small_tile_option_keeps_the_context_usable(golden, monkeypatch):
    """tile_nnz = 64 is below hex8's longest CSRC rows (up to 80 entries): the
    row tiles grow to the longest row instead of leaving the context without
    a layout, and the solve is unchanged within the FP64 bar."""
    from elastic_amd.solver import Context

    monkeypatch.setenv("ELASTIC_LAYOUT", "scalar")
    g = golden("hex8_5j")
    with Context(0) as c:
        c.set_structure(g["ia"], g["ja"])
        c.set_values(g["ad"], g["al"])
        x0, s0 = c.solve(g["b"], tol=1e-11)
        c.set_option("tile_nnz", 64)
        x1, s1 = c.solve(g["b"], tol=1e-11)
        assert abs(s1["iters"] - s0["iters"]) <= 2
        assert rel(x1, x0) <= 1e-10
        assert rel(x1, g["x"]) <= 1e-10
