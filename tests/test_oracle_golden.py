"""The oracle (C restatement) against the reference's own outputs.

Fixtures in tests/golden/ were written by the compiled reference (see
tests/golden/make_fixtures.py).  The serial paths of the oracle follow the
reference's arithmetic statement by statement, so everything here is compared
bit for bit: RCM permutation, equation numbers, CSRC ia/ja, assembled ad/al/b,
the symmetric SpMV, PCG/GMRES/BiCGSTAB solutions and iteration counts, the
3x3 block-Jacobi set-up and the nonlinear-loop residual.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import binding as O
from tests.cases import CASES, SMALL
from tests.conftest import REFERENCE_INPUT


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def systems():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = O.System(CASES[name][0]())
        return cache[name]

    return get


@pytest.mark.parametrize("name", SMALL)
def test_structure_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    assert np.array_equal(s.perm, g["inum"])
    assert np.array_equal(s.idn.ravel(), g["id"])
    assert np.array_equal(s.ia, g["ia"])
    assert np.array_equal(s.ja, g["ja"])


@pytest.mark.parametrize("name", SMALL)
def test_assembly_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    assert np.array_equal(s.ad, g["ad"])
    assert np.array_equal(s.al[: s.nad], g["al"])
    assert np.array_equal(s.b, g["b"])


@pytest.mark.parametrize("name", SMALL)
def test_spmv_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    y = O.matvec(s.ia, s.ja, s.ad, s.al, g["spmv_v"])
    assert np.array_equal(y, g["spmv_y"])


@pytest.mark.parametrize("name", SMALL)
def test_pcg_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    tol = float(g["tol"])
    x, st = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=tol)
    assert st["status"] == 0
    assert st["iters"] == int(g["pcg_iters"])
    assert np.array_equal(x, g["x"])
    # second pass of the nonlinear loop: b0 - K u
    assert np.array_equal(s.residual(x), g["resid"])


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 2])
def test_gmres_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    x, st = O.gmres(s.ia, s.ja, s.ad, s.al, s.b, k=50, tol=float(g["tol"]))
    assert st["iters"] == int(g["gmres_iters"])
    assert np.array_equal(x, g["x_gmres"])


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 4])
def test_bicgstab_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    x, st = O.pbicgstab(s.ia, s.ja, s.ad, s.al, s.b, tol=float(g["tol"]))
    assert st["iters"] == int(g["bicgstab_iters"])
    assert np.array_equal(x, g["x_bicgstab"])


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 8])
def test_block_jacobi_bit_exact(name, golden, systems):
    g, s = golden(name), systems(name)
    m = O.block_precond3(s.ad, s.al, s.ia, s.ja)
    assert np.array_equal(m, g["bdiag"])


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 16])
def test_colormsh_bit_exact(name, golden):
    """or_colormsh (Colormsh.f:65-247) with setcolor's ordering
    (Colormsh.f:255-308) gives the reference's elcolor and colorg."""
    g = golden(name)
    mesh = CASES[name][0]()
    ncol, ic = O.colormsh(mesh)
    assert ncol >= 2 and np.all(ic >= 1)
    elcolor, colorg = O.setcolor(ic, ncol)
    assert np.array_equal(elcolor, g["elcolor"])
    assert np.array_equal(colorg, g["colorg"])


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8"])
def test_colormsh_is_a_colouring(name):
    """No two elements of one colour share a node."""
    mesh = CASES[name][0]()
    ncol, ic = O.colormsh(mesh)
    nodes = mesh.ix[:, : mesh.nen]
    for c in range(1, ncol + 1):
        el = nodes[ic == c].ravel()
        el = el[el > 0]
        assert np.unique(el).size == el.size, c


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 8])
def test_op_block_precond_bit_exact(name, golden):
    """or_op_block_precond3 (Precond.f:419-540) applied to the reference's own
    m(6, .) and vector gives the reference's op_block_precond output."""
    g = golden(name)
    y = O.op_block_precond3(g["bdiag"], g["spmv_v"])
    assert np.array_equal(y, g["bdiag_y"])


@pytest.mark.parametrize("name", [n for n in SMALL if CASES[n][2] & 256])
def test_block_jacobi2_bit_exact(name, golden, systems):
    """block_precond and op_block_precond with nblock = 2 (Precond.f:163-358,
    419-540) restated: the reference's m(3, .) and its application."""
    g, s = golden(name), systems(name)
    assert np.array_equal(O.block_precond2(s.ad, s.al, s.ia, s.ja), g["bdiag2"])
    assert np.array_equal(O.op_block_precond2(g["bdiag2"], g["spmv_v"]), g["bdiag2_y"])


def dense(ia, ja, ad, al):
    n = ad.shape[0]
    A = np.diag(ad.copy())
    for i in range(n):
        for k in range(ia[i] - 1, ia[i + 1] - 1):
            A[i, ja[k] - 1] = A[ja[k] - 1, i] = al[k]
    return A


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8", "tria3_8"])
def test_block_precond_is_the_block_inverse(name, systems):
    """block_precond + op_block_precond = the inverse of each diagonal 3x3
    block of A (consecutive equation triples, rest rows by 1/ad), applied to a
    vector; a transposed or mis-slotted cofactor fails this."""
    s = systems(name)
    A = dense(s.ia, s.ja, s.ad, s.al[: s.nad])
    n = s.neq
    v = 1.0 + np.arange(n) % 7 / 8.0 + np.sin(np.arange(n))
    y = O.op_block_precond3(O.block_precond3(s.ad, s.al, s.ia, s.ja), v)
    z = np.empty(n)
    nb = n // 3
    for i in range(nb):
        sl = slice(3 * i, 3 * i + 3)
        z[sl] = np.linalg.solve(A[sl, sl], v[sl])
    z[3 * nb:] = v[3 * nb:] / s.ad[3 * nb:]
    assert np.linalg.norm(y - z) <= 1e-13 * np.linalg.norm(z)


@pytest.mark.parametrize("name", ["quad4_8", "tria3_8", "req1"])
def test_block_precond2_is_the_block_inverse(name, systems):
    """The 2x2 set-up and application = the inverse of each diagonal 2x2 block
    of consecutive equations (the node blocks of these 2-D meshes)."""
    s = systems(name)
    A = dense(s.ia, s.ja, s.ad, s.al[: s.nad])
    n = s.neq
    v = 1.0 + np.arange(n) % 7 / 8.0 + np.cos(np.arange(n))
    y = O.op_block_precond2(O.block_precond2(s.ad, s.al, s.ia, s.ja), v)
    z = np.empty(n)
    for i in range(n // 2):
        sl = slice(2 * i, 2 * i + 2)
        z[sl] = np.linalg.solve(A[sl, sl], v[sl])
    z[2 * (n // 2):] = v[2 * (n // 2):] / s.ad[2 * (n // 2):]
    assert np.linalg.norm(y - z) <= 1e-13 * np.linalg.norm(z)
    x, st = O.pcg_block(2, s.ia, s.ja, s.ad, s.al, s.b, tol=1e-12)
    assert st["status"] == 0
    assert np.linalg.norm(x - np.linalg.solve(A, s.b)) <= 1e-9 * np.linalg.norm(x)


@pytest.mark.parametrize("name", ["hex8_5j", "tet4_3j", "quad4_8"])
def test_pcg_bdiag3_solves_the_system(name, systems):
    """PCG with the block-Jacobi application reaches the dense solution; on
    3-dof systems, whose triples are node blocks, in fewer iterations than
    Jacobi PCG."""
    s = systems(name)
    A = dense(s.ia, s.ja, s.ad, s.al[: s.nad])
    xd = np.linalg.solve(A, s.b)
    x, st = O.pcg_bdiag3(s.ia, s.ja, s.ad, s.al, s.b, tol=1e-12)
    assert st["status"] == 0
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert st["norm_m_r"] < 3.16 * st["conv"]
    if s.mesh.ndf == 3:
        _, sj = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=1e-12)
        assert st["iters"] < sj["iters"]


@pytest.mark.parametrize("name", ["hex8_20", "quad_el"])
def test_large_case_digests(name, golden, systems):
    """Workload 1 (hex8 20^3) and the reference's quad deck, regenerated."""
    g, s = golden(name), systems(name)
    assert sha(s.perm) == str(g["sha_inum"])
    assert sha(s.idn.ravel()) == str(g["sha_id"])
    assert sha(s.ia) == str(g["sha_ia"])
    assert sha(s.ja) == str(g["sha_ja"])
    for k, v in (("ad", s.ad), ("al", s.al[: s.nad]), ("b", s.b)):
        n = g["norm_" + k]
        assert np.linalg.norm(v) == n[0] and np.abs(v).max() == n[1] and v.sum() == n[2]
    x, st = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=float(g["tol"]))
    assert st["iters"] == int(g["pcg_iters"])
    assert np.array_equal(x, g["x"])


@pytest.mark.skipif(not os.path.isdir(REFERENCE_INPUT), reason="reference decks not present")
@pytest.mark.parametrize("name", ["quad_el", "quad_nl", "quad_nf", "quad_el_ept"])
def test_reference_decks(name, golden):
    """Reading the reference's own decks with elastic_amd.deck reproduces the
    reference's results bit for bit."""
    from elastic_amd import deck
    from tests.cases import REF_DECKS

    g = golden(name)
    s = O.System(deck.read_deck(os.path.join(REFERENCE_INPUT, REF_DECKS[name])))
    assert sha(s.ia) == str(g["sha_ia"]) and sha(s.ja) == str(g["sha_ja"])
    x, st = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=float(g["tol"]))
    assert st["iters"] == int(g["pcg_iters"])
    assert np.array_equal(x, g["x"])


def test_quad_deck_energy_norm(golden):
    """Known answer of the reference deck: sqrt(x.Kx) of the exact problem is
    1.379745 (input/quad/exemplo1_apostila_el.dat:3); the FE value is below it."""
    g = golden("quad_el")
    e = np.sqrt(float(g["pcg_xkx"]))
    assert 1.378 < e < 1.379745


def test_pcg_omp_matches_serial(systems):
    """The OpenMP restatement (CPU baseline) converges to the serial solution."""
    s = systems("hex8_5j")
    x0, st0 = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=1e-10)
    x1, st1 = O.pcg(s.ia, s.ja, s.ad, s.al, s.b, tol=1e-10, nth=4)
    assert abs(st1["iters"] - st0["iters"]) <= 1
    assert np.linalg.norm(x1 - x0) <= 1e-9 * np.linalg.norm(x0)
