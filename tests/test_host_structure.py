"""The library's host set-up (RCM, equation numbers, CSRC pattern) and its
exported symbols, on the CPU.  The set-up is compared with the reference's own
outputs (golden fixtures) bit for bit."""
import ctypes
import hashlib

import numpy as np
import pytest

from elastic_amd import _lib
from tests.cases import CASES, SMALL


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("name", SMALL)
def test_structure_matches_reference(name, golden):
    g = golden(name)
    m = CASES[name][0]()
    perm = _lib.reord(m)
    idn, neq = _lib.numeq(m, perm)
    ia, ja = _lib.csrstruct(m, idn, neq)
    assert np.array_equal(perm, g["inum"])
    assert np.array_equal(idn.ravel(), g["id"])
    assert np.array_equal(ia, g["ia"])
    assert np.array_equal(ja, g["ja"])


@pytest.mark.parametrize("name", ["hex8_20", "quad_el"])
def test_structure_digests(name, golden):
    g = golden(name)
    m = CASES[name][0]()
    perm = _lib.reord(m)
    idn, neq = _lib.numeq(m, perm)
    ia, ja = _lib.csrstruct(m, idn, neq)
    assert sha(perm) == str(g["sha_inum"])
    assert sha(idn.ravel()) == str(g["sha_id"])
    assert sha(ia) == str(g["sha_ia"])
    assert sha(ja) == str(g["sha_ja"])


def test_structure_invariants_tet():
    """ja strictly ascending per row and below the diagonal; the node graph is
    symmetric; nad matches the closed form of the tet split stencil."""
    from elastic_amd import mesh as M
    m = M.tet4_cube(4, jitter=0.2, seed=5, per_element_E=False)
    perm = _lib.reord(m)
    assert np.array_equal(np.sort(perm), np.arange(1, m.nnode + 1))
    idn, neq = _lib.numeq(m, perm)
    ia, ja = _lib.csrstruct(m, idn, neq)
    for i in range(neq):
        row = ja[ia[i] - 1: ia[i + 1] - 1]
        assert np.all(np.diff(row) > 0) and (row.size == 0 or row[-1] <= i)
    # full symmetric pattern from the mesh: each element couples all its dofs
    dense = np.zeros((neq, neq), bool)
    for el in m.ix[:, :4]:
        eqs = idn[el - 1].ravel()
        eqs = eqs[eqs > 0] - 1
        dense[np.ix_(eqs, eqs)] = True
    low = np.tril(dense, -1)
    assert low.sum() == ia[-1] - 1
    rows = np.repeat(np.arange(neq), np.diff(ia))
    assert low[rows, ja - 1].all()


@pytest.mark.parametrize("name", ["cfg2_hex8_100", "cfg3_tet4_80", "cfg5_pull_60"])
def test_full_size_structure_digests(name, golden):
    """The full-size workloads (hex8 100^3 jittered, tet4 80^3 x 6, the
    plasticity mesh hex8 60^3): RCM order, equation numbers and the CSRC
    pattern are the reference's, bit for bit (digests of the reference's own
    set-up, Reord.f:1-434, Numeq.f:139-191, Csr.f:13-305)."""
    from tests.cases import DIGESTS

    g = golden(name)
    m = DIGESTS[name]()
    perm = _lib.reord(m)
    idn, neq = _lib.numeq(m, perm)
    ia, ja = _lib.csrstruct(m, idn, neq)
    assert [neq, ia[-1] - 1] == [int(g["dims"][5]), int(g["dims"][6])]
    assert sha(perm) == str(g["sha_inum"])
    assert sha(idn.ravel()) == str(g["sha_id"])
    assert sha(ia) == str(g["sha_ia"])
    assert sha(ja) == str(g["sha_ja"])
