"""elastic_spmv_csrc_noalloc: the symmetric CSRC product on caller-held device
arrays, built without any device allocation (a fixed scratch array per GPU).

Checked through the C-ABI on two calls in a row, the second on a smaller
system (so the scratch holds the first call's larger contents when the
second starts): each result equals the existing entry point's product
(elastic_spmv on a context with the same CSRC values) within 1e-13 relative
L2 (the two sum in different orders), and equals the oracle's
matvec_csrcsym1 (Matvec.f:528-621) bit for bit, the order this product keeps.
A size beyond the fixed scratch is refused before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


class _Hip:
    """Device memory and a stream from the HIP runtime the library already
    loaded (test plumbing: the entry point under test allocates nothing)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc):
        assert rc == 0, "HIP call failed (%d)" % rc

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))))
        self._ok(self.h.hipMemcpy(p, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1))
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        self._ok(self.h.hipMemcpy(C.c_void_p(out.ctypes.data), p, C.c_size_t(out.nbytes), 2))
        return out

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _product(L, ia, ja, ad, al, x):
    hip = _Hip()
    try:
        hip._ok(hip.h.hipSetDevice(0))
        stream = C.c_void_p()
        hip._ok(hip.h.hipStreamCreate(C.byref(stream)))
        args = [hip.put(a) for a in (ia, ja, ad, al, x)]
        y0 = np.full(x.shape, np.nan)
        dy = hip.put(y0)
        rc = L.elastic_spmv_csrc_noalloc(ad.shape[0], ja.shape[0], *args, dy, stream)
        assert rc == 0
        hip._ok(hip.h.hipStreamSynchronize(stream))
        y = hip.get(dy, y0)
        hip._ok(hip.h.hipStreamDestroy(stream))
        return y
    finally:
        hip.close()


@pytest.mark.parametrize("first,second", [("hex8_20", "hex8_5j"), ("quad4_8", "tria3_8")])
def test_noalloc_product_matches_the_context_product(golden, first, second):
    from elastic_amd import _lib
    from elastic_amd.solver import Context

    L = _lib.lib()
    for name in (first, second):  # the second system is the smaller one
        print("noalloc", name, flush=True)
        s = O.System(CASES[name][0]())
        ia, ja, ad, al = s.ia, s.ja, s.ad, s.al[: s.nad]
        x = 1.0 + (np.arange(1, s.neq + 1) % 7) * 0.125 + np.sin(np.arange(s.neq))
        y = _product(L, ia, ja, ad, al, x)
        with Context(0) as c:
            c.set_structure(ia, ja)
            c.set_values(ad, al)
            yc = c.spmv(x)
        assert rel(y, yc) <= 1e-13
        assert np.array_equal(y, O.matvec(ia, ja, ad, al, x))


def test_noalloc_refuses_sizes_beyond_the_scratch():
    from elastic_amd import _lib

    L = _lib.lib()
    one = np.zeros(1, np.int32)
    p = C.c_void_p(one.ctypes.data)  # never dereferenced: the size check comes first
    rc = L.elastic_spmv_csrc_noalloc(10_000_000, 10_000_000, p, p, p, p, p, p, None)
    assert rc == _lib.BAD_ARG
