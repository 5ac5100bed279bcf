"""Host-side mirror of the reference's solver interface on the GPU library.

``solv`` has the meaning of the reference's ``solv`` (Solv.f:13-440): it
overwrites ``m`` with the Jacobi inverse, solves from x0 = 0 and raises on the
conditions where the reference stops (zero diagonal, maxit).  ``Context`` is
the object interface over the C-ABI (structure once, values per assembly,
solves with host or resident vectors).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, PCG, GMRES, BICGSTAB, DIAG, BDIAG3, ElasticError, Stats  # noqa: F401


class Context:
    def __init__(self, device: int = 0, tile_nnz: int | None = None):
        self._L = _lib.lib()
        self._h = C.c_void_p()
        rc = self._L.elastic_create(C.byref(self._h), device)
        if rc != _lib.OK:
            raise ElasticError(rc, "elastic_create on device %d" % device)
        if tile_nnz:
            self._chk(self._L.elastic_set_tile_nnz(self._h, tile_nnz))
        self.neq = 0
        self.nad = 0

    def _chk(self, rc, allow=()):
        return _lib._check(self._h, rc, allow)

    def close(self):
        if self._h:
            self._L.elastic_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- system
    def set_structure(self, ia: np.ndarray, ja: np.ndarray):
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        self.neq = ia.shape[0] - 1
        self.nad = int(ia[-1] - 1)
        ja_arg = ja if ja.size else np.zeros(1, np.int32)
        self._chk(self._L.elastic_set_structure(self._h, self.neq, self.nad, ia, ja_arg))

    def set_values(self, ad: np.ndarray, al: np.ndarray):
        ad = np.ascontiguousarray(ad, np.float64)
        al = np.ascontiguousarray(al[: self.nad] if self.nad else np.zeros(1), np.float64)
        self._chk(self._L.elastic_set_values(self._h, ad, al))

    def get_values(self):
        ad = np.zeros(self.neq)
        al = np.zeros(max(self.nad, 1))
        self._chk(self._L.elastic_get_values(self._h, ad, al))
        return ad, al[: self.nad]

    def get_bdiag(self) -> np.ndarray:
        """m(6, neq/3 + mod(neq,3)) of the 3x3 block-Jacobi preconditioner,
        in block_precond's layout (Precond.f:163-358), flattened."""
        m = np.zeros(6 * (self.neq // 3 + self.neq % 3))
        self._chk(self._L.elastic_get_bdiag(self._h, m))
        return m

    def get_block_precond(self, bsz: int) -> np.ndarray:
        """m(bsz(bsz+1)/2, neq/bsz + mod(neq,bsz)) of block_precond with block
        size bsz (2 or 3), flattened (Precond.f:163-358)."""
        w = bsz * (bsz + 1) // 2
        m = np.zeros(w * (self.neq // bsz + self.neq % bsz))
        self._chk(self._L.elastic_get_block_precond(self._h, bsz, m))
        return m

    def layout_info(self):
        info = np.zeros(4, np.int64)
        self._chk(self._L.elastic_layout_info(self._h, info))
        # nb: node block size of the active layout (1 = scalar tiles, 0 = the
        # gather product); nnz_max: blocks (or nonzeros) of the largest tile
        return dict(ntiles=int(info[0]), nnz_max=int(info[1]), ext_partials=int(info[2]),
                    nb=int(info[3]))

    def set_option(self, name: str, value: int):
        """Tuning knob of the context (see elastic_set_option)."""
        self._chk(self._L.elastic_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        self._chk(self._L.elastic_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    def product_timing(self, reset: bool = True):
        """(average ms, count) of the PCG iteration products sampled under
        the "time_products" option since the last reset."""
        ms, n = C.c_double(), C.c_int64()
        self._chk(self._L.elastic_product_timing(self._h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def set_tile_blocks(self, tile_blocks: int):
        self._chk(self._L.elastic_set_tile_blocks(self._h, tile_blocks))

    def autotune(self, precision: int = FP64, reps: int = 32, tiles: bool = True):
        """Picks the fastest tile size (unless tiles=False) and lanes per tile
        for the resident values, timed inside a short PCG run (rebuilds the
        layout; the resident x is overwritten); returns (tile size, product
        time in ms)."""
        self.set_option("tune_tiles", 1 if tiles else 0)
        best, ms = C.c_int32(), C.c_double()
        self._chk(self._L.elastic_autotune(self._h, precision, reps, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    # ---------------------------------------------------------- products
    def spmv(self, x: np.ndarray, precision: int = FP64) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._chk(self._L.elastic_spmv(self._h, precision, x, y))
        return y

    def time_spmv(self, precision: int = FP64, reps: int = 20):
        tot, ker = C.c_double(), C.c_double()
        self._chk(self._L.elastic_time_spmv(self._h, precision, reps, C.byref(tot), C.byref(ker)))
        return tot.value, ker.value

    # ---------------------------------------------------------- solves
    def solve(self, b, solver=PCG, precond=DIAG, precision=FP64, tol=1e-11, maxit=10000, ngram=50,
              allow_maxit=False):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        st = Stats()
        rc = self._L.elastic_solve(self._h, solver, precond, precision, b, x, tol, maxit, ngram,
                                   C.byref(st))
        self._chk(rc, allow=(_lib.MAXIT,) if allow_maxit else ())
        return x, st.as_dict()

    def upload_rhs(self, b):
        self._chk(self._L.elastic_upload_rhs(self._h, np.ascontiguousarray(b, np.float64)))

    def solve_resident(self, solver=PCG, precond=DIAG, precision=FP64, tol=1e-11, maxit=10000,
                       ngram=50, allow_maxit=False):
        st = Stats()
        rc = self._L.elastic_solve_resident(self._h, solver, precond, precision, tol, maxit, ngram,
                                            C.byref(st))
        self._chk(rc, allow=(_lib.MAXIT,) if allow_maxit else ())
        return st.as_dict()

    def download_x(self):
        x = np.zeros(self.neq)
        self._chk(self._L.elastic_download_x(self._h, x))
        return x


_default_ctx: Context | None = None


def solv(neq, nad, ia, ja, ad, al, m, b, x, tol=1e-11, maxit=10000, ngram=50, unsym=False,
         solver=PCG, precision=FP64, precond=DIAG):
    """Solv.f:13-440 on the GPU: m <- 1/ad, x <- A^-1 b (x0 = 0); returns the
    solver statistics.  Raises ElasticError where the reference stops."""
    global _default_ctx
    if unsym:
        raise ElasticError(_lib.BAD_ARG, "Solver 1 nao disponivel para matriz nao-simetrica")
    if _default_ctx is None:
        _default_ctx = Context(0)
    c = _default_ctx
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja[:nad], np.int32)
    # the structure is rebuilt whenever its contents change, not only its sizes
    key = getattr(c, "_solv_pattern", None)
    if key is None or not (np.array_equal(key[0], ia) and np.array_equal(key[1], ja)):
        c.set_structure(ia, ja)
        c._solv_pattern = (ia.copy(), ja.copy())
    c.set_values(ad, al)
    m[:] = 1.0 / ad
    xs, st = c.solve(b, solver=solver, precond=precond, precision=precision, tol=tol, maxit=maxit,
                     ngram=ngram)
    x[:] = xs
    return st
