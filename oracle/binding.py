"""ctypes binding of the C restatement (oracle/_build/liboracle.so).

Test infrastructure: only tests/, the smoke check and the CPU-baseline leg of
bench.py import this module.  Functions take and return numpy arrays in the
reference's conventions (1-based index values); see oracle/oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Loads(C.Structure):
    _fields_ = [("nload", C.c_int), ("type", C.c_int * 200), ("nparc", C.c_int * 200),
                ("f", (C.c_double * 3) * 200)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("xkx", C.c_double), ("norm_x", C.c_double),
                ("norm_r", C.c_double), ("norm_m_r", C.c_double), ("conv", C.c_double),
                ("status", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_graph.argtypes = [C.c_int, C.c_int, C.c_int, I32P, I32P, C.POINTER(C.POINTER(C.c_int32))]
        L.or_graph.restype = C.c_int
        L.or_reord.argtypes = [I32P, I32P, C.c_int, C.c_int, C.c_int]
        L.or_numeq.argtypes = [I32P, I32P, C.c_int, C.c_int]
        L.or_numeq.restype = C.c_int
        L.or_csrstruct.argtypes = [I32P, I32P, I32P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   I32P, C.POINTER(C.POINTER(C.c_int32))]
        L.or_csrstruct.restype = C.c_int
        L.or_element.argtypes = [C.c_int, C.c_int, F64P, I32P, F64P, F64P, F64P, F64P, C.c_int,
                                 C.c_int, C.POINTER(Loads), F64P]
        L.or_element.restype = C.c_int
        L.or_pform.argtypes = [I32P, I32P, I32P, F64P, F64P, I32P, I32P, I32P, F64P, F64P, F64P,
                               F64P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.POINTER(Loads), F64P]
        L.or_pform.restype = C.c_int
        L.or_pload.argtypes = [I32P, F64P, F64P, F64P, I32P, C.c_int, C.c_int, C.POINTER(Loads)]
        L.or_update.argtypes = [C.c_int, C.c_int, I32P, F64P, F64P]
        L.or_pre_diag.argtypes = [F64P, F64P, C.c_int]
        L.or_block_precond3.argtypes = [F64P, F64P, I32P, I32P, F64P, C.c_int]
        L.or_op_block_precond3.argtypes = [F64P, F64P, F64P, C.c_int]
        L.or_matvec_csrcsym1.argtypes = [C.c_int, I32P, I32P, F64P, F64P, F64P, F64P]
        L.or_dot.argtypes = [F64P, F64P, C.c_int]
        L.or_dot.restype = C.c_double
        solver_args = [C.c_int, I32P, I32P, F64P, F64P, F64P, F64P, F64P]
        L.or_pcg.argtypes = solver_args + [C.c_double, C.c_int, C.POINTER(Stats)]
        L.or_pcg_omp.argtypes = solver_args + [C.c_double, C.c_int, C.c_int, C.POINTER(Stats)]
        L.or_gmres.argtypes = solver_args + [C.c_int, C.c_double, C.c_int, C.POINTER(Stats)]
        L.or_pbicgstab.argtypes = solver_args + [C.c_double, C.c_int, C.POINTER(Stats)]
        L.or_pcg_bdiag3.argtypes = solver_args + [C.c_double, C.c_int, C.POINTER(Stats)]
        L.or_pcg_block.argtypes = [C.c_int] + solver_args + [C.c_double, C.c_int, C.POINTER(Stats)]
        L.or_block_precond2.argtypes = [F64P, F64P, I32P, I32P, F64P, C.c_int]
        L.or_op_block_precond2.argtypes = [F64P, F64P, F64P, C.c_int]
        L.or_colormsh.argtypes = [I32P, C.c_int, C.c_int, C.c_int, C.c_int, I32P]
        L.or_tform.argtypes = [I32P, F64P, F64P, I32P, F64P, C.c_void_p, F64P, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_colormsh.restype = C.c_int
        _lib = L
    return _lib


def _free_ptr(p):
    C.CDLL(None).free(C.cast(p, C.c_void_p))


def _take(ptr, n):
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].copy()
    _free_ptr(ptr)
    return out


def loads_struct(loads) -> Loads:
    s = Loads()
    s.nload = len(loads)
    for i, (ty, vals) in enumerate(loads):
        s.type[i] = ty
        s.nparc[i] = len(vals)
        for q, v in enumerate(vals[:3]):
            s.f[i][q] = v
    return s


GRAVITY = np.array([0.0, 0.0, -9.81])


# ------------------------------------------------------------------ host steps
def graph(mesh):
    ip = np.zeros(mesh.nnode + 1, np.int32)
    p = C.POINTER(C.c_int32)()
    n = lib().or_graph(mesh.nnode, mesh.numel, mesh.nen, np.ascontiguousarray(mesh.ix), ip, C.byref(p))
    return ip, _take(p, n)


def reord(mesh):
    perm = np.zeros(mesh.nnode, np.int32)
    lib().or_reord(np.ascontiguousarray(mesh.ix), perm, mesh.nnode, mesh.numel, mesh.nen)
    return perm


def numeq(mesh, perm):
    idn = np.ascontiguousarray(mesh.id, dtype=np.int32).copy()
    neq = lib().or_numeq(idn, perm, mesh.nnode, mesh.ndf)
    return idn, neq


def csrstruct(mesh, idn, perm, neq):
    ia = np.zeros(neq + 1, np.int32)
    p = C.POINTER(C.c_int32)()
    nad = lib().or_csrstruct(idn, np.ascontiguousarray(mesh.ix), perm, mesh.nnode, mesh.numel,
                             mesh.nen, mesh.ndf, neq, ia, C.byref(p))
    return ia, _take(p, nad)


def element(iel, isw, e, iq, xl, ul, ndm, nst, loads=(), gravity=GRAVITY):
    p = np.zeros(nst)
    s = np.zeros(nst * nst)
    ls = loads_struct(list(loads))
    rc = lib().or_element(iel, isw, np.ascontiguousarray(e, np.float64),
                          np.ascontiguousarray(iq, np.int32), np.ascontiguousarray(xl, np.float64),
                          np.ascontiguousarray(ul, np.float64), p, s, ndm, nst, C.byref(ls),
                          np.ascontiguousarray(gravity, np.float64))
    return rc, s.reshape(nst, nst).T.copy(), p  # s returned as s[i, j] = s(i,j)


class System:
    """The reference's set-up (mesh -> RCM -> numeq -> CSRC -> loads -> K, b)."""

    def __init__(self, mesh, u=None):
        self.mesh = mesh
        self.perm = reord(mesh)
        self.idn, self.neq = numeq(mesh, self.perm)
        self.ia, self.ja = csrstruct(mesh, self.idn, self.perm, self.neq)
        self.nad = int(self.ia[-1] - 1)
        self.ad = np.zeros(self.neq)
        self.al = np.zeros(max(self.nad, 1))
        self.b0 = np.zeros(self.neq)
        self.u = np.zeros((mesh.nnode, mesh.ndf)) if u is None else u.copy()
        f = mesh.f.copy()
        ls = loads_struct(mesh.loads)
        L = lib()
        L.or_pload(self.idn, f, self.u, self.b0, np.ascontiguousarray(mesh.nload, np.int32),
                   mesh.nnode, mesh.ndf, C.byref(ls))
        self._pform(self.b0, lhs=0, rhs=1, isw=4)
        self.b = self.b0.copy()
        self._pform(self.b, lhs=1, rhs=1, isw=2)

    def _pform(self, b, lhs, rhs, isw):
        m = self.mesh
        ls = loads_struct(m.loads)
        rc = lib().or_pform(np.ascontiguousarray(m.ix), np.ascontiguousarray(m.eload),
                            np.ascontiguousarray(m.ie), np.ascontiguousarray(m.e),
                            np.ascontiguousarray(m.x), self.idn, self.ia, self.ja, self.al, self.ad,
                            b, self.u, m.numel, m.nen, m.nen, m.ndf, m.ndm, self.neq, lhs, rhs, isw,
                            C.byref(ls), GRAVITY)
        if rc:
            raise ValueError("non-positive Jacobian in element %d" % rc)

    def displacements(self, x):
        """u(ndf, nnode) of update_mec (Predict.f:183-209): x on the free dofs,
        the prescribed values elsewhere; returned as (nnode, ndf)."""
        u = self.u.copy()
        lib().or_update(self.mesh.nnode, self.mesh.ndf, self.idn, u, np.ascontiguousarray(x))
        return u

    def residual(self, x):
        """Second pass of the nonlinear loop (Main.f:789-815): b0 - K u(x)."""
        u = self.u.copy()
        lib().or_update(self.mesh.nnode, self.mesh.ndf, self.idn, u, np.ascontiguousarray(x))
        saved, self.u = self.u, u
        b = self.b0.copy()
        self._pform(b, lhs=1, rhs=1, isw=2)
        self.u = saved
        return b


def tform(mesh, u, tx0=None):
    """Nodal stresses of tform_mec (Pform.f:258-421, isw = 3): u (nnode, ndf)
    -> tx (nnode, ntn), ntn = 4 in 2-D, 6 in 3-D."""
    ntn = 6 if mesh.ndm == 3 else 4
    tx = np.zeros((mesh.nnode, ntn))
    t0 = None if tx0 is None else np.ascontiguousarray(tx0, np.float64)
    rc = lib().or_tform(np.ascontiguousarray(mesh.ix), np.ascontiguousarray(mesh.x),
                        np.ascontiguousarray(mesh.e), np.ascontiguousarray(mesh.ie),
                        np.ascontiguousarray(u, np.float64),
                        None if t0 is None else t0.ctypes.data_as(C.c_void_p), tx,
                        mesh.nnode, mesh.numel, mesh.nen, mesh.ndm, mesh.ndf, ntn)
    if rc:
        raise ValueError("non-positive Jacobian at a node of element %d" % rc)
    return tx


# ------------------------------------------------------------------- solvers
def pre_diag(ad):
    m = np.zeros_like(ad)
    lib().or_pre_diag(m, np.ascontiguousarray(ad), ad.shape[0])
    return m


def block_precond3(ad, al, ia, ja):
    neq = ad.shape[0]
    m = np.zeros(6 * (neq // 3 + neq % 3))
    lib().or_block_precond3(ad, al, ia, ja, m, neq)
    return m


def block_precond2(ad, al, ia, ja):
    neq = ad.shape[0]
    m = np.zeros(3 * (neq // 2 + neq % 2))
    lib().or_block_precond2(ad, al, ia, ja, m, neq)
    return m


def op_block_precond2(m, x):
    y = np.zeros_like(x)
    lib().or_op_block_precond2(m, np.ascontiguousarray(x), y, x.shape[0])
    return y


def pcg_block(nblock, ia, ja, ad, al, b, tol=1e-11, maxit=10000, bm=None):
    """PCG with the block-Jacobi application of op_block_precond, nblock 2 or 3."""
    if bm is None:
        bm = (block_precond2 if nblock == 2 else block_precond3)(ad, al, ia, ja)
    x = np.zeros_like(b)
    st = Stats()
    lib().or_pcg_block(nblock, b.shape[0], ia, ja, ad, al, bm, b, x, tol, maxit, C.byref(st))
    return x, st.as_dict()


def op_block_precond3(m, x):
    y = np.zeros_like(x)
    lib().or_op_block_precond3(m, np.ascontiguousarray(x), y, x.shape[0])
    return y


def matvec(ia, ja, ad, al, x):
    y = np.zeros_like(x)
    lib().or_matvec_csrcsym1(x.shape[0], ia, ja, ad, al, np.ascontiguousarray(x), y)
    return y


def pcg(ia, ja, ad, al, b, tol=1e-11, maxit=10000, m=None, nth=0):
    m = pre_diag(ad) if m is None else m
    x = np.zeros_like(b)
    st = Stats()
    if nth:
        lib().or_pcg_omp(b.shape[0], ia, ja, ad, al, m, b, x, tol, maxit, nth, C.byref(st))
    else:
        lib().or_pcg(b.shape[0], ia, ja, ad, al, m, b, x, tol, maxit, C.byref(st))
    return x, st.as_dict()


def pcg_bdiag3(ia, ja, ad, al, b, tol=1e-11, maxit=10000, bm=None):
    """PCG with the 3x3 block-Jacobi application of op_block_precond."""
    bm = block_precond3(ad, al, ia, ja) if bm is None else bm
    x = np.zeros_like(b)
    st = Stats()
    lib().or_pcg_bdiag3(b.shape[0], ia, ja, ad, al, bm, b, x, tol, maxit, C.byref(st))
    return x, st.as_dict()


def gmres(ia, ja, ad, al, b, k=50, tol=1e-11, maxit=10000, m=None):
    m = pre_diag(ad) if m is None else m
    x = np.zeros_like(b)
    st = Stats()
    lib().or_gmres(b.shape[0], ia, ja, ad, al, m, b, x, k, tol, maxit, C.byref(st))
    return x, st.as_dict()


def pbicgstab(ia, ja, ad, al, b, tol=1e-11, maxit=10000, m=None):
    m = pre_diag(ad) if m is None else m
    x = np.zeros_like(b)
    st = Stats()
    lib().or_pbicgstab(b.shape[0], ia, ja, ad, al, m, b, x, tol, maxit, C.byref(st))
    return x, st.as_dict()


def colormsh(mesh):
    ic = np.zeros(mesh.numel, np.int32)
    n = lib().or_colormsh(np.ascontiguousarray(mesh.ix), mesh.nnode, mesh.numel, mesh.nen,
                          mesh.nen, ic)
    return n, ic


def setcolor(icolor, numcolors):
    """Colormsh.f:255-308: elcolor lists the elements colour by colour, in
    ascending element order within a colour; colorg(2, ncol) holds each
    colour's first and last position in elcolor (1-based)."""
    numel = icolor.shape[0]
    if numcolors <= 1:
        return np.arange(1, numel + 1, dtype=np.int32), np.array([1, numel], np.int32)
    el, cg, pos = [], [], 1
    for c in range(1, numcolors + 1):
        members = np.nonzero(icolor == c)[0] + 1
        el.append(members)
        cg += [pos, pos + len(members) - 1]
        pos += len(members)
    return np.concatenate(el).astype(np.int32), np.array(cg, np.int32)
