"""The reference's mesh -> solve path on the GPU library (Main.f:348-849).

``System(mesh)`` runs the host set-up of the ``mesh`` macro with the library's
own restatements (RCM order, equation numbers, CSRC pattern; bit-identical to
the reference), uploads the structure and the mesh, and assembles on the GPU
the way the ``solvm`` macro does: nodal loads (``pload_mec``), element loads
(``pform_mec`` isw 4) and the stiffness with b = b0 - K u (isw 2).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import Context


def pload(mesh, idn, neq):
    """Pload.f:2-57 at t = 0: b(k) = f for free dofs, u = f for restrained
    ones; nodal load numbers (types 1 and 2) set f first."""
    f = mesh.f.copy()
    for (node, dof) in zip(*np.nonzero(mesh.nload)):
        ty, vals = mesh.loads[mesh.nload[node, dof] - 1]
        f[node, dof] = vals[0] if ty in (1, 2) else 0.0
    b = np.zeros(neq)
    u = np.zeros((mesh.nnode, mesh.ndf))
    free = idn > 0
    b[idn[free] - 1] = f[free]
    u[~free] = f[~free]
    return b, u


class System:
    def __init__(self, mesh, ctx: Context | None = None, device: int = 0, tile_nnz=None,
                 assemble: bool = True):
        self.mesh = mesh
        self.ctx = ctx or Context(device, tile_nnz=tile_nnz)
        self.perm = _lib.reord(mesh)
        self.idn, self.neq = _lib.numeq(mesh, self.perm)
        self.ia, self.ja = _lib.csrstruct(mesh, self.idn, self.neq)
        self.nad = int(self.ia[-1] - 1)
        self.ctx.set_structure(self.ia, self.ja)
        lt = np.array([ty for ty, _ in mesh.loads] or [0], np.int32)
        ln = np.array([len(v) for _, v in mesh.loads] or [0], np.int32)
        lv = np.zeros((max(len(mesh.loads), 1), 3))
        for i, (_, v) in enumerate(mesh.loads):
            lv[i, : min(3, len(v))] = v[:3]
        L = self.ctx._L
        self.ctx._chk(L.elastic_set_mesh(
            self.ctx._h, mesh.nnode, mesh.numel, mesh.nen, mesh.ndf, mesh.ndm, mesh.numat,
            np.ascontiguousarray(mesh.ix, np.int32), np.ascontiguousarray(mesh.x, np.float64),
            self.idn, np.ascontiguousarray(mesh.ie, np.int32), np.ascontiguousarray(mesh.e, np.float64),
            np.ascontiguousarray(mesh.eload, np.int32), len(mesh.loads), lt, ln, lv.ravel().copy()))
        self.b0, self.u = pload(mesh, self.idn, self.neq)
        self.b = None
        if assemble:
            self.assemble_loads()
            self.assemble()

    def _assemble(self, isw, lhs, rhs, u, b):
        L = self.ctx._L
        up = u.ctypes.data_as(C.c_void_p) if u is not None else None
        bp = b.ctypes.data_as(C.c_void_p) if b is not None else None
        self.ctx._chk(L.elastic_assemble(self.ctx._h, isw, lhs, rhs, up, bp))

    def assemble_loads(self):
        """b0 += element loads (pform_mec isw 4, Main.f:756-767)."""
        self._assemble(4, 0, 1, None, self.b0)

    def assemble(self, u=None):
        """K on the device and b = b0 - K u (pform_mec isw 2, Main.f:791-802)."""
        u = self.u if u is None else u
        self.b = self.b0.copy()
        uu = np.ascontiguousarray(u, np.float64) if np.any(u) else None
        self._assemble(2, 1, 1, uu, self.b)
        return self.b

    def residual(self, x):
        """b0 - K u(x) without touching the device K (second nonlinear pass)."""
        u = self.u.copy()
        free = self.idn > 0
        u[free] = x[self.idn[free] - 1]
        b = self.b0.copy()
        self._assemble(2, 0, 1, np.ascontiguousarray(u), b)
        return b

    def displacements(self, x):
        """u(nnode, ndf) of update_mec (Predict.f:183-209): x on the free dofs,
        the prescribed values elsewhere."""
        u = self.u.copy()
        free = self.idn > 0
        u[free] = x[self.idn[free] - 1]
        return u

    def stress(self, u=None, from_x: bool = False, tx0=None):
        """Nodal stresses tx (nnode, ntn) of tform_mec on the GPU (Pform.f:258-421):
        ntn = 4 in 2-D (sxx syy szz sxy), 6 in 3-D (sxx syy szz sxy syz sxz).
        u (nnode, ndf) defaults to the prescribed displacements; from_x puts
        the context's resident solution on the free dofs first (update_mec)."""
        m = self.mesh
        ntn = 6 if m.ndm == 3 else 4
        uu = np.ascontiguousarray(self.u if u is None else u, np.float64)
        t0 = None if tx0 is None else np.ascontiguousarray(tx0, np.float64)
        tx = np.zeros((m.nnode, ntn))
        self.ctx._chk(self.ctx._L.elastic_stress(
            self.ctx._h, uu.ctypes.data_as(C.c_void_p), 1 if from_x else 0,
            None if t0 is None else t0.ctypes.data_as(C.c_void_p), tx.ctypes.data_as(C.c_void_p)))
        return tx

    def values(self):
        return self.ctx.get_values()

    def time_assembly(self, reps=3):
        ms = C.c_double()
        self.ctx._chk(self.ctx._L.elastic_time_assembly(self.ctx._h, reps, C.byref(ms)))
        return ms.value

    def solve(self, **kw):
        return self.ctx.solve(self.b, **kw)
