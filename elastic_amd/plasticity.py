"""J2 plasticity load paths on the GPU: Newton iterations that re-assemble the
consistent tangent and the residual on the device (elastic_j2_* of
libelastic_hip) and solve the linear systems with the resident PCG, one path
after the other or many paths together in a batch.

The reference has no called plasticity (Plasticity01.f is not reached from
Main.f); the material update is the closed-form radial return with linear
isotropic hardening, current yield stress sy + H * equivalent plastic strain
(the law of Plasticity01.f:322).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, ElasticError


class PlasticPath:
    """One load path: a hex8 mesh with prescribed displacements ubar(nnode, 3)
    at load factor 1 on restrained dofs, yield stress sy, hardening H."""

    def __init__(self, mesh, ubar, sy: float, H: float, device: int = 0):
        from .system import System

        self.sys = System(mesh, device=device, assemble=False)
        self.ctx = self.sys.ctx
        self.mesh = mesh
        self.sy, self.H = sy, H
        ub = np.ascontiguousarray(ubar, np.float64)
        self._L = self.ctx._L
        self.ctx._chk(self._L.elastic_j2_setup(self.ctx._h, sy, H, ub.ctypes.data_as(C.c_void_p)))

    def increment(self, lam: float):
        self.ctx._chk(self._L.elastic_j2_increment(self.ctx._h, lam))

    def assemble(self, lhs: bool = True) -> float:
        r = C.c_double()
        self.ctx._chk(self._L.elastic_j2_assemble(self.ctx._h, 1 if lhs else 0, C.byref(r)))
        return r.value

    def correct(self):
        self.ctx._chk(self._L.elastic_j2_correct(self.ctx._h))

    def commit(self):
        a, n = C.c_double(), C.c_int64()
        self.ctx._chk(self._L.elastic_j2_commit(self.ctx._h, C.byref(a), C.byref(n)))
        return a.value, n.value

    def u(self) -> np.ndarray:
        u = np.zeros(self.mesh.nnode * 3)
        self.ctx._chk(self._L.elastic_j2_get_u(self.ctx._h, u))
        return u.reshape(-1, 3)

    def residual(self) -> np.ndarray:
        r = np.zeros(self.sys.neq)
        self.ctx._chk(self._L.elastic_j2_get_residual(self.ctx._h, r))
        return r

    def state(self, trial: bool = False):
        ngp = self.mesh.numel * 8
        ep, al = np.zeros(6 * ngp), np.zeros(ngp)
        self.ctx._chk(self._L.elastic_j2_get_state(self.ctx._h, 1 if trial else 0,
                                                   ep.ctypes.data_as(C.c_void_p),
                                                   al.ctypes.data_as(C.c_void_p)))
        return ep.reshape(-1, 8, 6), al.reshape(-1, 8)

    def close(self):
        self.ctx.close()


def run_paths(paths, increments: int = 10, rtol: float = 1e-8, max_newton: int = 30,
              precision: int = MIXED, tol: float = 1e-11, batch=None, maxit: int = 50000):
    """Drives every path through `increments` equal load steps.  Per Newton
    iteration each unconverged path re-assembles K_t and R = -f_int on the
    GPU; the systems K_t du = R are solved together in `batch` (a
    serving.Batch) when given, else one after the other; u += du.  A path's
    increment has converged when ||R|| <= rtol * ||R|| of its first iteration.
    Returns per path the list of (increment, iteration, ||R||)."""
    hist = [[] for _ in paths]
    for inc in range(1, increments + 1):
        lam = inc / increments
        for p in paths:
            p.increment(lam)
        active = list(range(len(paths)))
        r0 = {}
        it = 0
        while active:
            need = []
            for i in active:
                r = paths[i].assemble(lhs=True)
                r0.setdefault(i, r)
                hist[i].append((inc, it, r))
                if r <= rtol * r0[i] or r == 0.0:
                    paths[i].commit()
                else:
                    need.append(i)
            if not need:
                break
            if it >= max_newton:
                raise ElasticError(_lib.MAXIT, "Newton did not converge in increment %d" % inc)
            if batch is not None:
                queue, done = list(need), set()
                while len(done) < len(need):
                    while queue and batch.submit(paths[queue[0]].ctx, precision, tol, maxit,
                                                 queue[0]):
                        queue.pop(0)
                    batch.step()
                    for tag, st in batch.poll():
                        if st["status"] != 0:
                            raise ElasticError(st["status"], "linear solve of path %d" % tag)
                        done.add(tag)
            else:
                for i in need:
                    paths[i].ctx.solve_resident(precision=precision, tol=tol, maxit=maxit)
            for i in need:
                paths[i].correct()
            active = need
            it += 1
    return hist
