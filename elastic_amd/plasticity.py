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

    def step(self, s: float):
        """u += s du along the last correction (line search)."""
        self.ctx._chk(self._L.elastic_j2_step(self.ctx._h, float(s)))

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


def run_paths(paths, increments: int = 10, rtol: float = 1e-8, max_newton: int = 60,
              precision: int = MIXED, tol: float = 1e-11, batch=None, maxit: int = 50000,
              counts: dict | None = None, max_halvings: int = 6, last: int | None = None,
              progress=None, first: int = 1):
    """Drives every path through `increments` equal load steps.  Per Newton
    iteration each unconverged path re-assembles K_t and R = -f_int on the
    GPU; the systems K_t du = R are solved together in `batch` (a
    serving.Batch) when given, else one after the other; u += du.  A path's
    increment has converged when ||R|| <= rtol * ||R|| of its first iteration.

    Backtracking line search: a full step is kept when ||R|| decreases;
    otherwise the step along du is halved (re-assembling only) up to
    `max_halvings` times, and the last trial is kept.  Full Newton steps can
    cycle once the set of yielding points changes from one iterate to the
    next under a large load step; the halving restores descent.
    Returns per path the list of (increment, assembly, ||R||); `counts`
    (optional dict) receives the numbers of linear solves and of halvings.
    `last` stops after that increment (the load still split in `increments`);
    `first` resumes a run whose increments before it are committed.
    `progress` (a file, optional) receives one line per finished increment."""
    hist = [[] for _ in paths]
    nsolve = nhalf = 0
    for inc in range(first, (last or increments) + 1):
        lam = inc / increments
        for p in paths:
            p.increment(lam)
        active = list(range(len(paths)))
        r0, r_acc, frac = {}, {}, {}
        it = 0
        while active:
            need, retry = [], []
            for i in active:
                r = paths[i].assemble(lhs=True)
                r0.setdefault(i, r)
                hist[i].append((inc, it, r))
                if r <= rtol * r0[i] or r == 0.0:
                    paths[i].commit()
                elif (i in frac and not (r < (1.0 - 1e-4 * frac[i]) * r_acc[i])
                      and frac[i] > 0.5 ** max_halvings):
                    paths[i].step(-0.5 * frac[i])   # back to half the last trial step
                    frac[i] *= 0.5
                    nhalf += 1
                    retry.append(i)
                else:
                    r_acc[i] = r                     # accepted point: new Newton direction
                    frac.pop(i, None)
                    need.append(i)
            if not need and not retry:
                break
            if it >= max_newton:
                raise ElasticError(_lib.MAXIT, "Newton did not converge in increment %d" % inc)
            if batch is not None and need:
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
                frac[i] = 1.0
            nsolve += len(need)
            active = need + retry
            it += 1
        if progress is not None:
            print("increment %d done: %d assemblies, %d linear solves so far"
                  % (inc, it, nsolve), file=progress, flush=True)
    if counts is not None:
        counts["linear_solves"] = nsolve
        counts["halvings"] = nhalf
    return hist


def compare_solves(paths, tol: float = 1e-10, maxit: int = 50000, fp32_sums: bool = True) -> dict:
    """Solve-only comparison on the paths' systems as last assembled (K_t du =
    R of each path, resident): plain FP64 PCG one system after the other (the
    drop-in solver call), then all systems together in a batch, in FP64 and in
    mixed precision (FP64 sums in the FP32 operator, and with `fp32_sums`
    also FP32 sums).  Returns wall times, solved DOF/s, iteration totals and
    the largest relative L2 difference of du against the sequential FP64
    solutions.  Each context's FP32-sum option is put back afterwards."""
    import time

    from .serving import Batch

    saved = [p.ctx.get_option("accum32") for p in paths]
    dof = float(sum(p.sys.neq for p in paths))
    a = time.perf_counter()  # the assemblies before this returned on completion
    it64 = 0
    for p in paths:
        it64 += p.ctx.solve_resident(precision=FP64, tol=tol, maxit=maxit)["iters"]
    t_seq = time.perf_counter() - a
    x64 = [p.ctx.download_x() for p in paths]
    out = {"systems": len(paths), "dof": dof, "tol": tol,
           "sequential_fp64": {"s": t_seq, "dof_per_s": dof / t_seq, "iters": it64}}
    runs = [("batched_fp64", FP64, 0), ("batched_mixed", MIXED, 0)]
    if fp32_sums:
        runs.append(("batched_mixed_fp32_sums", MIXED, 1))
    for name, prec, acc in runs:
        for p in paths:
            p.ctx.set_option("accum32", acc)
        a = time.perf_counter()
        done, itb = {}, 0
        with Batch(0, max_slots=max(1, min(len(paths), 255)), iters_per_step=16) as b:
            queue = list(range(len(paths)))
            while len(done) < len(paths):
                while queue and b.submit(paths[queue[0]].ctx, prec, tol, maxit, queue[0]):
                    queue.pop(0)
                b.step()
                for tag, st in b.poll():
                    if st["status"] != 0:
                        raise ElasticError(st["status"], "linear solve of path %d" % tag)
                    done[tag] = st
                    itb += st["iters"]
        t_bat = time.perf_counter() - a
        err = max(float(np.linalg.norm(p.ctx.download_x() - x64[i]) / np.linalg.norm(x64[i]))
                  for i, p in enumerate(paths))
        out[name] = {"s": t_bat, "dof_per_s": dof / t_bat, "iters": itb,
                     "speedup_vs_sequential_fp64": t_seq / t_bat, "max_rel_err_vs_fp64": err}
    for p, a32 in zip(paths, saved):
        p.ctx.set_option("accum32", a32)
    return out


def load_paths(n: int = 60, npaths: int = 32, stretch: float = 5e-3, H: float = 0.05,
               seed: int = 20261207, device: int = 0):
    """The plasticity workload: hex8 n^3 held between grips (z=0 clamped, z=1
    pulled by `stretch` at load factor 1), E = 1, nu = 0.3, hardening H, and
    `npaths` load paths whose yield stresses are drawn from U[5e-4, 2e-3]
    (counter-based, so every run draws the same values)."""
    from .mesh import hex8_block_pull, uniform01

    m, ubar = hex8_block_pull(n, stretch)
    u = uniform01(seed, np.arange(npaths, dtype=np.uint64))
    sys_y = 5e-4 + u * (2e-3 - 5e-4)
    return [PlasticPath(m, ubar, float(sy), H, device=device) for sy in sys_y], sys_y


def main(argv=None) -> int:
    """python -m elastic_amd.plasticity [--n 60] [--paths 32] [--increments 10]
    [--mode batched|sequential] [--precision mixed|fp64] [--sums auto|fp64|fp32]: runs the load paths
    and prints one JSON line (wall time, Newton iterations, solved DOF/s)."""
    import argparse
    import json
    import sys
    import time

    from .serving import Batch

    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=60)
    ap.add_argument("--paths", type=int, default=32)
    ap.add_argument("--increments", type=int, default=10)
    ap.add_argument("--stretch", type=float, default=5e-3)
    ap.add_argument("--mode", choices=["batched", "sequential"], default="batched")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--sums", choices=["auto", "fp64", "fp32"], default="auto",
                    help="sums of the FP32 operator in mixed-precision solves (auto: the "
                         "library's choice, FP32 for 3-dof systems of >= 32,768 nodes)")
    a = ap.parse_args(argv)
    prec = MIXED if a.precision == "mixed" else FP64
    t0 = time.perf_counter()
    paths, sys_y = load_paths(a.n, a.paths, a.stretch)
    if a.sums != "auto":
        for p in paths:
            p.ctx.set_option("accum32", 1 if a.sums == "fp32" else 0)
    t_setup = time.perf_counter() - t0
    neq = paths[0].sys.neq
    t1 = time.perf_counter()
    counts: dict = {}
    if a.mode == "batched":
        with Batch(0, max_slots=max(1, min(a.paths, 255)), iters_per_step=16) as b:
            hist = run_paths(paths, increments=a.increments, rtol=a.rtol, precision=prec, batch=b,
                             counts=counts, progress=sys.stderr)
    else:
        hist = run_paths(paths, increments=a.increments, rtol=a.rtol, precision=prec,
                         counts=counts, progress=sys.stderr)
    wall = time.perf_counter() - t1
    solves = counts["linear_solves"]
    newton = [max(it for (inc, it, _) in h if inc == k) for h in hist
              for k in range(1, a.increments + 1)]  # assemblies per increment
    plastic = [p.commit()[1] for p in paths]
    out = {"workload": "J2 hex8 %d^3 pulled %g, %d paths, sy U[5e-4,2e-3], H 0.05, %d increments"
                       % (a.n, a.stretch, a.paths, a.increments),
           "mode": a.mode, "precision": a.precision, "operator_sums": a.sums,
           "neq_per_path": neq,
           "setup_s": t_setup, "wall_s": wall, "linear_solves": solves,
           "line_search_halvings": counts["halvings"],
           "assemblies_per_increment_max": max(newton), "solved_dof_per_s": neq * solves / wall,
           "plastic_points_min": int(min(plastic)), "plastic_points_max": int(max(plastic))}
    print(json.dumps(out), flush=True)
    for p in paths:
        p.close()
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
