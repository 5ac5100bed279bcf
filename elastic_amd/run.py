"""Runs one input deck the way the reference's driver runs a linear-elastic
deck (Main.f: ``mesh`` macro 348-534, ``solvm`` 718-849, ``pgeo`` 855-963),
with the set-up, assembly, solve and nodal stresses on the GPU library:

    python -m elastic_amd.run deck.dat out/prefix [--legacy] [--binary]
           [--solver pcg|gmres|bicgstab] [--precision fp64|mixed]
           [--tol 1e-11] [--maxit 10000]

The deck is read by elastic_amd.deck (the ``mesh`` subset of Rdata.f);
the result file is ``<prefix>_istep_0.vtu`` (or ``.vtk`` with --legacy), in
the reference's format, holding the displacements and the nodal stresses.
"""
from __future__ import annotations

import argparse
import sys
import time

from . import deck, vtk
from .solver import BICGSTAB, FP64, GMRES, MIXED, PCG
from .system import System

SOLVERS = {"pcg": PCG, "gmres": GMRES, "bicgstab": BICGSTAB}
PRECISIONS = {"fp64": FP64, "mixed": MIXED}


def run(deck_path: str, prefix: str, solver: int = PCG, precision: int = FP64,
        tol: float = 1e-11, maxit: int = 10000, legacy: bool = False, binary: bool = False,
        stress: bool = True) -> dict:
    """Deck -> GPU set-up and assembly -> solve -> update_mec -> tform_mec ->
    result file.  Returns the solve statistics, u (nnode, ndf), the nodal
    stresses tx (nnode, ntn) and the path of the file written."""
    t0 = time.perf_counter()
    mesh = deck.read_deck(deck_path)
    sysm = System(mesh)
    t_setup = time.perf_counter() - t0
    x, st = sysm.solve(solver=solver, precision=precision, tol=tol, maxit=maxit)
    u = sysm.displacements(x)                       # update_mec (Main.f:837)
    tx = sysm.stress(from_x=True) if stress else None  # tform_mec (Main.f:946), resident x
    path = vtk.write_result(prefix, mesh.ix, mesh.x, u, tx, mesh.nen, mesh.ndm, mesh.ndf, 0,
                            legacy, binary)
    return {"stats": st, "u": u, "tx": tx, "path": path, "neq": sysm.neq, "nad": sysm.nad,
            "setup_s": t_setup, "total_s": time.perf_counter() - t0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("deck")
    ap.add_argument("prefix")
    ap.add_argument("--legacy", action="store_true", help="legacy .vtk instead of XML .vtu")
    ap.add_argument("--binary", action="store_true", help="binary legacy file")
    ap.add_argument("--solver", choices=sorted(SOLVERS), default="pcg")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fp64")
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--maxit", type=int, default=10000)
    ap.add_argument("--no-stress", action="store_true")
    a = ap.parse_args(argv)
    r = run(a.deck, a.prefix, SOLVERS[a.solver], PRECISIONS[a.precision], a.tol, a.maxit,
            a.legacy, a.binary, not a.no_stress)
    st = r["stats"]
    print("neq %d nad %d: %s it %d, x.Kx %.10e, ||x|| %.10e, rel res %.3e; wrote %s (%.2f s)"
          % (r["neq"], r["nad"], a.solver, st["iters"], st["xkx"], st["norm_x"], st["rel_res"],
             r["path"], r["total_s"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
