"""Writes the golden fixtures of tests/golden/ by running the reference.

The reference (Fortran, /root/reference/mef/src) is compiled by
oracle/Makefile.ref into oracle/_ref/ref_harness, a test-only driver that runs
the reference's own mesh/solvm path (RCM, equation numbering, CSRC structure,
element stiffness + load assembly, PCG through ``solv``) and additionally calls
``gmres``, ``pbicgstab`` and ``block_precond`` (3x3 and 2x2 blocks) with their
own signatures, and
``tform_mec`` + ``write_mesh_res_mec`` for the nodal stresses and result files.
For every case this script writes a deck with elastic_amd.deck, runs the
harness, and stores its outputs as a small .npz fixture (large cases keep
SHA-256 digests of the arrays and summaries of the floating-point ones; the
full-size workloads of DIGESTS run the reference's set-up only).

Run from the repository root (needs /root/reference, so only in the build
container):  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from elastic_amd import deck  # noqa: E402
from tests.cases import CASES, DIGESTS, REF_DECKS  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))
REF_INPUT = "/root/reference/input"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_harness(deck_path, tol, maxit, mode, workdir, nth=1, maxmem=400000000):
    r = subprocess.run([HARNESS, deck_path, workdir, repr(tol), str(maxit), str(nth), str(mode),
                        str(maxmem)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
    meta = open(os.path.join(workdir, "meta.txt")).read().split()
    nnode, numel, nen, ndf, ndm, neq, nad = (int(v) for v in meta[:7])
    rd = lambda n, dt: np.fromfile(os.path.join(workdir, n), dtype=dt)  # noqa: E731
    out = dict(inum=rd("inum.i4", np.int32), id=rd("id.i4", np.int32), ia=rd("ia.i4", np.int32),
               ja=rd("ja.i4", np.int32), ad=rd("ad.f8", np.float64), al=rd("al.f8", np.float64),
               b=rd("b.f8", np.float64))
    for name in ("elcolor", "colorg"):
        out[name] = rd(name + ".i4", np.int32)
    for name in ("x", "spmv_v", "spmv_y", "resid", "x_gmres", "x_bicgstab", "bdiag", "bdiag_y",
                 "bdiag2", "bdiag2_y", "tx", "u"):
        p = os.path.join(workdir, name + ".f8")
        if os.path.exists(p):
            out[name] = np.fromfile(p, dtype=np.float64)
    # result files of write_mesh_res_mec, kept byte for byte
    for key, fname in (("vtk_ascii", "res_istep_0.vtk"), ("vtk_binary", "resb_istep_0.vtk"),
                       ("vtu_ascii", "res_istep_0.vtu")):
        p = os.path.join(workdir, fname)
        if os.path.exists(p):
            out[key] = np.frombuffer(open(p, "rb").read(), dtype=np.uint8)
    lp = os.path.join(workdir, "solver.log")
    log = open(lp).read() if os.path.exists(lp) else ""
    stats = {}
    for key, pat in (("pcg", r"PCG:\s+it\s+(\d+) x \* Kx\s+(\S+) \|\|x\|\|\s+(\S+)"),
                     ("gmres", r"GMRES:\s+it\s+(\d+) x \* Kx\s+(\S+) \|\|x\|\|\s+(\S+)"),
                     ("bicgstab", r"BICGSTAB:\s+it\s+(\d+) x \* Kx\s+(\S+) \|\|x\|\|\s+(\S+)")):
        mt = re.search(pat, log)
        if mt:
            stats[key + "_iters"] = int(mt.group(1))
            stats[key + "_xkx"] = float(mt.group(2).replace("D", "E"))
            stats[key + "_normx"] = float(mt.group(3).replace("D", "E"))
    out["dims"] = np.array([nnode, numel, nen, ndf, ndm, neq, nad], np.int64)
    return out, stats


def summarize(out, stats, keep_full):
    res = {k: np.asarray(v) for k, v in stats.items()}
    res["dims"] = out["dims"]
    if keep_full:
        for k, v in out.items():
            res[k] = v
        return res
    for k in ("inum", "id", "ia", "ja", "elcolor", "colorg"):
        res["sha_" + k] = np.array(sha(out[k]))
    for k in ("ad", "al", "b", "spmv_y", "resid"):
        if k not in out:
            continue
        v = out[k]
        res["sha_" + k] = np.array(sha(v))
        res["norm_" + k] = np.array([np.linalg.norm(v), np.abs(v).max(), v.sum()])
        res["sample_" + k] = v[np.linspace(0, v.size - 1, 64).astype(np.int64)]
    for k in ("x", "x_gmres", "x_bicgstab", "bdiag"):
        if k in out:
            res[k] = out[k]
    return res


def main(only=None):
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref"], cwd=ROOT, check=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, build in DIGESTS.items():
            if only and name not in only:
                continue
            # full-size workloads: the reference's set-up only (mode 64), with
            # the coloured element loop (mode 16) whose sums the GPU assembly
            # reproduces; digests of every array
            m = build()
            dpath = os.path.join(tmp, name + ".dat")
            deck.write_deck(m, dpath)
            wd = os.path.join(tmp, name)
            os.makedirs(wd)
            out, stats = run_harness(dpath, 1e-10, 1, 16 | 64, wd, maxmem=1500000000)
            res = summarize(out, stats, False)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
            print(name, out["dims"].tolist(), flush=True)
            for f in os.listdir(wd):
                os.remove(os.path.join(wd, f))
            os.remove(dpath)
        for name, (build, tol, mode, full) in CASES.items():
            if (only and name not in only) or name in REF_DECKS:
                continue
            maxit = 10000
            m = build()
            dpath = os.path.join(tmp, name + ".dat")
            deck.write_deck(m, dpath)
            wd = os.path.join(tmp, name)
            os.makedirs(wd)
            out, stats = run_harness(dpath, tol, maxit, mode, wd)
            res = summarize(out, stats, full)
            res["tol"] = np.array(tol)
            if full:
                # the reference's coloured (OpenMP) element loop: same K and b
                # up to the order in which element contributions are added
                wd2 = os.path.join(tmp, name + "_omp")
                os.makedirs(wd2)
                out2, _ = run_harness(dpath, tol, maxit, 1 | 16, wd2)
                for k in ("ad", "al", "b"):
                    res[k + "_omp"] = out2[k]
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
            print(name, out["dims"].tolist(), stats)
        for name, rel in REF_DECKS.items():
            if only and name not in only:
                continue
            wd = os.path.join(tmp, name)
            os.makedirs(wd)
            out, stats = run_harness(os.path.join(REF_INPUT, rel), 1e-11, 10000, 1, wd)
            res = summarize(out, stats, False)
            res["tol"] = np.array(1e-11)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
            print(name, out["dims"].tolist(), stats)


if __name__ == "__main__":
    main(set(sys.argv[1:]) or None)
