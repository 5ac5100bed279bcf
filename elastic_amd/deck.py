"""Reader and writer for the subset of the reference's ``.dat`` input decks
that describes one linear-elastic mesh (Rdata.f:1-743 ``rdat``):

``mesh`` / ``nnode .. numel .. numat .. maxno .. ndf .. dim ..`` (Rdata.f:1272)
/ ``materials`` (``mate``, Rdata.f:830) / ``coordinates`` (``coord``, :791) /
``tria3|quad4|tetra4|hexa8`` (``elconn``, :747) / ``constraindisp`` (``bound``,
:941) / ``nodalforces`` (``forces``, :879) / ``elmtloads`` / ``nodalloads`` /
``loads`` (``rload``, :1020) / ``insert <file>`` / ``end mesh``.

Lines starting with ``//`` are comments.  Macro lines outside ``mesh`` (e.g.
``solvm``, ``stop``, or the stale inline ``config`` blocks of some example
decks) are skipped.  Numbers are written with ``repr`` so that every double
survives a round trip through the reference's list-directed reads.
"""
from __future__ import annotations

import os
import numpy as np

from .mesh import Mesh, PROP

_ELEMENTS = {"tria3": 3, "quad4": 4, "tetra4": 4, "hexa8": 8}


def _tokens(path):
    """Yield (tokens, path) per non-empty line, expanding ``insert``."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path, "r", errors="replace") as fh:
        for raw in fh:
            line = raw.replace("\r", "").strip()
            if not line or line.startswith("//"):
                continue
            tok = line.split()
            if tok[0] == "insert" and len(tok) > 1:
                yield from _tokens(os.path.join(base, tok[1]))
                continue
            yield tok


def _num(s: str) -> float:
    return float(s.lower().replace("d", "e"))


def read_deck(path: str) -> Mesh:
    it = _tokens(path)
    for tok in it:
        if tok[0] == "mesh":
            break
    else:
        raise ValueError("no 'mesh' macro in %s" % path)
    params = next(it)
    kv = {params[q]: int(params[q + 1]) for q in range(0, len(params) - 1, 2)}
    nnode, numel, numat = kv["nnode"], kv["numel"], kv["numat"]
    nen, ndf, ndm = kv["maxno"], kv["ndf"], kv["dim"]
    ix = np.zeros((numel, nen + 1), np.int32)
    x = np.zeros((nnode, ndm))
    idf = np.zeros((nnode, ndf), np.int32)
    ie = np.zeros(numat, np.int32)
    e = np.zeros((numat, PROP))
    eload = np.zeros((numel, 7), np.int32)
    nload = np.zeros((nnode, ndf), np.int32)
    f = np.zeros((nnode, ndf))
    loads: dict[int, tuple] = {}
    for tok in it:
        key = tok[0]
        if key == "end" and len(tok) > 1 and tok[1] == "mesh":
            break
        if key == "materials":
            for t in it:
                if t[0] == "end":
                    break
                ma = int(t[0])
                ie[ma - 1] = int(t[1])
                for q, v in enumerate(t[2:]):
                    e[ma - 1, q] = _num(v)
        elif key == "coordinates":
            for t in it:
                if t[0] == "end":
                    break
                k = int(t[0])
                x[k - 1] = [_num(v) for v in t[1:1 + ndm]]
        elif key in _ELEMENTS:
            nv = _ELEMENTS[key]
            for t in it:
                if t[0] == "end":
                    break
                k = int(t[0])
                ix[k - 1, :nv] = [int(v) for v in t[1:1 + nv]]
                ix[k - 1, nen] = int(t[1 + nv])
        elif key in ("constraindisp", "nodalloads"):
            arr = idf if key == "constraindisp" else nload
            for t in it:
                if t[0] == "end":
                    break
                k = int(t[0])
                vals = [int(v) for v in t[1:1 + ndf]]
                arr[k - 1, :len(vals)] = vals
        elif key == "nodalforces":
            for t in it:
                if t[0] == "end":
                    break
                k = int(t[0])
                f[k - 1] = [_num(v) for v in t[1:1 + ndf]]
        elif key == "elmtloads":
            for t in it:
                if t[0] == "end":
                    break
                k = int(t[0])
                vals = [int(v) for v in t[1:8]]
                eload[k - 1, :len(vals)] = vals
        elif key == "loads":
            for t in it:
                if t[0] == "end":
                    break
                i, ty = int(t[0]), int(t[1])
                if ty == 40:
                    npar = int(t[2])
                    loads[i] = (40, [_num(v) for v in t[3:3 + npar]])
                elif ty in (1, 2):
                    loads[i] = (ty, [_num(v) for v in t[2:]])
                else:
                    raise ValueError("load type %d is outside the supported subset" % ty)
    lst = [loads.get(i, (1, [0.0])) for i in range(1, (max(loads) if loads else 0) + 1)]
    name = os.path.splitext(os.path.basename(path))[0]
    return Mesh(nnode, numel, nen, ndf, ndm, ix, x, idf, ie, e, eload, nload, f, lst, name=name)


def write_deck(mesh: Mesh, path: str, macros=("solvm", "stop")) -> None:
    """Write ``mesh`` as a deck the reference driver reads unchanged."""
    el_key = {3: "tria3", 4: "quad4" if mesh.ndm == 2 else "tetra4", 8: "hexa8"}[mesh.nen]
    nfaces = {"tria3": 3, "quad4": 4, "tetra4": 4, "hexa8": 6}[el_key]
    r = repr
    with open(path, "w") as fh:
        fh.write("mesh\n")
        fh.write("nnode %d numel %d numat %d maxno %d ndf %d dim %d\n"
                 % (mesh.nnode, mesh.numel, mesh.numat, mesh.nen, mesh.ndf, mesh.ndm))
        fh.write("materials\n")
        for ma in range(mesh.numat):
            vals = mesh.e[ma]
            nz = int(np.max(np.nonzero(vals)[0])) + 1 if np.any(vals) else 1
            fh.write("%d %d %s\n" % (ma + 1, mesh.ie[ma], " ".join(r(float(v)) for v in vals[:max(nz, 2)])))
        fh.write("end materials\ncoordinates\n")
        fh.writelines("%d %s\n" % (k + 1, " ".join(r(float(v)) for v in mesh.x[k]))
                      for k in range(mesh.nnode))
        fh.write("end coordinates\n%s\n" % el_key)
        fh.writelines("%d %s\n" % (k + 1, " ".join(str(int(v)) for v in mesh.ix[k]))
                      for k in range(mesh.numel))
        fh.write("end %s\nconstraindisp\n" % el_key)
        for k in np.nonzero(mesh.id.any(axis=1))[0]:
            fh.write("%d %s\n" % (k + 1, " ".join(str(int(v)) for v in mesh.id[k])))
        fh.write("end constraindisp\n")
        if np.any(mesh.f):
            fh.write("nodalforces\n")
            for k in np.nonzero((mesh.f != 0).any(axis=1))[0]:
                fh.write("%d %s\n" % (k + 1, " ".join(r(float(v)) for v in mesh.f[k])))
            fh.write("end nodalforces\n")
        if np.any(mesh.nload):
            fh.write("nodalloads\n")
            for k in np.nonzero(mesh.nload.any(axis=1))[0]:
                fh.write("%d %s\n" % (k + 1, " ".join(str(int(v)) for v in mesh.nload[k])))
            fh.write("end nodalloads\n")
        if np.any(mesh.eload):
            fh.write("elmtloads\n")
            for k in np.nonzero(mesh.eload.any(axis=1))[0]:
                fh.write("%d %s\n" % (k + 1, " ".join(str(int(v)) for v in mesh.eload[k, :nfaces])))
            fh.write("end elmtloads\n")
        if mesh.loads:
            fh.write("loads\n")
            for i, (ty, vals) in enumerate(mesh.loads):
                if ty == 40:
                    fh.write("%d 40 %d %s\n" % (i + 1, len(vals), " ".join(r(float(v)) for v in vals)))
                else:
                    fh.write("%d %d %s\n" % (i + 1, ty, " ".join(r(float(v)) for v in vals)))
            fh.write("end loads\n")
        fh.write("end mesh\n")
        for m in macros:
            fh.write(m + "\n")
