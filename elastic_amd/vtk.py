"""Result files of a solve in the reference's formats (Write_vtk.f:742-984
``write_mesh_res_mec`` with the writers of Vtk.f), so that a run of a deck
through this library leaves the same files the reference's ``pgeo``/``solvm``
path writes: legacy VTK (ASCII or big-endian binary, ``<prefix>_istep_<n>.vtk``)
or XML VTU (ASCII, ``<prefix>_istep_<n>.vtu``), holding the mesh, the material
and global id of every cell, the global id of every node, the displacements
(``desloc``) and the nodal stress tensor (``stress``).

The text is formatted as the reference's Fortran edit descriptors format it
(``Ew.d``, ``Ew.dEe``, ``Iw`` and flang's list-directed output), so the files
match the reference's byte for byte; the meshes here carry one element type
(the element ranges of elementos.fi are then (numel, 1) for that type).
"""
from __future__ import annotations

import math

import numpy as np

# VTK cell type and node count per element kind (Vtk.f:167-446)
_CELL = {"tria3": (5, 3), "quad4": (9, 4), "tetra4": (10, 4), "hexa8": (12, 8)}
_NUMET = {"tria3": 4, "quad4": 5, "tetra4": 5, "hexa8": 9}  # node count + 1 (Vtk.f:203-204)


def element_kind(nen: int, ndm: int) -> str:
    if nen == 3:
        return "tria3"
    if nen == 4:
        return "quad4" if ndm == 2 else "tetra4"
    if nen == 8:
        return "hexa8"
    raise ValueError("no VTK cell for nen=%d ndm=%d" % (nen, ndm))


# ------------------------------------------------------- Fortran edit descriptors
def f_e(v: float, w: int, d: int, e: int | None = None) -> str:
    """Fortran ``Ew.d`` (e None) or ``Ew.dEe``: 0.ddddd mantissa, rounded to
    nearest; ``Ew.d`` drops the letter E for exponents beyond 99."""
    v = float(v)
    neg = math.copysign(1.0, v) < 0
    if v == 0.0:
        digits, ex = "0" * d, 0
    else:
        m, x = ("%.*e" % (d - 1, abs(v))).split("e")
        digits, ex = m.replace(".", ""), int(x) + 1
    if e is None:
        expo = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
    else:
        expo = "E" + ("-" if ex < 0 else "+") + str(abs(ex)).zfill(e)
    s = ("-" if neg else "") + "0." + digits + expo
    return s.rjust(w) if len(s) <= w else "*" * w


def f_i(v: int, w: int) -> str:
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def _rows(vals, per: int, fmt) -> str:
    """One record per ``per`` values (a repeated descriptor like 10i10)."""
    vals = list(vals)
    return "".join("".join(fmt(v) for v in vals[k:k + per]) + "\n"
                   for k in range(0, max(len(vals), 1), per)) if vals else "\n"


def _pad(s: str, n: int) -> str:
    return s[:n].ljust(n)


# ------------------------------------------------------------------ writer
def result_name(prefix: str, istep: int, legacy: bool) -> str:
    """File name of ``name(prename, istep, 2|3)`` (Filenames.f:57-64, 141-148):
    the prefix up to its first '.', then ``_istep_<n>.vtk|.vtu``."""
    base = prefix.split(" ")[0]
    if "." in base:
        base = base[:base.index(".")]
    return base + "_istep_%d.%s" % (istep, "vtk" if legacy else "vtu")


def write_result(prefix: str, ix: np.ndarray, x: np.ndarray, u: np.ndarray,
                 tx: np.ndarray | None, nen: int, ndm: int, ndf: int, istep: int = 0,
                 legacy: bool = False, binary: bool = False) -> str:
    """Writes the result file of ``write_mesh_res_mec`` and returns its path.

    ix: (numel, nen+1) 1-based connectivity + material; x: (nnode, ndm);
    u: (nnode, ndf) displacements; tx: (nnode, ntn) nodal stresses (ntn = 4
    in 2-D: sxx syy szz sxy; 6 in 3-D: sxx syy szz sxy syz sxz) or None (no
    stress field, as with the reference's ``print_flag(3)`` off).  The XML
    format exists only in ASCII (the reference stops for binary VTU)."""
    if binary and not legacy:
        raise ValueError("binary VTU is not implemented by the reference (Vtk.f:131-133)")
    path = result_name(prefix, istep, legacy)
    if legacy:
        data = _legacy_binary(ix, x, u, tx, nen, ndm, ndf) if binary else \
            _legacy_ascii(ix, x, u, tx, nen, ndm, ndf).encode()
    else:
        data = _vtu_ascii(ix, x, u, tx, nen, ndm, ndf).encode()
    with open(path, "wb") as fh:
        fh.write(data)
    return path


def full_tensor(tx: np.ndarray) -> np.ndarray:
    """The 3x3 tensor rows of ``make_full_tensor`` (Write_vtk.f:1166-1243)."""
    n, ntn = tx.shape
    t = np.zeros((n, 9))
    if ntn == 4:
        t[:, 0], t[:, 1], t[:, 3], t[:, 4], t[:, 8] = tx[:, 0], tx[:, 3], tx[:, 3], tx[:, 1], tx[:, 2]
    else:
        t[:, 0], t[:, 1], t[:, 2] = tx[:, 0], tx[:, 3], tx[:, 5]
        t[:, 3], t[:, 4], t[:, 5] = tx[:, 3], tx[:, 1], tx[:, 4]
        t[:, 6], t[:, 7], t[:, 8] = tx[:, 5], tx[:, 4], tx[:, 2]
    return t


def _xyz(x: np.ndarray, ndm: int) -> np.ndarray:
    out = np.zeros((x.shape[0], 3))
    out[:, :ndm] = x[:, :ndm]
    return out


def _legacy_ascii(ix, x, u, tx, nen, ndm, ndf) -> str:
    kind = element_kind(nen, ndm)
    ctype, nn = _CELL[kind]
    numel, nnode = ix.shape[0], x.shape[0]
    e155 = lambda v: f_e(v, 15, 5)          # noqa: E731
    e1553 = lambda v: f_e(v, 15, 5, 3)      # noqa: E731
    i10 = lambda v: f_i(v, 10)              # noqa: E731
    o = []
    # head_vtk (Vtk.f:1845-1918), title 'Malha Res' in a character*30
    o += ["# vtk DataFile Version 3.0\n", _pad("Malha Res", 30) + "\n", "ASCII\n",
          "DATASET UNSTRUCTURED_GRID\n"]
    # coor_vtk (Vtk.f:24-111): (3e15.5) with z = 0 in 2-D
    o.append("POINTS " + f_i(nnode, 10) + " double\n")
    o += ["".join(e155(v) for v in p) + "\n" for p in _xyz(x, ndm)]
    # elm_vtk (Vtk.f:167-446): (10i10) node count + 0-based nodes, types (i3)
    o.append("CELLS " + f_i(numel, 10) + f_i(_NUMET[kind] * numel, 10) + "\n")
    for el in range(numel):
        o.append(_rows([nn] + [int(v) - 1 for v in ix[el, :nn]], 10, i10))
    o.append("CELL_TYPES " + f_i(numel, 10) + "\n")
    o += [f_i(ctype, 3) + "\n"] * numel
    # cell_data_vtk (Vtk.f:1972-2008) + cell_prop_vtk, list-directed header
    o.append("CELL_DATA " + f_i(numel, 9) + "\n")
    for name, vals in (("mat", ix[:, nen]), ("elIdG", np.arange(1, numel + 1))):
        o.append(" SCALARS " + _pad(name, 15) + " int  1\n")
        o.append("LOOKUP_TABLE default\n")
        o += [i10(v) + "\n" for v in vals]
    # point_data_vtk + point_prop_vtk (Vtk.f:1237-1498, 2139-2175)
    o.append("POINT_DATA " + f_i(nnode, 9) + "\n")
    # (a,1x,a15,1x,a8,1x,i3): an A field wider than its string right-justifies it
    o.append("SCALARS " + _pad("noIdG", 15) + " " + "int    ".rjust(8) + " " + f_i(1, 3) + "\n")
    o.append("LOOKUP_TABLE default\n")
    o += [i10(k) + "\n" for k in range(1, nnode + 1)]
    o.append("VECTORS " + _pad("desloc", 15) + " double\n")
    o += ["".join(e1553(v) for v in p) + "\n" for p in _xyz(u, min(ndm, ndf))]
    if tx is not None:
        o.append("TENSORS " + _pad("stress", 15) + " double\n")
        o += ["".join(e1553(v) for v in t) + "\n" for t in full_tensor(tx)]
    return "".join(o)


def _legacy_binary(ix, x, u, tx, nen, ndm, ndf) -> bytes:
    """Big-endian stream (access='stream', convert='big_endian'); text lines
    are the reference's buffers with trailing blanks trimmed."""
    kind = element_kind(nen, ndm)
    ctype, nn = _CELL[kind]
    numel, nnode = ix.shape[0], x.shape[0]
    lf = "\n"
    o = []
    txt = lambda s: o.append(s.encode())    # noqa: E731
    txt("# vtk DataFile Version 3.0" + lf)
    txt(_pad("Malha Res", 30))
    txt(lf + "BINARY" + lf)
    txt("DATASET UNSTRUCTURED_GRID" + lf)
    txt("POINTS " + f_i(nnode, 15) + " double" + lf)
    o.append(_xyz(x, ndm).astype(">f8").tobytes())
    txt(lf + "CELLS " + f_i(numel, 15) + f_i(_NUMET[kind] * numel, 15) + lf)
    conn = np.empty((numel, nn + 1), np.int64)
    conn[:, 0] = nn
    conn[:, 1:] = ix[:, :nn].astype(np.int64) - 1
    o.append(conn.astype(">i4").tobytes())
    txt(lf + "CELL_TYPES " + f_i(numel, 15) + lf)
    o.append(np.full(numel, ctype, ">i4").tobytes())
    txt(lf + "CELL_DATA " + f_i(numel, 8) + lf)
    for name, vals in (("mat", ix[:, nen]), ("elIdG", np.arange(1, numel + 1))):
        txt("SCALARS " + _pad(name, 15) + " int" + f_i(1, 8) + lf)
        txt("LOOKUP_TABLE default" + lf)
        o.append(np.asarray(vals).astype(">i4").tobytes())
    txt(lf + " POINT_DATA " + f_i(nnode, 8) + lf)
    txt(" SCALARS " + _pad("noIdG", 15) + " int " + f_i(1, 6) + lf)
    txt(" LOOKUP_TABLE default " + lf)
    o.append(np.arange(1, nnode + 1).astype(">i4").tobytes())
    txt(" VECTORS " + _pad("desloc", 15) + " double" + lf)
    o.append(_xyz(u, min(ndm, ndf)).astype(">f8").tobytes())
    if tx is not None:
        txt(" TENSORS " + _pad("stress", 15) + " double" + lf)
        o.append(full_tensor(tx).astype(">f8").tobytes())
    return b"".join(o)


def _vtu_ascii(ix, x, u, tx, nen, ndm, ndf) -> str:
    kind = element_kind(nen, ndm)
    ctype, nn = _CELL[kind]
    numel, nnode = ix.shape[0], x.shape[0]
    e155 = lambda v: f_e(v, 15, 5)          # noqa: E731
    e1553 = lambda v: f_e(v, 15, 5, 3)      # noqa: E731
    i10 = lambda v: f_i(v, 10)              # noqa: E731
    o = []
    # head_vtu (Vtk.f:1919-1971): the counts keep the leading blanks of (i15);
    # every tag is trim(part1)//trim(part2)..., so attributes are one blank apart
    o += ['<?xml version="1.0"?>\n',
          '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">\n',
          "<UnstructuredGrid>\n",
          '<Piece NumberOfPoints="%s" NumberOfCells="%s">\n' % (f_i(nnode, 15), f_i(numel, 15))]
    # coor_vtu (Vtk.f:112-166)
    o += ["<Points>\n", '<DataArray type="Float64" NumberOfComponents ="3" format="ascii">\n']
    o += ["".join(e155(v) for v in p) + "\n" for p in _xyz(x, ndm)]
    o += ["</DataArray>\n", "</Points>\n"]
    # elm_vtu (Vtk.f:861-1236)
    o += ["<Cells> \n", '<DataArray type="Int32" Name="connectivity" format="ascii">\n']
    for el in range(numel):
        o.append(_rows([int(v) - 1 for v in ix[el, :nn]], 10, i10))
    o.append("</DataArray>\n")
    o.append('<DataArray type="Int64" Name="offsets" format="ascii">\n')
    o += [f_i(nn * (k + 1), 20) + "\n" for k in range(numel)]
    o.append("</DataArray>\n")
    o.append('<DataArray type="UInt8" Name="types" format="ascii">\n')
    o += [f_i(ctype, 3) + "\n"] * numel
    o += ["</DataArray>\n", "</Cells>\n"]
    # cell_data_vtu + cell_prop_vtu (Vtk.f:1771-1844, 2009-2072)
    o.append('<CellData Scalars="scalars">\n')
    for name, vals in (("mat", ix[:, nen]), ("elIdG", np.arange(1, numel + 1))):
        o.append('<DataArray type="Int32" NumberOfComponents="1" Name="%s" format="ascii">\n' % name)
        o += [f_i(v, 9) + "\n" for v in vals]
        o.append("</DataArray>\n")
    o.append("</CellData>\n")
    # point_data_vtu + point_prop_vtu (Vtk.f:1499-1664, 2176-2217)
    o.append("<PointData>\n")
    o.append('<DataArray type="Int32" NumberOfComponents=" 1" Name="noIdG" format="ascii">\n')
    o += [i10(k) + "\n" for k in range(1, nnode + 1)]
    o.append("</DataArray>\n")
    o.append('<DataArray type="Float64" NumberOfComponents="%s" Name="desloc" format="ascii">\n'
             % f_i(ndf, 2))
    o += [_rows(p, 10, e1553) for p in np.asarray(u)[:, :ndf]]
    o.append("</DataArray>\n")
    if tx is not None:
        o.append('<DataArray type="Float64" NumberOfComponents="9" Name="stress" format="ascii">\n')
        o += [_rows(t, 10, e1553) for t in full_tensor(tx)]
        o.append("</DataArray>\n")
    o += ["</PointData>\n", "</Piece>\n", "</UnstructuredGrid>\n", "</VTKFile>\n"]
    return "".join(o)


def read_result_header(path: str) -> str:
    """First line of a result file (for tools that sniff the format)."""
    with open(path, "rb") as fh:
        return fh.readline().decode(errors="replace").rstrip("\n")


__all__ = ["write_result", "result_name", "full_tensor", "f_e", "f_i", "element_kind",
           "read_result_header"]
