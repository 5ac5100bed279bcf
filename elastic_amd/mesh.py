"""Meshes in the reference's array conventions, and the synthetic workloads.

Array layout follows the reference driver (Rdata.f:1-743, Main.f:348-534):

* ``ix``    int32 (numel, nen+1): element nodes (1-based) then the material
            number; the memory order equals Fortran ``ix(nen+1,numel)``.
* ``x``     float64 (nnode, ndm): coordinates, Fortran ``x(ndm,nnode)``.
* ``id``    int32 (nnode, ndf): 1 = restrained, 0 = free (``constraindisp``).
* ``ie``    int32 (numat,): element type per material (2 tria3 plane strain,
            3 tria3 plane stress, 4 quad4 plane strain, 5 quad4 plane stress,
            6 tet4, 7 hex8; Elmlib.f:1-166).
* ``e``     float64 (numat, 15): material constants (E, nu, density|thickness).
* ``eload`` int32 (numel, 7): per face/side load number (``elmtloads``).
* ``nload`` int32 (nnode, ndf): nodal load numbers (``nodalloads``).
* ``f``     float64 (nnode, ndf): nodal forces / prescribed values.
* ``loads`` list of (type, [values]) (``loads``; type 40 = traction vector).

The synthetic generators are the workloads the benchmarks and parity tests run.
Node numbering is ``1 + i + (N+1)(j + (N+1)k)`` with i fastest; element
local node orders follow the shape functions of the reference (hex8:
``sfhexa8_m`` node 1 = (+,+,+), 2 = (-,+,+), 3 = (-,-,+), 4 = (+,-,+), 5..8
the same at t = -1, Elmlib.f:1151-1184; quad4: counter-clockwise from the
lower-left corner as in the reference's own quad deck).  Pseudo-random values
(jitter, materials, tractions) come from a counter-based splitmix64 stream so
that every value depends only on (seed, counter).
"""
from __future__ import annotations

import dataclasses
import numpy as np

PROP = 15  # termprop.fi: material constants per material

IEL_TRIA3_STRAIN, IEL_TRIA3_STRESS, IEL_QUAD4_STRAIN, IEL_QUAD4_STRESS = 2, 3, 4, 5
IEL_TET4, IEL_HEX8 = 6, 7


@dataclasses.dataclass
class Mesh:
    nnode: int
    numel: int
    nen: int
    ndf: int
    ndm: int
    ix: np.ndarray
    x: np.ndarray
    id: np.ndarray
    ie: np.ndarray
    e: np.ndarray
    eload: np.ndarray
    nload: np.ndarray
    f: np.ndarray
    loads: list
    name: str = ""

    @property
    def numat(self) -> int:
        return int(self.ie.shape[0])

    @property
    def nenv(self) -> int:
        return self.nen

    @property
    def nst(self) -> int:
        return self.nen * self.ndf

    def element_type(self) -> int:
        types = np.unique(self.ie[self.ix[:, -1] - 1])
        if types.size != 1:
            raise ValueError("mixed element types in one mesh are not supported")
        return int(types[0])


# --------------------------------------------------------------------------- RNG
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform01(seed: int, counter) -> np.ndarray:
    """Counter-based uniform [0,1) doubles: splitmix64 finaliser of
    seed + (counter+1) * golden, top 53 bits."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (c + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ------------------------------------------------------------------- 3-D cubes
def _cube_nodes(N: int, jitter: float, seed: int) -> np.ndarray:
    n1 = N + 1
    k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    h = 1.0 / N
    x = np.stack([i / N, j / N, k / N], axis=1).astype(np.float64)
    if jitter > 0.0:
        interior = (i > 0) & (i < N) & (j > 0) & (j < N) & (k > 0) & (k < N)
        node0 = np.nonzero(interior)[0].astype(np.uint64)
        for d in range(3):
            u = uniform01(seed, node0 * np.uint64(3) + np.uint64(d))
            x[interior, d] += (2.0 * u - 1.0) * jitter * h
    return x


def _hex_cells(N: int) -> np.ndarray:
    """(N^3, 8) 1-based node numbers in sfhexa8_m order, cell (i,j,k) -> row
    i + N(j + N k)."""
    n1 = N + 1
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def nid(a, b, c):
        return 1 + a + n1 * (b + n1 * c)

    return np.stack(
        [nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1), nid(i, j, k + 1), nid(i + 1, j, k + 1),
         nid(i + 1, j + 1, k), nid(i, j + 1, k), nid(i, j, k), nid(i + 1, j, k)],
        axis=1,
    ).astype(np.int32)


def hex8_cube(N: int, jitter: float = 0.0, seed: int = 20261202, E: float = 1.0,
              nu: float = 0.3, traction=(0.0, 0.0, -1e-3)) -> Mesh:
    """Unit cube of N^3 trilinear hexahedra: z=0 face clamped, uniform traction
    on the z=1 face (face 1 of the top layer, Elmlib02.f:1655-1660), one
    material of type 7.  Interior nodes jittered by U[-jitter, jitter]*h."""
    nnode, numel = (N + 1) ** 3, N ** 3
    x = _cube_nodes(N, jitter, seed)
    ix = np.empty((numel, 9), np.int32)
    ix[:, :8] = _hex_cells(N)
    ix[:, 8] = 1
    idf = np.zeros((nnode, 3), np.int32)
    idf[x[:, 2] == 0.0] = 1
    eload = np.zeros((numel, 7), np.int32)
    eload[numel - N * N:, 0] = 1  # k = N-1 layer, face 1 (t = +1)
    e = np.zeros((1, PROP))
    e[0, 0], e[0, 1] = E, nu
    return Mesh(nnode, numel, 8, 3, 3, ix, x, idf, np.array([IEL_HEX8], np.int32), e, eload,
                np.zeros((nnode, 3), np.int32), np.zeros((nnode, 3)),
                [(40, list(map(float, traction)))], name=f"hex8_{N}")


# hex-local node quadruples of the six-tetrahedron split (Pdata.f:499-515);
# conforming across cube faces and positively oriented for jtetra4.
_TET_SPLIT = np.array([[2, 5, 6, 7], [2, 5, 7, 3], [2, 5, 3, 1], [3, 5, 7, 8], [3, 1, 8, 4],
                       [3, 1, 5, 8]]) - 1
_TET_FACES = np.array([[2, 3, 4], [1, 4, 3], [1, 2, 4], [1, 3, 2]]) - 1  # Elmlib02.f:1296


def tet_jacobian_det(x4: np.ndarray) -> np.ndarray:
    """jtetra4 determinant (Elmlib.f:2661-2681) for (n,4,3) node coordinates."""
    a = x4[:, 0] - x4[:, 3]
    b = x4[:, 1] - x4[:, 3]
    c = x4[:, 2] - x4[:, 3]
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def tet4_cube(N: int, jitter: float = 0.0, seed: int = 20261203, logE=(-1.0, 1.0),
              nu: float = 0.45, traction=(0.0, 0.0, -1e-3), per_element_E: bool = True) -> Mesh:
    """N^3 cubes, each split into six tetrahedra; z=0 clamped, traction on z=1.
    With per_element_E every element has its own material (type 6) with
    E = 10**U[logE]."""
    nnode = (N + 1) ** 3
    x = _cube_nodes(N, jitter, seed)
    hexes = _hex_cells(N)
    tets = hexes[:, _TET_SPLIT].reshape(-1, 4)
    numel = tets.shape[0]
    det = tet_jacobian_det(x[tets - 1])
    # resample the jitter of the nodes of tetrahedra whose jtetra4 determinant
    # is not above 1e-14 (Elmlib.f:2685), round by round, from counters no
    # earlier draw used; meshes without such tetrahedra are left unchanged
    n1, h = N + 1, 1.0 / N
    rnd = 0
    while np.any(det <= 1e-14):
        rnd += 1
        if rnd > 64 or jitter <= 0.0:
            raise ValueError("tetrahedron with non-positive Jacobian after resampling")
        nodes = np.unique(tets[det <= 1e-14].ravel() - 1)
        i, j, k = nodes % n1, (nodes // n1) % n1, nodes // (n1 * n1)
        inner = (i > 0) & (i < N) & (j > 0) & (j < N) & (k > 0) & (k < N)
        nodes, i, j, k = nodes[inner], i[inner], j[inner], k[inner]
        base = np.stack([i * h, j * h, k * h], axis=1)
        for d in range(3):
            ctr = (np.uint64(rnd) * np.uint64(nnode) + nodes.astype(np.uint64)) * np.uint64(3) + np.uint64(d)
            u = uniform01(seed + 7919, ctr)
            x[nodes, d] = base[:, d] + (2.0 * u - 1.0) * jitter * h
        det = tet_jacobian_det(x[tets - 1])
    ix = np.empty((numel, 5), np.int32)
    ix[:, :4] = tets
    top = x[:, 2] == 1.0
    eload = np.zeros((numel, 7), np.int32)
    for f in range(4):
        on = top[tets[:, _TET_FACES[f]] - 1].all(axis=1)
        eload[on, f] = 1
    if per_element_E:
        u = uniform01(seed + 1, np.arange(numel, dtype=np.uint64))
        E = 10.0 ** (logE[0] + (logE[1] - logE[0]) * u)
        e = np.zeros((numel, PROP))
        e[:, 0], e[:, 1] = E, nu
        ix[:, 4] = np.arange(1, numel + 1, dtype=np.int32)
        ie = np.full(numel, IEL_TET4, np.int32)
    else:
        e = np.zeros((1, PROP))
        e[0, 0], e[0, 1] = 10.0 ** logE[0], nu
        ix[:, 4] = 1
        ie = np.array([IEL_TET4], np.int32)
    idf = np.zeros((nnode, 3), np.int32)
    idf[x[:, 2] == 0.0] = 1
    return Mesh(nnode, numel, 4, 3, 3, ix, x, idf, ie, e, eload, np.zeros((nnode, 3), np.int32),
                np.zeros((nnode, 3)), [(40, list(map(float, traction)))], name=f"tet4_{N}")


def hex8_block_pull(N: int, stretch: float, E: float = 1.0, nu: float = 0.3):
    """Unit cube of N^3 hex8 held between two rigid grips: z=0 clamped, z=1
    restrained in all directions and moved by (0, 0, stretch) at load factor
    1 (every node fully free or fully restrained, so the node-blocked
    layout applies).  Returns (mesh, ubar(nnode, 3))."""
    m = hex8_cube(N, E=E, nu=nu)
    top = m.x[:, 2] == 1.0
    m.id[top] = 1
    m.eload[:] = 0
    m.loads = []
    ubar = np.zeros((m.nnode, 3))
    ubar[top, 2] = stretch
    m.name = f"hex8_pull_{N}"
    return m, ubar


# ------------------------------------------------------------------ 2-D squares
def _square(n: int, tria: bool, E: float, nu: float, traction, edge: str, plane_stress: bool,
            thickness: float = 0.0) -> Mesh:
    n1 = n + 1
    j, i = np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    x = np.stack([i / n, j / n], axis=1).astype(np.float64)
    nnode = n1 * n1
    cj, ci = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ci, cj = ci.ravel(), cj.ravel()

    def nid(a, b):
        return 1 + a + n1 * b

    p1, p2, p3, p4 = nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)
    if tria:
        # quad split along the (i,j)-(i+1,j+1) diagonal: A = (1,2,3), B = (1,3,4)
        a = np.stack([p1, p2, p3], 1)
        b = np.stack([p1, p3, p4], 1)
        conn = np.stack([a, b], 1).reshape(-1, 3)
        nen, iel = 3, (IEL_TRIA3_STRESS if plane_stress else IEL_TRIA3_STRAIN)
    else:
        conn = np.stack([p1, p2, p3, p4], 1)
        nen, iel = 4, (IEL_QUAD4_STRESS if plane_stress else IEL_QUAD4_STRAIN)
    numel = conn.shape[0]
    ix = np.empty((numel, nen + 1), np.int32)
    ix[:, :nen] = conn
    ix[:, nen] = 1
    eload = np.zeros((numel, 7), np.int32)
    if edge == "right":
        sel = ci == n - 1
        if tria:
            eload[np.nonzero(sel)[0] * 2, 1] = 1  # triangle A, side 2 = nodes (2,3)
        else:
            eload[sel, 1] = 1  # quad side 2 = nodes (2,3)
    elif edge == "top":
        sel = cj == n - 1
        if tria:
            eload[np.nonzero(sel)[0] * 2 + 1, 1] = 1  # triangle B, side 2 = nodes (3,4)
        else:
            eload[sel, 2] = 1  # quad side 3 = nodes (3,4)
    else:
        raise ValueError(edge)
    idf = np.zeros((nnode, 2), np.int32)
    idf[x[:, 0] == 0.0] = 1
    e = np.zeros((1, PROP))
    e[0, 0], e[0, 1], e[0, 2] = E, nu, thickness
    return Mesh(nnode, numel, nen, 2, 2, ix, x, idf, np.array([iel], np.int32), e, eload,
                np.zeros((nnode, 2), np.int32), np.zeros((nnode, 2)),
                [(40, list(map(float, traction)))], name=("tria3_%d" if tria else "quad4_%d") % n)


def quad4_square(n: int, E: float = 1.0, nu: float = 0.3, traction=(0.0, -1.0),
                 edge: str = "right", plane_stress: bool = False, thickness: float = 0.0) -> Mesh:
    """Unit square of n x n quad4 (plane strain type 4 unless plane_stress):
    x=0 edge clamped, traction on the given edge."""
    return _square(n, False, E, nu, traction, edge, plane_stress, thickness)


def tria3_square(n: int, E: float = 1.0, nu: float = 0.3, traction=(0.0, -1.0),
                 edge: str = "right", plane_stress: bool = False, thickness: float = 0.0) -> Mesh:
    """The quad4 square with each quad split along its (i,j)-(i+1,j+1) diagonal."""
    return _square(n, True, E, nu, traction, edge, plane_stress, thickness)


def tria3_fan(nring: int, E: float = 1.0, nu: float = 0.3, force=(0.3, -1.0)) -> Mesh:
    """A high-valence mesh: one hub node joined to a ring of `nring` nodes by
    `nring` tria3 elements (plane strain).  Three consecutive ring nodes are
    clamped and the ring node opposite them carries a nodal force.  RCM puts
    the hub near the end, so its CSRC row holds about 2 x nring lower
    entries (nring lower node blocks)."""
    t = 2.0 * np.pi * np.arange(nring) / nring
    x = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(t), np.sin(t)], 1)]).astype(np.float64)
    k = np.arange(nring)
    conn = np.stack([np.ones(nring, np.int64), 2 + k, 2 + (k + 1) % nring], 1)
    ix = np.empty((nring, 4), np.int32)
    ix[:, :3] = conn
    ix[:, 3] = 1
    nnode = nring + 1
    idf = np.zeros((nnode, 2), np.int32)
    idf[1:4] = 1
    f = np.zeros((nnode, 2))
    f[1 + nring // 2] = force
    e = np.zeros((1, PROP))
    e[0, 0], e[0, 1] = E, nu
    return Mesh(nnode, nring, 3, 2, 2, ix, x, idf, np.array([IEL_TRIA3_STRAIN], np.int32), e,
                np.zeros((nring, 7), np.int32), np.zeros((nnode, 2), np.int32), f, [],
                name="tria3_fan_%d" % nring)


# ------------------------------------------------------------------ workloads
BASE_SEED = 20261202


def config_mesh(cfg: int, scale: int | None = None) -> Mesh:
    """The five benchmark workloads (sizes overridable for parity tests):
    1 hex8 20^3; 2 hex8 100^3 jittered 0.15h; 3 tet4 80^3 x 6 jittered 0.2h with
    E = 10^U[-1,1], nu = 0.45; 5 hex8 60^3 (plasticity mesh)."""
    if cfg == 1:
        return hex8_cube(scale or 20)
    if cfg == 2:
        return hex8_cube(scale or 100, jitter=0.15, seed=BASE_SEED + 2)
    if cfg == 3:
        return tet4_cube(scale or 80, jitter=0.2, seed=BASE_SEED + 3)
    if cfg == 5:
        return hex8_cube(scale or 60)
    raise ValueError("config %d has no single mesh (4 is a request stream)" % cfg)


def serving_request(k: int, seed: int = BASE_SEED + 4, nmin: int = 32, nmax: int = 512) -> Mesh:
    """Request k of the plane-strain request stream: even k quad4, odd k tria3;
    side n ~ U{nmin..nmax}; E = 10^U[log10 .5, log10 2]; nu ~ U[.2,.4];
    x=0 edge clamped; traction with components U[-1,1] on the x=1 edge."""
    u = uniform01(seed, np.arange(5 * k, 5 * k + 5, dtype=np.uint64))
    n = nmin + min(int(u[0] * (nmax - nmin + 1)), nmax - nmin)
    E = 10.0 ** (np.log10(0.5) + (np.log10(2.0) - np.log10(0.5)) * u[1])
    nu = 0.2 + 0.2 * u[2]
    tr = (2.0 * u[3] - 1.0, 2.0 * u[4] - 1.0)
    gen = tria3_square if (k % 2) else quad4_square
    m = gen(n, E=float(E), nu=float(nu), traction=tr, edge="right")
    m.name = "%s_req%d" % (m.name, k)
    return m
