"""CPU restatement of J2 (von Mises) plasticity with linear isotropic
hardening, the checker of the GPU plastic assembly (test infrastructure:
only tests/ may import it).

The reference has no called plasticity path (Plasticity01.f is not reached
from Main.f; plasticst3d, Plasticity01.f:277-402, integrates with explicit
sub-steps).  This is the textbook closed-form radial return with its
consistent tangent (Simo & Hughes, Computational Inelasticity, 1998, Box 3.2
and eq. 3.3.12), with the reference's hardening law: current yield stress =
sy + hl * (equivalent plastic strain) (Plasticity01.f:322, 389).  Voigt order
xx, yy, zz, xy, yz, zx with engineering shear strains.  It is pinned by the
uniaxial closed form and by finite differences (tests/test_j2_oracle.py), not
by the reference: parity unpinned.
"""
from __future__ import annotations

import numpy as np

SQ23 = np.sqrt(2.0 / 3.0)
# deviatoric projector on Voigt stresses / tensor strains
I6 = np.eye(6)
ONE = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
PDEV = np.diag([1.0, 1.0, 1.0, 0.5, 0.5, 0.5]) - np.outer(ONE, ONE) / 3.0


def elastic_d(E: float, nu: float) -> np.ndarray:
    """Isotropic D for Voigt stresses and engineering shear strains."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = lam * np.outer(ONE, ONE) + 2 * mu * np.diag([1.0, 1.0, 1.0, 0.5, 0.5, 0.5])
    return D


def _tensor_strain(eps: np.ndarray) -> np.ndarray:
    e = eps.copy()
    e[3:] *= 0.5
    return e


def return_map(eps, ep_n, alpha_n, E, nu, sy, H):
    """Stress, tangent and the new internal variables at one point.

    eps: total strain (Voigt, engineering shear); ep_n: plastic strain
    (same convention); alpha_n: equivalent plastic strain.
    Returns (sigma[6], C[6,6], ep[6], alpha, plastic?)."""
    mu = E / (2 * (1 + nu))
    kappa = E / (3 * (1 - 2 * nu))
    D = elastic_d(E, nu)
    sig_tr = D @ (eps - ep_n)
    p = sig_tr[:3].sum() / 3.0
    s = sig_tr - p * ONE
    # ||s|| of the symmetric tensor (shear components count twice)
    snorm = np.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2))
    f = snorm - SQ23 * (sy + H * alpha_n)
    if f <= 0.0:
        return sig_tr, D, ep_n.copy(), alpha_n, False
    dg = f / (2 * mu + 2.0 * H / 3.0)
    n = s / snorm
    sig = sig_tr - 2 * mu * dg * n
    ep = ep_n + dg * np.concatenate([n[:3], 2 * n[3:]])
    alpha = alpha_n + SQ23 * dg
    theta = 1 - 2 * mu * dg / snorm
    thetab = 1.0 / (1 + H / (3 * mu)) - (1 - theta)
    # C = kappa 1x1 + 2 mu theta Pdev - 2 mu thetab n x n (engineering shear columns)
    C = kappa * np.outer(ONE, ONE) + 2 * mu * theta * PDEV - 2 * mu * thetab * np.outer(n, n)
    return sig, C, ep, alpha, True


# ------------------------------------------------------------ hex8 element
GP = np.array([-0.577350269189626, 0.577350269189626])
# natural coordinates of the 8 nodes (Elmlib.f:1177-1184 ordering)
RN = np.array([1, -1, -1, 1, 1, -1, -1, 1], float)
SN = np.array([1, 1, -1, -1, 1, 1, -1, -1], float)
TN = np.array([1, 1, 1, 1, -1, -1, -1, -1], float)


def gauss_points():
    """The 8 points in the order t (outer), s, r (inner), weight 1."""
    return [(GP[lx], GP[ly], GP[lz]) for lz in range(2) for ly in range(2) for lx in range(2)]


def b_matrix(xl: np.ndarray, r: float, s: float, t: float):
    """B (6 x 24, engineering shear) and det J of a hex8 at (r, s, t)."""
    dr = 0.125 * RN * (1 + s * SN) * (1 + t * TN)
    ds = 0.125 * SN * (1 + r * RN) * (1 + t * TN)
    dt = 0.125 * TN * (1 + r * RN) * (1 + s * SN)
    dN = np.stack([dr, ds, dt])           # 3 x 8 natural derivatives
    J = dN @ xl                            # 3 x 3, xl is 8 x 3
    det = np.linalg.det(J)
    dX = np.linalg.solve(J, dN)            # 3 x 8 global derivatives
    B = np.zeros((6, 24))
    for a in range(8):
        hx, hy, hz = dX[:, a]
        c = 3 * a
        B[0, c] = hx
        B[1, c + 1] = hy
        B[2, c + 2] = hz
        B[3, c], B[3, c + 1] = hy, hx
        B[4, c + 1], B[4, c + 2] = hz, hy
        B[5, c], B[5, c + 2] = hz, hx
    return B, det


def element(xl, ue, state_ep, state_al, E, nu, sy, H):
    """Internal force, tangent and trial state of one hex8 element.
    state_ep (8, 6), state_al (8,) at the 8 Gauss points."""
    f = np.zeros(24)
    K = np.zeros((24, 24))
    ep_new = np.zeros((8, 6))
    al_new = np.zeros(8)
    for g, (r, s, t) in enumerate(gauss_points()):
        B, det = b_matrix(xl, r, s, t)
        eps = B @ ue
        sig, C, ep, al, _ = return_map(eps, state_ep[g], state_al[g], E, nu, sy, H)
        f += B.T @ sig * det
        K += B.T @ C @ B * det
        ep_new[g], al_new[g] = ep, al
    return f, K, ep_new, al_new
