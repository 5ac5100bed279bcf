"""Pinning of the J2 plasticity oracle (oracle/j2.py) - it has no counterpart
in the reference's called code, so it is pinned by what the mathematics
fixes:

* uniaxial stress under monotonic loading matches the closed form
  sigma = sy + E H / (E + H) (eps - sy / E) and alpha = (sigma - sy) / H;
* the consistent tangent equals the finite-difference derivative of the
  return-mapped stress (rel. 1e-6), in the plastic and the elastic regime;
* the element tangent equals the finite-difference derivative of the element
  internal force (rel. 1e-5) on a distorted, yielding hex8;
* in the elastic regime the element stiffness equals the reference's hex8
  stiffness (elmt07_mec, through the oracle's bit-exact restatement) to 1e-12.
"""
import numpy as np
import pytest

from oracle import j2

E, NU, SY, H = 1.0, 0.3, 1e-3, 0.05


def uniaxial_path(emax, steps):
    ep = np.zeros(6)
    al = 0.0
    el = 0.0
    out = []
    for k in range(1, steps + 1):
        ex = emax * k / steps
        for _ in range(50):  # lateral strain such that sigma_yy = sigma_zz = 0
            eps = np.array([ex, el, el, 0, 0, 0])
            sig, C, _, _, _ = j2.return_map(eps, ep, al, E, NU, SY, H)
            d = C[1, 1] + C[1, 2]
            step = -sig[1] / d
            el += step
            if abs(step) < 1e-16:
                break
        eps = np.array([ex, el, el, 0, 0, 0])
        sig, C, ep, al, pl = j2.return_map(eps, ep, al, E, NU, SY, H)
        out.append((ex, sig, al, pl))
    return out


def test_uniaxial_closed_form():
    for ex, sig, al, pl in uniaxial_path(5e-3, 40):
        assert abs(sig[1]) < 1e-13 and abs(sig[2]) < 1e-13
        if ex <= SY / E:
            assert not pl and abs(sig[0] - E * ex) < 1e-15
        else:
            s_cf = SY + E * H / (E + H) * (ex - SY / E)
            assert abs(sig[0] - s_cf) < 1e-12 * s_cf
            assert abs(al - (sig[0] - SY) / H) < 1e-10


@pytest.mark.parametrize("scale", [0.2, 3.0])
def test_tangent_matches_finite_differences(scale):
    rng = np.random.default_rng(7)
    eps = scale * SY * rng.standard_normal(6)
    ep_n = 0.1 * SY * rng.standard_normal(6)
    ep_n[:3] -= ep_n[:3].mean()
    al_n = 0.5 * SY
    sig, C, _, _, pl = j2.return_map(eps, ep_n, al_n, E, NU, SY, H)
    assert pl == (scale > 1)
    h = 1e-8
    Cfd = np.zeros((6, 6))
    for j in range(6):
        d = np.zeros(6)
        d[j] = h
        sp = j2.return_map(eps + d, ep_n, al_n, E, NU, SY, H)[0]
        sm = j2.return_map(eps - d, ep_n, al_n, E, NU, SY, H)[0]
        Cfd[:, j] = (sp - sm) / (2 * h)
    assert np.allclose(C, C.T, atol=1e-12)
    assert np.abs(C - Cfd).max() <= 1e-6 * np.abs(C).max()


def distorted_hex():
    xl = np.array([[1, 1, 1], [0, 1, 1], [0, 0, 1], [1, 0, 1],
                   [1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0]], float)
    rng = np.random.default_rng(3)
    return xl + 0.1 * rng.uniform(-1, 1, xl.shape)


def test_element_tangent_matches_finite_differences():
    xl = distorted_hex()
    rng = np.random.default_rng(5)
    ue = 4 * SY * rng.standard_normal(24)
    st_ep, st_al = np.zeros((8, 6)), np.zeros(8)
    f, K, ep, al = j2.element(xl, ue, st_ep, st_al, E, NU, SY, H)
    assert (al > 0).any()  # yielding
    h = 1e-9
    Kfd = np.zeros((24, 24))
    for j in range(24):
        d = np.zeros(24)
        d[j] = h
        Kfd[:, j] = (j2.element(xl, ue + d, st_ep, st_al, E, NU, SY, H)[0] -
                     j2.element(xl, ue - d, st_ep, st_al, E, NU, SY, H)[0]) / (2 * h)
    assert np.abs(K - Kfd).max() <= 1e-5 * np.abs(K).max()


def test_elastic_limit_is_the_reference_hex8():
    from oracle import binding as O

    xl = distorted_hex()
    ue = np.zeros(24)
    f, K, _, _ = j2.element(xl, ue, np.zeros((8, 6)), np.zeros(8), E, NU, 1e30, H)
    e = np.zeros(15)
    e[0], e[1] = E, NU
    rc, s, _ = O.element(7, 2, e, np.zeros(8, np.int32), xl, ue, 3, 24)
    assert rc == 0
    assert np.abs(K - s).max() <= 1e-12 * np.abs(s).max()
