"""Pins of the oracle's pointwise nonlinear products (modules/nonlinear.F90
:71-146, 266-272, 335-341, 420-430).

The products are derived here from vector geometry, not from the
reference's expressions:
  * advection u x (curl u) by the determinant form of the cross product in
    the right-handed (r, theta, phi) frame;
  * the Coriolis term -(2 Pm/E) z x u with z = cos(theta) r - sin(theta)
    theta, theta from the Gauss node x = cos(theta);
  * the Lorentz term (2 Pm/E) (curl B) x B and the induction E = u x B;
  * the advection of codensity and composition -(Pr/Pm) u.grad C,
    -(Sc/Pm) u.grad xi.
They are applied to the physical fields the oracle's transform produced
and projected with the oracle's drop-in analysis (pinned by
test_oracle_sht.py).  The projections compared are the ones that are pure
definitions: N_T = -t / sqrt(l(l+1)) of the velocity (variables.F90:762-780),
N_Pol = -t / sqrt(l(l+1)) of E (nonlinear.F90:441-443), and the scalar
analyses.  A dropped term, a wrong sign or a swapped component in any
product changes these.
"""
import numpy as np
import pytest

import oracle_lib as O
from lsd import config_small, seeded_state
from lsd.initial import harmonics


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


@pytest.fixture(scope="module")
def evaluated():
    cfg = config_small(comp=1, mag=1, cod=1)
    o = O.Oracle(cfg)
    o.set_state(seeded_state(cfg, 41, 1e-2))
    o.transform()
    phys = [o.get_phys(k) for k in range(18)]
    o.nonlinear()
    return cfg, o, phys


def _ll(cfg):
    l, _ = harmonics(cfg.L, cfg.M, cfg.Mp)
    ll = np.sqrt(l * (l + 1.0))
    rt = np.zeros_like(ll)
    rt[l > 0] = 1.0 / ll[l > 0]
    return l, rt


def _frame(o, cfg):
    x = o.table(0)                         # cos(theta_j), Gauss nodes
    s = np.sqrt(1.0 - x * x)
    shape = (1, 1, len(x))
    return x.reshape(shape), s.reshape(shape)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_velocity_products_toroidal_projection(evaluated):
    cfg, o, P = evaluated
    cosv, sinv = _frame(o, cfg)
    u, w, B, J = P[0:3], P[3:6], P[6:9], P[9:12]
    z = (cosv + 0.0 * u[0], -sinv + 0.0 * u[0], 0.0 * u[0])
    cor = 2.0 * cfg.Pm / cfg.E
    adv = cross(u, w)
    zxu = cross(z, u)
    lor = cross(J, B)
    n = [adv[c] - cor * zxu[c] + cor * lor[c] for c in range(3)]
    _, _, t = o.rtp2qst(*n)
    l, rt = _ll(cfg)
    nt_re = (-t[0] * rt[None, :]).T
    nt_im = (-t[1] * rt[None, :]).T
    got = o.get_coll(6)
    mask = l > 0
    assert _rel(got[0][mask], nt_re[mask]) < 1e-12
    assert _rel(got[1][mask], nt_im[mask]) < 1e-12


def test_induction_poloidal_projection(evaluated):
    cfg, o, P = evaluated
    u, B = P[0:3], P[6:9]
    e = cross(u, B)
    _, _, t = o.rtp2qst(*e)
    l, rt = _ll(cfg)
    got = o.get_coll(11)  # N of the poloidal magnetic field
    mask = l > 0
    assert _rel(got[0][mask], (-t[0] * rt[None, :]).T[mask]) < 1e-12
    assert _rel(got[1][mask], (-t[1] * rt[None, :]).T[mask]) < 1e-12


@pytest.mark.parametrize("which", ["cod", "comp"])
def test_scalar_advection(evaluated, which):
    cfg, o, P = evaluated
    u = P[0:3]
    g = P[12:15] if which == "cod" else P[15:18]
    k = cfg.Pr / cfg.Pm if which == "cod" else cfg.Sc / cfg.Pm
    prod = -k * (u[0] * g[0] + u[1] * g[1] + u[2] * g[2])
    sre, sim = o.phys2spec(prod)
    got = o.get_coll(8 if which == "cod" else 9)
    assert _rel(got[0], sre.T) < 1e-12
    assert _rel(got[1], sim.T) < 1e-12
