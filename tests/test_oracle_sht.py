"""Pins of the oracle's spherical-harmonic transforms (modules/transform.fftw3.F90
tra_spec2phys, tra_phys2spec, tra_grad, tra_qst2rtp, tra_rtp2qst) against closed
forms and identities."""
import numpy as np
import pytest

import oracle_lib as O
from lsd import config_small
from lsd.initial import harmonics


def _grid(o):
    x, _ = O.gauwts(o.Th)
    theta = np.arccos(x)
    phi = 2 * np.pi * np.arange(o.Ph) / o.Ph
    return theta, phi


def _schmidt(l, m, x):
    pa, _ = O.schnla(x, l)
    return pa[l * (l + 1) // 2 + m]


def _spec_random(cfg, seed, lmin=0):
    H = cfg.sizes()[0]
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    rng = np.random.default_rng(seed)
    re = rng.standard_normal((cfg.N, H))
    im = rng.standard_normal((cfg.N, H))
    im[:, m == 0] = 0.0
    re[:, l < lmin] = 0.0
    im[:, l < lmin] = 0.0
    return re, im


@pytest.fixture(scope="module")
def small():
    return O.Oracle(config_small(), threads=2)


def test_single_harmonic_closed_form(small):
    # f = 2 Re(c P_l^m e^{i m phi}) for m > 0 (FFTW c2r convention, Appendix A)
    o = small
    cfg = o.cfg
    l_, m_ = harmonics(cfg.L, cfg.M, cfg.Mp)
    theta, phi = _grid(o)
    for (l, m, c) in ((3, 0, 0.7 + 0.0j), (5, 2, 0.3 - 1.1j), (11, 11, -0.4 + 0.25j), (7, 6, 1.0 + 2.0j)):
        h = int(np.where((l_ == l) & (m_ == m))[0][0])
        re = np.zeros((cfg.N, o.H))
        im = np.zeros((cfg.N, o.H))
        re[:, h] = c.real
        im[:, h] = c.imag
        p = o.spec2phys(re, im)[0]  # shell 0: [k, j]
        P = np.array([_schmidt(l, m, np.cos(t)) for t in theta])
        if m == 0:
            exact = c.real * P[None, :] * np.ones((o.Ph, 1))
        else:
            exact = 2 * np.real(c * np.exp(1j * m * phi))[:, None] * P[None, :]
        assert np.abs(p - exact).max() < 1e-13 * max(1.0, np.abs(exact).max())


def test_scalar_round_trip(small):
    o = small
    re, im = _spec_random(o.cfg, 1)
    p = o.spec2phys(re, im)
    re2, im2 = o.phys2spec(p)
    assert np.abs(re2 - re).max() < 1e-12 * np.abs(re).max()
    assert np.abs(im2 - im).max() < 1e-12 * np.abs(im).max()


def test_vector_round_trip(small):
    o = small
    q = _spec_random(o.cfg, 2)
    s = _spec_random(o.cfg, 3, lmin=1)
    t = _spec_random(o.cfg, 4, lmin=1)
    r, th, ph = o.qst2rtp(q, s, t)
    q2, s2, t2 = o.rtp2qst(r, th, ph)
    for a, b in ((q, q2), (s, s2), (t, t2)):
        assert np.abs(a[0] - b[0]).max() < 1e-12 * np.abs(a[0]).max()
        assert np.abs(a[1] - b[1]).max() < 1e-12 * max(1.0, np.abs(a[1]).max())


def test_toroidal_l1_is_solid_body_rotation(small):
    # t = -sqrt(2) T for T = Y_1^0 gives u_phi = sin(theta), u_r = u_theta = 0
    o = small
    cfg = o.cfg
    l_, m_ = harmonics(cfg.L, cfg.M, cfg.Mp)
    h = int(np.where((l_ == 1) & (m_ == 0))[0][0])
    z = np.zeros((cfg.N, o.H))
    t = (z.copy(), z.copy())
    t[0][:, h] = -np.sqrt(2.0)
    r, th, ph = o.qst2rtp((z, z), (z, z), t)
    theta, _ = _grid(o)
    assert np.abs(r).max() < 1e-14 and np.abs(th).max() < 1e-14
    assert np.abs(ph[0] - np.sin(theta)[None, :]).max() < 1e-13


def test_grad_matches_spheroidal_synthesis(small):
    # tra_grad(c) (tables P_, sin1mP, factor 1/r) equals the theta/phi part of
    # tra_qst2rtp with s = sqrt(l(l+1)) c / r (tables llP_, llsin1mP)
    o = small
    cfg = o.cfg
    re, im = _spec_random(cfg, 5)
    l_, _ = harmonics(cfg.L, cfg.M, cfg.Mp)
    rinv = 1.0 / o.mesh(0)
    f = np.sqrt(l_ * (l_ + 1.0))[None, :] * rinv[:, None]
    z = np.zeros_like(re)
    gt, gp = o.grad(re, im)
    _, th, ph = o.qst2rtp((z, z), (re * f, im * f), (z, z))
    assert np.abs(gt - th).max() < 1e-12 * np.abs(th).max()
    assert np.abs(gp - ph).max() < 1e-12 * np.abs(ph).max()
