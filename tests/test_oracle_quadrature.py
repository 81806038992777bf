"""Pins of the oracle's Gauss-Legendre nodes, Schmidt functions and Legendre
tables (auxiliary/gauwts.f, auxiliary/schnla.f, modules/legendre.F90)."""
import os

import numpy as np
import pytest

import oracle_lib as O
from lsd import config1, config_small

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_aux_vectors.npz"))


@pytest.mark.parametrize("n", [18, 48, 256, 384])
def test_gauss_bitexact_vs_reference_routine(n):
    # outputs of the reference's own GAUWTS, compiled from auxiliary/gauwts.f
    x, w = O.gauwts(n)
    assert np.array_equal(x, GOLD["gauss_x_%d" % n])
    assert np.array_equal(w, GOLD["gauss_w_%d" % n])


@pytest.mark.parametrize("n", [18, 48, 256, 384])
def test_gauss_vs_numpy_and_exactness(n):
    # GAUWTS stops Newton at |dz| <= 1e-13 and forms the weights from the
    # derivative of the previous iterate, so its weights carry ~1e-13
    # relative error by construction (the golden test above pins that).
    x, w = O.gauwts(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.abs(x - xn).max() < 2e-15
    assert np.abs(w - wn).max() < 1e-12
    assert abs(w.sum() - 2.0) < 1e-12
    # exact for polynomials of degree < 2n
    for k in (0, 2, 10, 2 * n - 2):
        exact = 2.0 / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-12 * max(1.0, exact)


@pytest.mark.parametrize("tag", ["first", "mid", "last"])
def test_schnla_bitexact_vs_reference_routine(tag):
    x = float(GOLD["schnla31_x_%s" % tag][0])
    pa, dpa = O.schnla(x, 31)
    assert np.array_equal(pa, GOLD["schnla31_p_%s" % tag])
    assert np.array_equal(dpa, GOLD["schnla31_dp_%s" % tag])


def test_schnla_high_degree_vs_reference_routine():
    x = float(GOLD["schnla169_x"][0])
    pa, dpa = O.schnla(x, 169)
    idx = GOLD["schnla169_idx"]
    assert np.array_equal(pa[idx], GOLD["schnla169_p"])
    assert np.array_equal(dpa[idx], GOLD["schnla169_dp"])


def test_schnla_closed_form_mpmath():
    # Schmidt semi-normalised P_l^m without the Condon-Shortley phase:
    # P = sqrt((2 - d_m0) (l-m)!/(l+m)!) (-1)^m P_lm (mpmath's Ferrers
    # function carries the phase), and dP/dtheta = -sin(theta) dP/dx with
    # (x^2-1) dP_l^m/dx = l x P_l^m - (l+m) P_{l-1}^m.
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 20
    lh = 16
    for x in (-0.93, 0.41):
        pa, dpa = O.schnla(x, lh)
        X = mp.mpf(x)
        s = mp.sqrt(1 - X * X)
        for l in range(lh + 1):
            for m in range(l + 1):
                norm = mp.sqrt((2 - (m == 0)) * mp.factorial(l - m) / mp.factorial(l + m)) * (-1) ** m
                p_l = mp.legenp(l, m, X, type=2)
                p_lm1 = mp.legenp(l - 1, m, X, type=2) if l - 1 >= m else 0
                dpdx = (l * X * p_l - (l + m) * p_lm1) / (X * X - 1)
                val = float(norm * p_l)
                der = float(-s * norm * dpdx)
                k = l * (l + 1) // 2 + m
                assert abs(pa[k] - val) <= 1e-13 * max(1.0, abs(val))
                assert abs(dpa[k] - der) <= 1e-12 * max(1.0, abs(der))


@pytest.mark.parametrize("cfg", [config1(), config_small()], ids=["config1", "small"])
def test_tables_are_dual_under_gauss_quadrature(cfg):
    # sum over all Th nodes of gwP_l^m P_l'^m = delta_ll' (parity gives the
    # northern half), the identity tra_phys2spec o tra_spec2phys = 1 rests on
    o = O.Oracle(cfg, threads=1)
    P = o.table(2)
    gwP = o.table(7)
    from lsd.initial import harmonics
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    for mm in np.unique(m):
        hs = np.where(m == mm)[0]
        G = gwP[hs] @ P[hs].T
        par = (l[hs][:, None] + l[hs][None, :]) % 2
        full = np.where(par == 0, 2.0 * G, 0.0)
        assert np.abs(full - np.eye(len(hs))).max() < 1e-12
    # derivative tables: l(l+1) normalisation of the vector weights,
    # sum_j (gwllP_ llP_ + gwllsin1mP llsin1mP) over the sphere = delta for l >= 1
    llP, lls = o.table(5), o.table(6)
    gl, gs = o.table(8), o.table(9)
    for mm in np.unique(m):
        hs = np.where((m == mm) & (l > 0))[0]
        A = gl[hs] @ llP[hs].T + gs[hs] @ lls[hs].T
        par = (l[hs][:, None] + l[hs][None, :]) % 2
        full = np.where(par == 0, 2.0 * A, 0.0)
        assert np.abs(full - np.eye(len(hs))).max() < 1e-11
