"""Pins of the oracle's energies (program/io.F90:1799-1884 with
var_coll_norm / var_coll_normlm, modules/variables.F90:901-957, 1267-1313).

Parseval: the spectral energy Efac * sum_h w_h int |qst_h|^2 r^2 dr must
equal Efac times the same integral done in physical space on the
transform grid -- Gauss-Legendre in theta, the phi sum, the radial int
r^2 dr weights -- of |u|^2 (or |B|^2) from the oracle's own synthesis.  The
grid is dealiased (Th = 3L/2 nodes, Ph >= 3M points), so the quadrature is
exact for the squared band-limited fields and only rounding separates the
two.  This pins the weights 4 pi/(2l+1) (x4 for m /= 0), the qst form and
Efac.  The spectra must sum to the total.
"""
import numpy as np
import pytest

import oracle_lib as O
from lsd import config_small, config1, seeded_state


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("mk", ["small", "config1"])
def test_energy_parseval(which, mk):
    cfg = config_small() if mk == "small" else config1()
    o = O.Oracle(cfg)
    o.set_state(seeded_state(cfg, 17, 1e-2))
    E, El, Em = o.energy(which)
    o.transform()
    base = 6 if which else 0
    comps = [o.get_phys(base + k) for k in range(3)]        # [n, k(phi), j(theta)]
    sq = sum(c * c for c in comps)
    gw = o.table(1)                                         # Gauss weights (theta)
    ang = (sq * gw[None, None, :]).sum(axis=(1, 2)) * (2.0 * np.pi / o.Ph)
    intr2dr = o.mesh(5)
    vol = (ang * intr2dr).sum()
    efac = 0.5 if which == 0 else cfg.Pm / cfg.E
    assert abs(E - efac * vol) <= 1e-11 * abs(E)
    assert abs(El.sum() - E) <= 1e-12 * abs(E)
    assert abs(Em.sum() - E) <= 1e-12 * abs(E)
