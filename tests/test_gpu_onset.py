"""Onset of convection against the published value: at E = 1e-3, Pr = 1,
r_i/r_o = 0.35 with fixed codensity and no-slip walls, Christensen et al.
(2001) give the critical modified Rayleigh number 55.91 (the benchmark of
Benchmark_files/Boussinesq/Christensen_case0_benchmark runs at 100).  The
linear growth rate of a small perturbation, stepped by the GPU timestep
(lsd.onset.growth_rate), must be negative 5 % below it and positive 5 %
above it (measured -2.0 and +1.9 at L = M = 32, N = 32; the growth rate
moves by about 0.7 per unit of Ra_mod there, so a buoyancy or Coriolis term
off by a few per cent flips one of the signs).
"""
import pytest

from lsd.onset import growth_rate, SCAN

pytestmark = pytest.mark.gpu


def test_onset_brackets_published_value():
    p = dict(SCAN[1e-3])
    p.pop("factors")
    p.pop("fine")
    p["t_end"] = 1.5
    lo, _ = growth_rate(1e-3, 0.95 * 55.91, **p)
    hi, _ = growth_rate(1e-3, 1.05 * 55.91, **p)
    print("sigma at 0.95 and 1.05 x 55.91: %.4g %.4g" % (lo, hi), flush=True)
    assert lo < -1.0
    assert hi > 1.0
