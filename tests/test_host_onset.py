"""Host logic of the onset scan and of the Rayleigh numbers built on it
(CPU only): the zero finder of lsd.onset, the measured-onset power law
lsd.config.ra_crit and config 3's Rayleigh number of 5 x the onset."""
import math

import pytest

from lsd.config import RA_MOD_ONSET, config3, ra_crit, unified_ra
from lsd.onset import onset


def test_zero_between_bracketing_points():
    # unsorted input; the straight line through (2, -1) and (4, 3) is zero at 2.5
    assert onset([4.0, 1.0, 2.0, 6.0], [3.0, -4.0, -1.0, 5.0]) == pytest.approx(2.5)
    assert onset([1.0, 2.0], [1.0, 2.0]) is None


def test_ra_crit_is_the_measured_onset():
    assert ra_crit(1e-3) == pytest.approx(unified_ra(55.91, 1e-3), rel=1e-14)
    assert ra_crit(1e-4) == pytest.approx(unified_ra(RA_MOD_ONSET[1e-4], 1e-4), rel=1e-14)
    # d_Ra = Ra_mod / (r_o E): the benchmark's 100 -> 6.5e4
    assert unified_ra(100.0, 1e-3) == pytest.approx(6.5e4, rel=1e-14)
    # power law between the two: log-linear midpoint
    mid = math.sqrt(55.91 * RA_MOD_ONSET[1e-4])
    assert ra_crit(10 ** -3.5) == pytest.approx(unified_ra(mid, 10 ** -3.5), rel=1e-12)
    for E in (5e-5, 2e-3):
        with pytest.raises(ValueError):
            ra_crit(E)


def test_config3_is_five_times_onset():
    c = config3()
    assert c.Ra == pytest.approx(5 * ra_crit(1e-4), rel=1e-14)
    assert c.Ra * c.E / (1 - c.rratio) == pytest.approx(5 * RA_MOD_ONSET[1e-4], rel=1e-14)
