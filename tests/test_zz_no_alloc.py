"""lsd_energy_noalloc (the energy diagnostic with its scratch in a static
device array, for callers that cannot allocate) against lsd_energy through
the C-ABI: equal bit for bit on two calls in a row, the second on a smaller
case, for the kinetic and the magnetic energy and their spectra; sizes past
the static array are refused."""
import numpy as np
import pytest

from lsd import make_config, grid_for, seeded_state, PLAIN_F64, MIXED_F32
from lsd.gpu import Context, LsdError

pytestmark = pytest.mark.gpu


def case(L, N, precision, seed):
    th, ph = grid_for(L, L)
    cfg = make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=1, precision=precision)
    g = Context(cfg)
    g.set_state(seeded_state(cfg, seed, 1e-3))
    return g


def test_noalloc_energy_equals_allocating_entry_point():
    big = case(64, 48, PLAIN_F64, 31)
    small = case(16, 24, MIXED_F32, 32)
    big.step(2)
    small.step(2)
    for g in (big, small):  # second call on the smaller case: the array keeps the larger call's contents
        print("case L=%d N=%d H=%d" % (g.cfg.L, g.N, g.H), flush=True)
        for which in (0, 1):
            E1, El1, Em1 = g.energy(which)
            E2, El2, Em2 = g.energy_noalloc(which)
            assert E1 > 0 and E1 == E2, (which, E1, E2)
            assert np.array_equal(El1, El2) and np.array_equal(Em1, Em2), which
            assert abs(El2.sum() - E2) <= 1e-12 * E2 and abs(Em2.sum() - E2) <= 1e-12 * E2
    big.close()
    small.close()


def test_noalloc_energy_refuses_sizes_past_its_array():
    # H = 400 * 400 - 399 * 400 / 2 = 80200 > 65536
    cfg = make_config(L=400, M=400, N=16, Th=600, Ph=1024, comp=0, precision=PLAIN_F64)
    g = Context(cfg)
    with pytest.raises(LsdError, match="static scratch"):
        g.energy_noalloc(0)
    E, _, _ = g.energy(0)
    assert E == 0.0
    g.close()
