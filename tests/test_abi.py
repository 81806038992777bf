"""The C-ABI library: it loads, exports every entry point include/lsd.h
declares, and checks configurations the way par_precompute does
(Scripts/parameters.F90_template:274-288), all without touching a GPU."""
import ctypes
import os
import re

import pytest

from lsd import config1, config3, make_config
from lsd.gpu import load, config_sizes, LsdError, EXPORTS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "lsd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|void|const char\*)\s+(lsd_\w+)\s*\(", src, re.M)))


def test_every_declared_symbol_is_exported():
    lib = load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(EXPORTS)


def test_config_sizes_follow_the_reference():
    H, Th, Ph = config_sizes(make_config(L=64, M=16, Mp=4, N=64))
    assert H == 64 * 16 - 4 * 15 * 16 // 2
    assert (Th, Ph) == (96, 48)            # 3L/2, 3M (i_Th, i_Ph)
    assert config_sizes(config1()) == (528, 48, 96)
    assert config_sizes(config3()) == (14535, 256, 512)


@pytest.mark.parametrize("kw,msg", [
    (dict(M=15), "M even"),
    (dict(L=8, M=16), "L > (M-1)*Mp"),
    (dict(Th=47), "Th must be even"),
    (dict(Ph=40), "Ph must be even and >= 2M"),
    (dict(vel_bc=5), "i_vel_bc must be 1..4"),
    (dict(Nic=16, vel_bc=3), "no-slip inner boundary"),
    (dict(Nic=4), "8 <= i_Nic"),
    (dict(Nic=16, mag=0), "needs the magnetic field"),
])
def test_config_errors(kw, msg):
    with pytest.raises(LsdError, match=re.escape(msg)):
        config_sizes(make_config(**kw))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_fails_loudly_without_a_gpu():
    from lsd.gpu import Context
    with pytest.raises(LsdError):
        Context(config1())
