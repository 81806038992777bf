"""Conducting inner core (i_Nic > 1) and inner-core rotation (b_rot_tstep)
on the GPU against the CPU oracle, which the reference built with i_Nic > 1
and b_rot_tstep pins (test_oracle_ref_step.py; the GPU against those
fixtures directly: test_gpu_ref_step.py, cases ic and case2).

Cases here: the two-pass solve kernels' radial sizes (N = 32) for the
fluid and the coupled inner-core solves of the magnetic field (8 inner-core
points, band of 39 rows), 20 plain steps from seeded fields and a rotation
rate of 0.3, fields to 1e-11 relative L2 per field, the rotation rate and
the torques to 1e-10 relative, same corrector counts; and the mixed form
(FP32 transforms) against the plain form on the same case, 10 steps to
1e-6 per field and in the rotation rate.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from lsd import make_config, seeded_state, MIXED_F32
from lsd.gpu import Context
from lsd.config import christensen_case2
from lsd.initial import seeded_ic, ic_extend

HERE = os.path.dirname(os.path.abspath(__file__))

pytestmark = pytest.mark.gpu

NAMES = ("uT", "uP", "C", "BT", "BP", "icBT", "icBP")


def _cfg(**kw):
    base = dict(L=16, M=16, N=32, Th=24, Ph=48, comp=0, Nic=8, rot=1, dt=1e-4)
    base.update(kw)
    return make_config(**base)


def _state(cfg):
    st = seeded_state(cfg, 31, 1e-2)
    st.update(seeded_ic(cfg, 32, 1e-2))
    st["omega"] = 0.3
    return st


def _rel(a, b):
    d = np.sqrt(((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2).sum())
    n = np.sqrt((b[0] ** 2 + b[1] ** 2).sum())
    return d / n if n > 0 else d


def test_innercore_steps_equal_oracle():
    cfg = _cfg()
    print("case L=M=16 N=32 Nic=8 rot", cfg.sizes(), flush=True)
    st = _state(cfg)
    g = Context(cfg)
    o = O.Oracle(cfg)
    try:
        g.set_state(st)
        o.set_state(st)
        for k in range(1, 21):
            sg = g.step(1)
            so = o.step(1)
            assert sg.evaluations == so.evaluations, k
            rg, ro = g.rotation(), o.rotation()
            assert np.all(np.abs(rg - ro) <= 1e-10 * np.abs(ro) + 1e-300), (k, rg, ro)
            if k in (1, 5, 20):
                errs = {n: _rel(g.get_named(n), o.get_named(n)) for n in NAMES}
                print("step", k, {n: "%.1e" % e for n, e in errs.items()}, "omega", rg[0], flush=True)
                assert max(errs.values()) < 1e-11, (k, errs)
    finally:
        g.close()


def test_innercore_mixed_form():
    cfg = _cfg()
    cm = _cfg(precision=MIXED_F32)
    st = _state(cfg)
    gp, gm = Context(cfg), Context(cm)
    try:
        gp.set_state(st)
        gm.set_state(st)
        gp.step(10)
        gm.step(10)
        errs = {n: _rel(gm.get_named(n), gp.get_named(n)) for n in NAMES}
        om_p, om_m = gp.rotation()[0], gm.rotation()[0]
        print("mixed vs plain after 10 steps", {n: "%.1e" % e for n, e in errs.items()},
              "omega %.9g %.9g" % (om_m, om_p), flush=True)
        assert max(errs.values()) < 1e-6
        assert abs(om_m - om_p) <= 1e-6 * abs(om_p)
    finally:
        gp.close()
        gm.close()


def test_state_file_inner_core(tmp_path):
    """lsd_load_state on a file without an inner core (the reference's
    case-1 test state) extends the outer core's field into the inner core
    as io_ic_extend does (program/io.F90:1438-1477, here restated on the
    host from the library's own radial weights); lsd_save_state writes
    icBT, icBP and icOmega, and loading that file back restores them."""
    cfg = christensen_case2()
    g = Context(cfg)
    try:
        g.load_state(os.path.join(HERE, "golden", "state_case1.cdf.in"))
        r, dr1 = g.mesh(0), g.mesh(1)
        for k, name in ((0, "BT"), (1, "BP")):
            want = ic_extend(cfg, g.get_coll(4 + k), r, dr1)
            got = g.get_ic(k)
            assert _rel(got, want) < 1e-14, name
        g.set_rotation(-1.25)
        g.step(3)
        path = str(tmp_path / "state_ic.cdf")
        g.save_state(path)
        h = Context(cfg)
        try:
            h.load_state(path)
            for k in range(6):
                a, b = g.get_coll(k), h.get_coll(k)
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), k
            for k in range(2):
                a, b = g.get_ic(k), h.get_ic(k)
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), k
            assert h.rotation()[0] == g.rotation()[0]
        finally:
            h.close()
    finally:
        g.close()
