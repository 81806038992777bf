"""The Christensen et al. (2001) dynamo benchmark, case 1, as the
reference's own test runs it (tests/parameters-case1.F90 at the resolution
of its state file: N = 64, L = 64, M = 16, Mp = 4; E = 1e-3, Ra = 6.5e4,
Pr = 1, Pm = 5, insulating inner core, adaptive step under
i_time_step_control = 1), started from the reference's test state
tests/state.cdf.in (tests/golden/state_case1.cdf.in) through lsd_load_state
and stepped by the plain FP64 GPU path for 0.65 magnetic diffusion times.

Checks:
  * the kinetic and magnetic energies follow the reference's own case-1
    series (Benchmark_files/Boussinesq/Christensen_case1_benchmark/
    vel_energy.dat, mag_energy.dat; tests/golden/christensen_case1_series.npz)
    at every time it records;
  * their means over the last 500 records (10000 steps), as
    tests/check_output.py:15-21 averages them, lie within the benchmark
    tolerances of tests/check_output.py:25-27: KE 11231.222 +- 7.2994,
    ME 228620.5 +- 145.98.
The file's B is in the normalisation of the version that wrote it and the
series (ME = (Pm / 2E) int B^2); the present code's is B / sqrt(2)
(ref_fixtures.B_SCALE_OLD_STATE), so B is rescaled after loading.
"""
import os

import numpy as np
import pytest

import ref_fixtures as rf
from lsd.config import christensen_case1, christensen_case2
from lsd.gpu import Context

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KE_BENCH, KE_TOL = 11231.222, 7.2994
ME_BENCH, ME_TOL = 228620.5, 145.98


def test_christensen_case1_benchmark():
    ser = np.load(os.path.join(HERE, "golden", "christensen_case1_series.npz"))
    cfg = christensen_case1()
    g = Context(cfg)
    g.load_state(os.path.join(HERE, "golden", "state_case1.cdf.in"))
    for k in (4, 5):  # BT, BP
        re, im = g.get_coll(k)
        g.set_coll(k, re * rf.B_SCALE_OLD_STATE, im * rf.B_SCALE_OLD_STATE)
    t_end = float(ser["t"][-1])
    ts, kes, mes = [0.0], [g.energy(0)[0]], [g.energy(1)[0]]
    print("t=0 KE %.6f ME %.4f (series %.6f %.4f)" % (kes[0], mes[0], ser["KE"][0], ser["ME"][0]), flush=True)
    assert abs(kes[0] - ser["KE"][0]) <= 1e-9 * ser["KE"][0]
    assert abs(mes[0] - ser["ME"][0]) <= 1e-9 * ser["ME"][0]
    steps = 0
    while ts[-1] < t_end:
        st = g.step(20)  # the series records every i_save_rate2 = 20 steps
        steps += 20
        ts.append(st.t)
        kes.append(g.energy(0)[0])
        mes.append(g.energy(1)[0])
        if steps % 2000 == 0:
            print("step %d t %.5f dt %.3e KE %.4f ME %.3f" % (steps, st.t, g.get_dt()[0], kes[-1], mes[-1]),
                  flush=True)
    ts, kes, mes = np.array(ts), np.array(kes), np.array(mes)
    # against the reference's series at its own record times
    ke_i = np.interp(ser["t"], ts, kes)
    me_i = np.interp(ser["t"], ts, mes)
    dke = np.abs(ke_i / ser["KE"] - 1).max()
    dme = np.abs(me_i / ser["ME"] - 1).max()
    print("steps %d, t %.4f; largest relative difference from the series: KE %.2e ME %.2e"
          % (steps, ts[-1], dke, dme), flush=True)
    assert dke < 1e-3 and dme < 1e-3
    # the benchmark means, averaged as tests/check_output.py does
    ke_m, me_m = kes[-500:].mean(), mes[-500:].mean()
    print("mean of the last 500 records: KE %.3f (bench %.3f +- %.4f), ME %.2f (bench %.1f +- %.2f)"
          % (ke_m, KE_BENCH, KE_TOL, me_m, ME_BENCH, ME_TOL), flush=True)
    assert abs(ke_m - KE_BENCH) < KE_TOL
    assert abs(me_m - ME_BENCH) < ME_TOL
    g.close()


# Christensen case 2 at the resolution of the reference's benchmark run
# (Benchmark_files/Boussinesq/Christensen_case2_benchmark/INFO): the series
# of that directory are steady (vel_energy.dat, mag_energy.dat,
# rot_torque.dat: KE 15474.553, ME 308651.156, rot_omega -13.297956, each
# constant to 1e-7 relative over its 797 records).
KE2, ME2, OM2 = 15474.553, 308651.156, -13.297956


def test_christensen_case2_benchmark():
    """Conducting, freely rotating inner core of 16 points: from the case-1
    test state (its field extended into the inner core by lsd_load_state as
    io_ic_extend does, rot_omega = 0) to t = 1.4 under the adaptive step;
    the means of KE, ME and rot_omega over t in [1.2, 1.4] must match the
    reference's steady values to 1e-4 relative (measured: 3e-6, 2e-6 and
    2e-5 over t in [1, 9], with a residual oscillation of 3e-5 in KE that
    the averaging window spans 2.4 periods of)."""
    cfg = christensen_case2()
    g = Context(cfg)
    try:
        g.load_state(os.path.join(HERE, "golden", "state_case1.cdf.in"))
        s = rf.B_SCALE_OLD_STATE
        for k in (4, 5):
            re, im = g.get_coll(k)
            g.set_coll(k, re * s, im * s)
        for k in (0, 1):
            re, im = g.get_ic(k)
            g.set_ic(k, re * s, im * s)
        t, steps = 0.0, 0
        kes, mes, oms = [], [], []
        while t < 1.4:
            st = g.step(100)
            steps += 100
            t = st.t
            if t >= 1.2:
                kes.append(g.energy(0)[0])
                mes.append(g.energy(1)[0])
                oms.append(g.rotation()[0])
            if steps % 10000 == 0:
                print("step %d t %.4f KE %.3f ME %.2f omega %.5f" % (steps, t, g.energy(0)[0], g.energy(1)[0],
                                                                     g.rotation()[0]), flush=True)
        ke, me, om = np.mean(kes), np.mean(mes), np.mean(oms)
        print("means over t in [1.2, 1.4] (%d samples): KE %.3f (ref %.3f) ME %.2f (ref %.2f) omega %.6f (ref %.6f)"
              % (len(kes), ke, KE2, me, ME2, om, OM2), flush=True)
        assert abs(ke / KE2 - 1) < 1e-4
        assert abs(me / ME2 - 1) < 1e-4
        assert abs(om / OM2 - 1) < 1e-4
    finally:
        g.close()
