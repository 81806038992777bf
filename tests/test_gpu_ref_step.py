"""The plain FP64 GPU step against the reference's own FFTW/LAPACK CPU path:
golden per-step vectors of the reference's unmodified modules
(tests/golden/ref_step_<case>.npz; bounds in ref_fixtures.py), through the
C-ABI.  The same vectors pin the C oracle (test_oracle_ref_step.py).

Every step's corrector count (tim_check_cgce) and tim_t must equal the
reference's; the fields must agree to the per-step bound times the step
count (plus the reference's own 1/dt boundary-row spread,
ref_fixtures.step_bound).  Cases: L = M = 12, N = 16, 18 x 36 (LDS-column
solves), config 1 (L = M = 32, N = 32, 48 x 96; 100 steps, seven of them
with two correctors), config 3 (L = M = 170, N = 96, 256 x 512), the
adaptive step on the small grid (d_timestep < 0: 40 steps chosen by the
CFL limit non_maxtstep and tim_new_tstep, the operators refactorised in
FP64 at every change), the Christensen case-1 dynamo, a conducting
inner core of 8 points that rotates (the coupled inner/outer-core solves,
its induction, the torques and the rotation rate, also compared) and the
Christensen case-2 benchmark (16 inner-core points, N = 64) and the
stress-free walls (i_vel_bc = 2, 3, 4; with both walls stress-free
vel_angmomclr removes the rigid rotations), replayed from graphs and
launched eagerly.
"""
import numpy as np
import pytest

import ref_fixtures as rf
from lsd.gpu import Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,nsteps,graph", [("small", 20, True), ("small", 20, False), ("cfg1", 100, True),
                                               ("cfg3", 2, True), ("adapt", 40, True), ("adapt", 40, False),
                                               ("adapt1", 40, True), ("case1", 30, True), ("ic", 20, True),
                                               ("ic", 20, False), ("case2", 10, True), ("sf2", 20, True),
                                               ("sf3", 20, True), ("sf4", 20, True), ("sf4", 20, False)])
def test_plain_steps_equal_reference(case, nsteps, graph):
    fx = rf.load(case)
    cfg = rf.config(fx)
    print("case", case, cfg.sizes(), "graph", graph, flush=True)
    g = Context(cfg)
    try:
        g.set_graph(graph)
        g.set_state(rf.initial_state(fx, cfg))
        keep = {int(k.split("_")[0][1:]) for k in fx if k[0] in "sq" and k[1].isdigit()}
        worst = {}
        for k in range(1, nsteps + 1):
            st = g.step(1)
            assert st.evaluations == fx["evaluations"][k - 1], "step %d: corrector count differs" % k
            t_ref = fx["t"][k - 1]
            assert abs(st.t - t_ref) <= 1e-12 * t_ref, (k, st.t, t_ref)
            assert rf.dterr_ok(fx, k, st.last_dterr), (k, st.last_dterr, fx["dterr"][k - 1])
            if "omega" in fx:
                bad = rf.rotation_errors(fx, k, g.rotation())
                assert not bad, (k, bad)
            if k in keep:
                errs = rf.compare(fx, case, k, g.get_named, k)
                worst[k] = max(errs.values())
                print("step", k, {n: "%.2e" % e for n, e in errs.items()}, flush=True)
                assert worst[k] <= rf.step_bound(fx, case, k), "step %d: %s" % (k, errs)
    finally:
        g.close()
