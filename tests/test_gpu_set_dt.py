"""lsd_set_dt (the *_matrices rebuild of program/main.F90:126-135): a run
created at dt1 whose operators are refactorised for dt2 between steps must
step exactly as a run created at dt2, in the plain and the mixed forms; a
mid-run change of dt must match the CPU oracle's orc_set_dt step by step."""
import numpy as np
import pytest

from lsd import config_small, make_config, grid_for, seeded_state, MIXED_F32, PLAIN_F64
from lsd.gpu import Context, LsdError
from lsd.initial import FIELD_ID

pytestmark = pytest.mark.gpu


def cfg_for(precision, dt, N=32):
    th, ph = grid_for(16, 16)
    return make_config(L=16, M=16, N=N, Th=th, Ph=ph, comp=1, dt=dt, precision=precision)


@pytest.mark.parametrize("precision", [PLAIN_F64, MIXED_F32])
def test_set_dt_equals_fresh_context(precision):
    c1, c2 = cfg_for(precision, 1e-4), cfg_for(precision, 3e-5)
    st = seeded_state(c1, 11, 1e-3)
    a = Context(c1)
    a.set_state(st)
    a.set_dt(3e-5)
    b = Context(c2)
    b.set_state(st)
    sa, sb = a.step(3), b.step(3)
    assert sa.evaluations == sb.evaluations
    for k in range(6):
        x, y = a.get_coll(k), b.get_coll(k)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]), k
    a.close()
    b.close()


def test_set_dt_inside_a_step_is_rejected():
    c = cfg_for(PLAIN_F64, 1e-4)
    g = Context(c)
    g.set_state(seeded_state(c, 12, 1e-3))
    g.transform()
    g.nonlinear()
    g.predictor()  # tim_it = 1: the step is open
    with pytest.raises(LsdError):
        g.set_dt(5e-5)
    g.close()


@pytest.mark.parametrize("N", [32, 20])
def test_set_dt_mid_run_matches_oracle(N):
    """dt 1e-4 for two steps, then 4e-5 for two, against the oracle doing
    the same (N = 32: two-pass solve kernels; N = 20: LDS-column kernels)."""
    import oracle_lib
    c = cfg_for(PLAIN_F64, 1e-4, N=N)
    st = seeded_state(c, 13, 1e-3)
    g = Context(c)
    g.set_state(st)
    o = oracle_lib.Oracle(c)
    o.set_state(st)
    for k, dt in enumerate((1e-4, 1e-4, 4e-5, 4e-5)):
        if k == 2:
            g.set_dt(dt)
            o.set_dt(dt)
        sg, so = g.step(1), o.step(1)
        assert sg.evaluations == so.evaluations, k
        worst = 0.0
        for f in range(6):
            a, b = g.get_coll(f), o.get_coll(f)
            den = np.sqrt(sum((y ** 2).sum() for y in b))
            if den > 0:
                worst = max(worst, np.sqrt(sum(((x - y) ** 2).sum() for x, y in zip(a, b))) / den)
        print("step", k + 1, "dt", dt, "worst rel %.2e" % worst, flush=True)
        assert worst < 1e-11 * (k + 1), (k, worst)
    g.close()
