"""Ensembles (lsd_step_many): independent runs of different truncations
and parameters stepped together must each reproduce the run stepped alone,
bit for bit, and keep their own corrector counts.

The ensemble of the reference's use case draws L = M from {31, 47, 63, 85}
and N from {32, 48, 64}; the harmonic count the survey gives for L = M = 31
(H = 496) counts degrees 0..30 with M odd, which par_precompute rejects
(M must be even), so the maximal degrees are kept and L = M is rounded up to
the next even value here (32, 48, 64, 86), on the grids of lsd.grid_for.
"""
import numpy as np
import pytest

from lsd import make_config, seeded_state, MIXED_F32, PLAIN_F64, ra_crit, grid_for
from lsd.gpu import Context, step_many

pytestmark = pytest.mark.gpu


def member(k, rng):
    L = int(rng.choice([32, 48, 64, 86]))
    N = int(rng.choice([32, 48, 64]))
    E = float(10 ** rng.uniform(-4, -3))
    ra = float(rng.uniform(2, 10)) * ra_crit(E)
    pm = float(rng.uniform(1, 5))
    th, ph = grid_for(L, L)
    return make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=0, Ra=ra, Pm=pm, E=E, dt=2e-6, precision=MIXED_F32)


def test_ensemble_matches_members():
    rng = np.random.default_rng(20261204)
    cfgs = [member(k, rng) for k in range(6)]
    print("ensemble", [(c.L, c.N) for c in cfgs], flush=True)
    states = [seeded_state(c, 100 + k, 1e-3) for k, c in enumerate(cfgs)]
    alone = []
    for c, s in zip(cfgs, states):
        g = Context(c)
        g.set_state(s)
        st = g.step(5)
        alone.append(([g.get_coll(k) for k in range(6)], st.evaluations))
        g.close()
    ens = [Context(c) for c in cfgs]
    for g, s in zip(ens, states):
        g.set_state(s)
    stats = step_many(ens, 5)
    for i, g in enumerate(ens):
        assert stats[i].steps == 5 and stats[i].evaluations == alone[i][1]
        for k in range(6):
            a, b = g.get_coll(k), alone[i][0][k]
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (i, k)
        g.close()


def test_ensemble_member_that_stops():
    """A member whose corrector cannot converge (d_dterr = 1e-30, so
    tim_check_cgce stops it within its first step) is drained and reported;
    the other members still finish every step and equal their solo runs bit
    for bit."""
    from lsd.gpu import StepManyError
    rng = np.random.default_rng(20261205)
    cfgs = [member(k, rng) for k in range(3)]
    cfgs[1].dterr = 1e-30
    states = [seeded_state(c, 200 + k, 1e-3) for k, c in enumerate(cfgs)]
    print("ensemble", [(c.L, c.N, c.dterr) for c in cfgs], flush=True)
    alone = {}
    for i in (0, 2):
        g = Context(cfgs[i])
        g.set_state(states[i])
        st = g.step(4)
        alone[i] = ([g.get_coll(k) for k in range(6)], st.evaluations)
        g.close()
    ens = [Context(c) for c in cfgs]
    for g, s in zip(ens, states):
        g.set_state(s)
    with pytest.raises(StepManyError) as ei:
        step_many(ens, 4)
    stats = ei.value.stats
    print(str(ei.value), flush=True)
    assert "member 1" in str(ei.value) and "tim_check_cgce" in str(ei.value)
    assert stats[1].status == -5 and stats[1].steps == 0
    for i in (0, 2):
        assert stats[i].status == 0 and stats[i].steps == 4 and stats[i].evaluations == alone[i][1]
        for k in range(6):
            a, b = ens[i].get_coll(k), alone[i][0][k]
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (i, k)
    for g in ens:
        g.close()
