"""Kernel variants (lsd_set_variant) and the autotuner (lsd_tune).

Each variant computes the same step: the FP32 synthesis shapes (4 or 8 m
per workgroup, static or persistent) bit for bit, the FP32 analysis forms (wide-load,
static or persistent, bit for bit; LDS-staged to rounding) and the two solve forms (two-pass register-column FP64, and
LDS-column: FP64 in the plain form, FP32 factors with one FP64-residual
refinement sweep in the mixed form) to rounding.  At L = M = 170, N = 96 the
mixed form with FP32 refined solves must stay within 1e-6 of the plain
FP64 form after 1 and 100 steps (the bound of the mixed form)."""
import os

import numpy as np
import pytest

from lsd import config1, config3, seeded_state, MIXED_F32, PLAIN_F64, SEEDS, AMPS
from lsd.gpu import Context, LsdError, V_SYN_SCALAR, V_SYN_VECTOR, V_SYN_GRAD, V_ANA, V_SOLVE

pytestmark = pytest.mark.gpu


def run(cfg, state, steps, **variants):
    g = Context(cfg)
    for k, v in variants.items():
        g.set_variant({"syn0": V_SYN_SCALAR, "syn1": V_SYN_VECTOR, "syn2": V_SYN_GRAD, "ana": V_ANA,
                       "solve": V_SOLVE}[k], v)
    g.set_state(state)
    st = g.step(steps)
    out = [g.get_coll(k) for k in range(6)]
    g.close()
    return out, st.evaluations


def rel(a, b):
    worst = 0.0
    for x, y in zip(a, b):
        den = np.sqrt((y[0] ** 2).sum() + (y[1] ** 2).sum())
        if den > 0:
            worst = max(worst, np.sqrt(((x[0] - y[0]) ** 2).sum() + ((x[1] - y[1]) ** 2).sum()) / den)
    return worst


def test_synthesis_shapes_bit_for_bit():
    cfg = config1(MIXED_F32)
    st = seeded_state(cfg, 3, 1e-3)
    base, ev = run(cfg, st, 3, syn0=4, syn1=8, syn2=8)
    other, ev2 = run(cfg, st, 3, syn0=8, syn1=4, syn2=4)
    assert ev == ev2
    for (a, b), (c, d) in zip(base, other):
        assert np.array_equal(a, c) and np.array_equal(b, d)
    # the persistent forms (5, 9; 6, 10 with the next item's loads issued
    # before each epilogue): the same work items pulled from queues
    for shapes in ((5, 9, 9), (9, 5, 5), (6, 10, 10), (10, 6, 6)):
        pers, ev3 = run(cfg, st, 3, syn0=shapes[0], syn1=shapes[1], syn2=shapes[2])
        assert ev3 == ev
        for (a, b), (c, d) in zip(base, pers):
            assert np.array_equal(a, c) and np.array_equal(b, d)


def test_persistent_synthesis_at_config3():
    """The persistent synthesis at L = M = 170, N = 96 (about 2 work items
    per CU, uneven slot lengths): bit for bit the static form, over steps
    replayed from graphs (the queues reset themselves between launches)."""
    cfg = config3(MIXED_F32)
    st = seeded_state(cfg, SEEDS[3], AMPS[3])
    base, ev = run(cfg, st, 3, syn0=4, syn1=8, syn2=8)
    pers, ev2 = run(cfg, st, 3, syn0=5, syn1=9, syn2=9)
    assert ev == ev2
    for (a, b), (c, d) in zip(base, pers):
        assert np.array_equal(a, c) and np.array_equal(b, d)
    pref, ev3 = run(cfg, st, 3, syn0=10, syn1=10, syn2=10)
    assert ev == ev3
    for (a, b), (c, d) in zip(base, pref):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_analysis_forms_agree():
    cfg = config1(MIXED_F32)  # Th/2 = 24: both forms apply
    st = seeded_state(cfg, 4, 1e-3)
    a, _ = run(cfg, st, 10, ana=0)
    b, _ = run(cfg, st, 10, ana=1)
    e = rel(a, b)
    print("analysis forms differ by %.2e after 10 steps" % e, flush=True)
    assert e < 1e-6
    # the persistent wide-load form: the same work items, bit for bit
    p, _ = run(cfg, st, 10, ana=2)
    for (x, y), (u, v) in zip(a, p):
        assert np.array_equal(x, u) and np.array_equal(y, v)


def test_persistent_analysis_at_config3():
    """The persistent wide-load analysis at L = M = 170, N = 96, with the
    persistent synthesis beside it (both sets of queues in use), over steps
    replayed from graphs: bit for bit the static forms."""
    cfg = config3(MIXED_F32)
    st = seeded_state(cfg, SEEDS[3], AMPS[3])
    base, ev = run(cfg, st, 3, syn0=4, syn1=8, syn2=8, ana=0)
    pers, ev2 = run(cfg, st, 3, syn0=5, syn1=9, syn2=9, ana=2)
    assert ev == ev2
    for (a, b), (c, d) in zip(base, pers):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_solve_forms_agree_plain():
    cfg = config1(PLAIN_F64)
    st = seeded_state(cfg, 5, 1e-3)
    a, ea = run(cfg, st, 10, solve=0)
    b, eb = run(cfg, st, 10, solve=1)
    e = rel(a, b)
    print("plain solve forms differ by %.2e after 10 steps" % e, flush=True)
    assert ea == eb and e < 1e-11


def test_fp32_refined_solves_at_config3():
    cp, cm = config3(PLAIN_F64), config3(MIXED_F32)
    st = seeded_state(cp, SEEDS[3], AMPS[3])
    gp, gm = Context(cp), Context(cm)
    gm.set_variant(V_SOLVE, 1)
    gp.set_state(st)
    gm.set_state(st)
    for steps, mark in ((1, 1), (99, 100)):
        sp, sm = gp.step(steps), gm.step(steps)
        assert sp.evaluations == sm.evaluations
        e = rel([gm.get_coll(k) for k in range(6)], [gp.get_coll(k) for k in range(6)])
        print("mixed with FP32 refined solves vs plain after %d steps: %.2e" % (mark, e), flush=True)
        assert e < 1e-6
    gp.close()
    gm.close()


def test_fp64_lds_column_solves_in_mixed_form():
    """LSD_V_SOLVE = 2: the LDS-column kernels with FP64 factors in a mixed
    context, against the two-pass FP64 kernels of the same context (which
    keep the saved N in FP32): the same corrector counts and 1e-6 after 10
    steps."""
    cfg = config1(MIXED_F32)
    st = seeded_state(cfg, 6, 1e-3)
    a, ea = run(cfg, st, 10, solve=0)
    b, eb = run(cfg, st, 10, solve=2)
    assert ea == eb
    e = rel(a, b)
    print("two-pass and FP64 LDS-column solves (mixed form) differ by %.2e after 10 steps" % e, flush=True)
    assert e < 1e-6


def test_variant_checks():
    cfg = config1(MIXED_F32)
    g = Context(cfg)
    with pytest.raises(LsdError):
        g.set_variant(V_SYN_SCALAR, 7)
    with pytest.raises(LsdError):
        g.set_variant(99, 0)
    g.close()


def test_tune_writes_and_reuses_cache(tmp_path):
    cache = str(tmp_path / "tuning_cache.jsonl")
    cfg = config1(MIXED_F32)
    g = Context(cfg)
    g.set_state(seeded_state(cfg, 6, 1e-3))
    before = [g.get_coll(k) for k in range(6)]
    times = g.tune(3, cache)
    after = [g.get_coll(k) for k in range(6)]
    for (a, b), (c, d) in zip(before, after):  # the state is restored
        assert np.array_equal(a, c) and np.array_equal(b, d)
    chosen = g.variants()
    print("tuned", chosen, times, flush=True)
    assert chosen["syn_scalar"] in (4, 8) and chosen["solve"] in (0, 1)
    with open(cache) as f:
        assert f.read().count('"L": 32') == 1
    g.close()
    old = os.environ.get("LSD_TUNING_CACHE")
    os.environ["LSD_TUNING_CACHE"] = cache
    try:
        h = Context(cfg)
        assert h.variants() == chosen
        h.close()
    finally:
        if old is None:
            del os.environ["LSD_TUNING_CACHE"]
        else:
            os.environ["LSD_TUNING_CACHE"] = old
