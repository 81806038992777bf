"""Ensemble throughput at the reference's ensemble use case: 64 independent
runs with L = M drawn from {32, 48, 64, 86} (the maximal degrees 31, 47, 63,
85, rounded up to even M as par_precompute requires) and N from {32, 48, 64},
each with its own Ekman number (1e-4 .. 1e-3), Rayleigh number (2 to 10 times
the measured onset, lsd.config.ra_crit) and Pm (1 .. 5), mixed precision, stepped together by
lsd_step_many for a fixed number of timesteps from seeded states.

Reports the aggregate rate (member timesteps per second of wall time, over
all members), the same members stepped one after another with lsd_step for
comparison, and the member mix.  Prints one JSON line.

    python -m lsd.ensemble_bench [--members 64] [--steps 500] [--seed S]
"""
import argparse
import json
import sys
import time

import numpy as np


def members(n, seed):
    from .config import make_config, grid_for, ra_crit, MIXED_F32
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.choice([32, 48, 64, 86]))
        N = int(rng.choice([32, 48, 64]))
        E = float(10 ** rng.uniform(-4, -3))
        ra = float(rng.uniform(2, 10)) * ra_crit(E)
        pm = float(rng.uniform(1, 5))
        th, ph = grid_for(L, L)
        out.append(make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=0, Ra=ra, Pm=pm, E=E, dt=2e-6,
                               precision=MIXED_F32))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--seed", type=int, default=20261206)
    ap.add_argument("--solo-steps", type=int, default=50, help="steps per member in the one-at-a-time run")
    args = ap.parse_args(argv)
    from .gpu import Context, step_many
    from .initial import seeded_state
    cfgs = members(args.members, args.seed)
    t0 = time.perf_counter()
    ctxs = [Context(c) for c in cfgs]
    for k, (g, c) in enumerate(zip(ctxs, cfgs)):
        g.set_state(seeded_state(c, 1000 + k, 1e-3))
    setup_s = time.perf_counter() - t0
    # warm-up: graph capture on the first step of every member
    step_many(ctxs, 2)
    for g in ctxs:
        g.sync()
    print("members created and warmed up in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    t = time.perf_counter()
    stats = step_many(ctxs, args.steps)
    for g in ctxs:
        g.sync()
    ens_s = time.perf_counter() - t
    evals = int(sum(s.evaluations for s in stats))
    print("ensemble: %d x %d steps in %.2f s" % (args.members, args.steps, ens_s), file=sys.stderr, flush=True)
    # the same members one after another (fewer steps, rate per step)
    t = time.perf_counter()
    for g in ctxs:
        g.step(args.solo_steps)
        g.sync()
    solo_s = time.perf_counter() - t
    mix = {}
    for c in cfgs:
        key = "L%d_N%d" % (c.L, c.N)
        mix[key] = mix.get(key, 0) + 1
    rec = {
        "workload": "ensemble of %d mixed-precision runs, L=M in {32,48,64,86}, N in {32,48,64}" % args.members,
        "members": args.members,
        "steps_per_member": args.steps,
        "aggregate_timesteps_per_s": args.members * args.steps / ens_s,
        "nonlinear_evaluations": evals,
        "wall_s": ens_s,
        "one_at_a_time_timesteps_per_s": args.members * args.solo_steps / solo_s,
        "setup_s": setup_s,
        "mix": mix,
    }
    for g in ctxs:
        g.close()
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
