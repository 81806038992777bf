"""Christensen et al. (2001) dynamo benchmarks on the GPU timestep, at the
resolution of the reference's own runs (Benchmark_files/Boussinesq/
Christensen_case1_benchmark, Christensen_case2_benchmark: N = 64, L = 64,
M = 16, Mp = 4, adaptive step under i_time_step_control = 1).

Starts from a state file (the reference's test state tests/state.cdf.in for
both cases; for case 2 lsd_load_state extends its field into the conducting
inner core as io_ic_extend does and the inner core starts at rest), scales
B by --b-scale (that file holds B in the normalisation of the version that
wrote it, sqrt(2) larger than the present code's), steps to --t-end and
prints KE, ME (and rot_omega for case 2) every --every time units, then one
JSON line with the means and standard deviations over t >= --t-avg.

    python -m lsd.christensen --case 2 --state tests/golden/state_case1.cdf.in --t-end 9
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", type=int, choices=(1, 2), default=2)
    ap.add_argument("--state", required=True)
    ap.add_argument("--b-scale", type=float, default=2.0 ** -0.5)
    ap.add_argument("--t-end", type=float, default=9.0)
    ap.add_argument("--t-avg", type=float, default=1.0)
    ap.add_argument("--every", type=float, default=0.005)
    ap.add_argument("--chunk", type=int, default=100)
    args = ap.parse_args(argv)
    from .config import christensen_case1, christensen_case2
    from .gpu import Context
    cfg = christensen_case2() if args.case == 2 else christensen_case1()
    g = Context(cfg)
    g.load_state(args.state)
    for k in (4, 5):
        re, im = g.get_coll(k)
        g.set_coll(k, re * args.b_scale, im * args.b_scale)
    if cfg.Nic > 1:
        for k in (0, 1):
            re, im = g.get_ic(k)
            g.set_ic(k, re * args.b_scale, im * args.b_scale)
    t0 = time.time()
    t, steps, nxt = 0.0, 0, args.every
    rows = []
    while t < args.t_end:
        st = g.step(args.chunk)
        steps += args.chunk
        t = st.t
        if t >= nxt:
            nxt = t + args.every
            ke, me = g.energy(0)[0], g.energy(1)[0]
            om = g.rotation()[0] if cfg.rot else 0.0
            rows.append((t, ke, me, om))
            if len(rows) % 100 == 0:
                print("t %.4f steps %d dt %.3e KE %.4f ME %.3f omega %.6f (%.0f s)"
                      % (t, steps, g.get_dt()[0], ke, me, om, time.time() - t0), file=sys.stderr, flush=True)
    g.close()
    a = np.array(rows)
    m = a[:, 0] >= args.t_avg
    rec = {"case": args.case, "t_end": t, "steps": steps, "wall_s": time.time() - t0,
           "samples": int(m.sum()), "t_avg_from": args.t_avg,
           "KE": [float(a[m, 1].mean()), float(a[m, 1].std())],
           "ME": [float(a[m, 2].mean()), float(a[m, 2].std())]}
    if cfg.rot:
        rec["omega"] = [float(a[m, 3].mean()), float(a[m, 3].std())]
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
