"""Onset scan: the critical Rayleigh number of thermal convection in the
rotating shell, measured with the GPU timestep instead of taken from a
scaling law.

A non-magnetic run (mag = 0, composition off) starts from the conduction
profile plus a perturbation of 1e-6, small enough that the products stay
below rounding, so the step is linear in the perturbation.  After the
decaying modes have gone, the kinetic energy grows or decays as
exp(2 sigma t) with sigma the growth rate of the fastest mode; sigma is the
least-squares slope of log(KE) / 2 over the second half of the run.  The
scan takes sigma at a few Rayleigh numbers around an estimate and puts the
onset where the linear fit of sigma(Ra) through the points nearest zero
crosses it.  Results are modified Rayleigh numbers Ra_mod = d_Ra r_o E, the
scaling of Christensen et al. (2001), whose benchmark quotes Ra_mod = 55.91
at E = 1e-3, Pr = 1, r_i / r_o = 0.35
(Benchmark_files/Boussinesq/Christensen_case0_benchmark/INFO:1-5 gives the
conversion d_Ra = Ra_mod / (r_o E)); the E = 1e-3 scan checks the method
against that figure.

    python -m lsd.onset [--E 1e-3 1e-4]

prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

# Resolution, step and run length per Ekman number, with the scanned
# multiples of the E^(-1/3) estimate: the onset mode has m ~ 4 at E = 1e-3
# and a radial scale ~ E^(1/3), so these truncations resolve it with room to
# spare.  `fine` repeats two scanned points, the second below and the
# second above the onset, with finer resolution and half the step; the zero
# of the line through those two is the result, and its distance from the
# first shows the resolution error.
SCAN = {
    1e-3: dict(L=32, M=32, N=32, dt=1e-4, t_end=2.0, chunk=100,
               factors=(0.9, 0.95, 1.0, 1.05, 1.1, 1.2), fine=dict(L=48, M=48, N=64, dt=5e-5)),
    1e-4: dict(L=64, M=64, N=48, dt=1e-5, t_end=0.6, chunk=100,
               factors=(0.5, 0.56, 0.57, 0.58, 0.59, 0.6, 0.65, 1.0), fine=dict(L=96, M=96, N=72, dt=5e-6)),
}

# The fit stops at this kinetic energy: the saturated states of these runs
# have KE ~ 1e3, so the products are still 1e-6 of the linear terms there.
KE_LINEAR = 1e-9


def cfg_for(E, ra_mod, L, M, N, dt, Pm=1.0):
    from .config import make_config, grid_for, unified_ra
    th, ph = grid_for(L, M)
    return make_config(L=L, M=M, N=N, Th=th, Ph=ph, comp=0, mag=0, E=E, Ra=unified_ra(ra_mod, E),
                       Pr=1.0, Pm=Pm, dt=dt)


def growth_rate(E, ra_mod, L, M, N, dt, t_end, chunk, seed=11, amp=1e-10):
    """sigma (per unit time of the code) of the kinetic energy at Ra_mod:
    the slope of log(KE) / 2 over the last half of the samples taken before
    t_end or before KE reaches KE_LINEAR."""
    from .gpu import Context
    from .initial import seeded_state
    cfg = cfg_for(E, ra_mod, L, M, N, dt)
    t, ke = [], []
    with Context(cfg) as g:
        g.set_state(seeded_state(cfg, seed, amp))
        for k in range(int(round(t_end / (chunk * dt)))):
            g.step(chunk)
            e, _, _ = g.energy(0)
            if e > KE_LINEAR:
                break
            t.append((k + 1) * chunk * dt)
            ke.append(e)
    if len(t) < 16:
        raise RuntimeError("onset scan: only %d linear samples at Ra_mod = %g; lower amp or chunk" % (len(t), ra_mod))
    t, ke = np.array(t), np.array(ke)
    h = len(t) // 2
    slope = np.polyfit(t[h:], np.log(ke[h:]), 1)[0]
    return 0.5 * slope, float(t[-1])


def onset(ra, sig):
    """Zero of the straight line through the two points on either side of
    sigma = 0 (the last below and the first above); None without a sign
    change."""
    ra, sig = np.asarray(ra), np.asarray(sig)
    o = np.argsort(ra)
    ra, sig = ra[o], sig[o]
    for i in range(len(ra) - 1):
        if sig[i] < 0 <= sig[i + 1]:
            return float(ra[i] - sig[i] * (ra[i + 1] - ra[i]) / (sig[i + 1] - sig[i]))
    return None


def main(argv=None):
    from .config import RA_MOD_CRIT_E1EM3
    ap = argparse.ArgumentParser()
    ap.add_argument("--E", type=float, nargs="+", default=[1e-3, 1e-4])
    args = ap.parse_args(argv)
    t0 = time.time()
    out = {"scan": "onset of convection, Pr = 1, r_i/r_o = 0.35, fixed-codensity and no-slip walls, mag = 0"}
    for E in args.E:
        p = dict(SCAN[E])
        factors, fine = p.pop("factors"), p.pop("fine")
        est = RA_MOD_CRIT_E1EM3 * (1e-3 / E) ** (1.0 / 3.0)
        pts = []
        for f in factors:
            s, tl = growth_rate(E, f * est, **p)
            pts.append((f * est, s))
            print("[%.0f s] E=%g Ra_mod=%.3f sigma=%.4g (linear to t=%.3g)" % (time.time() - t0, E, f * est, s, tl),
                  file=sys.stderr, flush=True)
        rc = onset([a for a, _ in pts], [b for _, b in pts])
        if rc is None:
            raise RuntimeError("onset scan: no sign change of sigma at E = %g" % E)
        pf = dict(p, **fine)
        ras = sorted(a for a, _ in pts)
        below = [a for a in ras if a < rc]
        above = [a for a in ras if a > rc]
        lo = below[-2] if len(below) > 1 else below[-1]
        hi = above[1] if len(above) > 1 else above[0]
        fpts = []
        for ra in (lo, hi):
            s, _ = growth_rate(E, ra, **pf)
            fpts.append((ra, s))
            print("[%.0f s] E=%g fine Ra_mod=%.3f sigma=%.4g" % (time.time() - t0, E, ra, s), file=sys.stderr, flush=True)
        (a0, s0), (a1, s1) = fpts
        rf = float(a0 - s0 * (a1 - a0) / (s1 - s0))
        out["E=%g" % E] = {"estimate": est, "onset": rf, "onset_coarse": rc,
                           "resolution": p, "points": [[a, b] for a, b in pts],
                           "fine_resolution": pf, "fine_points": [[a, b] for a, b in fpts]}
    out["wall_s"] = time.time() - t0
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
