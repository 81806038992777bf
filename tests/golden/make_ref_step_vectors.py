#!/usr/bin/env python3
"""Writes tests/golden/ref_step_<case>.npz: per-step golden vectors of the
reference's own FFTW/LAPACK CPU path, built from its unmodified sources by
oracle/ref_build.py and driven through program/main.F90:100-164 by
oracle/ref_step.F90 (run in the build container, where the reference
sources are; the .npz files are data and travel to every machine).

Each case starts from the seeded state of lsd.initial.seeded_state (the
seed and amplitude are stored) and records, for every step, the number of
nonlinear evaluations (1 predictor + the correctors tim_check_cgce asked
for), tim_dterr of the accepted corrector and tim_t.  Fields are kept
  * in full at the steps in FULL[case];
  * as a fixed sample of (harmonic, radius) entries, the per-degree norms
    sqrt(sum |c|^2) and the total norm at the steps in SAMPLED[case]
    (full fields at L = M = 170, N = 96 would be 134 MB per step).
Step 0 also keeps the radial mesh arrays (r, dr(1..3), radLap, r1dr,
intr2dr, modules/meshs.F90:292-347) and the quadrature nodes and weights
(leg_gp, leg_gw) with the tables leg_P and leg_gwP (full below L = 64).

Usage: python tests/golden/make_ref_step_vectors.py [case ...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "leeds-code_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ref_io  # noqa: E402
from lsd.config import config1, config3, config_small, christensen_case1, christensen_case2  # noqa: E402
from lsd.initial import AMPS, SEEDS, harmonics, seeded_state  # noqa: E402

CASES = {
    # name: (config, seed, amplitude, steps)
    "small": (config_small, 7, 1e-3, 20),
    "cfg1": (config1, SEEDS[1], AMPS[1], 100),
    "cfg3": (config3, SEEDS[3], AMPS[3], 2),
    # adaptive step (d_timestep < 0): non_maxtstep and tim_new_tstep choose
    # dt at every step; tim_t records the steps taken
    "adapt": (lambda: config_small(dt=-1e-4), 9, 1e-3, 40),
    # the same with i_time_step_control = 1 (tim_corr_dt = tim_dterr, dt
    # halved above d_corr_upper and grown by 1.273974 below d_corr_lower)
    "adapt1": (lambda: config_small(dt=-1e-4, step_control=1), 10, 1e-3, 40),
    # the Christensen case-1 dynamo (N = 64, L = 64, M = 16, Mp = 4) from the
    # reference's test state tests/state.cdf.in (tests/golden/state_case1.cdf.in),
    # adaptive step under control 1; seed -1 marks a state read from a file
    "case1": (christensen_case1, -1, 0.0, 30),
    # a conducting inner core of 8 points that rotates (i_Nic = 8,
    # b_rot_tstep): seeded inner-core fields (seed + 1) and rotation rate
    # OMEGA0[case] at the start
    "ic": (lambda: config_small(Nic=8, rot=1), 5, 1e-2, 20),
    # the Christensen case-2 benchmark (conducting, rotating inner core of
    # 16 points) from the case-1 test state, its field extended into the
    # inner core as io_ic_extend does (program/io.F90:1438-1477)
    "case2": (christensen_case2, -1, 0.0, 10),
    # stress-free walls (vel_bc_Tor, vel_bc_Pol): outer (2), inner (3), both
    # (4, with vel_angmomclr removing the rigid rotations)
    "sf2": (lambda: config_small(vel_bc=2), 12, 1e-2, 20),
    "sf3": (lambda: config_small(vel_bc=3), 13, 1e-2, 20),
    "sf4": (lambda: config_small(vel_bc=4), 14, 1e-2, 20),
}
OMEGA0 = {"ic": 0.3, "case2": 0.0}
FULL = {"small": (1, 2, 20), "cfg1": (1,), "cfg3": (), "adapt": (1, 10, 40), "adapt1": (1, 20, 40),
        "case1": (), "ic": (1, 2, 20), "case2": (), "sf2": (1, 20), "sf3": (1, 20), "sf4": (1, 2, 20)}
SAMPLED = {"small": (), "cfg1": (10, 100), "cfg3": (1, 2), "adapt": (), "adapt1": (), "case1": (1, 10, 30),
           "ic": (), "case2": (1, 10), "sf2": (), "sf3": (), "sf4": ()}
STATE_FILE = {"case1": "state_case1.cdf.in", "case2": "state_case1.cdf.in"}
# cases whose fixtures also record the reference's own per-step spread
SPREAD = ("adapt", "adapt1", "case1", "case2")
NSAMPLE = 4096


def sample_index(H, N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    flat = rng.choice(H * N, size=min(NSAMPLE, H * N), replace=False)
    flat.sort()
    return (flat // N).astype(np.int32), (flat % N).astype(np.int32)


def make(case):
    mk, seed, amp, nsteps = CASES[case]
    cfg = mk()
    H, Th, Ph, Th2 = cfg.sizes()
    out = {"config": np.array(json.dumps(cfg.as_dict())), "seed": np.array(seed), "amp": np.array(amp),
           "nsteps": np.array(nsteps)}
    sys.path.insert(0, os.path.dirname(HERE))
    import ref_fixtures
    if case in STATE_FILE:
        out["state_file"] = np.array(STATE_FILE[case])
        out["b_scale"] = np.array(ref_fixtures.B_SCALE_OLD_STATE)
    if case in OMEGA0:
        out["omega0"] = np.array(OMEGA0[case])
    state = ref_fixtures.initial_state(out, cfg)
    recs, _ = ref_io.run(case, cfg, state, nsteps, 1)
    if case in SPREAD:
        # the reference's own rounding spread: the same run on MKL's
        # COMPATIBLE code path, per step the largest relative difference of
        # a field from the run above
        env = os.environ.get("MKL_CBWR")
        os.environ["MKL_CBWR"] = "COMPATIBLE"
        try:
            alt, _ = ref_io.run(case, cfg, state, nsteps, 1)
        finally:
            if env is None:
                del os.environ["MKL_CBWR"]
            else:
                os.environ["MKL_CBWR"] = env
        spread = []
        for a, b in zip(recs[1:], alt[1:]):
            worst = 0.0
            for name, (re, im) in a["fields"].items():
                re2, im2 = b["fields"][name]
                den = np.sqrt((re ** 2).sum() + (im ** 2).sum())
                if den > 0:
                    worst = max(worst, np.sqrt(((re - re2) ** 2).sum() + ((im - im2) ** 2).sum()) / den)
            spread.append(worst)
        out["spread"] = np.array(spread)
        out["dterr_spread"] = np.array([abs(a["dterr"] - b["dterr"]) for a, b in zip(recs[1:], alt[1:])])
        if ref_io.has_ic(cfg):
            for key in ("omega", "velTorq", "magTorq"):
                out[key + "_spread"] = np.array([abs(a[key] - b[key]) for a, b in zip(recs[1:], alt[1:])])
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    st = recs[1:]
    out["step"] = np.array([r["step"] for r in st], dtype=np.int32)
    out["evaluations"] = np.array([r["evaluations"] for r in st], dtype=np.int32)
    out["dterr"] = np.array([r["dterr"] for r in st])
    out["t"] = np.array([r["t"] for r in st])
    if ref_io.has_ic(cfg):
        for key in ("omega", "velTorq", "magTorq"):
            out[key] = np.array([r[key] for r in st])
    r0 = recs[0]
    for k, v in r0["mesh"].items():
        out["mesh_" + k] = v
    out["gp"] = r0["tables"]["gp"]
    out["gw"] = r0["tables"]["gw"]
    if cfg.L < 64:
        out["P"] = r0["tables"]["P"]
        out["gwP"] = r0["tables"]["gwP"]
    hs, ns = sample_index(H, cfg.N, seed + 1)
    out["sample_h"] = hs
    out["sample_n"] = ns
    by_step = {r["step"]: r for r in st}
    for s in FULL[case]:
        for name, (re, im) in by_step[s]["fields"].items():
            out["s%d_%s_re" % (s, name)] = re
            out["s%d_%s_im" % (s, name)] = im
    for s in SAMPLED[case]:
        for name, (re, im) in by_step[s]["fields"].items():
            if name.startswith("ic"):
                # the inner-core fields are small: kept whole
                out["s%d_%s_re" % (s, name)] = re
                out["s%d_%s_im" % (s, name)] = im
                continue
            out["q%d_%s_re" % (s, name)] = re[hs, ns]
            out["q%d_%s_im" % (s, name)] = im[hs, ns]
            p = (re ** 2 + im ** 2).sum(axis=1)
            out["q%d_%s_norm" % (s, name)] = np.array(np.sqrt(p.sum()))
            out["q%d_%s_lnorm" % (s, name)] = np.sqrt(np.bincount(l, weights=p, minlength=cfg.L))
    path = os.path.join(HERE, "ref_step_%s.npz" % case)
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes;", "evaluations per step", out["evaluations"].tolist())


def main(argv):
    for c in argv or list(CASES):
        if not ref_io.available(c):
            raise SystemExit("oracle/_ref/%s is not built (python oracle/ref_build.py %s)" % (c, c))
        make(c)


if __name__ == "__main__":
    main(sys.argv[1:])
