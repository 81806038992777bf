#!/usr/bin/env python3
"""Timesteps/s of the Leeds Spherical Dynamo per-timestep path on one MI355X.

Workload: the full IMEX timestep (predictor + correctors until
tim_check_cgce accepts) of a Boussinesq MHD shell with codensity and
composition at L = M = 170, N = 96, Th x Ph = 256 x 512, E = 1e-4, Pm = 2,
dt = 2e-6, from a seeded synthetic state resident in HBM.

``value`` is the mixed-precision form (FP32 Legendre on MFMA, FP32 FFTs and
products, FP32 solves with an FP64-residual refinement sweep, FP64 state);
the plain FP64 form runs the same steps from the same state for the ratio,
and the per-field relative error of the mixed form against it after all
steps is reported.  ``roofline`` reports the mixed form's dominant kernel
kernel, the fused physical-space pass (k_physfuse: phi FFTs, products and
the equatorial fold in one kernel, the largest single kernel of the step;
the other classes each span several kernels), timed live with HIP events on
the context's stream over the timed region; ``traffic`` is its HBM bytes
per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE counters recorded in
profiles/traffic.json.  ``cpu_baseline`` times the CPU oracle on the host
for one timestep.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "leeds-code_amd"))

PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, state, threads):
    """One timestep of the CPU oracle (the checker) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle_lib
        o = oracle_lib.Oracle(cfg, threads=threads)
    except Exception as e:  # oracle not built on this host
        log("cpu_baseline skipped:", e)
        return None
    o.set_state(state)
    t0 = time.perf_counter()
    st = o.step(1)
    dt = time.perf_counter() - t0
    return {"value": st.steps / dt, "unit": "timesteps/s", "cores": threads, "kind": "port",
            "sample": "1 timestep (%d nonlinear evaluations) of the same L=M=170, N=96 workload, "
                      "FP64 C oracle, OpenMP over shells and harmonics" % st.evaluations}


def load_traffic(kernel_class):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_class)
    except (OSError, ValueError):
        return None


def per_class(ctx, names):
    """Kernel time of each class over one step each."""
    out = {}
    for k, name in enumerate(names):
        ctx.profile_begin(k)
        ctx.step(1)
        ms, n, w = ctx.profile_end()
        out[name] = {"ms": ms, "launches": n, "work": w}
    return out


def timed(ctx, steps, cls, sync_dev):
    ctx.sync()
    if sync_dev:
        sync_dev()
    if cls is not None:
        ctx.profile_begin(cls)
    t0 = time.perf_counter()
    st = ctx.step(steps)
    ctx.sync()
    if sync_dev:
        sync_dev()
    t1 = time.perf_counter()
    prof = ctx.profile_end() if cls is not None else None
    return t1 - t0, st, prof


def field_errors(a, b, cfg):
    import numpy as np
    from lsd.initial import FIELD_ID
    out = {}
    for name in ("uT", "uP", "C", "X", "BT", "BP"):
        x, y = a.get_coll(FIELD_ID[name]), b.get_coll(FIELD_ID[name])
        num = np.sqrt(sum(((p - q) ** 2).sum() for p, q in zip(x, y)))
        den = np.sqrt(sum((q ** 2).sum() for q in y))
        out[name] = float(num / den) if den > 0 else float(num)
    return out


def copy_calibration():
    """Five copies of 1 GiB between two device buffers: each dispatch of the
    copy kernel reads and writes exactly 2^30 bytes, which scales what the
    hardware counters report for the other kernels."""
    import torch
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 5}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 1 GiB (the reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the per-timestep path runs on a single MI355X "
              "(multi-GPU sharding is design-only); rerun with --gpus 1", file=sys.stderr)
        return 2

    from lsd import config3, seeded_state, SEEDS, AMPS, PLAIN_F64, MIXED_F32
    from lsd.gpu import Context, K_NAMES, K_PHYS

    try:
        import torch
        sync_dev = torch.cuda.synchronize if torch.cuda.is_available() else None
    except Exception:
        sync_dev = None

    cfg = config3(MIXED_F32)
    H, Th, Ph, Th2 = cfg.sizes()
    state = seeded_state(cfg, SEEDS[3], AMPS[3])

    mixed = Context(cfg)
    mixed.set_state(state)
    mixed.step(args.warmup)
    pc = per_class(mixed, K_NAMES)
    dom = K_PHYS
    log("mixed per-class kernel ms per step:", {k: round(v["ms"], 4) for k, v in pc.items()})
    elapsed, st, (kms, nl, work) = timed(mixed, args.steps, dom, sync_dev)
    value = args.steps / elapsed
    gel, _, _ = timed(mixed, args.steps, None, sync_dev)
    log("mixed: %.2f timesteps/s with per-launch events (eager), %.2f from graph replay" %
        (value, args.steps / gel))
    nsteps_total = args.warmup + len(K_NAMES) + 2 * args.steps

    plain_out, errs = None, None
    if not args.no_plain:
        plain = Context(config3(PLAIN_F64))
        plain.set_state(state)
        plain.step(args.warmup)
        ppc = per_class(plain, K_NAMES)
        log("plain per-class kernel ms per step:", {k: round(v["ms"], 4) for k, v in ppc.items()})
        pel, pst, _ = timed(plain, args.steps, None, sync_dev)
        pgel, _, _ = timed(plain, args.steps, None, sync_dev)
        errs = field_errors(mixed, plain, cfg)
        log("plain: %.2f and %.2f timesteps/s in two graph-replay runs" % (args.steps / pel, args.steps / pgel))
        log("mixed vs plain FP64 after %d steps:" % nsteps_total, {k: "%.2e" % v for k, v in errs.items()})
        plain_out = {"value": args.steps / pel, "ms_per_step": 1e3 * pel / args.steps,
                     "evaluations_per_step": pst.evaluations / max(pst.steps, 1),
                     "kernel_ms_per_step": {k: v["ms"] for k, v in ppc.items()}}
        plain.close()

    avg_s = (kms / 1e3) / max(nl, 1)
    per_launch = work / max(nl, 1)  # algorithmic bytes: Fourier fields in, folded fields out
    achieved = per_launch / avg_s / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS}
    roof["kernel"] = "k_physfuse (class %s)" % K_NAMES[dom]
    roof["algorithmic_bytes_per_launch"] = per_launch
    roof["launches"] = nl
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["traffic"] = load_traffic(K_NAMES[dom])

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(config3(PLAIN_F64), state, threads)

    out = {
        "metric": "timesteps/s at L=170,N=96: mixed-precision vs plain FP64 GPU, <=1e-6 rel err",
        "value": value,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms and solves, f64 state and refinement",
        "data": "synthetic: seeded (l+1)^-2 spectrum, smooth radial shapes, resident in HBM",
        "config": {"workload": "Boussinesq MHD IMEX timestep with codensity and composition, mixed FP32",
                   "L": cfg.L, "M": cfg.M, "N": cfg.N, "Th": Th, "Ph": Ph, "H": H, "E": cfg.E,
                   "Pm": cfg.Pm, "dt": cfg.dt, "evaluations_per_step": st.evaluations / max(st.steps, 1),
                   "parallelism": "single GPU"},
        "plain_f64": plain_out,
        "speedup_vs_plain_f64": (value / plain_out["value"]) if plain_out else None,
        "rel_err_vs_plain_f64": errs,
        "rel_err_after_steps": nsteps_total if errs else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: v["ms"] for k, v in pc.items()},
    }
    print(json.dumps(out), flush=True)
    mixed.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
