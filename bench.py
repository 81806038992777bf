#!/usr/bin/env python3
"""Timesteps/s of the Leeds Spherical Dynamo per-timestep path on one MI355X.

Workload: the full IMEX timestep (predictor + correctors until
tim_check_cgce accepts) of a Boussinesq MHD shell with codensity and
composition at L = M = 170, N = 96, Th x Ph = 256 x 512, E = 1e-4, Pm = 2,
dt = 2e-6, from a seeded synthetic state resident in HBM.

Prints one JSON line.  ``roofline`` reports the dominant kernel class,
timed live with HIP events on the context's stream over the timed region;
``cpu_baseline`` times the CPU oracle on the host for one timestep.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "leeds-code_amd"))

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the per-timestep path runs on a single MI355X "
              "(multi-GPU sharding is design-only); rerun with --gpus 1", file=sys.stderr)
        return 2

    from lsd import config3, seeded_state, SEEDS, AMPS, PLAIN_F64
    from lsd.gpu import Context, K_NAMES, K_SYNTH, K_ANAL

    cfg = config3(PLAIN_F64)
    H, Th, Ph, Th2 = cfg.sizes()
    state = seeded_state(cfg, SEEDS[3], AMPS[3])
    ctx = Context(cfg)
    ctx.set_state(state)

    ctx.step(args.warmup)
    ctx.sync()
    # per-class kernel time over one step each, to pick the dominant class
    per_class = {}
    for k, name in enumerate(K_NAMES):
        ctx.profile_begin(k)
        ctx.step(1)
        ms, n, w = ctx.profile_end()
        per_class[name] = {"ms": ms, "launches": n, "work": w}
    dom = max(range(len(K_NAMES)), key=lambda k: per_class[K_NAMES[k]]["ms"])
    log("per-class kernel ms per step:", {k: round(v["ms"], 4) for k, v in per_class.items()})

    try:
        import torch
        sync_dev = torch.cuda.synchronize if torch.cuda.is_available() else None
    except Exception:
        sync_dev = None

    ctx.sync()
    if sync_dev:
        sync_dev()
    ctx.profile_begin(dom)
    t0 = time.perf_counter()
    st = ctx.step(args.steps)
    ctx.sync()
    if sync_dev:
        sync_dev()
    t1 = time.perf_counter()
    kms, nl, work = ctx.profile_end()
    elapsed = t1 - t0
    value = args.steps / elapsed

    compute_bound = dom in (K_SYNTH, K_ANAL)
    avg_s = (kms / 1e3) / max(nl, 1)
    per_launch = work / max(nl, 1)
    if compute_bound:
        achieved = per_launch / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS}
    else:
        achieved = per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS}
    roof["kernel"] = K_NAMES[dom]
    roof["launches"] = nl
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["traffic"] = load_traffic(K_NAMES[dom])

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(cfg, state, threads)

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
        "dtype": "f64",
        "data": "synthetic: seeded (l+1)^-2 spectrum, smooth radial shapes, resident in HBM",
        "config": {"workload": "Boussinesq MHD IMEX timestep with codensity and composition, plain FP64",
                   "L": cfg.L, "M": cfg.M, "N": cfg.N, "Th": Th, "Ph": Ph, "H": H, "E": cfg.E,
                   "Pm": cfg.Pm, "dt": cfg.dt, "evaluations_per_step": st.evaluations / max(st.steps, 1),
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: v["ms"] for k, v in per_class.items()},
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
