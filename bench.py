#!/usr/bin/env python3
"""Timesteps/s of the Leeds Spherical Dynamo per-timestep path on one MI355X.

Workload: the full IMEX timestep (predictor + correctors until
tim_check_cgce accepts) of a Boussinesq MHD shell with codensity and
composition at L = M = 170, N = 96, Th x Ph = 256 x 512, E = 1e-4, Pm = 2,
dt = 2e-6, from a seeded synthetic state resident in HBM.

``value`` is the mixed-precision form replayed from its captured HIP graphs:
FP32 Legendre contractions on MFMA, FP32 phi FFTs and nonlinear products;
the state, the radial stencils, the banded solves (at N = 96 the two-pass
FP64 solve kernels) and the time stepping stay FP64.  The plain FP64 form
runs the same number of steps by graph replay for the ratio.  The per-field
relative error of the mixed form against the plain FP64 form is measured
after 1 and after 1000 steps from the same seeded state.  ``roofline``
reports the largest single kernel of the mixed step, the fused
physical-space pass (k_physfuse: phi FFTs, products and the equatorial
fold), timed with HIP events on the context's stream in a separate eager
run of the same steps; ``traffic`` is its HBM bytes per launch from the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters recorded in
profiles/traffic.json.  ``headline_roofline`` is t_roof / t_step with
t_roof = Legendre flops / matrix peak + compulsory bytes / HBM peak;
``legendre_mfma`` is the Legendre kernels' flops / time against the FP32
matrix peak.  ``cpu_baseline`` times the reference's own FFTW/LAPACK CPU
path (its unmodified Fortran modules built from source with MKL, serial as
the reference's _Nth = _Nr = 1 build) on the host, with the C oracle on all
host threads beside it.
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
PEAK_FP32_TF = 157.3      # MI355X dense FP32 matrix
PEAK_FP64_TF = 78.6       # MI355X dense FP64 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable": avail, "model": model}


def cpu_baseline(cfg, state, threads, ref_steps):
    """The reference's own CPU path (oracle/_ref, built from its sources
    against MKL's FFTW3/LAPACK) timed over steps 2..ref_steps+1 on one core,
    and the C oracle on `threads` threads for one timestep."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    out = None
    try:
        import ref_io
        if ref_io.available("cfg3"):
            _, sec = ref_io.run("cfg3", cfg, state, ref_steps + 1, every=ref_steps + 1, timing=True, threads=1)
            out = {"value": 1.0 / sec, "unit": "timesteps/s", "cores": 1, "kind": "reference",
                   "sample": "timed steps 2..%d of the same L=M=170, N=96 workload from the same state: the reference's "
                             "unmodified Fortran modules (serial build) with MKL FFTW3/LAPACK on 1 thread"
                             % (ref_steps + 1)}
        else:
            log("reference executable not present")
    except Exception as e:
        log("reference CPU baseline skipped:", e)
    try:
        import oracle_lib
        o = oracle_lib.Oracle(cfg, threads=threads)
        o.set_state(state)
        t0 = time.perf_counter()
        st = o.step(1)
        dt = time.perf_counter() - t0
        port = {"value": st.steps / dt, "unit": "timesteps/s", "cores": threads, "kind": "port",
                "sample": "1 timestep (%d nonlinear evaluations), FP64 C oracle, OpenMP over shells and harmonics"
                          % st.evaluations}
    except Exception as e:  # oracle not built on this host
        log("oracle CPU baseline skipped:", e)
        port = None
    if out is None:
        out = port
    elif port is not None:
        out["port_all_threads"] = port
    if out is not None:
        out["host"] = host_cpu()
    return out


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


def headline(cfg, ms_per_step, evals, f32):
    """t_roof / t_step: Legendre flops at the matrix peak of the precision
    plus the fusion-invariant compulsory bytes (48 coll-field passes of
    state and N history, 8 Legendre tables per evaluation) at the HBM
    peak."""
    H, Th, Ph, Th2 = cfg.sizes()
    unit = 4.0 * H * Th2 * cfg.N
    flops = evals * 38 * unit
    word = 4 if f32 else 8
    bmin = 48 * H * cfg.N * 2 * word + evals * 8 * H * Th2 * word
    peak = PEAK_FP32_TF if f32 else PEAK_FP64_TF
    t_roof = flops / (peak * 1e12) + bmin / (PEAK_HBM_GBS * 1e9)
    return {"t_roof_ms": 1e3 * t_roof, "t_step_ms": ms_per_step, "frac": 1e3 * t_roof / ms_per_step,
            "legendre_flops": flops, "compulsory_bytes": bmin, "matrix_peak_tflops": peak}


def errors_after(cfg_m, cfg_p, state, marks):
    """Per-field relative error of the mixed form against the plain form
    after each step count in marks (fresh contexts from `state`)."""
    from lsd.gpu import Context
    gm, gp = Context(cfg_m), Context(cfg_p)
    gm.set_state(state)
    gp.set_state(state)
    out, done = {}, 0
    for k in marks:
        sm, sp = gm.step(k - done), gp.step(k - done)
        if sm.evaluations != sp.evaluations:
            log("corrector counts differ by step", k)
        done = k
        out[str(k)] = field_errors(gm, gp, cfg_m)
        log("mixed vs plain FP64 after %d steps:" % k, {n: "%.2e" % v for n, v in out[str(k)].items()})
    gm.close()
    gp.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--err-steps", type=int, default=1000,
                    help="steps after which the mixed-vs-plain error is reported (beside 1 step); 0 skips")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ref-steps", type=int, default=1, help="timed steps of the reference CPU path")
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
    from lsd.gpu import Context, K_NAMES, K_PHYS, K_SYNTH, K_ANAL

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
    log("mixed per-class kernel ms per step:", {k: round(v["ms"], 4) for k, v in pc.items()})
    elapsed, st, _ = timed(mixed, args.steps, None, sync_dev)
    value = args.steps / elapsed
    dom = K_PHYS
    eel, _, (kms, nl, work) = timed(mixed, args.steps, dom, sync_dev)
    log("mixed: %.2f timesteps/s from graph replay; %.2f with per-launch events on the %s kernel (eager)" %
        (value, args.steps / eel, K_NAMES[dom]))

    plain_out = None
    if not args.no_plain:
        plain = Context(config3(PLAIN_F64))
        plain.set_state(state)
        plain.step(args.warmup)
        ppc = per_class(plain, K_NAMES)
        log("plain per-class kernel ms per step:", {k: round(v["ms"], 4) for k, v in ppc.items()})
        pel, pst, _ = timed(plain, args.steps, None, sync_dev)
        log("plain: %.2f timesteps/s from graph replay" % (args.steps / pel))
        plain_out = {"value": args.steps / pel, "ms_per_step": 1e3 * pel / args.steps,
                     "evaluations_per_step": pst.evaluations / max(pst.steps, 1),
                     "kernel_ms_per_step": {k: v["ms"] for k, v in ppc.items()},
                     "headline_roofline": headline(cfg, 1e3 * pel / args.steps,
                                                   pst.evaluations / max(pst.steps, 1), False)}
        plain.close()

    errs = None
    if args.err_steps > 0 and not args.no_plain:
        marks = [1] + ([args.err_steps] if args.err_steps > 1 else [])
        errs = errors_after(cfg, config3(PLAIN_F64), state, marks)

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

    evals = st.evaluations / max(st.steps, 1)
    leg_flops = pc["synth"]["work"] + pc["anal"]["work"]
    leg_ms = pc["synth"]["ms"] + pc["anal"]["ms"]
    leg = {"flops_per_step": leg_flops, "ms_per_step": leg_ms,
           "achieved_tflops": leg_flops / (leg_ms * 1e-3) / 1e12 if leg_ms > 0 else None,
           "peak_tflops": PEAK_FP32_TF}
    if leg["achieved_tflops"]:
        leg["frac"] = leg["achieved_tflops"] / PEAK_FP32_TF
        for name in ("synth", "anal"):
            v = pc[name]
            leg[name] = {"ms": v["ms"], "tflops": v["work"] / (v["ms"] * 1e-3) / 1e12,
                         "frac": v["work"] / (v["ms"] * 1e-3) / 1e12 / PEAK_FP32_TF}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(config3(PLAIN_F64), state, threads, args.ref_steps)

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
        "dtype": "f32 Legendre (MFMA), phi FFTs and products; f64 state, radial stencils, solves (N=96) "
                 "and time stepping",
        "data": "synthetic: seeded (l+1)^-2 spectrum, smooth radial shapes, resident in HBM",
        "config": {"workload": "Boussinesq MHD IMEX timestep with codensity and composition, mixed FP32, "
                               "graph replay",
                   "L": cfg.L, "M": cfg.M, "N": cfg.N, "Th": Th, "Ph": Ph, "H": H, "E": cfg.E,
                   "Pm": cfg.Pm, "dt": cfg.dt, "evaluations_per_step": evals,
                   "parallelism": "single GPU"},
        "plain_f64": plain_out,
        "speedup_vs_plain_f64": (value / plain_out["value"]) if plain_out else None,
        "rel_err_vs_plain_f64": errs,
        "roofline": roof,
        "headline_roofline": headline(cfg, 1e3 * elapsed / args.steps, evals, True),
        "legendre_mfma": leg,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: v["ms"] for k, v in pc.items()},
    }
    print(json.dumps(out), flush=True)
    mixed.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
