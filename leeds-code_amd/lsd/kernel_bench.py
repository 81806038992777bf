"""Isolated-kernel comparison of the precision forms (the workload of the
survey's second configuration): FP64 vs FP32 vs BF16x3 Legendre transforms,
and FP64 vs FP32-with-FP64-residual-refinement banded solves.

  * SHT: a scalar field and the (q, s, t) of a vector field, seeded Gaussian
    coefficients with (l + 1)^-2 decay, synthesis then analysis
    (lsd_sht_roundtrip) at L = M = 256 (lmax 255), 384 x 768, N = 128; kernel
    time per round trip from the library's per-class timing, and the
    relative error of the round trip (sqrt(sum |c_out - c_in|^2 / sum
    |c_in|^2) over all (l, m, r)).
  * Banded solves: the predictor solve x* = X^-1 (Y x + N) of every field at
    L = M = 256, N = 128, dt = 1e-5, d_implicit = 0.5 (operators
    c1 / dt - c2 c (radLap - l(l+1)/r^2) with the reference's boundary rows),
    with the same state and N in every form; the FP64 form against the mixed
    form's FP32 factors with one FP64-residual refinement sweep.

    python -m lsd.kernel_bench [--reps R]

Prints one JSON line.
"""
import argparse
import json
import sys

import numpy as np


def seeded_coeffs(cfg, seed, nf):
    from .initial import harmonics
    H = cfg.sizes()[0]
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(nf):
        w = ((l + 1.0) ** -2)[:, None]
        re = rng.standard_normal((H, cfg.N)) * w
        im = rng.standard_normal((H, cfg.N)) * w
        im[m == 0] = 0.0
        re[l == 0] = 0.0 if k > 0 else re[l == 0]  # s and t have no l = 0 part
        im[l == 0] = 0.0
        out.append((re, im))
    return out


def rel(a, b):
    num = sum(((x[0] - y[0]) ** 2).sum() + ((x[1] - y[1]) ** 2).sum() for x, y in zip(a, b))
    den = sum((y[0] ** 2).sum() + (y[1] ** 2).sum() for y in b)
    return float(np.sqrt(num / den))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    from .config import make_config, PLAIN_F64, MIXED_F32, MIXED_BF16X3
    from .gpu import Context, K_SYNTH, K_ANAL, K_SOLVE, V_SOLVE
    from .initial import seeded_state

    def cfg(prec):
        return make_config(L=256, M=256, N=128, Th=384, Ph=768, comp=1, E=1e-4, Pm=2.0, dt=1e-5,
                           precision=prec)

    c0 = cfg(PLAIN_F64)
    fields = {0: seeded_coeffs(c0, 20261202, 1), 1: seeded_coeffs(c0, 20261203, 3)}
    out = {"workload": "L=M=256 (lmax 255), 384 x 768, N=128", "sht": {}, "solve": {}}
    for prec, name in ((PLAIN_F64, "fp64"), (MIXED_F32, "fp32"), (MIXED_BF16X3, "bf16x3")):
        g = Context(cfg(prec))
        rec = {}
        for kind, kname in ((0, "scalar"), (1, "vector")):
            res = g.sht_roundtrip(kind, fields[kind])  # warm-up
            times = {}
            for cls, cn in ((K_SYNTH, "synthesis_ms"), (K_ANAL, "analysis_ms")):
                g.profile_begin(cls)
                for _ in range(args.reps):
                    res = g.sht_roundtrip(kind, fields[kind])
                ms, n, w = g.profile_end()
                times[cn] = ms / args.reps
                times[cn.replace("_ms", "_tflops")] = w / (ms * 1e-3) / 1e12 if ms > 0 else None
            rec[kname] = dict(times, rel_err=rel(res, fields[kind]))
            print(name, kname, rec[kname], file=sys.stderr, flush=True)
        out["sht"][name] = rec
        g.close()
    # banded solves: the predictor with the same state and N in both forms
    st = seeded_state(c0, 20261202, 1e-2)
    nf = seeded_state(c0, 20261212, 1.0)
    res = {}
    for prec, name in ((PLAIN_F64, "fp64"), (MIXED_F32, "fp32_refined")):
        g = Context(cfg(prec))
        if prec == MIXED_F32:
            g.set_variant(V_SOLVE, 1)  # the FP32 factors (the mixed default at N = 128 is FP64)
        g.set_state(st)
        names = ("uT", "uP", "C", "X", "BT", "BP")
        for k, n in enumerate(names):
            g.set_coll(6 + k, *nf[n])
        g.profile_begin(K_SOLVE)
        g.predictor()
        ms, nl, w = g.profile_end()
        res[name] = ([g.get_coll(k) for k in range(6)], ms, w)
        g.close()
    out["solve"] = {"fp64_ms": res["fp64"][1], "fp32_refined_ms": res["fp32_refined"][1],
                    "fp64_gbs": res["fp64"][2] / (res["fp64"][1] * 1e-3) / 1e9,
                    "fp32_refined_gbs": res["fp32_refined"][2] / (res["fp32_refined"][1] * 1e-3) / 1e9,
                    "rel_err_fp32_refined_vs_fp64": rel(res["fp32_refined"][0], res["fp64"][0])}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
