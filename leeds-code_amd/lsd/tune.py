"""Autotuner driver: runs lsd_tune (the library's timing of its kernel
variants, include/lsd.h) for the configurations this build is used at and
appends the choices to a tuning cache, which lsd_create reads.

    python -m lsd.tune [--out PATH] [--reps R] [--only NAME ...]

Default output: tuning_cache.jsonl beside liblsdgpu.so (the file lsd_create
looks for when LSD_TUNING_CACHE is unset).  Run on the GPU the library will
run on; each configuration takes a few seconds.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def configurations():
    from .config import config1, config3, config_small, christensen_case1, make_config, grid_for, MIXED_F32, PLAIN_F64
    out = {}
    for prec, tag in ((MIXED_F32, "mixed"), (PLAIN_F64, "plain")):
        out["config3_" + tag] = config3(prec)
        out["config1_" + tag] = config1(prec)
        # the ensemble's truncations (L = M even, N from {32, 48, 64})
        for L in (32, 48, 64, 86):
            th, ph = grid_for(L, L)
            for N in (32, 48, 64):
                out["ens_L%d_N%d_%s" % (L, N, tag)] = make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=0,
                                                                  precision=prec)
    out["case1_plain"] = christensen_case1()
    out["small_mixed"] = config_small(precision=MIXED_F32)
    # the largest grid of the mixed-form tests (lmax 255, 384 x 768, N = 128)
    base = config3()
    out["L256_mixed"] = make_config(L=256, M=256, N=128, Th=384, Ph=768, comp=1, Ra=base.Ra, Ra_comp=base.Ra_comp,
                                    Pr=1.0, Sc=1.0, Pm=2.0, E=1e-4, dt=2e-6, precision=MIXED_F32)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(HERE), "tuning_cache.jsonl"))
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args(argv)
    from .gpu import Context
    from .initial import seeded_state
    cfgs = configurations()
    names = args.only or list(cfgs)
    for name in names:
        cfg = cfgs[name]
        g = Context(cfg)
        g.set_state(seeded_state(cfg, 1, 1e-3))
        g.step(1)
        t = g.tune(args.reps, args.out)
        print(name, g.variants(), {k: (round(a, 1), round(b, 1)) for k, (a, b) in t.items()}, flush=True)
        g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
