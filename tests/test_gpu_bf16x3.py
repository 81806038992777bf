"""The BF16x3 split-precision Legendre form (LSD_MIXED_BF16X3): products as
hi hi + hi lo + lo hi of bf16 splits on v_mfma_f32_16x16x32_bf16 with FP32
accumulation (about 16 significant bits per product).

  * Transforms alone (lsd_sht_roundtrip: synthesis then analysis) on
    seeded (l + 1)^-2 coefficients: the FP64 round trip equals the oracle's
    tra_spec2phys -> tra_phys2spec (tra_qst2rtp -> tra_rtp2qst) to 1e-13.
    Neither recovers the input exactly: the reference's Gauss nodes stop
    their Newton iteration at 1e-13, which leaves about 1e-12 in the m = 0
    modes (8e-12 at L = 64, 1e-12 at L = 30, the same in both).  The FP32
    round trip recovers the input to 2e-6 (measured 8.6e-7 at L = 64) and
    BF16x3 to 1e-4 (measured 2.4e-5), less accurately than FP32.
  * Timesteps: BF16x3 against the plain FP64 form after 1 and 10 steps at
    L = M = 32, N = 32, and after 1 and 1000 steps at L = M = 170, N = 96,
    within 1e-3 (measured values printed; the mixed-form bound 1e-6 is the
    FP32 form's)."""
import numpy as np
import pytest

from lsd import config1, config3, make_config, grid_for, seeded_state, PLAIN_F64, MIXED_F32, MIXED_BF16X3
from lsd import SEEDS, AMPS
from lsd.gpu import Context
from lsd.kernel_bench import seeded_coeffs, rel
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,N", [(64, 32), (30, 40)])
def test_roundtrip_precision_forms(L, N):
    th, ph = grid_for(L, L)
    errs = {}
    outs = {}
    for prec in (PLAIN_F64, MIXED_F32, MIXED_BF16X3):
        cfg = make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=0, precision=prec)
        g = Context(cfg)
        for kind in (0, 1):
            f = seeded_coeffs(cfg, 7 + kind, 3 if kind else 1)
            outs[(prec, kind)] = g.sht_roundtrip(kind, f)
            errs[(prec, kind)] = rel(outs[(prec, kind)], f)
        g.close()
    print("round-trip errors", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    o = Oracle(make_config(L=L, M=L, N=N, Th=th, Ph=ph, comp=0, precision=PLAIN_F64))
    for kind in (0, 1):
        f = seeded_coeffs(o.cfg, 7 + kind, 3 if kind else 1)
        spec = [(np.ascontiguousarray(re.T), np.ascontiguousarray(im.T)) for re, im in f]
        if kind == 0:
            back = [o.phys2spec(o.spec2phys(*spec[0]))]
        else:
            back = o.rtp2qst(*o.qst2rtp(*spec))
        ref = [(re.T, im.T) for re, im in back]
        d = rel(outs[(PLAIN_F64, kind)], ref)
        print("kind %d: FP64 round trip vs oracle %.2e (oracle vs input %.2e)" % (kind, d, rel(ref, f)),
              flush=True)
        assert d < 1e-13
        assert errs[(PLAIN_F64, kind)] < 1e-10
        assert errs[(MIXED_F32, kind)] < 2e-6
        assert errs[(MIXED_BF16X3, kind)] < 1e-4
        assert errs[(MIXED_BF16X3, kind)] > errs[(MIXED_F32, kind)]


def _errors(gm, gp):
    return rel([gm.get_coll(k) for k in range(6)], [gp.get_coll(k) for k in range(6)])


@pytest.mark.parametrize("which", ["config1", "config3"])
def test_bf16x3_timesteps(which):
    fn = config1 if which == "config1" else config3
    seed, amp = (3, 1e-3) if which == "config1" else (SEEDS[3], AMPS[3])
    cp, cb = fn(PLAIN_F64), fn(MIXED_BF16X3)
    st = seeded_state(cp, seed, amp)
    gp, gb = Context(cp), Context(cb)
    gp.set_state(st)
    gb.set_state(st)
    marks = (1, 10) if which == "config1" else (1, 1000)
    done = 0
    for k in marks:
        sp, sb = gp.step(k - done), gb.step(k - done)
        done = k
        assert sp.evaluations == sb.evaluations
        e = _errors(gb, gp)
        print("%s: BF16x3 vs plain FP64 after %d steps: %.2e" % (which, k, e), flush=True)
        assert np.isfinite(e) and e < 1e-3
    gp.close()
    gb.close()
