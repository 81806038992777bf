"""The reference's own physics modules with the drop-in legendre and
transform modules: oracle/_ref/<case>/dropin_step is oracle/ref_step.F90
(program/main.F90:100-164 without file I/O) linked with the reference's
unmodified parameters, dyn_mpi, meshs, variables, timestep, velocity,
magnetic, codensity, composition, rotation and nonlinear modules and, in
the places of modules/legendre.F90 and modules/transform.fftw3.F90, the
repository's leeds-code_amd/fortran/legendre.hip.F90 and transform.hip.F90
over the C-ABI (built by `python oracle/ref_build.py --dropin cfg1 small`
in the build container; only the executable travels).

So the reference's CPU code forms the nonlinear products, the spectral
operators and the banded solves, and every Legendre table and transform
comes from the GPU.  Checked per step against
  * the library's own step (lsd_step, plain FP64) from the same seeded
    state: relative L2 difference of every field <= 1e-12 per step (k 1e-12
    after step k, as the reference fixtures are bounded) over 5 steps, with
    the same corrector counts;
  * the reference's own FFTW/LAPACK build (tests/golden/ref_step_<case>.npz):
    the tables leg_gp, leg_gw, leg_P, leg_gwP that the drop-in legendre
    module gave the reference's code to 1e-14 of their largest entry, and
    the full fields of step 1 to 1e-12.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ref_io  # noqa: E402
from lsd import config1, config_small, seeded_state, PLAIN_F64, SEEDS, AMPS  # noqa: E402
from lsd.gpu import Context  # noqa: E402

pytestmark = pytest.mark.gpu

FIELDS = ("uT", "uP", "C", "X", "BT", "BP")
CASES = {"cfg1": (config1, SEEDS[1], AMPS[1]), "small": (config_small, 7, 1e-3)}


def rel(a, b):
    num = ((a[0] - b[0]) ** 2).sum() + ((a[1] - b[1]) ** 2).sum()
    den = (b[0] ** 2).sum() + (b[1] ** 2).sum()
    return 0.0 if den == 0 else float(np.sqrt(num / den))


@pytest.mark.parametrize("case", ["cfg1", "small"])
def test_reference_modules_with_dropin_transforms(case):
    if not ref_io.available(case, "dropin_step"):
        pytest.skip("oracle/_ref/%s/dropin_step not built (python oracle/ref_build.py --dropin)" % case)
    mk, seed, amp = CASES[case]
    cfg = mk(precision=PLAIN_F64)
    state = seeded_state(cfg, seed, amp)
    nsteps = 5
    print("case", case, "dropin_step for", nsteps, "steps", flush=True)
    recs, _ = ref_io.run(case, cfg, state, nsteps, 1, name="dropin_step")
    g = Context(cfg)
    g.set_state(state)
    worst = 0.0
    for rec in recs[1:]:
        st = g.step(1)
        assert st.evaluations == rec["evaluations"], (rec["step"], st.evaluations, rec["evaluations"])
        for k, name in enumerate(FIELDS):
            if name == "X" and not cfg.comp:
                continue
            d = rel(rec["fields"][name], g.get_coll(k))
            worst = max(worst, d)
            assert d <= 1e-12 * rec["step"], (rec["step"], name, d)
    print("worst field difference from lsd_step over %d steps: %.2e" % (nsteps, worst), flush=True)
    g.close()
    # the tables the drop-in legendre module handed the reference's code,
    # and step 1, against the reference's own build
    gold = np.load(os.path.join(ROOT, "tests", "golden", "ref_step_%s.npz" % case))
    tabs = recs[0]["tables"]
    for key in ("gp", "gw", "P", "gwP"):
        if key in gold.files:
            a, b = tabs[key], gold[key]
            d = float(np.abs(a - b).max() / np.abs(b).max())
            print("legendre drop-in %s vs the reference's leg_precompute: %.2e" % (key, d), flush=True)
            assert d <= 1e-14, (key, d)
    for name in FIELDS:
        key = "s1_%s" % name
        if key + "_re" not in gold.files or (name == "X" and not cfg.comp):
            continue
        d = rel(recs[1]["fields"][name], (gold[key + "_re"], gold[key + "_im"]))
        print("step 1 %s vs the reference's FFTW/LAPACK build: %.2e" % (name, d), flush=True)
        assert d <= 1e-12, (name, d)
