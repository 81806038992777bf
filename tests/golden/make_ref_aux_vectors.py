"""Writes tests/golden/ref_aux_vectors.npz: outputs of the reference's own
auxiliary routines GAUWTS and SCHNLA (auxiliary/gauwts.f, auxiliary/schnla.f
with pmm*.f, plm.f, dp*.f), compiled from their sources by
`make -C oracle ref` into oracle/_ref/libref_aux.so.

Inputs: Gauss rules of 18, 48, 256 and 384 points; Schmidt functions with
LH = 31 at the first, a middle and the last node of the 48-point rule (all
(l, m)), and with LH = 169 at a middle node of the 256-point rule for the
orders m in SCHNLA_M (packed index l(l+1)/2 + m stored beside the values).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib  # noqa: E402


SCHNLA_M = (0, 1, 2, 3, 84, 85, 166, 167, 168, 169)


def main():
    lib = oracle_lib.ref_aux()
    if lib is None:
        raise SystemExit("oracle/_ref/libref_aux.so missing: run `make -C oracle ref` where the reference is present")
    out = {}
    for n in (18, 48, 256, 384):
        x, w = oracle_lib.ref_gauwts(lib, n)
        out["gauss_x_%d" % n] = x
        out["gauss_w_%d" % n] = w
    x = out["gauss_x_48"]
    for tag, j in (("first", 0), ("mid", 12), ("last", 23)):
        pa, dpa = oracle_lib.ref_schnla(lib, float(x[j]), 31)
        out["schnla31_x_%s" % tag] = np.array([x[j]])
        out["schnla31_p_%s" % tag] = pa
        out["schnla31_dp_%s" % tag] = dpa
    x = out["gauss_x_256"]
    pa, dpa = oracle_lib.ref_schnla(lib, float(x[64]), 169)
    idx = np.array([l * (l + 1) // 2 + m for m in SCHNLA_M for l in range(m, 170)], dtype=np.int64)
    out["schnla169_x"] = np.array([x[64]])
    out["schnla169_idx"] = idx
    out["schnla169_p"] = pa[idx]
    out["schnla169_dp"] = dpa[idx]
    np.savez_compressed(os.path.join(HERE, "ref_aux_vectors.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
