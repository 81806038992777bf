"""Seeded synthetic initial states shared by the GPU path, the oracle and the
benchmark, so every consumer sees identical bytes.

Coefficients: Re, Im ~ N(0,1) * A * (l+1)^-2 for every (l, m, r_n), drawn
with numpy PCG64(seed); Im = 0 for m = 0; l = 0 is zero for the toroidal and
poloidal scalars of u and B, while codensity and composition keep it.
Draw order: fields uT, uP, C, X, BT, BP, each as one (2, H, N) normal array
(Re then Im) in the reference coll layout re[h, n].
"""
import numpy as np

FIELD_ORDER = ("uT", "uP", "C", "X", "BT", "BP")
# coll-field ids of include/lsd.h
FIELD_ID = {"uT": 0, "uP": 1, "C": 2, "X": 3, "BT": 4, "BP": 5}


def harmonics(L, M, Mp=1):
    """Folded harmonic order (Scripts/parallel.h_template:65-80): returns
    arrays l[h], m[h] (m is the actual order m_ = m*Mp)."""
    H = L * M - Mp * (M - 1) * M // 2
    Hc = (M + 1) // 2
    Hr = H // Hc
    ls, ms = [], []
    for jj in range(1, Hc + 1):
        for ii in range(1, Hr + 1):
            m = jj - 1
            l = m * Mp + ii - 1
            if l >= L:
                m = 2 * Hc - jj
                l = L - Hr + ii - 1
            ls.append(l)
            ms.append(m * Mp)
    return np.array(ls, dtype=np.int64), np.array(ms, dtype=np.int64)


def seeded_state(cfg, seed, amp):
    """dict field -> (re[H,N], im[H,N]) float64 arrays."""
    H = cfg.sizes()[0]
    N = cfg.N
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    scale = amp * (l.astype(np.float64) + 1.0) ** -2
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for name in FIELD_ORDER:
        z = rng.standard_normal((2, H, N))
        re = z[0] * scale[:, None]
        im = z[1] * scale[:, None]
        im[m == 0, :] = 0.0
        if name in ("uT", "uP", "BT", "BP"):
            re[l == 0, :] = 0.0
            im[l == 0, :] = 0.0
        if name == "X" and not cfg.comp:
            re[:] = 0.0
            im[:] = 0.0
        if name in ("BT", "BP") and not cfg.mag:
            re[:] = 0.0
            im[:] = 0.0
        out[name] = (np.ascontiguousarray(re), np.ascontiguousarray(im))
    return out


SEEDS = {1: 20261201, 2: 20261202, 3: 20261203, 4: 20261204, 5: 20261205}
AMPS = {1: 1e-3, 2: 1.0, 3: 1e-2, 4: 1e-3}
