"""Seeded synthetic initial states shared by the GPU path, the oracle and the
benchmark, so every consumer sees identical bytes.

Coefficient of harmonic (l, m) at radius r_n:
    Re, Im = z_{lm} * A * (l+1)^-2 * phi_f(x_n),   x = r - r_i in [0, 1]
with z ~ N(0,1) drawn with numpy PCG64(seed) once per (l, m), Im = 0 for
m = 0, and a smooth radial shape per field:
    phi = sin(pi x)        toroidal u, toroidal B, codensity, composition
    phi = sin(pi x)^2      poloidal u (P = dP/dr = 0 at both walls)
    phi = sin(pi x)        poloidal B
l = 0 is zero for u and B; the codensity and composition l = 0 modes carry
the conductive profile r_i (r_o - r) / r (C = 1 at r_i, 0 at r_o, the
reference's i_cod_bc = 1 values) plus the random part.

Randomness is in (l, m) only: independent noise at every radial point of a
Chebyshev grid (spacing ~ 1/N^2 at the walls) gives curls of order N^4 A and
makes the fixed-dt predictor-corrector of tim_check_cgce diverge at the
E = 1e-4, dt = 2e-6 case, which the reference would abort.
Draw order: fields uT, uP, C, X, BT, BP, each one (2, H) normal array.
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


def radial_grid(N, rratio):
    """Chebyshev collocation points (modules/meshs.F90:135-143)."""
    n = np.arange(1, N + 1)
    ri = rratio / (1.0 - rratio)
    return 0.5 * (1.0 + np.cos(np.pi * (N - n) / (N - 1.0))) + ri, ri


def seeded_state(cfg, seed, amp):
    """dict field -> (re[H,N], im[H,N]) float64 arrays (reference coll layout)."""
    H = cfg.sizes()[0]
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    r, ri = radial_grid(cfg.N, cfg.rratio)
    ro = ri + 1.0
    x = r - ri
    s1 = np.sin(np.pi * x)
    shapes = {"uT": s1, "uP": s1 * s1, "C": s1, "X": s1, "BT": s1, "BP": s1}
    scale = amp * (l.astype(np.float64) + 1.0) ** -2
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for name in FIELD_ORDER:
        z = rng.standard_normal((2, H))
        re = (z[0] * scale)[:, None] * shapes[name][None, :]
        im = (z[1] * scale)[:, None] * shapes[name][None, :]
        im[m == 0, :] = 0.0
        if name in ("uT", "uP", "BT", "BP"):
            re[l == 0, :] = 0.0
            im[l == 0, :] = 0.0
        if name in ("C", "X"):
            re[(l == 0) & (m == 0), :] += ri * (ro - r) / r
        off = (name == "X" and not cfg.comp) or (name in ("BT", "BP") and not cfg.mag) or \
              (name == "C" and not cfg.cod)
        if off:
            re[:] = 0.0
            im[:] = 0.0
        out[name] = (np.ascontiguousarray(re), np.ascontiguousarray(im))
    return out


def ic_radii(cfg):
    """Inner-core points r = r_i cos(pi (2 Nic - n) / (2 Nic - 1)),
    n = Nic+1 .. 2 Nic (modules/meshs.F90:165-169)."""
    Ni = max(1, cfg.Nic)
    ri = cfg.rratio / (1.0 - cfg.rratio)
    N_ = 2 * Ni
    n = np.arange(Ni + 1, N_ + 1)
    return ri * np.cos(np.pi * (N_ - n) / (N_ - 1))


def seeded_ic(cfg, seed, amp):
    """Seeded inner-core fields {icBT, icBP: (re[H,Nic], im[H,Nic])}: smooth
    in r, zero at r_i (where the seeded outer-core field is zero too) and
    at l = 0, with the (l+1)^-2 spectrum of seeded_state."""
    H = cfg.sizes()[0]
    Ni = max(1, cfg.Nic)
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    ri = cfg.rratio / (1.0 - cfg.rratio)
    r = ic_radii(cfg)
    shape = np.sin(np.pi * r / ri) * (r / ri) ** 2
    scale = amp * (l.astype(np.float64) + 1.0) ** -2
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for name in ("icBT", "icBP"):
        z = rng.standard_normal((2, H))
        re = (z[0] * scale)[:, None] * shape[None, :]
        im = (z[1] * scale)[:, None] * shape[None, :]
        im[m == 0, :] = 0.0
        re[l == 0, :] = 0.0
        im[l == 0, :] = 0.0
        out[name] = (np.ascontiguousarray(re), np.ascontiguousarray(im))
    return out


def ic_extend(cfg, field, r, dr1):
    """io_ic_extend (program/io.F90:1438-1477): an inner-core field a r^3 +
    b r^2 that meets the outer-core field (re[H,N], im[H,N]) at r_i in value
    and radial slope, the slope from the first row of the outer-core weights
    dr1 (the library's or the reference's mesh array dr(1), [N, 7] band
    storage).  Returns (re[H,Nic], im[H,Nic])."""
    rc = ic_radii(cfg)
    r3 = rc * (rc * rc)
    r2 = rc * rc
    ri = float(r[0])
    ri2 = ri * ri
    ri3 = ri * ri2
    w = [float(dr1[j, 3 - j]) for j in range(4)]  # dr(1)%M(KL+1+1-j, j), j = 1..4
    out = []
    for part in field:
        f = part[:, 0]
        f1 = np.zeros_like(f)
        for j in range(4):
            f1 = f1 + w[j] * part[:, j]
        a = (f1 * ri - 2.0 * f) / ri3
        b = (f - a * ri3) / ri2
        out.append(np.ascontiguousarray(a[:, None] * r3[None, :] + b[:, None] * r2[None, :]))
    return out[0], out[1]


SEEDS = {1: 20261201, 2: 20261202, 3: 20261203, 4: 20261204, 5: 20261205}
AMPS = {1: 1e-3, 2: 1.0, 3: 1e-2, 4: 1e-3}
