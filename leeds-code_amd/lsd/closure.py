"""Learned closure for truncated runs: a small network supplies the part of
the nonlinear terms that the harmonics beyond the truncation would have
contributed.

Data: a plain high-resolution run; at each snapshot its state is truncated
to the low-resolution harmonics (same (l, m), l < L_lo, m < M_lo).  The
target is the difference between the high-resolution nonlinear terms
(truncated) and those the low-resolution model computes from the truncated
state, for the five time-stepped fields of a Boussinesq MHD run without
composition (u toroidal and poloidal, codensity, B toroidal and poloidal).

Network input per (radial shell n, order m, degree l): the real and
imaginary parts of the five resolved state coefficients at degrees
l-4 .. l+4 of the same (n, m) (zero outside m <= l' < L_lo), and l / L_lo,
m / M_lo -- 92 values; output: the real and imaginary parts of the five
nonlinear-term corrections at (n, l, m) -- 10 values.  Inputs and targets
are scaled per field by their root-mean-square over the training set.

The network runs on the GPU (lsd.gpu.Mlp: HIP kernels for inference and
training).  The low-resolution model without the network is the fallback.
"""
import numpy as np

from .initial import harmonics
from .gpu import Mlp

STATE = (0, 1, 2, 4, 5)       # uT, uP, C, BT, BP
NTERM = (6, 7, 8, 10, 11)     # their nonlinear terms
WIN = 4
N_IN = 2 * len(STATE) * (2 * WIN + 1) + 2
N_OUT = 2 * len(NTERM)


def truncation_map(cfg_hi, cfg_lo):
    """Index of each low-resolution harmonic in the high-resolution order."""
    lh, mh = harmonics(cfg_hi.L, cfg_hi.M, cfg_hi.Mp)
    ll, ml = harmonics(cfg_lo.L, cfg_lo.M, cfg_lo.Mp)
    pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(lh, mh))}
    return np.array([pos[(int(a), int(b))] for a, b in zip(ll, ml)])


def truncate(fields_hi, tmap):
    return {k: (re[tmap], im[tmap]) for k, (re, im) in fields_hi.items()}


def read_fields(ctx, ids):
    return {k: ctx.get_coll(k) for k in ids}


class Features:
    """Builds the (shell, harmonic)-row feature matrix of a truncated state."""

    def __init__(self, cfg_lo):
        self.cfg = cfg_lo
        l, m = harmonics(cfg_lo.L, cfg_lo.M, cfg_lo.Mp)
        self.l, self.m = l, m
        pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(l, m))}
        H = len(l)
        # neighbour index of (l + d, m) for d in -WIN..WIN, H (a zero row) if absent
        self.nbr = np.full((H, 2 * WIN + 1), H, dtype=np.int64)
        for h in range(H):
            for k, d in enumerate(range(-WIN, WIN + 1)):
                self.nbr[h, k] = pos.get((int(l[h]) + d, int(m[h])), H)
        self.lm = np.stack([l / cfg_lo.L, m / cfg_lo.M], axis=1).astype(np.float32)

    def __call__(self, state, in_scale):
        """state: {field id: (re[H, N], im[H, N])}; rows ordered (n, h)."""
        H, N = self.l.size, self.cfg.N
        cols = []
        for f, k in enumerate(STATE):
            re, im = state[k]
            for part in (re, im):
                a = np.concatenate([part, np.zeros((1, N))], axis=0) / in_scale[f]
                g = a[self.nbr]                      # [H, 9, N]
                cols.append(np.transpose(g, (2, 0, 1)))  # [N, H, 9]
        x = np.concatenate(cols, axis=2).reshape(N * H, -1)
        lm = np.broadcast_to(self.lm[None], (N, H, 2)).reshape(N * H, 2)
        return np.concatenate([x, lm], axis=1).astype(np.float32)


def targets(n_hi_trunc, n_lo, out_scale):
    """Rows (n, h) x 10 of (N_hi - N_lo) / scale, Re and Im per field."""
    cols = []
    for f, k in enumerate(NTERM):
        for p in (0, 1):
            d = (n_hi_trunc[k][p] - n_lo[k][p]) / out_scale[f]   # [H, N]
            cols.append(d.T.reshape(-1, 1))
    return np.concatenate(cols, axis=1).astype(np.float32)


def to_correction(y, out_scale, H, N):
    """Network output rows (n, h) -> {N-field id: (re[H, N], im[H, N])}."""
    y = y.reshape(N, H, N_OUT)
    out = {}
    for f, k in enumerate(NTERM):
        out[k] = (y[:, :, 2 * f].T * out_scale[f], y[:, :, 2 * f + 1].T * out_scale[f])
    return out


def snapshots(g_hi, g_lo, cfg_hi, cfg_lo, count, every):
    """(truncated state, truncated high-resolution N, low-resolution N) at
    `count` snapshots `every` high-resolution steps apart."""
    tmap = truncation_map(cfg_hi, cfg_lo)
    out = []
    for s in range(count):
        g_hi.step(every)
        state_hi = read_fields(g_hi, STATE)
        g_hi.transform()
        g_hi.nonlinear()
        n_hi = truncate(read_fields(g_hi, NTERM), tmap)
        st = truncate(state_hi, tmap)
        for k, (re, im) in st.items():
            g_lo.set_coll(k, re, im)
        g_lo.transform()
        g_lo.nonlinear()
        n_lo = read_fields(g_lo, NTERM)
        out.append((st, n_hi, n_lo))
    return out


def rms_scales(snaps):
    ins = [np.sqrt(np.mean([np.mean(s[0][k][0] ** 2 + s[0][k][1] ** 2) for s in snaps]) / 2) for k in STATE]
    outs = [np.sqrt(np.mean([np.mean((s[1][k][0] - s[2][k][0]) ** 2 + (s[1][k][1] - s[2][k][1]) ** 2)
                             for s in snaps]) / 2) for k in NTERM]
    return np.maximum(np.array(ins), 1e-300), np.maximum(np.array(outs), 1e-300)


class Closure:
    """Trained network + scales; `correction(state)` gives the N increments."""

    def __init__(self, cfg_lo, width=256, seed=0):
        self.cfg = cfg_lo
        self.feat = Features(cfg_lo)
        self.net = Mlp(N_IN, width, N_OUT, seed=seed)
        self.in_scale = np.ones(len(STATE))
        self.out_scale = np.ones(len(NTERM))
        # fields whose guarded increment is applied in time stepping, with
        # their weights: the toroidal-velocity term.  The others are
        # predicted (held-out skill) but, applied together with it, drive
        # the poloidal velocity away from the resolved run.
        self.weights = {6: 1.0}

    def dataset(self, snaps):
        X = np.concatenate([self.feat(s[0], self.in_scale) for s in snaps])
        Y = np.concatenate([targets(s[1], s[2], self.out_scale) for s in snaps])
        return X, Y

    def fit(self, snaps, steps, batch=8192, lr=1e-3, seed=0):
        self.in_scale, self.out_scale = rms_scales(snaps)
        X, Y = self.dataset(snaps)
        perm = np.random.default_rng(seed).permutation(X.shape[0])
        return self.net.train(X[perm], Y[perm], batch=batch, steps=steps, lr=lr)

    def correction(self, state, guard=True):
        """Nonlinear-term increments for a truncated state.  With `guard`
        (the form used in time stepping) the increment is zero where the
        reference keeps N = 0 (l = 0 of the velocity and magnetic
        potentials); in the top WIN degrees, whose neighbour window the
        truncation cuts, the toroidal and codensity increments are kept
        only where they remove energy (Re(conj(x) dN) <= 0, the sign of
        their contribution to d|x|^2/dt) and the poloidal ones are
        dropped; each field's increment is then scaled by `weights`
        (absent fields: zero)."""
        H, N = self.feat.l.size, self.cfg.N
        y = self.net.forward(self.feat(state, self.in_scale))
        c = to_correction(y, self.out_scale, H, N)
        if guard:
            top = (self.feat.l > self.cfg.L - 1 - WIN)[:, None]
            l0 = (self.feat.l == 0)[:, None]
            for f, k in enumerate(NTERM):
                re, im = c[k]
                if k in (7, 11):
                    drop = np.broadcast_to(top, re.shape)
                else:
                    xr, xi = state[STATE[f]]
                    drop = top & (xr * re + xi * im > 0.0)
                if k != 8:
                    drop = drop | l0
                c[k] = (np.where(drop, 0.0, re), np.where(drop, 0.0, im))
            w = self.weights
            c = {k: (v[0] * w.get(k, 0.0), v[1] * w.get(k, 0.0)) for k, v in c.items()}
        return c

    def explained(self, snaps, guard=False, fields=NTERM):
        """Held-out skill: 1 - |target - prediction|^2 / |target|^2 over the
        snapshots (in physical units, summed over `fields`)."""
        num = den = 0.0
        for st, n_hi, n_lo in snaps:
            c = self.correction(st, guard)
            for k in fields:
                for p in (0, 1):
                    t = n_hi[k][p] - n_lo[k][p]
                    num += np.sum((t - c[k][p]) ** 2)
                    den += np.sum(t ** 2)
        return 1.0 - num / den


def check_cgce(s, dterr_tol, dt):
    """tim_check_cgce, i_time_step_control = 2 (modules/timestep.F90:791-827),
    on the host state dict s; returns True when the run must stop."""
    stop = False
    if s["it"] == 1:
        s["lasterr"] = 1e99
    if s["it"] > 10:
        stop, s["it"] = True, 0
    elif s["dterr"] > s["lasterr"]:
        stop = s["dterr"] > 2.0 * dterr_tol
        s["it"] = 0
    elif s["dterr"] > dterr_tol:
        s["lasterr"] = s["dterr"]
        s["it"] += 1
    else:
        s["it"] = 0
    s["dterr"] = 0.0
    return stop


def step_with_closure(g, cfg, nsteps, closure=None):
    """The main loop of program/main.F90:100-164 on the phase entry points,
    with the network's correction added to the nonlinear terms when a
    closure is given (None: the plain truncated model).  The correction is
    treated explicitly: evaluated once per step, on the state the step
    starts from, and added unchanged in the predictor and every corrector
    pass, so it does not enter the corrector's convergence test
    (tim_check_cgce measures N* - N)."""
    s = {"it": 0, "dterr": 0.0, "lasterr": 1e99}
    for _ in range(nsteps):
        corr = closure.correction(read_fields(g, STATE)) if closure is not None else None
        while True:
            g.transform()
            g.nonlinear()
            if corr is not None:
                for k in NTERM:
                    if not (np.any(corr[k][0]) or np.any(corr[k][1])):
                        continue
                    re, im = g.get_coll(k)
                    g.set_coll(k, re + corr[k][0], im + corr[k][1])
            if s["it"] == 0:
                g.predictor()
                s["it"] = 1
            else:
                s["dterr"] += g.corrector()
                d = s["dterr"]
                if check_cgce(s, cfg.dterr, cfg.dt):
                    raise RuntimeError("tim_check_cgce: run stopped at step %d (tim_dterr %.3g)" % (_, d))
            if s["it"] == 0:
                break
