"""Pins of the oracle's time step against what the equations fix: boundary
conditions of every field after a step, the theta-scheme on a pure diffusion
mode, the buoyancy coupling, and the conduction steady state
(modules/velocity.F90, magnetic.F90, codensity.F90, timestep.F90,
nonlinear.F90)."""
import numpy as np
import pytest

import oracle_lib as O
from lsd import config_small, seeded_state
from lsd.initial import harmonics, FIELD_ID


def _apply(band, f):
    N = f.shape[-1]
    out = np.zeros_like(f)
    for j in range(N):
        for n in range(max(0, j - 3), min(N, j + 4)):
            out[..., n] += band[j, 3 + n - j] * f[..., j]
    return out


@pytest.fixture(scope="module")
def stepped():
    cfg = config_small()
    o = O.Oracle(cfg, threads=2)
    o.set_state(seeded_state(cfg, 3, 1e-3))
    st = o.step(2)
    return cfg, o, st


def test_velocity_no_slip_after_step(stepped):
    cfg, o, _ = stepped
    dr1 = o.mesh(1)
    for part in o.get_coll(FIELD_ID["uT"]):
        assert np.abs(part[:, [0, -1]]).max() < 1e-12 * np.abs(part).max()
    for part in o.get_coll(FIELD_ID["uP"]):
        scale = np.abs(part).max()
        assert np.abs(part[:, [0, -1]]).max() < 1e-12 * scale       # P = 0
        d = _apply(dr1, part)
        assert np.abs(d[:, [0, -1]]).max() < 1e-9 * np.abs(d).max()  # dP/dr = 0


def test_insulating_magnetic_bc_after_step(stepped):
    cfg, o, _ = stepped
    l, _ = harmonics(cfg.L, cfg.M, cfg.Mp)
    r = o.mesh(0)
    dr1 = o.mesh(1)
    for part in o.get_coll(FIELD_ID["BT"]):
        assert np.abs(part[:, [0, -1]]).max() < 1e-12 * np.abs(part).max()
    for part in o.get_coll(FIELD_ID["BP"]):
        d = _apply(dr1, part)
        inner = d[:, 0] - l / r[0] * part[:, 0]
        outer = d[:, -1] + (l + 1) / r[-1] * part[:, -1]
        scale = np.abs(d).max()
        assert np.abs(inner[l > 0]).max() < 1e-9 * scale
        assert np.abs(outer[l > 0]).max() < 1e-9 * scale


def test_codensity_dirichlet_after_step(stepped):
    cfg, o, _ = stepped
    # boundary rows come out of a pivoted banded solve, so they hold to
    # rounding of the operator's largest entries (~1/h^2), not to the ulp
    re, im = o.get_coll(FIELD_ID["C"])
    assert abs(re[0, 0] - 1.0) < 1e-12 and abs(re[0, -1]) < 1e-12
    assert np.abs(re[1:, [0, -1]]).max() < 1e-12
    assert np.abs(im[:, [0, -1]]).max() < 1e-12


def test_conduction_state_is_steady():
    cfg = config_small()
    o = O.Oracle(cfg, threads=2)
    r = o.mesh(0)
    ri, ro = r[0], r[-1]
    H, N = o.H, cfg.N
    re = np.zeros((H, N))
    re[0] = ri * (ro - r) / r  # grad^2 C = 0, C(ri) = 1, C(ro) = 0
    o.set_coll(FIELD_ID["C"], re, np.zeros((H, N)))
    st = o.step(3)
    assert st.evaluations == 6
    for name in ("uT", "uP", "BT", "BP"):
        for part in o.get_coll(FIELD_ID[name]):
            assert np.abs(part).max() == 0.0
    c, _ = o.get_coll(FIELD_ID["C"])
    assert np.abs(c[0] - re[0]).max() < 1e-5          # FD truncation only (N = 16)
    assert np.abs(c[1:]).max() == 0.0


def test_buoyancy_closed_form():
    # u = B = 0: N_Pol = (Ra Pm^2/Pr) C (q/r with q = Ra r Pm^2/Pr C, s = 0),
    # N_Tor = 0, N_C = 0 (nonlinear.F90:158-197)
    cfg = config_small()
    o = O.Oracle(cfg, threads=2)
    H, N = o.H, cfg.N
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    h = int(np.where((l == 4) & (m == 3))[0][0])
    rng = np.random.default_rng(0)
    re = np.zeros((H, N))
    im = np.zeros((H, N))
    re[h] = rng.standard_normal(N)
    im[h] = rng.standard_normal(N)
    o.set_coll(FIELD_ID["C"], re, im)
    o.transform()
    o.nonlinear()
    npr, npi = o.get_coll(7)
    k = cfg.Ra * cfg.Pm ** 2 / cfg.Pr
    assert np.abs(npr - k * re).max() < 1e-12 * k * np.abs(re).max()
    assert np.abs(npi - k * im).max() < 1e-12 * k * np.abs(im).max()
    for f in (6, 8, 10, 11):
        for part in o.get_coll(f):
            assert np.abs(part).max() < 1e-12


def test_theta_scheme_on_toroidal_diffusion():
    # an axisymmetric toroidal l = 1 flow feels no toroidal forcing (its
    # advection and Coriolis terms are meridional), so one step is the
    # theta-scheme for dT/dt = Pm (d^2/dr^2 + 2/r d/dr - l(l+1)/r^2) T with
    # T = 0 at both walls:
    #   (I/dt - c Pm L) T1 = (I/dt + (1-c) Pm L) T0
    # Rotation is switched off (E huge): with 2 Pm/E dt = O(1) the Coriolis
    # force would return a poloidal flow's torque to T within the step.
    cfg = config_small(E=1e12)
    o = O.Oracle(cfg, threads=2)
    H, N = o.H, cfg.N
    l, m = harmonics(cfg.L, cfg.M, cfg.Mp)
    h = int(np.where((l == 1) & (m == 0))[0][0])
    r = o.mesh(0)
    x = r - r[0]
    T0 = 1e-6 * np.sin(np.pi * x) * (1.0 + 0.3 * x)
    re = np.zeros((H, N))
    re[h] = T0
    o.set_coll(FIELD_ID["uT"], re, np.zeros((H, N)))
    o.step(1)
    T1 = o.get_coll(FIELD_ID["uT"])[0][h]
    lap = o.mesh(3)
    D = np.zeros((N, N))
    for j in range(N):
        for n in range(max(0, j - 3), min(N, j + 4)):
            D[n, j] = lap[j, 3 + n - j]
    Ll = D - np.diag(2.0 / r ** 2)
    c, dt, Pm = cfg.implicit, cfg.dt, cfg.Pm
    X = np.eye(N) / dt - c * Pm * Ll
    Y = np.eye(N) / dt + (1 - c) * Pm * Ll
    rhs = Y @ T0
    X[0, :] = 0.0
    X[0, 0] = 1.0
    X[-1, :] = 0.0
    X[-1, -1] = 1.0
    rhs[0] = rhs[-1] = 0.0
    expect = np.linalg.solve(X, rhs)
    assert np.abs(T1 - expect).max() < 1e-10 * np.abs(expect).max()
