"""State files (lsd_read_state / lsd_write_state, the C-ABI's
io_load_state / io_save_state, program/io.F90:349-483, 1502-1621), checked
on the CPU.

tests/golden/state_case1.cdf.in is the reference's own test state
(tests/state.cdf.in: the Christensen case-1 dynamo at N = 64, L = 64,
M = 16, Mp = 4, written by the reference).  Its kinetic and magnetic
energies are the first line of the reference's case-1 series
(Benchmark_files/Boussinesq/Christensen_case1_benchmark/vel_energy.dat:2,
mag_energy.dat:2): KE = 11225.9803716 and ME = 228799.504998, the latter
with the factor Pm/(2E) of that series.  Loading the file into a full
M = 64, Mp = 1 run and taking the oracle's energies reproduces both, which
pins the m-major io_state_nh order (a wrong degree or order would change
the 4 pi / (2l+1) and m weights of the sums).
"""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle_lib
from lsd import make_config
from lsd.gpu import read_state, write_state, LsdError

HERE = os.path.dirname(os.path.abspath(__file__))
STATE = os.path.join(HERE, "golden", "state_case1.cdf.in")
KE0, ME0 = 11225.9803716, 228799.504998


def case1(Mp=1, M=64, N=64, **kw):
    base = dict(L=64, M=M, Mp=Mp, N=N, Th=96, Ph=3 * M, comp=0, Ra=6.5e4, Pr=1.0, Pm=5.0, E=1e-3, dt=-5e-5,
                implicit=0.5, step_control=1)
    base.update(kw)
    return make_config(**base)


def test_reference_state_energies():
    cfg = case1()
    st, t, dt = read_state(STATE, cfg)
    assert t == 0.0 and abs(dt - 6.106330848710143e-05) < 1e-18
    o = oracle_lib.Oracle(cfg)
    o.set_state(st)
    ke = o.energy(0)[0]
    me = o.energy(1)[0] / 2.0
    print("KE %.10f ME/2 %.9f" % (ke, me))
    assert abs(ke - KE0) <= 1e-10 * KE0
    assert abs(me - ME0) <= 1e-10 * ME0
    # composition absent from the file: codensity is loaded in its place
    assert np.array_equal(st["X"][0], st["C"][0]) and np.array_equal(st["X"][1], st["C"][1])


def test_same_truncation_is_the_file_itself():
    """At the file's own L, M, Mp and N every value is copied: each (l, m)
    of the run, in the folded order, is the file's row io_state_nh(l, m)."""
    cfg = case1(Mp=4, M=16)
    st, _, _ = read_state(STATE, cfg)
    f = netcdf_file(STATE, "r", mmap=False)
    from lsd.initial import harmonics
    l, m = harmonics(64, 16, 4)
    mm = m // 4
    idx = 64 * mm - 4 * (mm - 1) * mm // 2 + l - m  # io_state_nh, program/io.F90:1256-1269
    for name in ("uT", "uP", "C", "BT", "BP"):
        v = f.variables[name][:]
        assert np.array_equal(st[name][0], v[0][idx]) and np.array_equal(st[name][1], v[1][idx]), name
    f.close()


def test_write_read_round_trip(tmp_path):
    cfg = case1(Mp=1, M=32, L=40, N=24, Th=60, Ph=96, comp=1)
    H = cfg.sizes()[0]
    rng = np.random.default_rng(5)
    st = {n: (rng.standard_normal((H, cfg.N)), rng.standard_normal((H, cfg.N)))
          for n in ("uT", "uP", "C", "X", "BT", "BP")}
    p = str(tmp_path / "state.cdf.dat")
    write_state(p, cfg, st, 1.25, 3e-5)
    back, t, dt = read_state(p, cfg)
    assert t == 1.25 and dt == 3e-5
    for n in st:
        assert np.array_equal(back[n][0], st[n][0]) and np.array_equal(back[n][1], st[n][1]), n
    # an independent reader sees the reference's layout: (ReIm, H, r), m-major
    f = netcdf_file(p, "r", mmap=False)
    assert f.version_byte == 1
    assert f.variables["uT"].dimensions == ("ReIm", "H", "r")
    assert int(f.variables["uT"].L) == 40 and int(f.variables["uT"].M) == 32 and int(f.variables["uT"].Mp) == 1
    from lsd.initial import harmonics
    l, m = harmonics(40, 32, 1)
    idx = 40 * m - (m - 1) * m // 2 + l - m
    assert np.array_equal(f.variables["Comp"][:][0][idx], st["X"][0])
    assert abs(float(f.t) - 1.25) == 0 and float(f.variables["dt"].getValue()) == 3e-5
    f.close()


def test_radial_regrid_is_exact_on_sextic_profiles(tmp_path):
    """A file on N = 40 read into N = 29: 7-point Lagrange interpolation
    (io_interp_wts / io_interp) reproduces any radial profile of degree
    <= 6 (up to rounding)."""
    src = case1(Mp=1, M=8, L=10, N=40, Th=16, Ph=24)
    dst = case1(Mp=1, M=8, L=10, N=29, Th=16, Ph=24)
    H = src.sizes()[0]
    ri = 0.35 / 0.65

    def grid(N):
        n = np.arange(1, N + 1)
        return 0.5 * (1 + np.cos(np.pi * (N - n) / (N - 1))) + ri

    rng = np.random.default_rng(6)
    coef = rng.standard_normal((6, 2, H, 7))

    def profiles(r):
        pw = np.stack([r ** k for k in range(7)])  # (7, N)
        return coef @ pw  # (6, 2, H, N)

    a = profiles(grid(40))
    names = ("uT", "uP", "C", "X", "BT", "BP")
    p = str(tmp_path / "s.cdf")
    write_state(p, src, {n: (a[k, 0], a[k, 1]) for k, n in enumerate(names)}, 0.0, 1e-4)
    back, _, _ = read_state(p, dst)
    b = profiles(grid(29))
    for k, n in enumerate(names):
        for part in (0, 1):
            err = np.abs(back[n][part] - b[k, part]).max() / np.abs(b[k, part]).max()
            assert err < 1e-11, (n, part, err)


def test_nonmagnetic_run_and_errors(tmp_path):
    st, _, _ = read_state(STATE, case1(mag=0))
    assert not st["BT"][0].any() and not st["BP"][0].any() and st["uT"][0].any()
    with pytest.raises(LsdError, match="not found"):
        read_state(str(tmp_path / "missing.cdf"), case1())
    bad = tmp_path / "bad.cdf"
    bad.write_bytes(b"HDF5" + b"\0" * 64)
    with pytest.raises(LsdError, match="netCDF classic"):
        read_state(str(bad), case1())
