"""Pins of the oracle's FFT, banded LU and radial finite differences
(modules/transform.fftw3.F90 FFTW calls, modules/meshs.F90, dgbtrf/dgbtrs
as called at modules/meshs.F90:452 and modules/timestep.F90:580)."""
import numpy as np
import pytest

import oracle_lib as O
from lsd import config1, config3


@pytest.mark.parametrize("n", [36, 96, 512, 768, 250])
def test_cfft_vs_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.cfft(z, -1) - np.fft.fft(z)).max() < 1e-12 * np.sqrt(n)
    assert np.abs(O.cfft(z, +1) - np.fft.ifft(z) * n).max() < 1e-12 * np.sqrt(n)


def _band_from_dense(A, kl=3, ku=3):
    n = A.shape[0]
    ab = np.zeros((n, 10))  # column j holds ab(:, j) of LAPACK storage (2*kl+ku+1 rows)
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[j, kl + ku + i - j] = A[i, j]
    return ab


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_banded_lu_vs_lapack(seed):
    lapack = pytest.importorskip("scipy.linalg.lapack")
    rng = np.random.default_rng(seed)
    n = 40
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - 3), min(n, i + 4)):
            A[i, j] = rng.standard_normal()
    A[0, :] = 0.0
    A[0, 0] = 1.0  # a Dirichlet-style boundary row, as in the reference's operators
    ab = _band_from_dense(A)
    lu, piv, info = O.gbtrf(ab)
    assert info == 0
    # scipy: ab in LAPACK layout (ldab, n) with kl extra rows
    abl = np.zeros((10, n))
    abl[:, :] = ab.T
    lub, pivb, infob = lapack.dgbtrf(abl, 3, 3)
    assert infob == 0
    assert np.array_equal(piv, pivb + 1)
    assert np.abs(lu.T - lub).max() <= 1e-13 * np.abs(lub).max()
    b = rng.standard_normal(n)
    x = O.gbtrs(lu, piv, b)
    xb, infos = lapack.dgbtrs(lub, 3, 3, b, pivb)
    assert np.abs(x - xb).max() <= 1e-13 * np.abs(xb).max()
    assert np.abs(A @ x - b).max() <= 1e-11 * np.abs(b).max() * np.abs(A).max()


def _fornberg(z, x, m):
    """Fornberg (1988) finite-difference weights at z for nodes x, in exact
    rationals via mpmath -- an algorithm independent of the Taylor-matrix
    inverse the reference uses."""
    import mpmath as mp
    n = len(x)
    c = [[mp.mpf(0)] * (m + 1) for _ in range(n)]
    c1 = mp.mpf(1)
    c4 = x[0] - z
    c[0][0] = mp.mpf(1)
    for i in range(1, n):
        mn = min(i, m)
        c2 = mp.mpf(1)
        c5 = c4
        c4 = x[i] - z
        for j in range(i):
            c3 = x[i] - x[j]
            c2 = c2 * c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i][k] = c1 * (k * c[i - 1][k - 1] - c5 * c[i - 1][k]) / c2
                c[i][0] = -c1 * c5 * c[i - 1][0] / c2
            for k in range(mn, 0, -1):
                c[j][k] = (c4 * c[j][k] - k * c[j][k - 1]) / c3
            c[j][0] = c4 * c[j][0] / c3
        c1 = c2
    return c


@pytest.mark.parametrize("cfg", [config1(), config3()], ids=["N32", "N96"])
def test_fd_weights_vs_fornberg(cfg):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    o = O.Oracle(cfg, threads=1)
    r = o.mesh(0)
    dr1 = o.mesh(1)
    dr2 = o.mesh(2)
    N = cfg.N
    R = [mp.mpf(v) for v in r]  # the same (FP64) grid points
    for n in (0, 1, 2, 3, N // 2, N - 2, N - 1):
        lo, hi = max(0, n - 3), min(N - 1, n + 3)
        w = _fornberg(R[n], R[lo:hi + 1], 2)
        for j in range(lo, hi + 1):
            a1 = dr1[j, 3 + n - j]
            a2 = dr2[j, 3 + n - j]
            e1 = float(w[j - lo][1])
            e2 = float(w[j - lo][2])
            s1 = max(abs(float(v[1])) for v in w)
            s2 = max(abs(float(v[2])) for v in w)
            assert abs(a1 - e1) <= 1e-12 * s1
            assert abs(a2 - e2) <= 1e-11 * s2


def test_radial_operators_on_polynomials():
    cfg = config1()
    o = O.Oracle(cfg, threads=1)
    r = o.mesh(0)
    N = cfg.N
    dr1, lap, r1dr = o.mesh(1), o.mesh(3), o.mesh(4)

    def apply(band, f):
        out = np.zeros(N)
        for j in range(N):
            for n in range(max(0, j - 3), min(N, j + 4)):
                out[n] += band[j, 3 + n - j] * f[j]
        return out
    # exact for cubics at every row (boundary stencils have 4 points)
    for k in (1, 2, 3):
        f = r ** k
        assert np.abs(apply(dr1, f) - k * r ** (k - 1)).max() < 1e-9 * np.abs(k * r ** (k - 1)).max()
        # radLap r^k = k(k+1) r^(k-2)
        assert np.abs(apply(lap, f) - k * (k + 1) * r ** (k - 2)).max() < 1e-7 * k * (k + 1)
        # (1/r + d/dr) r^k = (k+1) r^(k-1)
        assert np.abs(apply(r1dr, f) - (k + 1) * r ** (k - 1)).max() < 1e-9 * (k + 1) * 4


def test_integration_weights_converge():
    # intr2dr (modules/meshs.F90:311-336) integrates each point's local Taylor
    # polynomial over half the neighbouring intervals but weights it by r_n^2
    # at the stencil centre, so sum(w) -> int r^2 dr at low order in h.
    errs = []
    for N in (32, 96):
        cfg = config1()
        cfg.N = N
        o = O.Oracle(cfg, threads=1)
        r = o.mesh(0)
        w = o.mesh(5)
        exact = (r[-1] ** 3 - r[0] ** 3) / 3.0
        errs.append(abs(w.sum() - exact) / exact)
    assert errs[0] < 1e-3
    assert errs[1] < errs[0] / 4.0
