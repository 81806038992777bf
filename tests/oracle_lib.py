"""ctypes binding of the CPU oracle (oracle/liblsd_oracle.so) for tests.

The oracle is the checker: tests compare the HIP library against it.
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "leeds-code_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from lsd.config import LsdConfig  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liblsd_oracle.so")
REF_AUX_SO = os.path.join(ROOT, "oracle", "_ref", "libref_aux.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def load():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(ORACLE_SO)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.POINTER(LsdConfig)]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_last_error.restype = ctypes.c_char_p
    for name in ("orc_get_table", "orc_get_mesh", "orc_get_phys"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    for name in ("orc_set_coll", "orc_get_coll"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    for name in ("orc_transform", "orc_nonlinear", "orc_predictor"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.orc_corrector.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.orc_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_get_dt.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 3
    lib.orc_gauwts.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
    lib.orc_schnla.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int]
    lib.orc_harmonics.argtypes = [ctypes.c_int] * 3 + [_ip, _ip]
    lib.orc_gbtrf.argtypes = [ctypes.c_int, _dp, _ip]
    lib.orc_gbtrs.argtypes = [ctypes.c_int, _dp, _ip, _dp]
    lib.orc_cfft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_tra_spec2phys.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.orc_tra_phys2spec.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.orc_tra_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    lib.orc_tra_qst2rtp.argtypes = [ctypes.c_void_p] + [_dp] * 9
    lib.orc_tra_rtp2qst.argtypes = [ctypes.c_void_p] + [_dp] * 9
    lib.orc_coll_norm.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.orc_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    for name in ("orc_set_ic", "orc_get_ic"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.orc_get_rot.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_set_rot.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_get_icr.argtypes = [ctypes.c_void_p, _dp]
    return lib


# lsd_step_stats of include/lsd.h, which orc_step fills: one definition for
# the library and the oracle
from lsd.gpu import StepStats  # noqa: E402


class Oracle:
    """One oracle simulation (reference layouts throughout)."""

    def __init__(self, cfg, threads=None):
        self.lib = load()
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        self.lib.orc_set_threads(threads)
        self.cfg = cfg
        self.H, self.Th, self.Ph, self.Th2 = cfg.sizes()
        self.N = cfg.N
        self._c = LsdConfig.from_buffer_copy(cfg)
        self.h = self.lib.orc_create(ctypes.byref(self._c))
        if not self.h:
            raise RuntimeError("orc_create: " + self.lib.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, self.lib.orc_last_error().decode()))

    def table(self, t):
        n = self.Th if t in (0, 1) else self.H * self.Th2
        out = np.zeros(n)
        self._chk(self.lib.orc_get_table(self.h, t, _p(out)), "table")
        return out if t in (0, 1) else out.reshape(self.H, self.Th2)

    def mesh(self, w):
        n = self.N if w in (0, 5) else 7 * self.N
        out = np.zeros(n)
        self._chk(self.lib.orc_get_mesh(self.h, w, _p(out)), "mesh")
        return out if w in (0, 5) else out.reshape(self.N, 7)

    def set_coll(self, k, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        self._chk(self.lib.orc_set_coll(self.h, k, _p(re), _p(im)), "set_coll")

    def get_coll(self, k):
        re = np.zeros((self.H, self.N))
        im = np.zeros((self.H, self.N))
        self._chk(self.lib.orc_get_coll(self.h, k, _p(re), _p(im)), "get_coll")
        return re, im

    @property
    def Nic(self):
        return max(1, self.cfg.Nic)

    def set_ic(self, k, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        assert re.shape == (self.H, self.Nic)
        self._chk(self.lib.orc_set_ic(self.h, k, _p(re), _p(im)), "set_ic")

    def get_ic(self, k):
        re = np.zeros((self.H, self.Nic))
        im = np.zeros((self.H, self.Nic))
        self._chk(self.lib.orc_get_ic(self.h, k, _p(re), _p(im)), "get_ic")
        return re, im

    def rotation(self):
        out = np.zeros(3)
        self.lib.orc_get_rot(self.h, _p(out))
        return out

    def set_rotation(self, omega):
        self.lib.orc_set_rot(self.h, float(omega))

    def ic_radii(self):
        out = np.zeros(self.Nic)
        self.lib.orc_get_icr(self.h, _p(out))
        return out

    def get_phys(self, k):
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.orc_get_phys(self.h, k, _p(p)), "get_phys")
        return p

    def set_state(self, state):
        """Fields by name (uT .. BP), and the inner-core fields icBT, icBP
        and the rotation rate omega where the state has them."""
        from lsd.initial import FIELD_ID
        for name, v in state.items():
            if name == "omega":
                self.set_rotation(v)
            elif name in ("icBT", "icBP"):
                self.set_ic(0 if name == "icBT" else 1, *v)
            else:
                self.set_coll(FIELD_ID[name], *v)

    def get_named(self, name):
        if name in ("icBT", "icBP"):
            return self.get_ic(0 if name == "icBT" else 1)
        from lsd.initial import FIELD_ID
        return self.get_coll(FIELD_ID[name])

    def transform(self):
        self._chk(self.lib.orc_transform(self.h), "transform")

    def nonlinear(self):
        self._chk(self.lib.orc_nonlinear(self.h), "nonlinear")

    def predictor(self):
        self._chk(self.lib.orc_predictor(self.h), "predictor")

    def corrector(self):
        d = ctypes.c_double(0)
        self._chk(self.lib.orc_corrector(self.h, ctypes.byref(d)), "corrector")
        return d.value

    def set_dt(self, dt):
        self._chk(self.lib.orc_set_dt(self.h, dt), "orc_set_dt")

    def get_dt(self):
        """(tim_dt, tim_cfl_dt, tim_corr_dt)"""
        v = [ctypes.c_double() for _ in range(3)]
        self._chk(self.lib.orc_get_dt(self.h, *[ctypes.byref(x) for x in v]), "orc_get_dt")
        return tuple(x.value for x in v)

    def step(self, n):
        st = StepStats()
        self._chk(self.lib.orc_step(self.h, n, ctypes.byref(st)), "step")
        return st

    def norm(self, k):
        out = np.zeros(4)
        self._chk(self.lib.orc_coll_norm(self.h, k, _p(out)), "norm")
        return out

    def energy(self, which):
        E = np.zeros(1)
        El = np.zeros(self.cfg.L)
        Em = np.zeros(self.cfg.M)
        self._chk(self.lib.orc_energy(self.h, which, _p(E), _p(El), _p(Em)), "energy")
        return float(E[0]), El, Em

    # drop-in transforms (spec layout re[n, h]; phys [n, k, j])
    def spec2phys(self, sre, sim):
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.orc_tra_spec2phys(self.h, _p(sre), _p(sim), _p(p)), "spec2phys")
        return p

    def phys2spec(self, p):
        sre = np.zeros((self.N, self.H))
        sim = np.zeros((self.N, self.H))
        self._chk(self.lib.orc_tra_phys2spec(self.h, _p(p), _p(sre), _p(sim)), "phys2spec")
        return sre, sim

    def grad(self, sre, sim):
        t = np.zeros((self.N, self.Ph, self.Th))
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.orc_tra_grad(self.h, _p(sre), _p(sim), _p(t), _p(p)), "grad")
        return t, p

    def qst2rtp(self, q, s, t):
        out = [np.zeros((self.N, self.Ph, self.Th)) for _ in range(3)]
        args = [_p(a) for a in (q[0], q[1], s[0], s[1], t[0], t[1])] + [_p(o) for o in out]
        self._chk(self.lib.orc_tra_qst2rtp(self.h, *args), "qst2rtp")
        return out

    def rtp2qst(self, r, th, ph):
        out = [np.zeros((self.N, self.H)) for _ in range(6)]
        args = [_p(a) for a in (r, th, ph)] + [_p(o) for o in out]
        self._chk(self.lib.orc_tra_rtp2qst(self.h, *args), "rtp2qst")
        return (out[0], out[1]), (out[2], out[3]), (out[4], out[5])


def gauwts(n):
    lib = load()
    x = np.zeros(n)
    w = np.zeros(n)
    lib.orc_gauwts(-1.0, 1.0, _p(x), _p(w), n)
    return x, w


def schnla(x, lh):
    lib = load()
    n = (lh + 1) * (lh + 2) // 2
    pa = np.zeros(n)
    dpa = np.zeros(n)
    lib.orc_schnla(_p(pa), _p(dpa), float(x), lh)
    return pa, dpa


def cfft(z, sign):
    lib = load()
    re = np.ascontiguousarray(z.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(z.imag, dtype=np.float64).copy()
    lib.orc_cfft(len(re), _p(re), _p(im), sign)
    return re + 1j * im


def gbtrf(ab):
    """ab: (10, n) Fortran band storage as a (n, 10) C array of columns."""
    lib = load()
    ab = np.ascontiguousarray(ab, dtype=np.float64).copy()
    n = ab.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib.orc_gbtrf(n, _p(ab), piv.ctypes.data_as(_ip))
    return ab, piv, info


def gbtrs(ab, piv, b):
    lib = load()
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib.orc_gbtrs(len(b), _p(np.ascontiguousarray(ab)), piv.ctypes.data_as(_ip), _p(b))
    return b


def ref_aux():
    """The reference's auxiliary Legendre routines built by `make -C oracle ref`
    (None when they were not built)."""
    if not os.path.exists(REF_AUX_SO):
        return None
    lib = ctypes.CDLL(REF_AUX_SO)
    lib.gauwts_.argtypes = [_dp, _dp, _dp, _dp, _ip]
    lib.schnla_.argtypes = [_dp, _dp, _dp, _ip, _ip]
    return lib


def ref_gauwts(lib, n):
    x = np.zeros(n)
    w = np.zeros(n)
    a = ctypes.c_double(-1.0)
    b = ctypes.c_double(1.0)
    nn = ctypes.c_int(n)
    lib.gauwts_(ctypes.byref(a), ctypes.byref(b), _p(x), _p(w), ctypes.byref(nn))
    return x, w


def ref_schnla(lib, x, lh):
    n = (lh + 1) * (lh + 2) // 2
    pa = np.zeros(n)
    dpa = np.zeros(n)
    xx = np.array([x], dtype=np.float64)
    l = ctypes.c_int(lh)
    one = ctypes.c_int(1)
    lib.schnla_(_p(pa), _p(dpa), _p(xx), ctypes.byref(l), ctypes.byref(one))
    return pa, dpa
