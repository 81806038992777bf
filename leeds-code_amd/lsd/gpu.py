"""ctypes binding of liblsdgpu.so (include/lsd.h).

Mirrors the reference's module interfaces for the per-timestep path:
``transform`` / ``nonlinear`` / ``predictor`` / ``corrector`` are the phases
of program/main.F90:101-157, ``step`` is the whole loop body, and the
``tra_*`` methods are the drop-ins for modules/transform.fftw3.F90's
tra_spec2phys, tra_phys2spec, tra_grad, tra_qst2rtp and tra_rtp2qst.
Errors raise ``LsdError`` with the library's message (the reference stops
the program with the same text).

There is no CPU fallback: if the HIP library or a GPU is missing, every
call fails.
"""
import ctypes
import os

import numpy as np

from .config import LsdConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "liblsdgpu.so")

_dp = ctypes.POINTER(ctypes.c_double)

EXPORTS = (
    "lsd_last_error", "lsd_config_sizes", "lsd_create", "lsd_destroy", "lsd_set_coll", "lsd_get_coll",
    "lsd_get_phys", "lsd_get_table", "lsd_get_mesh", "lsd_transform", "lsd_nonlinear", "lsd_predictor",
    "lsd_corrector", "lsd_step", "lsd_set_dt", "lsd_get_dt",
    "lsd_read_state", "lsd_write_state", "lsd_load_state", "lsd_save_state",
    "lsd_set_variant", "lsd_get_variant", "lsd_tune", "lsd_sht_roundtrip", "lsd_tra_spec2phys", "lsd_tra_phys2spec", "lsd_tra_grad",
    "lsd_tra_qst2rtp", "lsd_tra_rtp2qst", "lsd_sync", "lsd_profile_begin", "lsd_profile_end", "lsd_set_graph", "lsd_energy", "lsd_energy_noalloc",
    "lsd_step_many", "lsd_set_ic", "lsd_get_ic", "lsd_get_rotation", "lsd_set_rotation",
    "lsd_mlp_create", "lsd_mlp_destroy", "lsd_mlp_num_params", "lsd_mlp_get_params", "lsd_mlp_set_params",
    "lsd_mlp_forward", "lsd_mlp_grad", "lsd_mlp_train",
)

# kernel variants (lsd_set_variant)
V_SYN_SCALAR, V_SYN_VECTOR, V_SYN_GRAD, V_ANA, V_SOLVE = range(5)
V_NAMES = ("syn_scalar", "syn_vector", "syn_grad", "ana", "solve")

# kernel classes for lsd_profile_begin
K_SYNTH, K_ANAL, K_FFT_INV, K_FFT_FWD, K_NONLIN, K_SPEC, K_SOLVE, K_PHYS = range(8)
K_NAMES = ("synth", "anal", "fft_inv", "fft_fwd", "nonlin", "spec", "solve", "phys")


class LsdError(RuntimeError):
    pass


class StepStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("evaluations", ctypes.c_int), ("last_its", ctypes.c_int),
                ("max_its", ctypes.c_int), ("last_dterr", ctypes.c_double), ("t", ctypes.c_double),
                ("status", ctypes.c_int)]


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LsdError("liblsdgpu.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    lib.lsd_last_error.restype = ctypes.c_char_p
    lib.lsd_config_sizes.argtypes = [ctypes.POINTER(LsdConfig)] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.lsd_create.argtypes = [ctypes.POINTER(LsdConfig), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.lsd_destroy.argtypes = [ctypes.c_void_p]
    lib.lsd_destroy.restype = None
    for n in ("lsd_set_coll", "lsd_get_coll", "lsd_set_ic", "lsd_get_ic"):
        getattr(lib, n).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.lsd_get_rotation.argtypes = [ctypes.c_void_p, _dp]
    lib.lsd_set_rotation.argtypes = [ctypes.c_void_p, ctypes.c_double]
    for n in ("lsd_get_phys", "lsd_get_table", "lsd_get_mesh"):
        getattr(lib, n).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    for n in ("lsd_transform", "lsd_nonlinear", "lsd_predictor", "lsd_sync"):
        getattr(lib, n).argtypes = [ctypes.c_void_p]
    lib.lsd_corrector.argtypes = [ctypes.c_void_p, _dp]
    lib.lsd_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(StepStats)]
    lib.lsd_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.lsd_read_state.argtypes = [ctypes.c_char_p, ctypes.POINTER(LsdConfig), _dp, _dp, _dp]
    lib.lsd_write_state.argtypes = [ctypes.c_char_p, ctypes.POINTER(LsdConfig), _dp, ctypes.c_double,
                                    ctypes.c_double]
    lib.lsd_load_state.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.lsd_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.lsd_get_variant.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.lsd_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, _dp]
    lib.lsd_sht_roundtrip.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.lsd_save_state.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.lsd_get_dt.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 3
    lib.lsd_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.lsd_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    lib.lsd_energy_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    lib.lsd_step_many.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(StepStats)]
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.lsd_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                                   ctypes.POINTER(ctypes.c_void_p)]
    lib.lsd_mlp_destroy.argtypes = [ctypes.c_void_p]
    lib.lsd_mlp_destroy.restype = None
    lib.lsd_mlp_num_params.argtypes = [ctypes.c_void_p]
    lib.lsd_mlp_num_params.restype = ctypes.c_long
    lib.lsd_mlp_get_params.argtypes = [ctypes.c_void_p, _fp]
    lib.lsd_mlp_set_params.argtypes = [ctypes.c_void_p, _fp]
    lib.lsd_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_long, _fp, _fp]
    lib.lsd_mlp_grad.argtypes = [ctypes.c_void_p, ctypes.c_long, _fp, _fp, _fp, _fp]
    lib.lsd_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_long, _fp, _fp, ctypes.c_long, ctypes.c_int,
                                  ctypes.c_float, _fp]
    lib.lsd_tra_spec2phys.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.lsd_tra_phys2spec.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.lsd_tra_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    lib.lsd_tra_qst2rtp.argtypes = [ctypes.c_void_p] + [_dp] * 9
    lib.lsd_tra_rtp2qst.argtypes = [ctypes.c_void_p] + [_dp] * 9
    lib.lsd_profile_begin.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.lsd_profile_end.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int), _dp]
    _lib = lib
    return lib


def _p(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise LsdError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def config_sizes(cfg):
    lib = load()
    H, Th, Ph = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    c = LsdConfig.from_buffer_copy(cfg)
    rc = lib.lsd_config_sizes(ctypes.byref(c), ctypes.byref(H), ctypes.byref(Th), ctypes.byref(Ph))
    if rc:
        raise LsdError(lib.lsd_last_error().decode())
    return H.value, Th.value, Ph.value


class Context:
    """One device-resident run (lsd_ctx)."""

    def __init__(self, cfg, device=0):
        self.lib = load()
        self.cfg = LsdConfig.from_buffer_copy(cfg)
        self.H, self.Th, self.Ph, self.Th2 = cfg.sizes()
        self.N = cfg.N
        h = ctypes.c_void_p()
        self._chk(self.lib.lsd_create(ctypes.byref(self.cfg), device, ctypes.byref(h)), "lsd_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.lsd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            raise LsdError("%s: %s (code %d)" % (what, self.lib.lsd_last_error().decode(), rc))

    # --- fields
    def set_coll(self, field, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        assert re.size == self.H * self.N and im.size == self.H * self.N
        self._chk(self.lib.lsd_set_coll(self.h, field, _p(re), _p(im)), "lsd_set_coll")

    def get_coll(self, field):
        re = np.zeros((self.H, self.N))
        im = np.zeros((self.H, self.N))
        self._chk(self.lib.lsd_get_coll(self.h, field, _p(re), _p(im)), "lsd_get_coll")
        return re, im

    def set_state(self, state):
        """Fields by name (uT .. BP), and where the state has them the
        inner-core fields icBT, icBP and the rotation rate omega."""
        from .initial import FIELD_ID
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
        from .initial import FIELD_ID
        return self.get_coll(FIELD_ID[name])

    # --- conducting inner core and its rotation
    @property
    def Nic(self):
        return max(1, self.cfg.Nic)

    def set_ic(self, which, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        assert re.size == self.H * self.Nic and im.size == self.H * self.Nic
        self._chk(self.lib.lsd_set_ic(self.h, which, _p(re), _p(im)), "lsd_set_ic")

    def get_ic(self, which):
        re = np.zeros((self.H, self.Nic))
        im = np.zeros((self.H, self.Nic))
        self._chk(self.lib.lsd_get_ic(self.h, which, _p(re), _p(im)), "lsd_get_ic")
        return re, im

    def rotation(self):
        """(rot_omega, rot_velTorq, rot_magTorq)"""
        out = np.zeros(3)
        self._chk(self.lib.lsd_get_rotation(self.h, _p(out)), "lsd_get_rotation")
        return out

    def set_rotation(self, omega):
        self._chk(self.lib.lsd_set_rotation(self.h, float(omega)), "lsd_set_rotation")

    def get_phys(self, field):
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.lsd_get_phys(self.h, field, _p(p)), "lsd_get_phys")
        return p

    def table(self, t):
        n = self.Th if t in (0, 1) else self.H * self.Th2
        out = np.zeros(n)
        self._chk(self.lib.lsd_get_table(self.h, t, _p(out)), "lsd_get_table")
        return out if t in (0, 1) else out.reshape(self.H, self.Th2)

    def mesh(self, w):
        n = self.N if w in (0, 5) else 7 * self.N
        out = np.zeros(n)
        self._chk(self.lib.lsd_get_mesh(self.h, w, _p(out)), "lsd_get_mesh")
        return out if w in (0, 5) else out.reshape(self.N, 7)

    # --- phases
    def transform(self):
        self._chk(self.lib.lsd_transform(self.h), "lsd_transform")

    def nonlinear(self):
        self._chk(self.lib.lsd_nonlinear(self.h), "lsd_nonlinear")

    def predictor(self):
        self._chk(self.lib.lsd_predictor(self.h), "lsd_predictor")

    def corrector(self):
        d = ctypes.c_double(0.0)
        self._chk(self.lib.lsd_corrector(self.h, ctypes.byref(d)), "lsd_corrector")
        return d.value

    def step(self, n=1):
        st = StepStats()
        self._chk(self.lib.lsd_step(self.h, n, ctypes.byref(st)), "lsd_step")
        return st

    def sync(self):
        self._chk(self.lib.lsd_sync(self.h), "lsd_sync")

    def profile_begin(self, kernel_class):
        self._chk(self.lib.lsd_profile_begin(self.h, kernel_class), "lsd_profile_begin")

    def profile_end(self):
        """(total kernel ms, launches, algorithmic work) since profile_begin."""
        ms, n, w = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(self.lib.lsd_profile_end(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(w)),
                  "lsd_profile_end")
        return ms.value, n.value, w.value

    def set_dt(self, dt):
        self._chk(self.lib.lsd_set_dt(self.h, dt), "lsd_set_dt")

    def load_state(self, path):
        """io_load_state: fields, tim_t (and tim_dt for an adaptive run)."""
        self._chk(self.lib.lsd_load_state(self.h, os.fsencode(path)), "lsd_load_state")

    def save_state(self, path):
        """io_save_state of the current fields, tim_t and tim_dt."""
        self._chk(self.lib.lsd_save_state(self.h, os.fsencode(path)), "lsd_save_state")

    def set_variant(self, which, value):
        self._chk(self.lib.lsd_set_variant(self.h, which, value), "lsd_set_variant")

    def get_variant(self, which):
        v = ctypes.c_int()
        self._chk(self.lib.lsd_get_variant(self.h, which, ctypes.byref(v)), "lsd_get_variant")
        return v.value

    def variants(self):
        return {n: self.get_variant(k) for k, n in enumerate(V_NAMES)}

    def tune(self, reps=10, cache_path=None):
        """lsd_tune: times each variant, keeps the fastest; returns
        {variant name: (time of variant 0 or shape 4, time of 1 or 8)} in us."""
        t = np.zeros(2 * len(V_NAMES))
        self._chk(self.lib.lsd_tune(self.h, reps, os.fsencode(cache_path) if cache_path else None, _p(t)),
                  "lsd_tune")
        return {n: (t[2 * k], t[2 * k + 1]) for k, n in enumerate(V_NAMES)}

    def sht_roundtrip(self, kind, fields):
        """lsd_sht_roundtrip: synthesis then analysis in the context's
        precision; fields: list of (re, im) [H][N] (1 scalar or q, s, t)."""
        nf = 3 if kind else 1
        buf = np.zeros((nf, 2, self.H, self.N))
        for k, (re, im) in enumerate(fields[:nf]):
            buf[k, 0], buf[k, 1] = re, im
        out = np.zeros_like(buf)
        self._chk(self.lib.lsd_sht_roundtrip(self.h, kind, _p(buf), _p(out)), "lsd_sht_roundtrip")
        return [(out[k, 0], out[k, 1]) for k in range(nf)]

    def get_dt(self):
        """(tim_dt, tim_cfl_dt, tim_corr_dt)"""
        v = [ctypes.c_double() for _ in range(3)]
        self._chk(self.lib.lsd_get_dt(self.h, *[ctypes.byref(x) for x in v]), "lsd_get_dt")
        return tuple(x.value for x in v)

    def energy(self, which):
        """(E, El[L], Em[M]) of the velocity (0) or magnetic field (1)."""
        E = ctypes.c_double()
        El = np.zeros(self.cfg.L)
        Em = np.zeros(self.cfg.M)
        self._chk(self.lib.lsd_energy(self.h, which, ctypes.byref(E), _p(El), _p(Em)), "lsd_energy")
        return E.value, El, Em

    def energy_noalloc(self, which):
        """As energy(), through lsd_energy_noalloc (no device allocation)."""
        E = ctypes.c_double()
        El = np.zeros(self.cfg.L)
        Em = np.zeros(self.cfg.M)
        self._chk(self.lib.lsd_energy_noalloc(self.h, which, ctypes.byref(E), _p(El), _p(Em)),
                  "lsd_energy_noalloc")
        return E.value, El, Em

    def set_graph(self, on):
        """lsd_step from captured HIP graphs (True, default) or eager launches."""
        self._chk(self.lib.lsd_set_graph(self.h, 1 if on else 0), "lsd_set_graph")

    # --- drop-in transforms (spec re[n, h]; phys [n, k, j])
    def spec2phys(self, sre, sim):
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.lsd_tra_spec2phys(self.h, _p(sre), _p(sim), _p(p)), "lsd_tra_spec2phys")
        return p

    def phys2spec(self, p):
        sre = np.zeros((self.N, self.H))
        sim = np.zeros((self.N, self.H))
        self._chk(self.lib.lsd_tra_phys2spec(self.h, _p(np.ascontiguousarray(p)), _p(sre), _p(sim)),
                  "lsd_tra_phys2spec")
        return sre, sim

    def grad(self, sre, sim):
        t = np.zeros((self.N, self.Ph, self.Th))
        p = np.zeros((self.N, self.Ph, self.Th))
        self._chk(self.lib.lsd_tra_grad(self.h, _p(sre), _p(sim), _p(t), _p(p)), "lsd_tra_grad")
        return t, p

    def qst2rtp(self, q, s, t):
        out = [np.zeros((self.N, self.Ph, self.Th)) for _ in range(3)]
        args = [_p(a) for a in (q[0], q[1], s[0], s[1], t[0], t[1])] + [_p(o) for o in out]
        self._chk(self.lib.lsd_tra_qst2rtp(self.h, *args), "lsd_tra_qst2rtp")
        return out

    def rtp2qst(self, r, th, ph):
        out = [np.zeros((self.N, self.H)) for _ in range(6)]
        args = [_p(np.ascontiguousarray(a)) for a in (r, th, ph)] + [_p(o) for o in out]
        self._chk(self.lib.lsd_tra_rtp2qst(self.h, *args), "lsd_tra_rtp2qst")
        return (out[0], out[1]), (out[2], out[3]), (out[4], out[5])


class StepManyError(LsdError):
    """lsd_step_many stopped at least one member; ``stats`` holds every
    member's StepStats (``status`` != 0 for the members that stopped)."""

    def __init__(self, msg, stats):
        super().__init__(msg)
        self.stats = stats


def read_state(path, cfg):
    """lsd_read_state (no GPU needed): ({name: (re, im)} in the coll layout
    for uT, uP, C, X, BT, BP; t; dt)."""
    lib = load()
    H = cfg.sizes()[0]
    buf = np.zeros((6, 2, H, cfg.N))
    t, dt = ctypes.c_double(), ctypes.c_double()
    c = LsdConfig.from_buffer_copy(cfg)
    rc = lib.lsd_read_state(os.fsencode(path), ctypes.byref(c), _p(buf), ctypes.byref(t), ctypes.byref(dt))
    if rc != 0:
        raise LsdError("lsd_read_state: %s (code %d)" % (lib.lsd_last_error().decode(), rc))
    names = ("uT", "uP", "C", "X", "BT", "BP")
    return {n: (buf[k, 0].copy(), buf[k, 1].copy()) for k, n in enumerate(names)}, t.value, dt.value


def write_state(path, cfg, state, t, dt):
    """lsd_write_state (no GPU needed) of {name: (re, im)}."""
    lib = load()
    H = cfg.sizes()[0]
    buf = np.zeros((6, 2, H, cfg.N))
    for k, n in enumerate(("uT", "uP", "C", "X", "BT", "BP")):
        if n in state:
            buf[k, 0], buf[k, 1] = state[n]
    c = LsdConfig.from_buffer_copy(cfg)
    rc = lib.lsd_write_state(os.fsencode(path), ctypes.byref(c), _p(buf), t, dt)
    if rc != 0:
        raise LsdError("lsd_write_state: %s (code %d)" % (lib.lsd_last_error().decode(), rc))


def step_many(contexts, nsteps):
    """lsd_step_many: steps the contexts together; returns their StepStats.
    Raises StepManyError (carrying every member's stats) if a member stops."""
    lib = load()
    n = len(contexts)
    hs = (ctypes.c_void_p * n)(*[c.h.value for c in contexts])
    st = (StepStats * n)()
    rc = lib.lsd_step_many(hs, n, nsteps, st)
    if rc != 0:
        raise StepManyError("lsd_step_many: %s (code %d)" % (lib.lsd_last_error().decode(), rc), list(st))
    return list(st)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The closure network on the GPU (lsd_mlp_*): 3 layers with GELU,
    BF16 weights on the matrix cores, FP32 accumulation, Adam training."""

    def __init__(self, n_in, width, n_out, seed=0):
        self.lib = load()
        self.n_in, self.width, self.n_out = n_in, width, n_out
        h = ctypes.c_void_p()
        rc = self.lib.lsd_mlp_create(n_in, width, n_out, seed, ctypes.byref(h))
        if rc != 0:
            raise LsdError("lsd_mlp_create: %s" % self.lib.lsd_last_error().decode())
        self.h = h
        self.n_params = self.lib.lsd_mlp_num_params(self.h)

    def _chk(self, rc, what):
        if rc != 0:
            raise LsdError("%s: %s (code %d)" % (what, self.lib.lsd_last_error().decode(), rc))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.lsd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self._chk(self.lib.lsd_mlp_get_params(self.h, _pf(p)), "lsd_mlp_get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        self._chk(self.lib.lsd_mlp_set_params(self.h, _pf(p)), "lsd_mlp_set_params")

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self._chk(self.lib.lsd_mlp_forward(self.h, x.shape[0], _pf(x), _pf(y)), "lsd_mlp_forward")
        return y

    def grad(self, x, target):
        x, target = _f32(x), _f32(target)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        self._chk(self.lib.lsd_mlp_grad(self.h, x.shape[0], _pf(x), _pf(target), _pf(loss), _pf(g)), "lsd_mlp_grad")
        return float(loss[0]), g

    def train(self, x, target, batch, steps, lr):
        x, target = _f32(x), _f32(target)
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        self._chk(self.lib.lsd_mlp_train(self.h, x.shape[0], _pf(x), _pf(target), batch, steps, lr, _pf(loss)),
                  "lsd_mlp_train")
        return loss[:steps]
