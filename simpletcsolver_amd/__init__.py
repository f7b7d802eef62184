"""MI355X-native RK3 time step of SimpleTCSolver (Taylor-Couette / channel DNS).

The product is the C-ABI library libtcstep.so (C host code + hand-written gfx950 HIP
kernels, headers in include/).  This module is a thin ctypes binding of that library for
tests, benchmarks and Python drivers; it never falls back to another implementation: if the
library is missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtcstep.so")
_LIB = None

FIELDS = ("ux", "uy", "uz", "p", "psi", "t")

# every symbol the C-ABI headers declare (include/*.h); tests check the library exports them
EXPORTED = (
    "tc_params_default", "tc_create", "tc_destroy", "tc_last_error", "tc_set_field",
    "tc_get_field", "tc_step", "tc_step_async", "tc_synchronize", "tc_get_dt",
    "tc_decide_dt", "tc_predict", "tc_potential", "tc_correct", "tc_pressure",
    "tc_divergence", "tc_injection", "tc_dissipation", "tc_total_energy",
    "tc_profile_enable", "tc_profile_read", "tc_device_bytes", "tc_ic_grid", "tc_ic_fields",
    "tc_ic_fields_noalloc",
    "tc_arm_diagnostics", "tc_probe_copy_bandwidth", "tc_set_field_ref", "tc_get_field_ref",
    "tc_set_dt_params", "tc_stat_arm", "tc_stat_collect", "tc_stat_reset", "tc_stat_get",
    "tc_tune_set", "tc_autotune", "tc_ensemble_create", "tc_ensemble_member",
    "tc_ensemble_autotune", "tc_ensemble_step", "tc_ensemble_synchronize",
    "tc_ensemble_invalidate", "tc_ensemble_destroy", "tc_cnn_create", "tc_cnn_destroy",
    "tc_cnn_set_layer", "tc_cnn_get_layer", "tc_cnn_predict", "tc_cnn_gradients",
    "tc_cnn_train_step", "tc_closure_attach",
)

# closure network layer shapes [27][co][ci] (include/tcstep.h)
CNN_SHAPES = ((27, 32, 16), (27, 32, 32), (27, 16, 32))

# tc_params_t.precision (include/tcstep.h)
PRECISIONS = {"fp64": 0, "mixed": 1, "mixed_refined": 2}

# running statistics of statistics.c, in its accumulation order
STATS = ("ux1", "ux2", "uy1", "uy2", "uz1", "uz2", "adv", "dif")


class TcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("is_curved", ctypes.c_int),
        ("lengths", ctypes.c_double * 3),
        ("Re", ctypes.c_double), ("Pr", ctypes.c_double),
        ("implicit", ctypes.c_int * 3),
        ("scalar", ctypes.c_int),
        ("uy_xm", ctypes.c_double), ("uy_xp", ctypes.c_double),
        ("t_xm", ctypes.c_double), ("t_xp", ctypes.c_double),
        ("coef_dt_adv", ctypes.c_double), ("coef_dt_dif", ctypes.c_double),
        ("dt_fixed", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class TcError(RuntimeError):
    pass


def load_library(path=None):
    """Load libtcstep.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TcError(f"{p} is missing: build it with __graft_entry__.build() (make -C simpletcsolver_amd/csrc)")
    lib = ctypes.CDLL(p)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    lib.tc_last_error.restype = ctypes.c_char_p
    lib.tc_params_default.argtypes = [ctypes.POINTER(TcParams)]
    lib.tc_create.argtypes = [ctypes.POINTER(TcParams), dp, dp, ctypes.POINTER(vp)]
    lib.tc_destroy.argtypes = [vp]
    lib.tc_set_field.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_get_field.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_step.argtypes = [vp, dp]
    lib.tc_step_async.argtypes = [vp, ctypes.c_int]
    lib.tc_synchronize.argtypes = [vp]
    lib.tc_get_dt.argtypes = [vp, dp]
    lib.tc_decide_dt.argtypes = [vp, dp]
    for name in ("tc_predict", "tc_potential", "tc_correct", "tc_pressure"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.tc_divergence.argtypes = [vp, dp]
    for name in ("tc_injection", "tc_dissipation", "tc_total_energy"):
        getattr(lib, name).argtypes = [vp, dp]
    lib.tc_profile_enable.argtypes = [vp, ctypes.c_int]
    lib.tc_arm_diagnostics.argtypes = [vp, ctypes.c_int]
    lib.tc_profile_read.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_longlong)]
    lib.tc_probe_copy_bandwidth.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, dp]
    lib.tc_device_bytes.restype = ctypes.c_size_t
    lib.tc_device_bytes.argtypes = [vp]
    lib.tc_ic_grid.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, dp, dp]
    lib.tc_ic_fields.argtypes = [vp, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_double]
    lib.tc_ic_fields_noalloc.argtypes = lib.tc_ic_fields.argtypes
    lib.tc_set_field_ref.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_get_field_ref.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_set_dt_params.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.tc_tune_set.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.tc_autotune.argtypes = [vp, ctypes.c_int, ip, ip, dp]
    lib.tc_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(TcParams), ctypes.POINTER(dp),
                                       ctypes.POINTER(dp), ctypes.POINTER(vp)]
    lib.tc_ensemble_member.restype = vp
    lib.tc_ensemble_member.argtypes = [vp, ctypes.c_int]
    lib.tc_ensemble_autotune.argtypes = [vp, ctypes.c_int, ip, ip]
    for name in ("tc_ensemble_synchronize", "tc_ensemble_invalidate", "tc_ensemble_destroy"):
        getattr(lib, name).argtypes = [vp]
    lib.tc_ensemble_step.argtypes = [vp, ctypes.c_int]
    fp = ctypes.POINTER(ctypes.c_float)
    lib.tc_cnn_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_ulonglong, ctypes.c_int, ctypes.POINTER(vp)]
    lib.tc_cnn_destroy.argtypes = [vp]
    lib.tc_cnn_set_layer.argtypes = [vp, ctypes.c_int, fp, fp]
    lib.tc_cnn_get_layer.argtypes = [vp, ctypes.c_int, fp, fp, fp, fp]
    lib.tc_cnn_predict.argtypes = [vp, fp, fp]
    lib.tc_cnn_gradients.argtypes = [vp, fp, fp, dp]
    lib.tc_cnn_train_step.argtypes = [vp, fp, fp, ctypes.c_float, dp]
    lib.tc_closure_attach.argtypes = [vp, vp, fp, fp, fp]
    lib.tc_stat_arm.argtypes = [vp, ctypes.c_int]
    lib.tc_stat_collect.argtypes = [vp]
    lib.tc_stat_reset.argtypes = [vp]
    lib.tc_stat_get.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_size_t)]
    if path is None:
        _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc, lib, what):
    if rc != 0:
        raise TcError(f"{what}: {lib.tc_last_error().decode(errors='replace')}")


def ic_grid(nx, lx, is_curved, stretch=2.0):
    """tanh-clustered radial grid: xf (nx+1), xc (nx+2)."""
    lib = load_library()
    xf = np.zeros(nx + 1)
    xc = np.zeros(nx + 2)
    _check(lib.tc_ic_grid(nx, float(lx), int(bool(is_curved)), float(stretch), _dp(xf), _dp(xc)), lib, "tc_ic_grid")
    return xf, xc


def npy_shape(name, nx, ny, nz):
    return (nz, ny, nx + 1) if name == "ux" else (nz, ny, nx + 2)


def ref_shape(name, nx, ny, nz):
    """shape of the reference's host array of a field (array_macros/fluid/*.h): y and z
    halos included"""
    return (nz + 2, ny + 2, nx + 1) if name == "ux" else (nz + 2, ny + 2, nx + 2)


def stat_shape(name, nx, ny, nz):
    """shape of a statistics accumulator (array_macros/statistics/*.h)"""
    return (nz, ny, nx + 1) if name in ("ux1", "ux2", "adv", "dif") else (nz, ny, nx + 2)


def make_params(nx, ny, nz, is_curved=True, lengths=(1.0, 1.0, 1.0), Re=1.0, Pr=1.0,
                implicit=(1, 0, 0), scalar=1, uy_walls=(1.0, 0.0), t_walls=(0.5, -0.5),
                coef_dt_adv=1.0, coef_dt_dif=1.0, dt_fixed=0.0, device=0, precision="fp64"):
    lib = load_library()
    p = TcParams()
    lib.tc_params_default(ctypes.byref(p))
    p.nx, p.ny, p.nz = nx, ny, nz
    p.is_curved = int(bool(is_curved))
    p.lengths[:] = [float(v) for v in lengths]
    p.Re, p.Pr = float(Re), float(Pr)
    p.implicit[:] = [int(bool(v)) for v in implicit]
    p.scalar = int(bool(scalar))
    p.uy_xm, p.uy_xp = [float(v) for v in uy_walls]
    p.t_xm, p.t_xp = [float(v) for v in t_walls]
    p.coef_dt_adv, p.coef_dt_dif = float(coef_dt_adv), float(coef_dt_dif)
    p.dt_fixed = float(dt_fixed)
    p.device = int(device)
    if precision not in PRECISIONS:
        raise ValueError(f"precision {precision!r}: one of {tuple(PRECISIONS)}")
    p.precision = PRECISIONS[precision]
    return p


class Solver:
    """One device-resident flow (tc_solver_t)."""

    @classmethod
    def _member(cls, lib, handle, params):
        """a view of an ensemble member: the ensemble owns the handle"""
        self = cls.__new__(cls)
        self._lib = lib
        self._h = ctypes.c_void_p(handle)
        self._owned = False
        self.params = params
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.precision = [k for k, v in PRECISIONS.items() if v == params.precision][0]
        return self

    def __init__(self, nx, ny, nz, xf, xc, is_curved=True, lengths=(1.0, 1.0, 1.0), Re=1.0,
                 Pr=1.0, implicit=(1, 0, 0), scalar=1, uy_walls=(1.0, 0.0), t_walls=(0.5, -0.5),
                 coef_dt_adv=1.0, coef_dt_dif=1.0, dt_fixed=0.0, device=0, precision="fp64"):
        self._owned = True
        lib = load_library()
        p = TcParams()
        lib.tc_params_default(ctypes.byref(p))
        p.nx, p.ny, p.nz = nx, ny, nz
        p.is_curved = int(bool(is_curved))
        p.lengths[:] = [float(v) for v in lengths]
        p.Re, p.Pr = float(Re), float(Pr)
        p.implicit[:] = [int(bool(v)) for v in implicit]
        p.scalar = int(bool(scalar))
        p.uy_xm, p.uy_xp = [float(v) for v in uy_walls]
        p.t_xm, p.t_xp = [float(v) for v in t_walls]
        p.coef_dt_adv, p.coef_dt_dif = float(coef_dt_adv), float(coef_dt_dif)
        p.dt_fixed = float(dt_fixed)
        p.device = int(device)
        if precision not in PRECISIONS:
            raise ValueError(f"precision {precision!r}: one of {tuple(PRECISIONS)}")
        p.precision = PRECISIONS[precision]
        self.precision = precision
        self.params = p
        self.nx, self.ny, self.nz = nx, ny, nz
        xf = np.ascontiguousarray(xf, dtype=np.float64)
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        h = ctypes.c_void_p()
        self._lib = lib
        _check(lib.tc_create(ctypes.byref(p), _dp(xf), _dp(xc), ctypes.byref(h)), lib, "tc_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and getattr(self, "_owned", True):
            self._lib.tc_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        _check(rc, self._lib, what)

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != npy_shape(name, self.nx, self.ny, self.nz):
            raise ValueError(f"{name}: shape {a.shape}")
        self._c(self._lib.tc_set_field(self._h, name.encode(), _dp(a)), "tc_set_field")

    def get_field(self, name):
        a = np.zeros(npy_shape(name, self.nx, self.ny, self.nz))
        self._c(self._lib.tc_get_field(self._h, name.encode(), _dp(a)), "tc_get_field")
        return a

    def load_fields(self, fields):
        for name, a in fields.items():
            self.set_field(name, a)

    def fields(self):
        return {n: self.get_field(n) for n in FIELDS}

    def step(self):
        dt = ctypes.c_double(0.0)
        self._c(self._lib.tc_step(self._h, ctypes.byref(dt)), "tc_step")
        return dt.value

    def step_async(self, n):
        self._c(self._lib.tc_step_async(self._h, int(n)), "tc_step_async")

    def synchronize(self):
        self._c(self._lib.tc_synchronize(self._h), "tc_synchronize")

    def dt(self):
        dt = ctypes.c_double(0.0)
        self._c(self._lib.tc_get_dt(self._h, ctypes.byref(dt)), "tc_get_dt")
        return dt.value

    def decide_dt(self):
        dt = ctypes.c_double(0.0)
        self._c(self._lib.tc_decide_dt(self._h, ctypes.byref(dt)), "tc_decide_dt")
        return dt.value

    def predict(self, rk, dt):
        self._c(self._lib.tc_predict(self._h, rk, dt), "tc_predict")

    def potential(self, rk, dt):
        self._c(self._lib.tc_potential(self._h, rk, dt), "tc_potential")

    def correct(self, rk, dt):
        self._c(self._lib.tc_correct(self._h, rk, dt), "tc_correct")

    def pressure(self, rk, dt):
        self._c(self._lib.tc_pressure(self._h, rk, dt), "tc_pressure")

    def divergence(self):
        v = ctypes.c_double(0.0)
        self._c(self._lib.tc_divergence(self._h, ctypes.byref(v)), "tc_divergence")
        return v.value

    def _vec(self, fn, n):
        out = np.zeros(n)
        self._c(fn(self._h, _dp(out)), fn.__name__)
        return out

    def injection(self):
        return self._vec(self._lib.tc_injection, 2)

    def dissipation(self):
        return self._vec(self._lib.tc_dissipation, 2)

    def total_energy(self):
        return self._vec(self._lib.tc_total_energy, 4)

    def ic(self, seed=1, amp=1e-2, t_amp=0.0, nmodes=0, mode_amp=0.0):
        self._c(self._lib.tc_ic_fields(self._h, int(seed), float(amp), float(t_amp), int(nmodes),
                                       float(mode_amp)), "tc_ic_fields")

    def arm_diagnostics(self, on=True):
        self._c(self._lib.tc_arm_diagnostics(self._h, int(bool(on))), "tc_arm_diagnostics")

    def profile(self, on=True):
        self._c(self._lib.tc_profile_enable(self._h, int(bool(on))), "tc_profile_enable")

    def profile_read(self, stage):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_longlong(0)
        self._c(self._lib.tc_profile_read(self._h, stage.encode(), ctypes.byref(ms), ctypes.byref(n)),
                "tc_profile_read")
        return ms.value, n.value

    def copy_bandwidth(self, nbytes=2 << 30, iters=10):
        """measured HBM copy bandwidth in GB/s (read + write bytes)"""
        v = ctypes.c_double(0.0)
        self._c(self._lib.tc_probe_copy_bandwidth(self._h, int(nbytes), int(iters), ctypes.byref(v)),
                "tc_probe_copy_bandwidth")
        return v.value

    def device_bytes(self):
        return self._lib.tc_device_bytes(self._h)

    def set_dt_params(self, coef_dt_adv, coef_dt_dif, dt_fixed=0.0):
        self._c(self._lib.tc_set_dt_params(self._h, float(coef_dt_adv), float(coef_dt_dif),
                                           float(dt_fixed)), "tc_set_dt_params")

    def set_field_ref(self, name, a):
        """replace a field from the reference's host layout (y and z halos included)"""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != ref_shape(name, self.nx, self.ny, self.nz):
            raise ValueError(f"{name}: shape {a.shape}")
        self._c(self._lib.tc_set_field_ref(self._h, name.encode(), _dp(a)), "tc_set_field_ref")

    def get_field_ref(self, name):
        a = np.zeros(ref_shape(name, self.nx, self.ny, self.nz))
        self._c(self._lib.tc_get_field_ref(self._h, name.encode(), _dp(a)), "tc_get_field_ref")
        return a

    def tune(self, kernel, kb, ty):
        """launch shape of a cell kernel: kb planes per block, ty j rows per block"""
        self._c(self._lib.tc_tune_set(self._h, kernel.encode(), int(kb), int(ty)), "tc_tune_set")

    def stat_arm(self, on=True):
        self._c(self._lib.tc_stat_arm(self._h, int(bool(on))), "tc_stat_arm")

    def stat_collect(self):
        self._c(self._lib.tc_stat_collect(self._h), "tc_stat_collect")

    def stat_reset(self):
        self._c(self._lib.tc_stat_reset(self._h), "tc_stat_reset")

    def stats(self):
        """the accumulators and the sample count"""
        out = {}
        num = ctypes.c_size_t(0)
        for name in STATS:
            a = np.zeros(stat_shape(name, self.nx, self.ny, self.nz))
            self._c(self._lib.tc_stat_get(self._h, name.encode(), _dp(a), ctypes.byref(num)), "tc_stat_get")
            out[name] = a
        return out, num.value

    def autotune(self, reps=3):
        """time the predictor over the candidate launch shapes and keep the fastest"""
        kb, ty, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        self._c(self._lib.tc_autotune(self._h, int(reps), ctypes.byref(kb), ctypes.byref(ty), ctypes.byref(ms)),
                "tc_autotune")
        return kb.value, ty.value, ms.value


class Ensemble:
    """n independent flows of any sizes on one device, stepped together through recorded
    graphs (tc_ensemble_t).  specs: dicts with nx, ny, nz, xf, xc and Solver keyword
    arguments."""

    def __init__(self, specs):
        lib = load_library()
        self._lib = lib
        n = len(specs)
        params = (TcParams * n)()
        self._keep = []
        xfp = (ctypes.POINTER(ctypes.c_double) * n)()
        xcp = (ctypes.POINTER(ctypes.c_double) * n)()
        for m, sp in enumerate(specs):
            sp = dict(sp)
            xf = np.ascontiguousarray(sp.pop("xf"), dtype=np.float64)
            xc = np.ascontiguousarray(sp.pop("xc"), dtype=np.float64)
            params[m] = make_params(**sp)
            self._keep += [xf, xc]
            xfp[m] = _dp(xf)
            xcp[m] = _dp(xc)
        h = ctypes.c_void_p()
        _check(lib.tc_ensemble_create(n, params, xfp, xcp, ctypes.byref(h)), lib, "tc_ensemble_create")
        self._h = h
        self.members = [Solver._member(lib, lib.tc_ensemble_member(h, m), params[m]) for m in range(n)]

    def _c(self, rc, what):
        _check(rc, self._lib, what)

    def autotune(self, reps=3):
        n = len(self.members)
        kb = (ctypes.c_int * n)()
        ty = (ctypes.c_int * n)()
        self._c(self._lib.tc_ensemble_autotune(self._h, int(reps), kb, ty), "tc_ensemble_autotune")
        return list(kb), list(ty)

    def step(self, n=1):
        self._c(self._lib.tc_ensemble_step(self._h, int(n)), "tc_ensemble_step")

    def synchronize(self):
        self._c(self._lib.tc_ensemble_synchronize(self._h), "tc_ensemble_synchronize")

    def invalidate(self):
        self._c(self._lib.tc_ensemble_invalidate(self._h), "tc_ensemble_invalidate")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for m in self.members:
                m._h = None
            self._lib.tc_ensemble_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cnn:
    """The closure network (tc_cnn_t) on batches of nx x ny x nz cells."""

    def __init__(self, nx, ny, nz, batch=1, seed=1, device=0):
        lib = load_library()
        self._lib = lib
        self.shape = (batch, nz, ny, nx)
        h = ctypes.c_void_p()
        _check(lib.tc_cnn_create(nx, ny, nz, batch, int(seed), int(device), ctypes.byref(h)), lib, "tc_cnn_create")
        self._h = h

    def _c(self, rc, what):
        _check(rc, self._lib, what)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.tc_cnn_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != self.shape + (5,):
            raise ValueError(f"inputs {x.shape}, expected {self.shape + (5,)}")
        return x

    def _y(self, y):
        y = np.ascontiguousarray(y, dtype=np.float32)
        if y.shape != self.shape + (4,):
            raise ValueError(f"targets {y.shape}, expected {self.shape + (4,)}")
        return y

    def set_layer(self, l, w=None, b=None):
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        b = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
        if w is not None and w.shape != CNN_SHAPES[l]:
            raise ValueError(f"layer {l} weights {w.shape}")
        self._c(self._lib.tc_cnn_set_layer(self._h, l, _fp(w), _fp(b)), "tc_cnn_set_layer")

    def layer(self, l):
        """weights, biases, and the gradients of the last gradients / train_step call"""
        w = np.zeros(CNN_SHAPES[l], dtype=np.float32)
        b = np.zeros(CNN_SHAPES[l][1], dtype=np.float32)
        gw = np.zeros_like(w)
        gb = np.zeros_like(b)
        self._c(self._lib.tc_cnn_get_layer(self._h, l, _fp(w), _fp(b), _fp(gw), _fp(gb)), "tc_cnn_get_layer")
        return w, b, gw, gb

    def predict(self, x):
        x = self._x(x)
        y = np.zeros(self.shape + (4,), dtype=np.float32)
        self._c(self._lib.tc_cnn_predict(self._h, _fp(x), _fp(y)), "tc_cnn_predict")
        return y

    def gradients(self, x, y):
        x, y = self._x(x), self._y(y)
        loss = ctypes.c_double(0.0)
        self._c(self._lib.tc_cnn_gradients(self._h, _fp(x), _fp(y), ctypes.byref(loss)), "tc_cnn_gradients")
        return loss.value

    def train_step(self, x, y, lr):
        x, y = self._x(x), self._y(y)
        loss = ctypes.c_double(0.0)
        self._c(self._lib.tc_cnn_train_step(self._h, _fp(x), _fp(y), float(lr), ctypes.byref(loss)),
                "tc_cnn_train_step")
        return loss.value


def attach_closure(solver, cnn, in_mean, in_std, out_scale):
    """couple a batch-1 network of the solver's grid (cnn None: detach)"""
    lib = solver._lib
    m = np.ascontiguousarray(in_mean, dtype=np.float32)
    sd = np.ascontiguousarray(in_std, dtype=np.float32)
    sc = np.ascontiguousarray(out_scale, dtype=np.float32)
    _check(lib.tc_closure_attach(solver._h, None if cnn is None else cnn._h, _fp(m), _fp(sd), _fp(sc)),
           lib, "tc_closure_attach")
    solver._closure = cnn
