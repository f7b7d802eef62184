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
    "tc_arm_diagnostics", "tc_probe_copy_bandwidth", "tc_set_field_ref", "tc_get_field_ref",
    "tc_set_dt_params", "tc_stat_arm", "tc_stat_collect", "tc_stat_reset", "tc_stat_get",
    "tc_tune_set",
)

# tc_params_t.precision (include/tcstep.h)
PRECISIONS = {"fp64": 0, "mixed": 1}

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
    lib.tc_set_field_ref.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_get_field_ref.argtypes = [vp, ctypes.c_char_p, dp]
    lib.tc_set_dt_params.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.tc_tune_set.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.tc_stat_arm.argtypes = [vp, ctypes.c_int]
    lib.tc_stat_collect.argtypes = [vp]
    lib.tc_stat_reset.argtypes = [vp]
    lib.tc_stat_get.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_size_t)]
    if path is None:
        _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


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


class Solver:
    """One device-resident flow (tc_solver_t)."""

    def __init__(self, nx, ny, nz, xf, xc, is_curved=True, lengths=(1.0, 1.0, 1.0), Re=1.0,
                 Pr=1.0, implicit=(1, 0, 0), scalar=1, uy_walls=(1.0, 0.0), t_walls=(0.5, -0.5),
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
        if getattr(self, "_h", None) is not None and self._h.value:
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
