"""ctypes binding of the CPU oracle (orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, as the checker.  The shipped library never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcConfig(ctypes.Structure):
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
        ("nthreads", ctypes.c_int),
    ]


def library():
    """Load liborc.so, building it with make when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liborc.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.check_call(["make", "-C", _HERE], stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ctypes.POINTER(OrcConfig), dp, dp]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_array.restype = dp
    lib.orc_array.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_hz.restype = ctypes.c_double
    lib.orc_hz.argtypes = [vp]
    for name in ("orc_update_boundaries", "orc_compute_tensors"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_decide_dt.argtypes = [vp, dp]
    lib.orc_integrate.argtypes = [vp, dp]
    for name in ("orc_predict", "orc_potential", "orc_correct", "orc_pressure"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, ctypes.c_double]
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_divergence.restype = ctypes.c_double
    lib.orc_divergence.argtypes = [vp]
    for name in ("orc_injection", "orc_dissipation", "orc_total_energy"):
        getattr(lib, name).argtypes = [vp, dp]
    lib.orc_fft_c2c.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
    lib.orc_fft_r2c.argtypes = [ctypes.c_int, dp, dp]
    lib.orc_fft_c2r.argtypes = [ctypes.c_int, dp, dp]
    for name in ("orc_gtsv", "orc_gtsv_complex", "orc_tdm_periodic"):
        getattr(lib, name).argtypes = [ctypes.c_int, dp, dp, dp, dp]
    lib.orc_rkcoefs.argtypes = [dp]
    lib.orc_set_mutation.argtypes = [ctypes.c_int]
    lib.orc_stat_collect.argtypes = [vp]
    lib.orc_stat_collect.restype = ctypes.c_size_t
    _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# stored shapes (z, y, x) of each named array in the reference host layout
def _shape(name, nx, ny, nz):
    if name in ("ux", "lxy", "lxz", "lyx0", "lyx1", "lzx"):
        return (nz + 2, ny + 2, nx + 1)
    if name in ("uy", "uz", "p", "psi", "t"):
        return (nz + 2, ny + 2, nx + 2)
    if name in ("lxx", "lyy0", "lyy1", "lzy", "lyz", "lzz"):
        return (nz + 2, ny + 2, nx)
    if name in ("ux1", "ux2", "adv", "dif"):
        return (nz, ny, nx + 1)
    if name in ("uy1", "uy2", "uz1", "uz2"):
        return (nz, ny, nx + 2)
    if name.startswith("srcux"):
        return (nz, ny, nx - 1)
    if name.startswith("src"):
        return (nz, ny, nx)
    return None


class Oracle:
    """One CPU restatement of the reference solver state (host layouts of the reference)."""

    def __init__(self, nx, ny, nz, xf, xc, is_curved=True, lengths=(1.0, 1.0, 1.0),
                 Re=100.0, Pr=1.0, implicit=(1, 0, 0), scalar=1, uy_walls=(1.0, 0.0),
                 t_walls=(0.5, -0.5), coef_dt_adv=1.0, coef_dt_dif=1.0, dt_fixed=0.0,
                 nthreads=0):
        lib = library()
        cfg = OrcConfig()
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.is_curved = int(bool(is_curved))
        cfg.lengths[:] = [float(v) for v in lengths]
        cfg.Re, cfg.Pr = float(Re), float(Pr)
        cfg.implicit[:] = [int(bool(v)) for v in implicit]
        cfg.scalar = int(bool(scalar))
        cfg.uy_xm, cfg.uy_xp = [float(v) for v in uy_walls]
        cfg.t_xm, cfg.t_xp = [float(v) for v in t_walls]
        cfg.coef_dt_adv, cfg.coef_dt_dif = float(coef_dt_adv), float(coef_dt_dif)
        cfg.dt_fixed = float(dt_fixed)
        cfg.nthreads = int(nthreads)
        self.cfg = cfg
        self.nx, self.ny, self.nz = nx, ny, nz
        xf = np.ascontiguousarray(xf, dtype=np.float64)
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        assert xf.shape == (nx + 1,) and xc.shape == (nx + 2,)
        self._h = lib.orc_create(ctypes.byref(cfg), _dp(xf), _dp(xc))
        if not self._h:
            raise ValueError("orc_create failed")
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, name):
        """numpy view (z, y, x) of a named array; writes go to the oracle's memory."""
        n = ctypes.c_size_t(0)
        ptr = self._lib.orc_array(self._h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        flat = np.ctypeslib.as_array(ptr, shape=(n.value,))
        shape = _shape(name, self.nx, self.ny, self.nz)
        return flat.reshape(shape) if shape else flat

    @property
    def hz(self):
        return self._lib.orc_hz(self._h)

    def load_fields(self, fields):
        """fields: dict of NPY-layout arrays (nz, ny, nx+1|nx+2) -> interior planes, then BCs."""
        for name in ("ux", "uy", "uz", "p", "psi", "t"):
            if name in fields:
                self.array(name)[1:-1, 1:-1, :] = fields[name]
        self._lib.orc_update_boundaries(self._h)

    def npy_fields(self):
        """current fields in the NPY layout of the reference (no y/z halos)."""
        return {n: self.array(n)[1:-1, 1:-1, :].copy() for n in ("ux", "uy", "uz", "p", "psi", "t")}

    def update_boundaries(self):
        self._lib.orc_update_boundaries(self._h)

    def compute_tensors(self):
        self._lib.orc_compute_tensors(self._h)

    def decide_dt(self):
        dt = ctypes.c_double(0.0)
        self._lib.orc_decide_dt(self._h, ctypes.byref(dt))
        return dt.value

    def integrate(self):
        dt = ctypes.c_double(0.0)
        if self._lib.orc_integrate(self._h, ctypes.byref(dt)) != 0:
            raise RuntimeError("orc_integrate failed")
        return dt.value

    def predict(self, rk, dt):
        self._lib.orc_predict(self._h, rk, dt)

    def potential(self, rk, dt):
        self._lib.orc_potential(self._h, rk, dt)

    def correct(self, rk, dt):
        self._lib.orc_correct(self._h, rk, dt)

    def pressure(self, rk, dt):
        self._lib.orc_pressure(self._h, rk, dt)

    def divergence(self):
        return self._lib.orc_divergence(self._h)

    def injection(self):
        out = np.zeros(2)
        self._lib.orc_injection(self._h, _dp(out))
        return out

    def dissipation(self):
        out = np.zeros(2)
        self._lib.orc_dissipation(self._h, _dp(out))
        return out

    def total_energy(self):
        out = np.zeros(4)
        self._lib.orc_total_energy(self._h, _dp(out))
        return out

    def stat_collect(self):
        """statistics.c collect (the dif term reads the tensors of the last predict)"""
        return self._lib.orc_stat_collect(self._h)

    def stats(self):
        return {n: self.array(n).copy() for n in ("ux1", "ux2", "uy1", "uy2", "uz1", "uz2", "adv", "dif")}


# planted mistakes of orc.h (ORC_MUT_*), in enum order; only the pin tests set one
MUTATIONS = ("none", "ux_x_f", "uy_y_f", "uz_z_f", "t_y_f", "ux_y_2f", "uz_x_2f", "pf_alpha",
             "ab_swap", "no_rotate", "p_f", "p_no_z", "p_x_index")


def set_mutation(name):
    """plant one of the mistakes of orc.h in every oracle of this process ("none": the
    reference's algorithm)"""
    library().orc_set_mutation(MUTATIONS.index(name))


def fft_c2c(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    library().orc_fft_c2c(x.size, int(sign), _dp(x.view(np.float64)), _dp(out.view(np.float64)))
    return out


def fft_r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.size // 2 + 1, dtype=np.complex128)
    library().orc_fft_r2c(x.size, _dp(x), _dp(out.view(np.float64)))
    return out


def fft_c2r(x, n):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(n)
    library().orc_fft_c2r(n, _dp(x.view(np.float64)), _dp(out))
    return out


def gtsv(l, c, u, q):
    l, c, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (l, c, u))
    if np.iscomplexobj(q):
        q = np.ascontiguousarray(q, dtype=np.complex128).copy()
        library().orc_gtsv_complex(len(c), _dp(l), _dp(c), _dp(u), _dp(q.view(np.float64)))
    else:
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        library().orc_gtsv(len(c), _dp(l), _dp(c), _dp(u), _dp(q))
    return q


def tdm_periodic(l, c, u, q):
    l, c, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (l, c, u))
    q = np.ascontiguousarray(q, dtype=np.float64).copy()
    library().orc_tdm_periodic(len(c), _dp(l), _dp(c), _dp(u), _dp(q))
    return q


def rkcoefs():
    out = np.zeros(9)
    library().orc_rkcoefs(_dp(out))
    return out.reshape(3, 3)
