"""The reference API as a library caller uses it, in the library's default host-coherent mode
(TC_COHERENCE unset = strict): domain_init and fluid_init on an initial-condition directory,
then per step decide_dt and the four stages of every substep, called one by one as
src/integrate.c:20-33 calls them, with the caller reading the host arrays after every step (as
src/logging/*.c and statistics.c:232-236 read fluid_t) and editing fluid->ux/uz/t between
steps.  After 10 steps the host arrays hold the oracle's fields, driven through the same
calls and the same edits (relative L2 <= 1e-10; p modulo its J-weighted mean).
"""
import ctypes
import os

import numpy as np
import pytest

from harness import make_case, modulo_mean, oracle_for, rel_l2
from test_gpu_refapi import write_ic_dir

pytestmark = pytest.mark.gpu

NDIMS = 3


class ArrayT(ctypes.Structure):
    _fields_ = [("size", ctypes.c_size_t), ("nadds", ctypes.c_void_p),
                ("datasize", ctypes.c_size_t), ("data", ctypes.c_void_p)]


class DomainT(ctypes.Structure):
    _fields_ = [("info", ctypes.c_void_p), ("is_curved", ctypes.c_bool),
                ("glsizes", ctypes.c_size_t * NDIMS), ("mysizes", ctypes.c_size_t * NDIMS),
                ("offsets", ctypes.c_size_t * NDIMS), ("lengths", ctypes.c_double * NDIMS),
                ("xf", ctypes.c_void_p), ("xc", ctypes.c_void_p),
                ("jdxf", ctypes.c_void_p), ("jdxc", ctypes.c_void_p),
                ("hxxf", ctypes.c_void_p), ("hxxc", ctypes.c_void_p),
                ("hyxf", ctypes.c_void_p), ("hyxc", ctypes.c_void_p),
                ("hz", ctypes.c_double)]


class FluidT(ctypes.Structure):
    _fields_ = ([(n, ArrayT) for n in ("ux", "uy", "uz", "p", "psi", "t")]
                + [("srcux", ArrayT * 3), ("srcuy", ArrayT * 3), ("srcuz", ArrayT * 3),
                   ("srct", ArrayT * 3)]
                + [(n, ArrayT) for n in ("lxx", "lyx0", "lyx1", "lxy", "lzx", "lxz", "lyy0",
                                         "lyy1", "lzy", "lyz", "lzz")]
                + [("Re", ctypes.c_double), ("Pr", ctypes.c_double)])


def host_view(arr, shape):
    """numpy view of a fluid_t host array (reference layout, y/z halos included)"""
    assert arr.datasize == 8 * int(np.prod(shape))
    buf = (ctypes.c_double * int(np.prod(shape))).from_address(arr.data)
    return np.ctypeslib.as_array(buf).reshape(shape)


def edit(a, k):
    """the caller's edit between steps: damp the interior and add a seeded perturbation"""
    rng = np.random.default_rng(k)
    a[1:-1, 1:-1, 1:-1] = 0.97 * a[1:-1, 1:-1, 1:-1] + 1e-3 * rng.standard_normal(a[1:-1, 1:-1, 1:-1].shape)


@pytest.mark.parametrize("implicit", [(1, 0, 0), (1, 1, 1)])
def test_strict_reference_api_with_host_edits(need_gpu, tmp_path, monkeypatch, implicit):
    from simpletcsolver_amd import load_library
    case = make_case(nx=16, ny=12, nz=8, Re=300.0, implicit=implicit, amp=0.1, t_amp=0.05,
                     t_walls=(1.0, 0.0))
    kw = case["kw"]
    icdir = str(tmp_path / "ic")
    write_ic_dir(icdir, case)
    monkeypatch.delenv("TC_COHERENCE", raising=False)
    monkeypatch.delenv("dt_fixed", raising=False)
    monkeypatch.delenv("TC_PRECISION", raising=False)
    env = {"Re": kw["Re"], "Pr": kw["Pr"], "coef_dt_adv": 0.7, "coef_dt_dif": 0.9,
           "implicit_x": implicit[0], "implicit_y": implicit[1], "implicit_z": implicit[2],
           "scalar": kw["scalar"], "uy_xm": kw["uy_walls"][0], "uy_xp": kw["uy_walls"][1],
           "t_xm": kw["t_walls"][0], "t_xp": kw["t_walls"][1]}
    for k, v in env.items():
        monkeypatch.setenv(k, repr(v) if isinstance(v, float) else str(v))

    lib = load_library()
    dp = ctypes.POINTER(ctypes.c_double)
    lib.domain_init.argtypes = [ctypes.c_char_p, ctypes.POINTER(DomainT)]
    lib.fluid_init.argtypes = [ctypes.c_char_p, ctypes.POINTER(DomainT), ctypes.POINTER(FluidT)]
    lib.decide_dt.argtypes = [ctypes.POINTER(DomainT), ctypes.POINTER(FluidT), dp]
    for name in ("fluid_predict_field", "fluid_compute_potential", "fluid_correct_velocity",
                 "fluid_update_pressure"):
        getattr(lib, name).argtypes = [ctypes.POINTER(DomainT), ctypes.c_size_t, ctypes.c_double,
                                       ctypes.POINTER(FluidT)]
    lib.tc_fluid_finalise.argtypes = [ctypes.POINTER(FluidT)]

    domain, fluid = DomainT(), FluidT()
    assert lib.domain_init(icdir.encode(), ctypes.byref(domain)) == 0
    assert lib.fluid_init(icdir.encode(), ctypes.byref(domain), ctypes.byref(fluid)) == 0
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    shapes = {n: (nz + 2, ny + 2, nx + 1) if n == "ux" else (nz + 2, ny + 2, nx + 2)
              for n in ("ux", "uy", "uz", "p", "psi", "t")}
    host = {n: host_view(getattr(fluid, n), shapes[n]) for n in shapes}

    o = oracle_for(dict(case, kw=dict(kw, coef_dt_adv=0.7, coef_dt_dif=0.9)))
    stages = (("fluid_predict_field", o.predict), ("fluid_compute_potential", o.potential),
              ("fluid_correct_velocity", o.correct), ("fluid_update_pressure", o.pressure))
    try:
        for step in range(10):
            dt = ctypes.c_double(0.0)
            assert lib.decide_dt(ctypes.byref(domain), ctypes.byref(fluid), ctypes.byref(dt)) == 0
            dto = o.decide_dt()
            assert abs(dt.value - dto) <= 1e-15 * dto, (step, dt.value, dto)
            for rk in range(3):
                for name, ostage in stages:
                    assert getattr(lib, name)(ctypes.byref(domain), rk, dto, ctypes.byref(fluid)) == 0
                    ostage(rk, dto)
            # the host arrays are current after every call: the caller reads them
            for n in ("ux", "uy", "uz", "t"):
                e = rel_l2(host[n][1:-1, 1:-1, :], o.array(n)[1:-1, 1:-1, :])
                assert e <= 1e-10, (step, n, e)
            if step in (2, 6):
                for n in ("ux", "uz", "t"):
                    edit(host[n], 100 * step + len(n))
                    edit(o.array(n), 100 * step + len(n))
                o.update_boundaries()
        of = o.npy_fields()
        for n in ("ux", "uy", "uz", "t", "p"):
            got = host[n][1:-1, 1:-1, :].copy()
            want = of[n]
            if n == "p":
                got, want = modulo_mean(got, case), modulo_mean(want, case)
            e = rel_l2(got, want)
            print(f"{n}: {e:.3e}")
            assert e <= 1e-10, n
    finally:
        lib.tc_fluid_finalise(ctypes.byref(fluid))
        o.close()
