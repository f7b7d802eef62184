"""Shared set-up for parity tests: one seeded case run through the oracle and the GPU."""
import numpy as np

import icgen
from oracle.orc import Oracle


def make_case(nx=16, ny=12, nz=8, is_curved=True, eta=0.714, ly=np.pi / 3, lz=None, Re=100.0,
              Pr=1.0, implicit=(1, 0, 0), scalar=1, stretch=2.0, seed=1, amp=1e-2, t_amp=0.0,
              dt_fixed=0.0, uy_walls=(1.0, 0.0), t_walls=(0.5, -0.5)):
    ri = 1.0
    lx = (1.0 / eta - 1.0) if is_curved else 1.0
    if lz is None:
        lz = 2.0 * np.pi * lx / 4.0
    xf, xc = icgen.grid(nx, lx, is_curved, stretch)
    kw = dict(is_curved=is_curved, lengths=(lx, ly, lz), Re=Re, Pr=Pr, implicit=implicit,
              scalar=scalar, uy_walls=uy_walls, t_walls=t_walls, dt_fixed=dt_fixed)
    fields = icgen.fields(nx, ny, nz, xf, xc, is_curved, seed=seed, amp=amp, uy_walls=uy_walls,
                          t_walls=t_walls, t_amp=t_amp)
    return dict(nx=nx, ny=ny, nz=nz, xf=xf, xc=xc, kw=kw, fields=fields, ri=ri)


def oracle_for(case):
    o = Oracle(case["nx"], case["ny"], case["nz"], case["xf"], case["xc"], **case["kw"])
    o.load_fields(case["fields"])
    return o


def gpu_for(case):
    from simpletcsolver_amd import Solver
    s = Solver(case["nx"], case["ny"], case["nz"], case["xf"], case["xc"], **case["kw"])
    s.load_fields(case["fields"])
    return s


def rel_l2(a, b):
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def jmean(a, case):
    """J-weighted mean of a centre field in the NPY layout (interior x cells).  J = hx hy hz
    depends on y and z only through the constant factors ly / ny and lz / nz, which cancel in
    a normalised mean, so the weights come from an oracle of 4 x 4 cells in y and z."""
    o = Oracle(case["nx"], 4, 4, case["xf"], case["xc"], **case["kw"])
    jd = o.array("jdxc")[1:-1].copy()
    o.close()
    w = np.broadcast_to(jd, a[:, :, 1:-1].shape)
    return float((a[:, :, 1:-1] * w).sum() / w.sum())


def modulo_mean(a, case):
    return a - jmean(a, case)
