"""Initial-condition directories in the reference's format, with the reference generator's
grid (initial_condition/main.py), so that its documented examples can be rerun.

The reference writes, into one directory, NPY files step (uint64), time, is_curved (bool),
glsizes (uint64), lengths, xf, xc and the fields ux (nz, ny, nx+1), uy, uz, p, t
(nz, ny, nx+2) (main.py:44-104); domain_init and fluid_init (include/domain.h,
include/fluid_solver.h) read exactly that.

Grid (main.py:55-83): nx+1 faces from a Chebyshev distribution whose nclip = 3 outermost
points on each side are cut off, x_i = -cos(pi (i + nclip) / (nx + 2 nclip)), i = 0..nx,
rescaled to [0, lx]; the cell centres are the face midpoints plus the two walls; the inner
radius 1 is added for a curved domain.  Fields (main.py:86-104): independent uniform values
on [0, 1), shifted by -1/2 except uy; the reference draws them unseeded, this writer from a
seeded generator.
"""
import os

import numpy as np


def clipped_chebyshev_grid(nx, lx, is_curved, nclip=3):
    """faces xf[0..nx] and centres xc[0..nx+1] of the reference generator's stretched grid"""
    i = np.arange(nx + 1, dtype=np.float64)
    xf = -np.cos(np.pi * (i + nclip) / (nx + 2.0 * nclip))
    xf = lx * (xf - xf.min()) / (xf.max() - xf.min())
    xc = np.concatenate([[0.0], 0.5 * xf[:-1] + 0.5 * xf[1:], [lx]])
    x0 = 1.0 if is_curved else 0.0
    return xf + x0, xc + x0


def random_fields(nx, ny, nz, seed):
    """the generator's initial fields: U[0, 1) - 1/2 for ux, uz, p, t and U[0, 1) for uy"""
    rng = np.random.default_rng(seed)
    out = {}
    for name in ("ux", "uy", "uz", "p", "t"):
        shape = (nz, ny, nx + 1) if name == "ux" else (nz, ny, nx + 2)
        a = rng.random(shape, dtype=np.float64)
        out[name] = a if name == "uy" else a - 0.5
    return out


def write(dest, is_curved, lengths, glsizes, seed=0, fields=None, xf=None, xc=None):
    """an initial-condition directory; returns (xf, xc, fields)"""
    nx, ny, nz = glsizes
    os.makedirs(dest, exist_ok=True)
    if xf is None:
        xf, xc = clipped_chebyshev_grid(nx, lengths[0], is_curved)
    if fields is None:
        fields = random_fields(nx, ny, nz, seed)
    np.save(os.path.join(dest, "step.npy"), np.array(0, dtype=np.uint64))
    np.save(os.path.join(dest, "time.npy"), np.array(0.0, dtype=np.float64))
    np.save(os.path.join(dest, "is_curved.npy"), np.array(bool(is_curved), dtype=np.bool_))
    np.save(os.path.join(dest, "xf.npy"), np.asarray(xf, dtype=np.float64))
    np.save(os.path.join(dest, "xc.npy"), np.asarray(xc, dtype=np.float64))
    np.save(os.path.join(dest, "glsizes.npy"), np.array(glsizes, dtype=np.uint64))
    np.save(os.path.join(dest, "lengths.npy"), np.array(lengths, dtype=np.float64))
    for name, a in fields.items():
        np.save(os.path.join(dest, f"{name}.npy"), np.asarray(a, dtype=np.float64))
    return xf, xc, fields


def laminar_torque_factor(Re, lengths, ri=1.0, ui=1.0):
    """the energy injection of circular Couette flow between radii ri and ri + lx over the
    domain lengths ly (radians) and lz: the normalisation that turns logged injection and
    dissipation into a Nusselt number (docs/source/example/typical/data/balance_main.py)"""
    ro = ri + lengths[0]
    return 2.0 / Re * ri ** 2 * ro ** 2 / (ro ** 2 - ri ** 2) * (ui / ri) ** 2 * lengths[1] * lengths[2]
