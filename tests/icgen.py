"""numpy statement of the seeded initial-condition definition used by the tests.

The library's generator (tc_ic_grid on the host, tc_ic_fields on the device, include/tcstep.h)
must reproduce these arrays bit for bit; tests/test_host_boundary.py and tests/test_gpu_ic.py
check that.  The definition:

* faces   xf[i] = ri + lx/2 (1 + tanh(s (2 i/nx - 1)) / tanh(s)), i = 0..nx (s = 0: uniform),
  with xf[0] = ri and xf[nx] = ri + lx set exactly; ri = 1 (curved) or 0 (planar);
* centres xc[0] = xf[0], xc[i] = xf[i-1]/2 + xf[i]/2 (i = 1..nx), xc[nx+1] = xf[nx]
  (the convention of the reference's initial_condition/main.py:52-83);
* laminar uy: circular Couette A r + B / r between the wall values (curved) or the linear
  profile (planar), evaluated at xc; ux = uz = p = 0;
* T: conduction profile between the wall values, log(r/ro)/log(ri/ro) (curved) or linear;
* noise: amp (2 u - 1) with u = (splitmix64(key) >> 11) 2^-53 and
  key = seed * 0x9E3779B97F4A7C15 + (field << 56) + linear NPY index (field: ux 0, uy 1, uz 2, t 3);
* modes (ux, uy, uz): mode_amp X(x) Y(y) Z(z), each factor the mean of nmodes seeded modes
  (mode_table below).
"""
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M64 = (1 << 64) - 1


def splitmix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, field, n):
    lin = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64((int(seed) * 0x9E3779B97F4A7C15) & M64) + np.uint64((int(field) << 56) & M64) + lin
    h = splitmix64(key)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def grid(nx, lx, is_curved, stretch=2.0):
    ri = 1.0 if is_curved else 0.0
    i = np.arange(nx + 1, dtype=np.float64)
    if stretch > 0.0:
        # math.tanh is the C library's tanh, as the C generator calls; numpy's vectorised tanh
        # may differ from it in the last bit
        th = np.array([math.tanh(stretch * (2.0 * v / nx - 1.0)) for v in i])
        xf = ri + lx * 0.5 * (1.0 + th / math.tanh(stretch))
    else:
        xf = ri + lx * i / nx
    xf[0] = ri
    xf[nx] = ri + lx
    xc = np.empty(nx + 2)
    xc[0] = xf[0]
    xc[1:nx + 1] = 0.5 * xf[:-1] + 0.5 * xf[1:]
    xc[nx + 1] = xf[nx]
    return xf, xc


def laminar_uy(r, ri, ro, ui, uo, is_curved):
    if is_curved:
        a = (uo * ro - ui * ri) / (ro * ro - ri * ri)
        b = ri * ro * (ui * ro - uo * ri) / (ro * ro - ri * ri)
        return a * r + b / r
    return uo + (ui - uo) * (ro - r) / (ro - ri)


def conduction_t(r, ri, ro, ti, to, is_curved):
    if is_curved:
        lr = np.array([math.log(v) for v in np.atleast_1d(r / ro)])   # the C library's log
        return to + (ti - to) * lr / math.log(ri / ro)
    return to + (ti - to) * (ro - r) / (ro - ri)


def _uniform_tag(seed, field, tag):
    with np.errstate(over="ignore"):
        key = np.uint64((int(seed) * 0x9E3779B97F4A7C15) & M64) + np.uint64((int(field) << 56) & M64) + np.uint64(tag)
    return float((splitmix64(key) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53))


def mode_table(seed, field, direction, nm, kind, pos):
    """mean of nm seeded modes: kind 0 sin(pi m xi) (vanishing at the walls), kind 1
    cos(2 pi m t + phase); amplitude and phase from the hash with field + 16 / field + 32 and
    tag (direction << 40) + (m << 8)"""
    out = np.zeros(len(pos))
    for m in range(1, nm + 1):
        tag = (direction << 40) + (m << 8)
        a = 2.0 * _uniform_tag(seed, 16 + field, tag) - 1.0
        phase = 2.0 * math.pi * _uniform_tag(seed, 32 + field, tag)
        for q, x in enumerate(pos):
            v = math.sin(math.pi * m * x) if kind == 0 else math.cos(2.0 * math.pi * m * x + phase)
            out[q] += a * v / nm
    return out


def fields(nx, ny, nz, xf, xc, is_curved, seed=1, amp=1e-2, uy_walls=(1.0, 0.0),
           t_walls=(0.5, -0.5), t_amp=0.0, nmodes=0, mode_amp=0.0):
    """NPY-layout fields: ux (nz, ny, nx+1), uy/uz/p/t (nz, ny, nx+2)."""
    ri, ro = xf[0], xf[-1]
    ux = np.zeros((nz, ny, nx + 1))
    uy = np.broadcast_to(laminar_uy(xc, ri, ro, uy_walls[0], uy_walls[1], is_curved), (nz, ny, nx + 2)).copy()
    uz = np.zeros((nz, ny, nx + 2))
    t = np.broadcast_to(conduction_t(xc, ri, ro, t_walls[0], t_walls[1], is_curved), (nz, ny, nx + 2)).copy()
    p = np.zeros((nz, ny, nx + 2))
    if amp != 0.0:
        ux += (amp * (2.0 * uniform(seed, 0, ux.size) - 1.0)).reshape(ux.shape)
        uy += (amp * (2.0 * uniform(seed, 1, uy.size) - 1.0)).reshape(uy.shape)
        uz += (amp * (2.0 * uniform(seed, 2, uz.size) - 1.0)).reshape(uz.shape)
    if t_amp != 0.0:
        t += (t_amp * (2.0 * uniform(seed, 3, t.size) - 1.0)).reshape(t.shape)
    if nmodes > 0 and mode_amp != 0.0:
        lx = xf[-1] - ri
        for f, a in enumerate((ux, uy, uz)):
            x = xf if f == 0 else xc
            mx = mode_table(seed, f, 0, nmodes, 0, [(v - ri) / lx for v in x])
            my = mode_table(seed, f, 1, nmodes, 1, [q / ny for q in range(ny)])
            mz = mode_table(seed, f, 2, nmodes, 1, [q / nz for q in range(nz)])
            a += ((mode_amp * mx[None, None, :]) * my[None, :, None]) * mz[:, None, None]
    return {"ux": ux, "uy": uy, "uz": uz, "p": p, "t": t}
