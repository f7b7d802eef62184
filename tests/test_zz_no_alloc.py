"""tc_ic_fields_noalloc, the allocation-free form of the device initial-condition fill: the
same fields as tc_ic_fields, bit for bit, over two calls in a row on one device, the second on
a smaller grid (the staging array keeps the first call's longer tables), and the same as the
numpy statement of the generator (tests/icgen.py)."""
import numpy as np
import pytest

import icgen

pytestmark = pytest.mark.gpu

KW = dict(seed=23, amp=0.03, t_amp=0.02, nmodes=5, mode_amp=0.15)


def _solver(nx, ny, nz, curved):
    from simpletcsolver_amd import Solver
    lx = 0.4005602240896359 if curved else 1.0
    xf, xc = icgen.grid(nx, lx, curved, 2.0)
    s = Solver(nx, ny, nz, xf, xc, is_curved=curved, lengths=(lx, 1.0, 0.7), Re=100.0,
               t_walls=(1.0, 0.0))
    return s, xf, xc, lx


def _noalloc(s):
    from simpletcsolver_amd import _check
    _check(s._lib.tc_ic_fields_noalloc(s._h, KW["seed"], KW["amp"], KW["t_amp"], KW["nmodes"],
                                       KW["mode_amp"]), s._lib, "tc_ic_fields_noalloc")


def test_noalloc_fill_equals_allocating_fill(need_gpu):
    names = ("ux", "uy", "uz", "p", "t")
    for nx, ny, nz, curved in ((60, 96, 72, True), (14, 12, 8, False)):
        print(f"noalloc IC fill: {nx}x{ny}x{nz} curved={curved}", flush=True)
        s, xf, xc, _ = _solver(nx, ny, nz, curved)
        _noalloc(s)
        got = {n: s.get_field(n) for n in names}
        s.ic(**KW)
        want = {n: s.get_field(n) for n in names}
        ref = icgen.fields(nx, ny, nz, xf, xc, curved, seed=KW["seed"], amp=KW["amp"],
                           t_walls=(1.0, 0.0), t_amp=KW["t_amp"], nmodes=KW["nmodes"],
                           mode_amp=KW["mode_amp"])
        for n in names:
            assert np.array_equal(got[n], want[n]), (nx, n)
            # interior (walls carry the boundary conditions) against the numpy statement
            assert np.array_equal(got[n][..., 1:-1], ref[n][..., 1:-1]), (nx, n)
        # and again after the allocating fill: the staging array is reused as it was left
        _noalloc(s)
        for n in names:
            assert np.array_equal(s.get_field(n), want[n]), (nx, n)
        s.close()
