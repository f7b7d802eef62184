"""The device initial-condition generator (tc_ic_fields) against its numpy statement
(tests/icgen.py): identical arrays, bit for bit, for noise, seeded modes and both geometries."""
import numpy as np
import pytest

import icgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("curved", [True, False])
@pytest.mark.parametrize("nmodes,mode_amp", [(0, 0.0), (4, 0.2)])
def test_device_ic_matches_numpy(need_gpu, curved, nmodes, mode_amp):
    from simpletcsolver_amd import Solver
    nx, ny, nz = 20, 12, 18
    lx = 0.4005602240896359 if curved else 1.0
    xf, xc = icgen.grid(nx, lx, curved, 2.0)
    s = Solver(nx, ny, nz, xf, xc, is_curved=curved, lengths=(lx, 1.0, 0.7), Re=100.0,
               t_walls=(1.0, 0.0))
    s.ic(seed=17, amp=0.05, t_amp=0.02, nmodes=nmodes, mode_amp=mode_amp)
    want = icgen.fields(nx, ny, nz, xf, xc, curved, seed=17, amp=0.05, t_walls=(1.0, 0.0),
                        t_amp=0.02, nmodes=nmodes, mode_amp=mode_amp)
    # the wall slots read back as the boundary conditions set them (fluid/boundary/*.c)
    want["ux"][..., 0] = want["ux"][..., -1] = 0.0
    want["uy"][..., 0], want["uy"][..., -1] = 1.0, 0.0
    want["uz"][..., 0] = want["uz"][..., -1] = 0.0
    want["t"][..., 0], want["t"][..., -1] = 1.0, 0.0
    want["p"][..., 0], want["p"][..., -1] = want["p"][..., 1], want["p"][..., -2]
    for name in ("ux", "uy", "uz", "p", "t"):
        got = s.get_field(name)
        assert np.array_equal(got, want[name]), name
    s.close()
