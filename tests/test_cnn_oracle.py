"""Pins of the closure network's numpy statement (tests/cnn_ref.py), on the CPU:

* the backward pass (flipped-kernel adjoint convolutions, ReLU masks, zero/periodic padding)
  against central finite differences of the loss, in float64 without BF16 rounding;
* the adjoint identity <conv(x, w), g> = <x, conv(g, flip(w))> for both paddings;
* BF16 rounding: ties to even, exact for representable values.
"""
import numpy as np
import pytest

import cnn_ref as cr


def test_bf16_rounding():
    # BF16 resolution at 1 is 2^-7: 1 + 2^-8 is a tie (to the even 1), 1 + 3 2^-8 a tie to
    # the even 1 + 2^-6
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.140625],
                 dtype=np.float32)
    got = cr.bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625], dtype=np.float32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("nx", [3, 5])
def test_adjoint_identity(nx):
    rng = np.random.default_rng(nx)
    x = rng.standard_normal((2, 3, 4, nx, 16))
    g = rng.standard_normal((2, 3, 4, nx, 32))
    w = rng.standard_normal((27, 32, 16))
    lhs = (cr.conv(x, w) * g).sum()
    rhs = (x * cr.conv(g, cr.flip(w))).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    nb, nz, ny, nx = 2, 3, 4, 5
    x = rng.standard_normal((nb, nz, ny, nx, 5)).astype(np.float64)
    y = rng.standard_normal((nb, nz, ny, nx, 4)).astype(np.float64)
    params = [(w.astype(np.float64), b.astype(np.float64)) for w, b in cr.init_params(1)]
    loss, grads, _ = cr.loss_and_grads(params, x, y, bf=False)
    h = 1e-6
    for l in range(3):
        w, b = params[l]
        gw, gb = grads[l]
        idx = [tuple(rng.integers(0, s) for s in w.shape) for _ in range(6)]
        if l == 0:
            idx = [(t, co, ci % 5) for t, co, ci in idx]
        if l == 2:
            idx = [(t, co % 4, ci) for t, co, ci in idx]
        for ix in idx:
            wp = w.copy(); wp[ix] += h
            wm = w.copy(); wm[ix] -= h
            pp = list(params); pp[l] = (wp, b)
            pm = list(params); pm[l] = (wm, b)
            fd = (cr.loss_and_grads(pp, x, y, bf=False)[0] - cr.loss_and_grads(pm, x, y, bf=False)[0]) / (2 * h)
            assert abs(fd - gw[ix]) <= 1e-6 + 1e-5 * abs(fd), (l, ix, fd, gw[ix])
        for co in range(2):
            bp = b.copy(); bp[co] += h
            bm = b.copy(); bm[co] -= h
            pp = list(params); pp[l] = (w, bp)
            pm = list(params); pm[l] = (w, bm)
            fd = (cr.loss_and_grads(pp, x, y, bf=False)[0] - cr.loss_and_grads(pm, x, y, bf=False)[0]) / (2 * h)
            assert abs(fd - gb[co]) <= 1e-6 + 1e-5 * abs(fd), (l, co, fd, gb[co])
