"""The closed-form pins of tests/pins.py on the device, through the C ABI (tc_predict,
tc_pressure, tc_step on fresh solvers): the RK3 combination and a/b rotation, the implicit
diffusion prefactors of every field and direction (the scalar's y quirk included), the
pressure update, and the third-order inviscid energy decay of the reference's energy
example.  Same closed forms and bars as the oracle's (tests/test_oracle_closed_forms.py)."""
import numpy as np
import pytest

import pins

pytestmark = pytest.mark.gpu

TOL = 1e-12
FIELDS = ("ux", "uy", "uz", "t")
IMPLICIT_SETS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1)]


@pytest.fixture(scope="module")
def planar(need_gpu):
    return pins.Setup(pins.DeviceFlow, curved=False)


@pytest.fixture(scope="module")
def curved(need_gpu):
    return pins.Setup(pins.DeviceFlow, curved=True)


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("field", FIELDS)
def test_device_explicit_diffusion_eigenvalue(request, geometry, field, d):
    st = request.getfixturevalue(geometry)
    err = pins.explicit_eigenvalue(st, field, d)
    print(f"{geometry} {field} d={d}: {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("rk", [1, 2])
def test_device_implicit_mode_is_crank_nicolson(request, geometry, field, d, rk):
    st = request.getfixturevalue(geometry)
    err = pins.implicit_mode(st, field, d, rk=rk)
    print(f"{geometry} {field} d={d} rk={rk}: {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("field", FIELDS)
def test_device_implicit_x_is_crank_nicolson(planar, field):
    err, _ = pins.implicit_x(planar, field)
    print(f"{field}: {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("implicit", IMPLICIT_SETS)
def test_device_rk3_step_closed_form(request, geometry, implicit):
    st = request.getfixturevalue(geometry)
    err = pins.rk3_step(st, implicit)
    print(f"{geometry} {implicit}: {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("implicit", IMPLICIT_SETS)
def test_device_pressure_update_closed_form(request, geometry, implicit):
    st = request.getfixturevalue(geometry)
    err = pins.pressure_update(st, implicit)
    print(f"{geometry} {implicit}: {err:.2e}")
    assert err <= TOL


def test_device_inviscid_energy_decay_is_third_order(need_gpu):
    losses = pins.energy_decay(pins.DeviceFlow)
    orders = pins.decay_orders(losses)
    print("losses", losses.tolist(), "orders", orders.tolist())
    assert np.all(losses > 0)
    assert np.all((orders > 2.7) & (orders < 3.3))
    assert np.all(np.abs(orders[-1] - 3.0) < 0.05)
