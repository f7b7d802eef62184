"""Closed-form pins (tests/pins.py) on the CPU oracle, and the planted mistakes they catch.

The pins cover what the energy identities of test_oracle_pins.py cannot see: the RK3
combination and the a/b rotation (predict/main.c:8-25, ux.c:476-494), the implicit diffusion
prefactors f / 2f of every field and direction including the scalar's y quirk (ux.c:495-542,
uy.c:522-562, uz.c:414-454, t.c:329-369), and the pressure update (update_pressure.c:23-173).
Each planted mistake (oracle/orc.h ORC_MUT_*) is a one-token slip of the kind a restatement
could make; the test shows the pin meant for it moves from round-off to O(1e-3..1).
"""
import numpy as np
import pytest

import pins
from oracle import orc

TOL = 1e-12

FIELDS = ("ux", "uy", "uz", "t")
IMPLICIT_SETS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1)]


@pytest.fixture(scope="module")
def planar():
    return pins.Setup(pins.OracleFlow, curved=False)


@pytest.fixture(scope="module")
def curved():
    return pins.Setup(pins.OracleFlow, curved=True)


@pytest.fixture
def mutation():
    yield orc.set_mutation
    orc.set_mutation("none")


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("field", FIELDS)
def test_explicit_diffusion_eigenvalue(request, geometry, field, d):
    st = request.getfixturevalue(geometry)
    assert pins.explicit_eigenvalue(st, field, d) <= TOL


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("rk", [0, 1, 2])
def test_implicit_mode_is_crank_nicolson(request, geometry, field, d, rk):
    st = request.getfixturevalue(geometry)
    assert pins.implicit_mode(st, field, d, rk=rk) <= TOL


@pytest.mark.parametrize("field", FIELDS)
def test_implicit_x_is_crank_nicolson(planar, field):
    err, mx = pins.implicit_x(planar, field)
    assert err <= TOL
    # the probed operator is diffusion with the walls at rest: every eigenvalue is negative
    assert np.all(np.linalg.eigvals(mx).real < 0)


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("implicit", IMPLICIT_SETS)
def test_rk3_step_closed_form(request, geometry, implicit):
    st = request.getfixturevalue(geometry)
    assert pins.rk3_step(st, implicit) <= TOL


@pytest.mark.parametrize("geometry", ["planar", "curved"])
@pytest.mark.parametrize("implicit", IMPLICIT_SETS)
def test_pressure_update_closed_form(request, geometry, implicit):
    st = request.getfixturevalue(geometry)
    assert pins.pressure_update(st, implicit) <= TOL


def test_inviscid_energy_decay_is_third_order():
    """the reference's energy example (docs/source/example/energy): 1 - E/E0 over a fixed
    time falls as dt^3 for the kinetic and the scalar energy alike"""
    losses = pins.energy_decay(pins.OracleFlow)
    orders = pins.decay_orders(losses)
    print("losses", losses.tolist(), "orders", orders.tolist())
    assert np.all(losses > 0)
    assert np.all((orders > 2.7) & (orders < 3.3))
    assert np.all(np.abs(orders[-1] - 3.0) < 0.05)


# ---------------------------------------------------------------------------------- mistakes
def _implicit_y(st):
    return max(pins.implicit_mode(st, f, 1) for f in FIELDS)


def _implicit_z(st):
    return max(pins.implicit_mode(st, f, 2) for f in FIELDS)


def _implicit_x(st):
    return max(pins.implicit_x(st, f)[0] for f in FIELDS)


def _rk3(st):
    return max(pins.rk3_step(st, imp) for imp in [(0, 0, 0), (1, 1, 1)])


def _pressure(st):
    return max(pins.pressure_update(st, imp) for imp in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)])


CAUGHT_BY = {
    "ux_x_f": _implicit_x,
    "uz_x_2f": _implicit_x,
    "uy_y_f": _implicit_y,
    "t_y_f": _implicit_y,
    "ux_y_2f": _implicit_y,
    "uz_z_f": _implicit_z,
    "pf_alpha": _implicit_y,
    "ab_swap": _rk3,
    "no_rotate": _rk3,
    "p_f": _pressure,
    "p_no_z": _pressure,
    "p_x_index": _pressure,
}


def test_every_planted_mistake_has_a_pin():
    assert set(CAUGHT_BY) == set(orc.MUTATIONS) - {"none"}


@pytest.mark.parametrize("name", sorted(CAUGHT_BY))
def test_planted_mistake_is_caught(planar, curved, mutation, name):
    st = planar if CAUGHT_BY[name] in (_implicit_x, _implicit_y, _implicit_z) else curved
    assert CAUGHT_BY[name](st) <= TOL
    mutation(name)
    err = CAUGHT_BY[name](st)
    print(name, err)
    assert err > 1e-3


@pytest.mark.parametrize("name", ["ab_swap", "no_rotate"])
def test_planted_rk_mistake_breaks_the_energy_order(mutation, name):
    mutation(name)
    losses = pins.energy_decay(pins.OracleFlow)
    orders = pins.decay_orders(losses)
    assert not (np.all(losses > 0) and np.all((orders > 2.7) & (orders < 3.3)))
