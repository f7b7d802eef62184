"""The learned closure at the coarse grid of its benchmark configuration: 64 x 128 x 128
Taylor-Couette cells (radius ratio 0.714, Re = 1e4, Pr = 1, uniform radial grid, implicit
radial diffusion, scalar on, CFL dt with coefficients 0.5), the 3-layer 32-channel network
coupled (seeded weights; what training changes is the weights, not the coupling checked here).

* the coupled run keeps max |div u| <= 1e-12 after every step (the correction enters the
  right-hand sides before the projection) and finite fields over 60 steps;
* the same 60 steps replayed from recorded graphs (a one-member ensemble) equal the eager
  steps bit for bit;
* the network changes the flow, and detaching it restores the plain method exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 128, 128
STEPS = 60
NAMES = ("ux", "uy", "uz", "t", "p")


def _spec():
    from simpletcsolver_amd import closure
    eta = 0.714
    lx = 1.0 / eta - 1.0
    xf, xc = closure.uniform_grid(NX, lx, True)
    return dict(nx=NX, ny=NY, nz=NZ, xf=xf, xc=xc, is_curved=True,
                lengths=(lx, np.pi / 3.0, 2.0 * np.pi * lx), Re=1.0e4, Pr=1.0, implicit=(1, 0, 0),
                scalar=1, t_walls=(1.0, 0.0), coef_dt_adv=0.5, coef_dt_dif=0.5)


def _solver(spec):
    from simpletcsolver_amd import Solver
    sp = dict(spec)
    return Solver(sp.pop("nx"), sp.pop("ny"), sp.pop("nz"), sp.pop("xf"), sp.pop("xc"), **sp)


def _net():
    from simpletcsolver_amd import Cnn
    return Cnn(NX, NY, NZ, batch=1, seed=4242)


NORM = (np.array([0.0, 0.3, 0.0, 0.5, 1.2]), np.array([0.05, 0.3, 0.05, 0.3, 0.12]),
        np.full(4, 5e-2))


def test_config_closure_coarse_64x128x128(need_gpu):
    from simpletcsolver_amd import Ensemble, attach_closure
    spec = _spec()
    base = _solver(spec)
    base.ic(seed=9, amp=1e-2, t_amp=1e-2, nmodes=8, mode_amp=0.1)
    f0 = base.fields()
    f0.pop("psi", None)

    # eager, closure coupled
    s = _solver(spec)
    s.load_fields(f0)
    net_s = _net()
    attach_closure(s, net_s, *NORM)
    for n in range(STEPS):
        s.step()
        div = s.divergence()
        assert div <= 1e-12, (n, div)
    eager = s.fields()
    for n in NAMES:
        assert np.isfinite(eager[n]).all(), n
    print(f"coupled: dt {s.dt():.4e}, max div {div:.2e}", flush=True)

    # recorded graphs, same network weights
    ens = Ensemble([spec])
    m = ens.members[0]
    m.load_fields(f0)
    net_m = _net()
    attach_closure(m, net_m, *NORM)
    ens.step(STEPS)
    ens.synchronize()
    replay = m.fields()
    for n in NAMES:
        assert np.array_equal(replay[n], eager[n]), n
    ens.close()
    net_m.close()

    # plain method, and the coupled solver with its network detached, from the same start
    for _ in range(STEPS):
        base.step()
    plain = base.fields()
    delta = max(float(np.abs(plain[n] - eager[n]).max()) for n in ("ux", "uy", "uz"))
    print(f"largest velocity change the closure made over {STEPS} steps: {delta:.3e}", flush=True)
    assert delta > 1e-8
    attach_closure(s, None, *NORM)
    s.load_fields(f0)
    for _ in range(STEPS):
        s.step()
    again = s.fields()
    for n in NAMES:
        assert np.array_equal(again[n], plain[n]), n
    s.close()
    net_s.close()
    base.close()
