"""The learned closure end to end on a small Taylor-Couette flow: data from the solver's own
fine run (box filter of width 4), training with the library's kernels, coupling to the coarse
solver.  The gain is measured a posteriori on snapshots the network never saw: one coarse
step with the closure against one without, both against the filtered fine solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_closure_reduces_the_coarse_step_error(need_gpu):
    from simpletcsolver_amd import closure
    eta, re = 0.714, 2000.0
    lx = 1.0 / eta - 1.0
    pair = closure.Pair(48, 64, 64, 4, lengths=(lx, np.pi / 3.0, 2.0 * lx), Re=re)
    fine = pair.fine
    fine.set_dt_params(0.5, 0.5)
    fine.ic(seed=5, amp=0.05, t_amp=0.02, nmodes=4, mode_amp=0.1)
    for _ in range(200):
        fine.step()
    dt = 4 * 0.8 * fine.dt()
    states = []
    x, y = closure.generate(pair, 28, 2, dt, states=states)
    ntr = 20
    base = closure.step_error(pair, states[ntr:], dt)
    net, norm, losses = closure.train(x[:ntr], y[:ntr], batch=4, iters=400, lr=2e-3, seed=3, log=print)
    held = closure.evaluate_loss(net, norm, x[ntr:], y[ntr:])
    one = closure.attach(pair.coarse, net, norm)
    with_closure = closure.step_error(pair, states[ntr:], dt)
    closure.detach(pair.coarse)
    again = closure.step_error(pair, states[ntr:], dt)
    print(f"training loss {losses[0]:.3e} -> {losses[-1]:.3e}; held-out loss {held:.3e}")
    print(f"held-out one-step error: no closure {base:.4e}, closure {with_closure:.4e} "
          f"({with_closure / base:.3f})")
    assert again == base                      # detaching restores the plain method exactly
    assert losses[-1] < 0.5 * losses[0]
    assert with_closure < 0.9 * base
    _graph_replay_matches_eager(pair, states[-1][0], net, norm, dt, lx, re)
    one.close()
    net.close()
    pair.close()


def _graph_replay_matches_eager(pair, f0, net, norm, dt, lx, re):
    """the closure's kernels recorded into the step graphs give the eager steps bit for bit"""
    from simpletcsolver_amd import Ensemble, closure
    spec = dict(nx=pair.nxc, ny=pair.nyc, nz=pair.nzc, xf=pair.xfc, xc=pair.xcc, is_curved=True,
                lengths=(lx, np.pi / 3.0, 2.0 * lx), Re=re, implicit=(1, 0, 0), scalar=1, dt_fixed=dt)
    ens = Ensemble([spec])
    m = ens.members[0]
    m.load_fields(f0)
    one = closure.attach(m, net, norm)
    pair.coarse.set_dt_params(0.5, 0.5, dt)
    pair.coarse.load_fields(f0)
    ref = closure.attach(pair.coarse, net, norm)
    for step in range(4):
        ens.step(1)
        ens.synchronize()
        pair.coarse.step()
        got, want = m.fields(), pair.coarse.fields()
        for n in ("ux", "uy", "uz", "t", "p"):
            assert np.array_equal(got[n], want[n]), (step, n)
    closure.detach(pair.coarse)
    ens.close()
    one.close()
    ref.close()
