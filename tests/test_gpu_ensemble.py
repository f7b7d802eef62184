"""The ensemble (tc_ensemble_t): flows of uneven sizes advanced together by recorded graphs.

Every member must give, step for step, exactly the fields the same flow gives when stepped
on its own by tc_step (the recorded kernels are the same launches).  The graphs are replayed
several times with the program's own eager work (reading fields back, diagnostics) between
the replays, and every replay is checked.  Launch shapes chosen by the autotuner change
nothing in the results.
"""
import numpy as np
import pytest

import icgen

pytestmark = pytest.mark.gpu

FIELDS = ("ux", "uy", "uz", "p", "t")


def _specs():
    out = []
    for m, (nx, ny, nz, eta, Re, curved, scalar, implicit) in enumerate([
            (16, 12, 8, 0.714, 300.0, True, 0, (0, 0, 0)),
            (21, 24, 12, 0.55, 800.0, True, 1, (0, 0, 0)),
            (12, 16, 16, 0.9, 500.0, False, 0, (0, 0, 1)),
            (18, 12, 24, 0.6, 200.0, True, 1, (1, 1, 0))]):
        lx = (1.0 / eta - 1.0) if curved else 1.0
        xf, xc = icgen.grid(nx, lx, curved, 1.5)
        f = icgen.fields(nx, ny, nz, xf, xc, curved, seed=10 + m, amp=0.05, t_amp=0.02)
        out.append((dict(nx=nx, ny=ny, nz=nz, xf=xf, xc=xc, is_curved=curved,
                         lengths=(lx, np.pi / 3, 1.0), Re=Re, implicit=implicit, scalar=scalar,
                         coef_dt_adv=0.5, coef_dt_dif=0.5), f))
    return out


def _solo(spec, fields):
    from simpletcsolver_amd import Solver
    sp = dict(spec)
    s = Solver(sp.pop("nx"), sp.pop("ny"), sp.pop("nz"), sp.pop("xf"), sp.pop("xc"), **sp)
    s.load_fields(fields)
    return s


@pytest.mark.parametrize("tune", [False, True])
def test_ensemble_replays_match_solo_steps(need_gpu, tune):
    from simpletcsolver_amd import Ensemble
    specs = _specs()
    ens = Ensemble([sp for sp, _ in specs])
    for m, (_, f) in enumerate(specs):
        ens.members[m].load_fields(f)
    if tune:
        kb, ty = ens.autotune(reps=2)
        print("autotuned (kb, ty):", list(zip(kb, ty)))
    solos = [_solo(sp, f) for sp, f in specs]
    for step in range(5):
        ens.step(1)
        ens.synchronize()
        for s in solos:
            s.step()
        for m, s in enumerate(solos):
            got = ens.members[m].fields()
            want = s.fields()
            for n in FIELDS:
                assert np.array_equal(got[n], want[n]), (step, m, n)
        # eager work between replays: diagnostics on the members
        for m in range(len(solos)):
            assert ens.members[m].divergence() <= 1e-12
    for s in solos:
        s.close()
    ens.close()


def test_ensemble_many_steps_in_one_call(need_gpu):
    from simpletcsolver_amd import Ensemble
    specs = _specs()[:3]
    ens = Ensemble([sp for sp, _ in specs])
    for m, (_, f) in enumerate(specs):
        ens.members[m].load_fields(f)
    solos = [_solo(sp, f) for sp, f in specs]
    ens.step(7)
    ens.synchronize()
    for s in solos:
        for _ in range(7):
            s.step()
    for m, s in enumerate(solos):
        got, want = ens.members[m].fields(), s.fields()
        for n in FIELDS:
            assert np.array_equal(got[n], want[n]), (m, n)
        assert abs(ens.members[m].dt() - s.dt()) == 0.0
    for s in solos:
        s.close()
    ens.close()


def test_ensemble_of_mixed_precision_members(need_gpu):
    """members of either precision side by side: each equals its solo run bit for bit"""
    from simpletcsolver_amd import Ensemble
    specs = _specs()[:3]
    specs[0][0]["precision"] = "mixed"
    specs[2][0]["precision"] = "mixed_refined"
    ens = Ensemble([sp for sp, _ in specs])
    for m, (_, f) in enumerate(specs):
        ens.members[m].load_fields(f)
    solos = [_solo(sp, f) for sp, f in specs]
    ens.step(4)
    ens.synchronize()
    for s in solos:
        for _ in range(4):
            s.step()
    for m, s in enumerate(solos):
        got, want = ens.members[m].fields(), s.fields()
        for n in ("ux", "uy", "uz", "p", "t"):
            assert np.array_equal(got[n], want[n]), (m, n)
    assert ens.members[2].divergence() <= 1e-12
    for s in solos:
        s.close()
    ens.close()


def test_ensemble_takes_up_member_changes(need_gpu):
    """changes the recorded graphs cannot see make the next step record again: a member's
    fields replaced, a member stepped on its own, a closure network coupled after the first
    step, then uncoupled and destroyed (its buffers freed).  Every step still equals the same
    flows stepped eagerly, bit for bit."""
    from simpletcsolver_amd import Cnn, Ensemble, attach_closure
    specs = _specs()[:2]
    ens = Ensemble([sp for sp, _ in specs])
    for m, (_, f) in enumerate(specs):
        ens.members[m].load_fields(f)
    solos = [_solo(sp, f) for sp, f in specs]

    def step_and_compare(n, what):
        ens.step(n)
        ens.synchronize()
        for s in solos:
            for _ in range(n):
                s.step()
        for m, s in enumerate(solos):
            got, want = ens.members[m].fields(), s.fields()
            for name in FIELDS:
                assert np.array_equal(got[name], want[name]), (what, m, name)

    step_and_compare(2, "recorded")
    # fields replaced from the host
    f0 = ens.members[0].fields()
    f0["ux"] *= 0.5
    ens.members[0].load_fields({"ux": f0["ux"]})
    solos[0].load_fields({"ux": f0["ux"]})
    step_and_compare(1, "fields replaced")
    # a member stepped on its own between ensemble steps (its buffer parity flips)
    ens.members[1].step()
    solos[1].step()
    step_and_compare(2, "solo step")
    # a closure coupled after the graphs were recorded, to member 0 and to its solo twin
    sp = specs[0][0]
    nets = [Cnn(sp["nx"], sp["ny"], sp["nz"], batch=1, seed=77) for _ in range(2)]
    mean, std, scale = np.zeros(5), np.ones(5), 1e-2 * np.ones(4)
    attach_closure(ens.members[0], nets[0], mean, std, scale)
    attach_closure(solos[0], nets[1], mean, std, scale)
    step_and_compare(2, "closure coupled")
    # uncoupled and destroyed: the replays must not touch the freed network
    attach_closure(ens.members[0], None, mean, std, scale)
    attach_closure(solos[0], None, mean, std, scale)
    for net in nets:
        net.close()
    step_and_compare(2, "closure destroyed")
    # destroyed while coupled: destroying the network uncouples it
    nets = [Cnn(sp["nx"], sp["ny"], sp["nz"], batch=1, seed=78) for _ in range(2)]
    attach_closure(ens.members[0], nets[0], mean, std, scale)
    attach_closure(solos[0], nets[1], mean, std, scale)
    step_and_compare(1, "closure coupled again")
    for net in nets:
        net.close()
    step_and_compare(2, "closure destroyed while coupled")
    for s in solos:
        s.close()
    ens.close()
