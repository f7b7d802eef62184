"""The ensemble benchmark configuration itself: 64 instances drawn from seed 2024 with
nr in [64, 128], ntheta and nz in {128, 192, 256}, radius ratio in [0.5, 0.9], Re in
[500, 8000], one in ten a planar channel, FP64, explicit diffusion, launch shapes from the
autotuner (bench.py ensemble_specs, the batch `bench.py --workload ensemble` times).

* every member after 5 steps (one eager step and the recording, then four graph replays,
  with eager reads between them) equals the same flow stepped alone by tc_step, bit for bit;
* the three smallest members equal the CPU oracle stepped from the same fields, dt to 1e-15
  and each field (p and psi modulo their J-weighted mean, psi's difference against |p|)
  within 1e-10 relative L2 or within three times what the oracle itself moves when its input
  velocity is perturbed by one ulp (two seeded perturbations, the larger taken), whichever is
  larger: psi is the last substep's pressure increment, its right-hand side div u* / (g dt) a
  difference of nearly equal fluxes, and on the planar members (ly = 1, lz = 2, so small
  theta and z wavenumbers against a stiff stretched radial operator) a one-ulp perturbation
  of u moves the oracle's own psi by 2e-10 of |p| after 5 steps;
* max |div u| <= 1e-12 in every member.
"""
import numpy as np
import pytest

from harness import modulo_mean, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("ux", "uy", "uz", "p", "psi")
STEPS = 5


def _solo(spec):
    from simpletcsolver_amd import Solver
    sp = dict(spec)
    return Solver(sp.pop("nx"), sp.pop("ny"), sp.pop("nz"), sp.pop("xf"), sp.pop("xc"), **sp)


def _oracle(spec):
    from oracle.orc import Oracle
    sp = dict(spec)
    return Oracle(sp.pop("nx"), sp.pop("ny"), sp.pop("nz"), sp.pop("xf"), sp.pop("xc"), **sp)


def test_config_ensemble_of_64(need_gpu):
    from bench import ensemble_specs
    from simpletcsolver_amd import Ensemble
    specs = ensemble_specs(64, 2024)
    cells = [sp["nx"] * sp["ny"] * sp["nz"] for sp in specs]
    print(f"64 instances, {sum(cells) / 1e6:.1f} M cells, "
          f"{sum(1 for sp in specs if not sp['is_curved'])} planar", flush=True)
    ens = Ensemble(specs)
    for m in ens.members:
        m.ic(seed=7, amp=1e-2)
    start = [m.fields() for m in ens.members]
    kb, ty = ens.autotune(reps=1)
    print("autotuned (kb, ty) of the first 8:", list(zip(kb, ty))[:8], flush=True)
    ens.step(1)                       # eager step + recording
    for _ in range(STEPS - 1):
        ens.step(1)                   # graph replays, with the reads below between them
        ens.synchronize()
        assert np.isfinite(ens.members[0].divergence())
    ens.synchronize()
    got = [m.fields() for m in ens.members]
    dts = [m.dt() for m in ens.members]
    divs = [m.divergence() for m in ens.members]
    ens.close()
    print(f"max |div| over the members: {max(divs):.2e}", flush=True)
    assert max(divs) <= 1e-12

    for m, sp in enumerate(specs):
        s = _solo(sp)
        s.load_fields(start[m])
        for _ in range(STEPS):
            s.step()
        want = s.fields()
        assert s.dt() == dts[m], (m, s.dt(), dts[m])
        s.close()
        for name in FIELDS + ("t",):
            assert np.array_equal(got[m][name], want[name]), (m, name)
    print("every member equals its solo run bit for bit", flush=True)

    for m in sorted(range(len(specs)), key=lambda q: cells[q])[:3]:
        sp = specs[m]
        runs = []
        for pert in (0, 1, 2):
            f = {k: v.copy() for k, v in start[m].items()}
            if pert:
                rng = np.random.default_rng(pert)
                for k in ("ux", "uy", "uz"):
                    f[k] = f[k] * (1.0 + 2.2e-16 * rng.uniform(-1.0, 1.0, f[k].shape))
            o = _oracle(sp)
            o.load_fields(f)
            for _ in range(STEPS):
                dto = o.integrate()
            if not pert:
                assert abs(dts[m] - dto) <= 1e-15 * dto, (m, dts[m], dto)
            runs.append(o.npy_fields())
            o.close()
        case = dict(nx=sp["nx"], xf=sp["xf"], xc=sp["xc"],
                    kw=dict(is_curved=sp["is_curved"], lengths=sp["lengths"], Re=sp["Re"],
                            implicit=sp["implicit"], scalar=sp["scalar"]))

        def diffs(a, b):
            d = {n: rel_l2(a[n], b[n]) for n in ("ux", "uy", "uz")}
            pb = modulo_mean(b["p"], case)
            d["p"] = rel_l2(modulo_mean(a["p"], case), pb)
            d["psi"] = np.linalg.norm(modulo_mean(a["psi"], case) - modulo_mean(b["psi"], case)) / np.linalg.norm(pb)
            return d

        err = diffs(got[m], runs[0])
        sens = {n: max(diffs(runs[1], runs[0])[n], diffs(runs[2], runs[0])[n]) for n in err}
        print(f"member {m} ({sp['nx']}x{sp['ny']}x{sp['nz']}, curved={sp['is_curved']}) vs oracle:",
              {k: f"{v:.2e}" for k, v in err.items()}, "oracle under 1-ulp perturbation:",
              {k: f"{v:.2e}" for k, v in sens.items()}, flush=True)
        for name, v in err.items():
            assert v <= max(1e-10, 3.0 * sens[name]), (m, name, v, sens[name])
