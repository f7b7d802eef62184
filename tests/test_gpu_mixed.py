"""The mixed-precision step (TC_MIXED): FP32 right-hand sides, RK sources, implicit-solve
increments, FFTs and potential, with the state (u, p), the divergence, the radial Poisson
solve, the projection and the diagnostics in FP64.

Its results are not the reference's to rounding; the tests bound its distance from the FP64
oracle per field and check what FP32 must not break: the projection still removes the
divergence down to the FP32 error of the potential, the manufactured potential comes back to
FP32 accuracy, and the torque (Nusselt number) stays within 0.1 % of the FP64 value.
"""
import numpy as np
import pytest

from harness import make_case, modulo_mean, oracle_for, rel_l2
from oracle import orc

pytestmark = pytest.mark.gpu


def _gpu(case, precision):
    from simpletcsolver_amd import Solver
    s = Solver(case["nx"], case["ny"], case["nz"], case["xf"], case["xc"], precision=precision, **case["kw"])
    s.load_fields(case["fields"])
    return s


@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 1, 1)])
def test_mixed_ten_steps_close_to_fp64_oracle(need_gpu, implicit):
    case = make_case(nx=24, ny=16, nz=12, Re=500.0, implicit=implicit, amp=0.1, t_amp=0.05)
    o = oracle_for(case)
    s = _gpu(case, "mixed")
    for _ in range(10):
        o.integrate()
        s.step()
    of, sf = o.npy_fields(), s.fields()
    errs = {}
    for n in ("ux", "uy", "uz", "t"):
        errs[n] = rel_l2(sf[n], of[n])
    errs["p"] = rel_l2(modulo_mean(sf["p"], case), modulo_mean(of["p"], case))
    print("mixed vs FP64 oracle after 10 steps, relative L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    for n in ("ux", "uy", "uz", "t"):
        assert errs[n] <= 1e-5, n
    assert errs["p"] <= 1e-3
    div = s.divergence()
    print(f"mixed max |div u| = {div:.2e}")
    assert div <= 1e-4
    s.close()


def test_mixed_manufactured_potential(need_gpu):
    nx, ny, nz = 33, 24, 48
    case = make_case(nx=nx, ny=ny, nz=nz, Re=100.0, amp=0.0, uy_walls=(0.0, 0.0))
    o = oracle_for(case)
    hxxf, hyxc, hz = o.array("hxxf").copy(), o.array("hyxc").copy(), o.hz
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((nz, ny, nx))
    rk, dt = 0, 0.3
    g = orc.rkcoefs()[rk, 2] * dt
    ux = np.zeros((nz, ny, nx + 1))
    ux[:, :, 1:nx] = g * (psi[:, :, 1:] - psi[:, :, :-1]) / hxxf[1:nx]
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=1)) / hyxc[1:nx + 1]
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=0)) / hz
    s = _gpu(case, "mixed")
    s.load_fields({"ux": ux, "uy": uy, "uz": uz})
    s.potential(rk, dt)
    got = s.get_field("psi")[:, :, 1:-1]
    d = got - psi
    err = np.abs(d - d.mean()).max() / np.abs(psi).max()
    print(f"mixed manufactured psi: {err:.2e}")
    assert err <= 1e-4
    s.close()


def test_mixed_torque_within_a_tenth_of_a_percent(need_gpu):
    case = make_case(nx=32, ny=24, nz=24, Re=300.0, implicit=(1, 0, 0), amp=0.05)
    s64 = _gpu(case, "fp64")
    s32 = _gpu(case, "mixed")
    s64.arm_diagnostics(True)
    s32.arm_diagnostics(True)
    for _ in range(20):
        s64.step()
        s32.step()
    i64, i32 = s64.injection()[0], s32.injection()[0]
    rel = abs(i32 - i64) / abs(i64)
    print(f"injection FP64 {i64:.12e} mixed {i32:.12e} rel {rel:.2e}")
    assert rel <= 1e-3
    s64.close()
    s32.close()


@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 1, 1)])
def test_mixed_refined_leaves_fp64_level_divergence(need_gpu, implicit):
    """TC_MIXED_REFINED: one defect-correction pass at the end of every step takes the
    divergence left by the FP32 potential down to the FP64 level (max |div u| <= 1e-10 after
    every step), and the fields stay as close to the FP64 oracle as the plain mixed step's"""
    case = make_case(nx=24, ny=16, nz=12, Re=500.0, implicit=implicit, amp=0.1, t_amp=0.05)
    o = oracle_for(case)
    s = _gpu(case, "mixed_refined")
    plain = _gpu(case, "mixed")
    worst = 0.0
    for _ in range(10):
        o.integrate()
        s.step()
        plain.step()
        worst = max(worst, s.divergence())
    of, sf = o.npy_fields(), s.fields()
    errs = {n: rel_l2(sf[n], of[n]) for n in ("ux", "uy", "uz", "t")}
    print(f"refined: max |div u| over 10 steps {worst:.2e} (plain mixed {plain.divergence():.2e}); "
          "relative L2 vs FP64:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert worst <= 1e-10
    for n in errs:
        assert errs[n] <= 1e-5, n
    s.close()
    plain.close()
