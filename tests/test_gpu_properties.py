"""Device path against what the mathematics fixes and against the oracle's statistics:

* a manufactured potential: u = gamma dt grad(psi_m) (the gradient of correct/*.c) must come
  back from tc_potential as psi_m up to a constant, at sizes where the FFT lengths take every
  radix of the device FFT (2, 3, 4, 5, 7, 8) and an odd radial length;
* a step leaves a solenoidal field (max |div u| at round-off, logging/divergence.c);
* plane Couette stays an exact steady state;
* running statistics (statistics.c) equal the oracle's after steps and samples in between;
* restart through the host copies is bit-exact on the device.
"""
import numpy as np
import pytest

from harness import gpu_for, make_case, oracle_for, rel_l2
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("curved", [True, False])
@pytest.mark.parametrize("shape", [(16, 12, 8), (9, 24, 6), (33, 16, 48), (12, 20, 14),
                                   (10, 70, 35), (7, 112, 50)])
def test_device_manufactured_potential(need_gpu, curved, shape):
    nx, ny, nz = shape
    case = make_case(nx=nx, ny=ny, nz=nz, Re=100.0, is_curved=curved, amp=0.0, uy_walls=(0.0, 0.0))
    o = oracle_for(case)
    hxxf, hyxc, hz = o.array("hxxf").copy(), o.array("hyxc").copy(), o.hz
    o.close()
    rng = np.random.default_rng(23)
    psi = rng.standard_normal((nz, ny, nx))
    rk, dt = 2, 0.21
    g = orc.rkcoefs()[rk, 2] * dt
    ux = np.zeros((nz, ny, nx + 1))
    ux[:, :, 1:nx] = g * (psi[:, :, 1:] - psi[:, :, :-1]) / hxxf[1:nx]
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=1)) / hyxc[1:nx + 1]
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=0)) / hz
    s = gpu_for(case)
    s.load_fields({"ux": ux, "uy": uy, "uz": uz})
    s.potential(rk, dt)
    got = s.get_field("psi")[:, :, 1:-1]
    d = got - psi
    err = np.abs(d - d.mean()).max() / np.abs(psi).max()
    print(f"manufactured psi {shape} curved={curved}: {err:.3e}")
    assert err <= 1e-11
    s.close()


def test_device_manufactured_potential_radix_5_7_long_lines(need_gpu):
    """lengths with radix 5 and 7 passes at a size where the transforms take their persistent
    kernels: 128 x 200 x 700 (theta: 100 = 4 * 25, z: 700 = 4 * 25 * 7)"""
    _manufactured_long(128, 200, 700, 31)


def _manufactured_long(nx, ny, nz, seed):
    from simpletcsolver_amd import Solver
    small = make_case(nx=nx, ny=ny, nz=4, Re=100.0, amp=0.0, uy_walls=(0.0, 0.0))
    o = oracle_for(small)
    hxxf, hyxc = o.array("hxxf").copy(), o.array("hyxc").copy()
    o.close()
    hz = small["kw"]["lengths"][2] / nz
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((nz, ny, nx))
    rk, dt = 1, 0.13
    g = orc.rkcoefs()[rk, 2] * dt
    ux = np.zeros((nz, ny, nx + 1))
    ux[:, :, 1:nx] = g * (psi[:, :, 1:] - psi[:, :, :-1]) / hxxf[1:nx]
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=1)) / hyxc[1:nx + 1]
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=0)) / hz
    s = Solver(nx, ny, nz, small["xf"], small["xc"], **small["kw"])
    s.load_fields({"ux": ux, "uy": uy, "uz": uz})
    s.potential(rk, dt)
    got = s.get_field("psi")[:, :, 1:-1]
    s.close()
    d = got - psi
    err = np.abs(d - d.mean()).max() / np.abs(psi).max()
    print(f"manufactured psi {(nx, ny, nz)}: {err:.3e}")
    assert err <= 1e-9


@pytest.mark.parametrize("shape,implicit,scalar", [((12, 20, 14), (1, 1, 1), 1),
                                                   ((10, 70, 35), (0, 0, 0), 0)])
def test_device_steps_with_radix_5_and_7_match_oracle(need_gpu, shape, implicit, scalar):
    """ten steps on grids whose transforms take radix 5 and 7 passes, against the oracle's
    mixed-radix transforms, to the ten-step bar of the parity tests"""
    from test_gpu_parity import VEL, _cmp
    nx, ny, nz = shape
    case = make_case(nx=nx, ny=ny, nz=nz, Re=200.0, implicit=implicit, scalar=scalar,
                     t_amp=1e-2 if scalar else 0.0)
    o, s = oracle_for(case), gpu_for(case)
    for n in range(10):
        dto = o.integrate()
        dts = s.step()
        assert abs(dts - dto) <= 1e-15 * dto, (n, dts, dto)
    _cmp(o, s, VEL + (("t",) if scalar else ()) + ("p", "psi"), case, 1e-10, gauge=("p", "psi"))
    assert s.divergence() <= 1e-12
    o.close()
    s.close()


def test_device_manufactured_potential_long_lines(need_gpu):
    """the same at 128 x 256 x 1024: the axial transforms (length 1024) take the persistent
    kernel with tiles of 8 lines and 1024 threads, the theta transforms twice the threads of a
    short tile.  The radial metrics come from the oracle on 4 planes (x and theta metrics do
    not depend on nz).  Tolerance: the smallest Poisson eigenvalues fall as 1/nz^2, so the
    round-off of the right-hand side reaches psi amplified by up to nz^2 eps."""
    from simpletcsolver_amd import Solver
    nx, ny, nz = 128, 256, 1024
    small = make_case(nx=nx, ny=ny, nz=4, Re=100.0, amp=0.0, uy_walls=(0.0, 0.0))
    o = oracle_for(small)
    hxxf, hyxc = o.array("hxxf").copy(), o.array("hyxc").copy()
    o.close()
    hz = small["kw"]["lengths"][2] / nz
    rng = np.random.default_rng(29)
    psi = rng.standard_normal((nz, ny, nx))
    rk, dt = 1, 0.13
    g = orc.rkcoefs()[rk, 2] * dt
    ux = np.zeros((nz, ny, nx + 1))
    ux[:, :, 1:nx] = g * (psi[:, :, 1:] - psi[:, :, :-1]) / hxxf[1:nx]
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=1)) / hyxc[1:nx + 1]
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=0)) / hz
    s = Solver(nx, ny, nz, small["xf"], small["xc"], **small["kw"])
    s.load_fields({"ux": ux, "uy": uy, "uz": uz})
    s.potential(rk, dt)
    got = s.get_field("psi")[:, :, 1:-1]
    s.close()
    d = got - psi
    err = np.abs(d - d.mean()).max() / np.abs(psi).max()
    print(f"manufactured psi {(nx, ny, nz)}: {err:.3e}")
    assert err <= 1e-9


def test_device_exact_transform_form_computes_the_same_bits(need_gpu):
    """at 128 x 256 x 1024 the persistent axial transforms (tiles of 8 lines cover nx = 128,
    1024 threads x 8 elements cover nz = 1024) and inverse theta transform (tiles of 16 lines,
    256 threads x 8 elements cover ny / 2 = 128) take their exact-tiling forms (twiddle factors
    in LDS, the next tile loaded in full before the current one's passes), and the divergence +
    theta transform its form with the twiddle tables in LDS; the general persistent forms
    compute the same potential bit for bit, each form alone and all three"""
    from simpletcsolver_amd import Solver
    nx, ny, nz = 128, 256, 1024
    small = make_case(nx=nx, ny=ny, nz=4, Re=100.0, amp=0.0, uy_walls=(0.0, 0.0))
    rng = np.random.default_rng(41)
    f = {"ux": rng.standard_normal((nz, ny, nx + 1)), "uy": rng.standard_normal((nz, ny, nx + 2)),
         "uz": rng.standard_normal((nz, ny, nx + 2))}
    f["ux"][:, :, 0] = 0.0
    f["ux"][:, :, -1] = 0.0
    out = []
    s = Solver(nx, ny, nz, small["xf"], small["xc"], **small["kw"])
    s.load_fields(f)
    for mask in (0, 1, 2, 4, 7, 7):
        s.tune("fft_exact", mask, 0)
        s.potential(1, 0.13)
        out.append(s.get_field("psi"))
    s.close()
    assert np.abs(out[0]).max() > 0.0
    for n in range(1, len(out)):
        assert np.array_equal(out[n], out[0]), n


def test_device_exact_fp32_transforms_compute_the_same_bits(need_gpu):
    """the mixed step's transforms at 128 x 256 x 1024: the axial ones (FP32 tiles of 16 lines,
    1024 threads x 16 elements covering nz = 1024) and the inverse theta one (tiles of 16 lines,
    256 threads x 8 elements covering ny / 2 = 128) take their persistent exact-tiling forms
    (twiddles in LDS, buffer-addressed, the whole next tile loaded before the passes), and the
    divergence + theta transform its persistent form (twiddle tables in LDS, uz's lower faces
    carried along a block's planes), where the one-tile forms ran before; each computes the same
    potential bit for bit"""
    from simpletcsolver_amd import Solver
    nx, ny, nz = 128, 256, 1024
    small = make_case(nx=nx, ny=ny, nz=4, Re=100.0, amp=0.0, uy_walls=(0.0, 0.0))
    rng = np.random.default_rng(43)
    f = {"ux": rng.standard_normal((nz, ny, nx + 1)), "uy": rng.standard_normal((nz, ny, nx + 2)),
         "uz": rng.standard_normal((nz, ny, nx + 2))}
    f["ux"][:, :, 0] = 0.0
    f["ux"][:, :, -1] = 0.0
    out = []
    s = Solver(nx, ny, nz, small["xf"], small["xc"], precision="mixed", **small["kw"])
    s.load_fields(f)
    for mask in (0, 1, 2, 4, 7, 7):
        s.tune("fft_exact", mask, 0)
        s.potential(1, 0.13)
        out.append(s.get_field("psi"))
    s.close()
    assert np.abs(out[0]).max() > 0.0
    for n in range(1, len(out)):
        assert np.array_equal(out[n], out[0]), n


@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 1, 1)])
def test_device_step_is_solenoidal(need_gpu, implicit):
    case = make_case(nx=24, ny=16, nz=12, Re=300.0, implicit=implicit, amp=0.3)
    s = gpu_for(case)
    assert s.divergence() > 1e-3
    for _ in range(3):
        s.step()
        assert s.divergence() <= 1e-13
    s.close()


def test_device_plane_couette_steady(need_gpu):
    case = make_case(nx=32, ny=8, nz=8, Re=50.0, is_curved=False, implicit=(1, 1, 1), amp=0.0,
                     stretch=2.5)
    s = gpu_for(case)
    u0 = s.fields()
    for _ in range(5):
        s.step()
    u1 = s.fields()
    for n in ("ux", "uy", "uz", "t"):
        assert np.abs(u1[n] - u0[n]).max() <= 1e-13, n
    s.close()


@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 0, 0)])
def test_statistics_match_oracle(need_gpu, implicit):
    case = make_case(nx=16, ny=12, nz=8, Re=200.0, implicit=implicit, amp=0.2, t_amp=0.05)
    o, s = oracle_for(case), gpu_for(case)
    s.stat_arm(True)
    for step in range(4):
        o.integrate()
        s.step()
        if step % 2 == 1:
            o.stat_collect()
            s.stat_collect()
    got, num = s.stats()
    want = o.stats()
    assert num == 2
    for n, a in want.items():
        e = rel_l2(got[n], a)
        print(f"{n}: {e:.3e}")
        assert e <= 1e-10, n
    s.close()


def test_device_restart_is_bit_exact(need_gpu):
    case = make_case(nx=16, ny=12, nz=8, Re=200.0, implicit=(1, 1, 1), amp=0.2, t_amp=0.05)
    s = gpu_for(case)
    for _ in range(4):
        s.step()
    full = s.fields()
    s.close()
    s = gpu_for(case)
    for _ in range(2):
        s.step()
    mid = s.fields()
    s.close()
    s = gpu_for(dict(case, fields=mid))
    for _ in range(2):
        s.step()
    again = s.fields()
    for n in ("ux", "uy", "uz", "p", "t"):
        assert np.array_equal(full[n], again[n]), n
    s.close()
