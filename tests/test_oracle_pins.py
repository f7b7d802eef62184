"""Pins of the CPU oracle (oracle/orc.c) by what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors for the time step, and its own build
needs FFTW3 and two empty submodules, so the oracle is pinned here by closed forms, library
routines (numpy.fft, numpy.linalg.solve), brute force on tiny inputs and the discrete
identities the scheme is built on (docs/source/numerical_method of the reference):

* RK coefficients: src/runge_kutta.c:8-15 as exact fractions;
* FFT: numpy.fft and an O(N^2) DFT (FFTW conventions of compute_potential.c:247-309);
* Thomas / cyclic Thomas: dense solves, and the singular-pivot gauge of tdm.c:38-44;
* decide_dt: the closed form of decide_dt.c:32-160;
* projection: div u after a substep (logging/divergence.c) at round-off, and a manufactured
  potential: u = grad(psi) must give back psi up to a constant;
* energy: the advective and pressure terms do no work on a solenoidal field, and the work of
  the diffusive terms equals injection - dissipation (logging/injection.c, dissipation.c);
* plane Couette is an exact discrete steady state with the exact wall stress; the circular
  Couette injection converges at first order to the closed form of the reference's wall
  stencil; the Cartesian limit zeroes the curvature tensors; restart is bit-exact.

A dropped term, a wrong sign or index or a transposed operand in the oracle breaks one of them.
"""
import numpy as np
import pytest

from harness import make_case, oracle_for
from oracle import orc


# ---------------------------------------------------------------------------------- RK3
def test_rkcoefs_exact_fractions():
    c = orc.rkcoefs()
    alpha = np.array([32.0, 25.0, 45.0]) / 60.0
    beta = np.array([0.0, -17.0, -25.0]) / 60.0
    assert np.array_equal(c[:, 0], alpha)
    assert np.array_equal(c[:, 1], beta)
    # gamma is written alpha + beta in the reference (runge_kutta.c:12-14): two roundings
    assert np.array_equal(c[:, 2], alpha + beta)
    assert np.allclose(c[:, 2], np.array([32.0, 8.0, 20.0]) / 60.0, rtol=1e-15, atol=0)
    # consistency of the low-storage scheme: the gamma weights sum to one
    assert abs(c[:, 2].sum() - 1.0) < 1e-15


# ---------------------------------------------------------------------------------- FFT
FFT_SIZES = (2, 3, 4, 5, 6, 8, 12, 16, 24, 48, 64, 96, 128, 192, 256, 1024)


@pytest.mark.parametrize("n", FFT_SIZES)
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tol = 1e-14 * np.sqrt(n) * np.log2(max(n, 2))
    fwd = orc.fft_c2c(x, -1)
    ref = np.fft.fft(x)
    assert np.linalg.norm(fwd - ref) <= tol * np.linalg.norm(ref)
    bwd = orc.fft_c2c(x, +1)
    ref = np.fft.ifft(x) * n
    assert np.linalg.norm(bwd - ref) <= tol * np.linalg.norm(ref)
    r = rng.standard_normal(n)
    spec = orc.fft_r2c(r)
    ref = np.fft.rfft(r)
    assert np.linalg.norm(spec - ref) <= tol * np.linalg.norm(ref)
    back = orc.fft_c2r(ref, n)
    assert np.linalg.norm(back - n * r) <= tol * np.linalg.norm(n * r)


@pytest.mark.parametrize("n", (2, 3, 4, 6, 8, 12))
def test_fft_brute_force_dft(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    m = np.arange(n)
    for sign in (-1, +1):
        w = np.exp(sign * 2j * np.pi * np.outer(m, m) / n)
        assert np.allclose(orc.fft_c2c(x, sign), w @ x, rtol=0, atol=1e-13 * n)


def test_c2r_ignores_imaginary_dc_and_nyquist():
    # FFTW's c2r reads a Hermitian half spectrum: Im of bins 0 and n/2 are not part of it
    n = 16
    rng = np.random.default_rng(3)
    s = np.fft.rfft(rng.standard_normal(n))
    t = s.copy()
    t[0] += 5j
    t[-1] -= 2j
    assert np.array_equal(orc.fft_c2r(s, n), orc.fft_c2r(t, n))


# ---------------------------------------------------------------------------------- Thomas
def _dense(l, c, u, periodic=False):
    n = len(c)
    a = np.diag(c) + np.diag(l[1:], -1) + np.diag(u[:-1], 1)
    if periodic:
        a[0, n - 1] += l[0]
        a[n - 1, 0] += u[n - 1]
    return a


@pytest.mark.parametrize("n", (2, 3, 5, 16, 64))
def test_gtsv_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    l, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    c = 2.5 + rng.uniform(0, 1, n)
    q = rng.standard_normal(n)
    x = orc.gtsv(l, c, u, q)
    assert np.allclose(x, np.linalg.solve(_dense(l, c, u), q), rtol=1e-13, atol=1e-14)
    qc = q + 1j * rng.standard_normal(n)
    xc = orc.gtsv(l, c, u, qc)
    assert np.allclose(xc, np.linalg.solve(_dense(l, c, u).astype(complex), qc), rtol=1e-13, atol=1e-14)


def test_gtsv_singular_neumann_gauge():
    # Neumann Laplacian: rows sum to zero, null space = constants; the last pivot vanishes
    # and tdm.c:38-44 sets the last unknown to zero, giving the particular solution
    n = 12
    l = np.ones(n)
    u = np.ones(n)
    c = -2.0 * np.ones(n)
    c[0] += l[0]
    c[-1] += u[-1]
    rng = np.random.default_rng(7)
    q = rng.standard_normal(n)
    q -= q.mean()
    x = orc.gtsv(l, c, u, q)
    assert x[-1] == 0.0
    assert np.allclose(_dense(l, c, u) @ x, q, atol=1e-12)


@pytest.mark.parametrize("n", (3, 4, 8, 33))
def test_tdm_periodic_matches_dense_solve(n):
    rng = np.random.default_rng(50 + n)
    l, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    c = 2.5 + rng.uniform(0, 1, n)
    q = rng.standard_normal(n)
    x = orc.tdm_periodic(l, c, u, q)
    assert np.allclose(x, np.linalg.solve(_dense(l, c, u, periodic=True), q), rtol=1e-13, atol=1e-14)


def test_tdm_periodic_singular_gauge():
    # periodic Laplacian (the y/z systems at zero diffusion number are never singular, but
    # the guard of tdm.c:93-95 must give a solution of the consistent singular system)
    n = 10
    l = np.ones(n)
    u = np.ones(n)
    c = -2.0 * np.ones(n)
    rng = np.random.default_rng(9)
    q = rng.standard_normal(n)
    q -= q.mean()
    x = orc.tdm_periodic(l, c, u, q)
    assert x[-1] == 0.0
    assert np.allclose(_dense(l, c, u, periodic=True) @ x, q, atol=1e-12)


# ---------------------------------------------------------------------------------- dt
@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 0, 0), (1, 1, 1)])
def test_decide_dt_closed_form(implicit):
    case = make_case(nx=16, ny=12, nz=8, Re=50.0, Pr=0.7, implicit=implicit, amp=0.5)
    o = oracle_for(case)          # coef_dt_adv = coef_dt_dif = 1
    nx = case["nx"]
    ux, uy, uz = o.array("ux"), o.array("uy"), o.array("uz")
    hxxf, hxxc, hyxc, hz = o.array("hxxf"), o.array("hxxc"), o.array("hyxc"), o.hz
    small = 1e-8
    adv = min(1.0,
              (hxxf[1:nx] / (np.abs(ux[1:-1, 1:-1, 1:nx]) + small)).min(),
              (hyxc[1:nx + 1] / (np.abs(uy[1:-1, 1:-1, 1:nx + 1]) + small)).min(),
              (hz / (np.abs(uz[1:-1, 1:-1, 1:nx + 1]) + small)).min())
    diff = max(1.0 / 50.0, 1.0 / 50.0 / 0.7)
    h = (hxxc.min(), hyxc[1], hz)
    dt = adv
    for d in range(3):
        if not implicit[d]:
            dt = min(dt, 1.0 / diff * 0.5 / 3 * h[d] ** 2)
    assert o.decide_dt() == dt


def test_decide_dt_uniform_flow():
    # a uniform axial flow U: dt = hz / (U + 1e-8) when that is the smallest bound
    case = make_case(nx=8, ny=8, nz=8, Re=1e6, implicit=(1, 1, 1), amp=0.0, uy_walls=(0.0, 0.0))
    f = case["fields"]
    f["uz"][...] = 3.0
    o = oracle_for(case)
    assert o.decide_dt() == o.hz / (3.0 + 1e-8)


# ---------------------------------------------------------------------------------- projection
@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 1, 1)])
@pytest.mark.parametrize("curved", [True, False])
def test_substep_leaves_solenoidal_field(implicit, curved):
    case = make_case(nx=16, ny=12, nz=8, Re=100.0, implicit=implicit, is_curved=curved, amp=0.3)
    o = oracle_for(case)
    assert o.divergence() > 1e-3
    o.integrate()
    assert o.divergence() <= 1e-13


@pytest.mark.parametrize("curved", [True, False])
@pytest.mark.parametrize("shape", [(16, 12, 8), (9, 6, 10)])
def test_manufactured_potential(curved, shape):
    """u = gamma dt grad(psi_m), with the gradient of correct/*.c, must give psi_m back (up to a
    constant) from compute_potential: the discrete div(grad) the Poisson solver inverts."""
    nx, ny, nz = shape
    case = make_case(nx=nx, ny=ny, nz=nz, Re=100.0, is_curved=curved, amp=0.0, uy_walls=(0.0, 0.0))
    o = oracle_for(case)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((nz, ny, nx))            # interior cells i = 1..nx
    rk, dt = 1, 0.37
    g = orc.rkcoefs()[rk, 2] * dt
    hxxf, hyxc, hz = o.array("hxxf"), o.array("hyxc"), o.hz
    ux = np.zeros((nz, ny, nx + 1))
    ux[:, :, 1:nx] = g * (psi[:, :, 1:] - psi[:, :, :-1]) / hxxf[1:nx]     # UX(i), i = 2..nx
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=1)) / hyxc[1:nx + 1]
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:nx + 1] = g * (psi - np.roll(psi, 1, axis=0)) / hz
    o.load_fields({"ux": ux, "uy": uy, "uz": uz})
    o.potential(rk, dt)
    got = o.array("psi")[1:-1, 1:-1, 1:-1]
    d = got - psi
    assert np.abs(d - d.mean()).max() <= 1e-11 * np.abs(psi).max()


# ---------------------------------------------------------------------------------- energy
def _work(o, u0, s):
    """sum of J u . s over the staggered positions (faces for ux, centres for uy, uz)"""
    nx = o.nx
    jf, jc = o.array("jdxf"), o.array("jdxc")
    ux0, uy0, uz0 = u0
    terms = [jf[1:nx] * ux0[1:-1, 1:-1, 1:nx] * s[0],
             jc[1:nx + 1] * uy0[1:-1, 1:-1, 1:nx + 1] * s[1],
             jc[1:nx + 1] * uz0[1:-1, 1:-1, 1:nx + 1] * s[2]]
    return sum(t.sum() for t in terms), sum(np.abs(t).sum() for t in terms)


def _solenoidal_case(curved, Re):
    case = make_case(nx=16, ny=12, nz=8, Re=100.0, implicit=(0, 0, 0), is_curved=curved, amp=0.3,
                     t_amp=0.3)
    o = oracle_for(case)
    o.integrate()
    fields = o.npy_fields()
    o.close()
    case = dict(case, fields=fields, kw=dict(case["kw"], Re=Re))
    o = oracle_for(case)
    assert o.divergence() <= 1e-13
    return o


@pytest.mark.parametrize("curved", [True, False])
def test_advection_and_pressure_do_no_work(curved):
    o = _solenoidal_case(curved, Re=1e300)          # diffusion ~1e-300: srca = advection
    u0 = tuple(o.array(n).copy() for n in ("ux", "uy", "uz"))
    t0 = o.array("t").copy()
    o.predict(0, 1e-3)
    a = [o.array(f"src{f}0") for f in ("ux", "uy", "uz")]
    g = [o.array(f"src{f}2") for f in ("ux", "uy", "uz")]   # pressure gradient (explicit diffusion)
    w, scale = _work(o, u0, a)
    assert abs(w) <= 1e-14 * scale
    w, scale = _work(o, u0, g)
    assert abs(w) <= 1e-14 * scale
    nx = o.nx
    tw = o.array("jdxc")[1:nx + 1] * t0[1:-1, 1:-1, 1:nx + 1] * o.array("srct0")
    assert abs(tw.sum()) <= 1e-14 * np.abs(tw).sum()


@pytest.mark.parametrize("curved", [True, False])
def test_diffusive_work_is_injection_minus_dissipation(curved):
    inviscid = _solenoidal_case(curved, Re=1e300)
    viscous = _solenoidal_case(curved, Re=80.0)
    u0 = tuple(viscous.array(n).copy() for n in ("ux", "uy", "uz"))
    t0 = viscous.array("t").copy()
    viscous.compute_tensors()
    inj, dis = viscous.injection(), viscous.dissipation()
    inviscid.predict(0, 1e-3)
    viscous.predict(0, 1e-3)
    d = [viscous.array(f"src{f}0") - inviscid.array(f"src{f}0") for f in ("ux", "uy", "uz")]
    w, scale = _work(viscous, u0, d)
    assert abs(w - (inj[0] - dis[0])) <= 1e-12 * scale
    nx = viscous.nx
    dt_ = viscous.array("srct0") - inviscid.array("srct0")
    tw = (viscous.array("jdxc")[1:nx + 1] * t0[1:-1, 1:-1, 1:nx + 1] * dt_).sum()
    assert abs(tw - (inj[1] - dis[1])) <= 1e-12 * abs(inj[1] - dis[1])


# ---------------------------------------------------------------------------------- flows
def test_plane_couette_is_a_discrete_steady_state():
    case = make_case(nx=16, ny=8, nz=8, Re=50.0, is_curved=False, implicit=(1, 0, 0), amp=0.0,
                     stretch=2.5, t_amp=0.0)
    o = oracle_for(case)
    u0 = o.npy_fields()
    for _ in range(5):
        o.integrate()
    u1 = o.npy_fields()
    for n in ("ux", "uy", "uz", "t"):
        assert np.abs(u1[n] - u0[n]).max() <= 1e-13, n
    assert o.divergence() <= 1e-14


def _couette_injection(nx, curved):
    eta = 0.714
    ly, lz = (np.pi / 3, 1.0) if curved else (1.0, 1.0)
    Re = 50.0
    case = make_case(nx=nx, ny=4, nz=4, Re=Re, implicit=(1, 0, 0), amp=0.0, ly=ly, lz=lz,
                     stretch=1.0, is_curved=curved, eta=eta)
    o = oracle_for(case)
    o.compute_tensors()
    return o.injection()[0], Re, ly, lz, eta


def test_plane_couette_injection_is_exact():
    # linear profile from 1 to 0 across lx = 1: the wall stress nu / lx over ly lz
    inj, Re, ly, lz, _ = _couette_injection(16, curved=False)
    assert abs(inj - ly * lz / Re) <= 1e-13 * inj


def test_circular_couette_injection_converges_to_its_wall_stencil():
    """Laminar circular Couette u = A r + B / r.  At the inner wall the reference forms
    LXY = (UY(1) - UY(0)) / HXXF(1), one-sided over half a cell, and LYX1 with the weight 0 on
    the wall value UY(0) (predict/lyx.c:80-81), which tends to -uy / (2 r), half the curvature
    term of the continuous stress.  The injection (logging/injection.c) therefore tends at
    first order to nu ri ly lz (-(A - B/ri^2) + (A + B/ri) / (2 ri)), not to the torque of the
    continuous flow; the planar case has no curvature term and is exact."""
    ri = 1.0
    vals = []
    for nx in (16, 32, 64):
        inj, Re, ly, lz, eta = _couette_injection(nx, curved=True)
        vals.append(inj)
    ro = ri / eta
    a = (0.0 * ro - 1.0 * ri) / (ro ** 2 - ri ** 2)
    b = ri * ro * (1.0 * ro - 0.0 * ri) / (ro ** 2 - ri ** 2)
    lim = 1.0 / Re * ri * ly * lz * (-(a - b / ri ** 2) + 0.5 * (a + b / ri) / ri)
    err = [abs(v - lim) / lim for v in vals]
    assert err[2] < 3e-3
    assert err[0] / err[1] > 1.8 and err[1] / err[2] > 1.8
    # the continuous torque differs by the missing half of the curvature term
    torque = 2.0 / Re * ri ** 2 * ro ** 2 / (ro ** 2 - ri ** 2) * ly * lz
    assert abs(lim / torque - 1.0) > 0.1


def test_cartesian_limit_zeroes_curvature_terms():
    case = make_case(nx=12, ny=8, nz=6, Re=100.0, is_curved=False, amp=0.3)
    o = oracle_for(case)
    o.compute_tensors()
    # the metric differences JDXC/HXXC vanish up to the rounding of hx hy hz / hx
    umax = max(np.abs(o.array(n)).max() for n in ("ux", "uy"))
    assert np.abs(o.array("lyx1")).max() <= 1e-13 * umax
    assert np.abs(o.array("lyy1")).max() <= 1e-13 * umax
    case = make_case(nx=12, ny=8, nz=6, Re=100.0, is_curved=True, amp=0.3)
    o = oracle_for(case)
    o.compute_tensors()
    assert o.array("lyx1").any() and o.array("lyy1").any()


@pytest.mark.parametrize("implicit", [(0, 0, 0), (1, 1, 1)])
def test_restart_is_bit_exact(implicit):
    case = make_case(nx=12, ny=8, nz=8, Re=200.0, implicit=implicit, amp=0.2, t_amp=0.1)
    o = oracle_for(case)
    for _ in range(4):
        o.integrate()
    full = o.npy_fields()
    o = oracle_for(case)
    for _ in range(2):
        o.integrate()
    mid = o.npy_fields()
    o = oracle_for(dict(case, fields=mid))
    for _ in range(2):
        o.integrate()
    again = o.npy_fields()
    for n in ("ux", "uy", "uz", "p", "t"):
        assert np.array_equal(full[n], again[n]), n


# ---------------------------------------------------------------------------------- statistics
def test_statistics_accumulate_samples():
    case = make_case(nx=10, ny=6, nz=4, Re=100.0, amp=0.3)
    o = oracle_for(case)
    o.integrate()
    assert o.stat_collect() == 1
    assert o.stat_collect() == 2
    st = o.stats()
    nx = o.nx
    ux = o.array("ux")[1:-1, 1:-1, :]
    uy = o.array("uy")[1:-1, 1:-1, :]
    assert np.array_equal(st["ux1"], 2.0 * ux)
    assert np.allclose(st["ux2"], 2.0 * ux * ux, rtol=1e-15, atol=0)
    assert np.array_equal(st["uy1"], 2.0 * uy)
    full_ux = o.array("ux")
    full_uy = o.array("uy")
    x = 0.5 * full_ux[1:-1, :-2, :] + 0.5 * full_ux[1:-1, 1:-1, :]
    y = 0.5 * full_uy[1:-1, 1:-1, 0:nx + 1] + 0.5 * full_uy[1:-1, 1:-1, 1:nx + 2]
    assert np.array_equal(st["adv"], 2.0 * (x * y))
    tens = o.array("lyx0") + o.array("lyx1") + o.array("lxy")
    assert np.allclose(st["dif"], 2.0 * tens[1:-1, 1:-1, :], rtol=1e-15, atol=1e-300)
