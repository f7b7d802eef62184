"""Closed-form pins of the RK3 time integration, the implicit diffusion prefactors and the
pressure update, written once and run on the oracle (tests/test_oracle_closed_forms.py) and on
the device through the C ABI (tests/test_gpu_closed_forms.py).

What the mathematics fixes, and what each pin states:

* Viscous term of an incompressible flow in stress form, d_j(nu (d_j u_i + d_i u_j)): the
  second difference of u_i along its own direction carries 2 nu, along the two others nu
  ("txx = 2 lxx", predict/ux.c:228-248, uy.c:276-296, uz.c:279-299); the scalar's is kappa in
  every direction (predict/t.c:178-239).  On a periodic Fourier mode cos(2 pi m j / n + phi)
  the second difference over a spacing h has the eigenvalue -4 sin^2(pi m / n) / h^2, so one
  explicit substep changes a mode by alpha dt c lambda per unit amplitude (c = 2 nu, nu or
  kappa).  Pin `explicit_eigenvalue`.
* Implicit treatment of a direction is Crank-Nicolson of the same operator: its explicit
  part moves to the gamma-weighted source and the increment is solved with
  (1 - gamma dt / 2 c L), which is 2f for the own direction and f for the others with
  f = gamma dt nu / 2 (ux.c:495-542, uy.c:522-562, uz.c:414-454, t.c:329-369).  The one
  exception, kept from the reference on purpose, is the scalar's y solve with 2f although its
  explicit y term carries kappa (t.c:346-349).  Pins `implicit_mode` (y, z: closed form) and
  `implicit_x` (x: against the explicit x operator probed from the same code on unit vectors).
* Wray's low-storage RK3 (runge_kutta.c:8-15): delta_s = dt (alpha_s a_s + beta_s a_{s-1} +
  gamma_s g_s), the a/b sources rotated every substep (predict/main.c:8-25).  On a linear
  problem (the scalar with the velocity at rest) a full step is a closed-form recursion.
  Pin `rk3_step`.
* Pressure update (update_pressure.c:23-173): P += psi - gamma dt nu L_d psi for every
  implicit direction d, with L_d psi the d-part of div(grad psi) (the gradient of
  correct/*.c, the divergence of logging/divergence.c).  Pin `pressure_update`.
* Energy: the advective and pressure terms do no work, so with the diffusivity switched off
  the kinetic and scalar energies change only through the time integration, whose error for
  RK3 is third order in dt (docs/source/example/energy: Re = 1e16, coef_dt_adv 0.1 .. 0.8;
  energy.py:70 draws the dt^3 guide).  Pin `energy_decay`.

Each pin returns the largest relative deviation from its closed form; rounding keeps it near
1e-14, and a slip of a factor 2, a swapped weight or a dropped term moves it to O(1e-2..1).
"""
import numpy as np

import icgen

RK_A = np.array([32.0, 25.0, 45.0]) / 60.0
RK_B = np.array([0.0, -17.0, -25.0]) / 60.0
RK_G = RK_A + RK_B

# explicit diffusion coefficient of each field along y and z, in units of the diffusivity
OWN = {"ux": 0, "uy": 1, "uz": 2, "t": None}
# implicit solve factor over the explicit coefficient: Crank-Nicolson (1) except the scalar's
# y solve (t.c:346-349: 2f against kappa)
CN_FACTOR = {("t", 1): 2.0}


def coef(field, d):
    """explicit diffusion coefficient of `field` along direction d (0 x, 1 y, 2 z), in units of
    its diffusivity: 2 along the own direction of a velocity component, 1 otherwise"""
    return 2.0 if OWN[field] == d else 1.0


class OracleFlow:
    """the oracle behind the small interface the pins use"""

    def __init__(self, nx, ny, nz, xf, xc, **kw):
        from oracle.orc import Oracle
        self.o = Oracle(nx, ny, nz, xf, xc, **kw)

    def load(self, fields):
        self.o.load_fields(fields)

    def predict(self, rk, dt):
        self.o.predict(rk, dt)

    def pressure(self, rk, dt):
        self.o.pressure(rk, dt)

    def get(self, name):
        return self.o.array(name)[1:-1, 1:-1, :].copy()

    def step(self):
        return self.o.integrate()

    def decide_dt(self):
        return self.o.decide_dt()

    def energy(self):
        return self.o.total_energy()

    def close(self):
        self.o.close()


class DeviceFlow:
    """a device solver (libtcstep.so through ctypes) behind the same interface"""

    def __init__(self, nx, ny, nz, xf, xc, **kw):
        from simpletcsolver_amd import Solver
        self.s = Solver(nx, ny, nz, xf, xc, **kw)

    def load(self, fields):
        self.s.load_fields(fields)

    def predict(self, rk, dt):
        self.s.predict(rk, dt)

    def pressure(self, rk, dt):
        self.s.pressure(rk, dt)

    def get(self, name):
        return self.s.get_field(name)

    def step(self):
        return self.s.step()

    def decide_dt(self):
        return self.s.decide_dt()

    def energy(self):
        return self.s.total_energy()

    def close(self):
        self.s.close()


class Setup:
    """a small box with walls at rest, the scalar's walls at 0, every field zero"""

    def __init__(self, flow_cls, nx=8, ny=8, nz=8, curved=False, Re=1.0, Pr=0.7, stretch=1.5):
        self.flow_cls = flow_cls
        self.nx, self.ny, self.nz = nx, ny, nz
        lx = 0.4 if curved else 1.0
        self.xf, self.xc = icgen.grid(nx, lx, curved, stretch)
        self.kw = dict(is_curved=curved, lengths=(lx, 0.9 if curved else 1.1, 0.8), Re=Re, Pr=Pr,
                       scalar=1, uy_walls=(0.0, 0.0), t_walls=(0.0, 0.0))
        self.nu = 1.0 / Re
        self.kappa = 1.0 / (Re * Pr)
        from oracle.orc import Oracle
        o = Oracle(nx, ny, nz, self.xf, self.xc, **self.kw)
        # face metrics: nx+1 entries for i = 1..nx+1; centre metrics: nx+2 for i = 0..nx+1
        self.hyxf = o.array("hyxf").copy()
        self.hyxc = o.array("hyxc").copy()
        self.hxxf = o.array("hxxf").copy()
        self.jdxf = o.array("jdxf").copy()
        self.jdxc = o.array("jdxc").copy()
        self.hz = self.kw["lengths"][2] / nz
        o.close()

    def flow(self, implicit=(0, 0, 0), **extra):
        return self.flow_cls(self.nx, self.ny, self.nz, self.xf, self.xc,
                             **dict(self.kw, implicit=implicit, **extra))

    def zeros(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return {n: np.zeros((nz, ny, nx + 1) if n == "ux" else (nz, ny, nx + 2))
                for n in ("ux", "uy", "uz", "p", "t")}

    def interior(self, field):
        """NPY x slots of the unknowns: ux faces i = 2..nx, centres i = 1..nx"""
        return slice(1, self.nx) if field == "ux" else slice(1, self.nx + 1)

    def spacing(self, field, d):
        """grid spacing along y (per x slot of the NPY layout) or z"""
        if d == 2:
            return np.full(self.nx + (1 if field == "ux" else 2), self.hz)
        return self.hyxf if field == "ux" else self.hyxc

    def diffusivity(self, field):
        return self.kappa if field == "t" else self.nu

    def shape(self, field, profile, my=0, mz=0, phase=(0.3, 1.1)):
        """profile(x) cos(2 pi my j / ny + phase) cos(2 pi mz k / nz + phase) in NPY layout"""
        ny, nz = self.ny, self.nz
        y = np.cos(2 * np.pi * my * np.arange(ny) / ny + phase[0]) if my else np.ones(ny)
        z = np.cos(2 * np.pi * mz * np.arange(nz) / nz + phase[1]) if mz else np.ones(nz)
        return profile[None, None, :] * y[None, :, None] * z[:, None, None]

    def profile(self, field, seed=5):
        """a smooth random x profile, zero on the wall slots"""
        nxs = self.nx + (1 if field == "ux" else 2)
        sl = self.interior(field)
        v = np.zeros(nxs)
        v[sl] = np.random.default_rng(seed).uniform(0.5, 1.5, len(range(nxs)[sl]))
        return v


def substep_increment(st, field, implicit, rk, dt, u, amp=1e-3):
    """linear part of one predictor substep's increment of `field` started from amp u (every
    other field zero) on a fresh flow: (F(+amp u) - F(-amp u)) / (2 amp) - u.  The advective
    terms are bilinear in the velocity, so the central difference removes them exactly."""
    out = []
    for sign in (1.0, -1.0):
        fl = st.flow(implicit)
        f = st.zeros()
        f[field] = sign * amp * u
        fl.load(f)
        fl.predict(rk, dt)
        out.append(fl.get(field))
        fl.close()
    return (out[0] - out[1]) / (2.0 * amp) - u


def _relmax(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


def explicit_eigenvalue(st, field, d, m=1, rk=1, dt=0.1):
    """one explicit substep on profile(x) times a Fourier mode along d (1 y, 2 z), minus the
    same on the profile alone, is alpha dt c diffusivity lambda_d profile mode"""
    v = st.profile(field)
    kw = dict(my=m) if d == 1 else dict(mz=m)
    u = st.shape(field, v, **kw)
    u0 = st.shape(field, v)
    mode = st.shape(field, np.ones_like(v), **kw)
    dm = substep_increment(st, field, (0, 0, 0), rk, dt, u)
    d0 = substep_increment(st, field, (0, 0, 0), rk, dt, u0)
    n = st.ny if d == 1 else st.nz
    lam = -4.0 * np.sin(np.pi * m / n) ** 2 / st.spacing(field, d) ** 2
    want = RK_A[rk] * dt * coef(field, d) * st.diffusivity(field) * lam[None, None, :] * u
    sl = st.interior(field)
    got = dm - d0 * mode
    return _relmax(got[:, :, sl], want[:, :, sl])


def implicit_mode(st, field, d, m=1, rk=1, dt=0.1):
    """the increment with direction d (1 y, 2 z) implicit equals Crank-Nicolson of the explicit
    operator along d: (delta_expl + (gamma - alpha) dt L_d u) / (1 - gamma dt / 2 k L_d) with
    L_d u = c diffusivity lambda_d u and k = 1 (2 for the scalar's y solve)"""
    v = st.profile(field)
    kw = dict(my=m) if d == 1 else dict(mz=m)
    u = st.shape(field, v, **kw)
    implicit = tuple(int(q == d) for q in range(3))
    de = substep_increment(st, field, (0, 0, 0), rk, dt, u)
    di = substep_increment(st, field, implicit, rk, dt, u)
    n = st.ny if d == 1 else st.nz
    lam = -4.0 * np.sin(np.pi * m / n) ** 2 / st.spacing(field, d) ** 2
    ld = coef(field, d) * st.diffusivity(field) * lam[None, None, :]
    k = CN_FACTOR.get((field, d), 1.0)
    want = (de + (RK_G[rk] - RK_A[rk]) * dt * ld * u) / (1.0 - 0.5 * RK_G[rk] * dt * k * ld)
    sl = st.interior(field)
    return _relmax(di[:, :, sl], want[:, :, sl])


def probe_x_operator(st, field, rk=0, dt=0.1):
    """the explicit operator along x on profiles uniform in y and z, column by column from
    unit vectors (only the x stencils see such a profile; on a planar grid no curvature term
    does either)"""
    sl = st.interior(field)
    nxs = st.nx + (1 if field == "ux" else 2)
    idx = range(nxs)[sl]
    m = np.zeros((len(idx), len(idx)))
    for c, i in enumerate(idx):
        e = np.zeros(nxs)
        e[i] = 1.0
        inc = substep_increment(st, field, (0, 0, 0), rk, dt, st.shape(field, e))
        m[:, c] = inc[0, 0, sl] / (RK_A[rk] * dt)
    return m


def implicit_x(st, field, rk=1, dt=0.1):
    """with x implicit, a profile uniform in y and z is advanced by Crank-Nicolson of the
    probed explicit x operator: (I - gamma dt / 2 M_x) delta = gamma dt M_x u"""
    mx = probe_x_operator(st, field)
    v = st.profile(field, seed=9)
    u = st.shape(field, v)
    di = substep_increment(st, field, (1, 0, 0), rk, dt, u)
    sl = st.interior(field)
    n = mx.shape[0]
    want = np.linalg.solve(np.eye(n) - 0.5 * RK_G[rk] * dt * mx, RK_G[rk] * dt * (mx @ v[sl]))
    return _relmax(di[0, 0, sl], want), mx


def rk3_step(st, implicit, dt=0.05, my=1, mz=2):
    """three predictor substeps of the scalar (velocity at rest: a linear problem) against the
    low-storage recursion in closed form.  The state is profile(x) times one Fourier mode in y
    and z; per x slot the y and z operators are kappa lambda(i), the x operator is probed."""
    field = "t"
    mx = probe_x_operator(st, field)
    v = st.profile(field, seed=13)
    u = st.shape(field, v, my=my, mz=mz)
    sl = st.interior(field)
    lam_y = -4.0 * np.sin(np.pi * my / st.ny) ** 2 / st.spacing(field, 1)[sl] ** 2
    lam_z = -4.0 * np.sin(np.pi * mz / st.nz) ** 2 / st.hz ** 2
    ops = [mx, np.diag(st.kappa * lam_y), st.kappa * lam_z * np.eye(len(v[sl]))]
    expl = sum((ops[q] for q in range(3) if not implicit[q]), np.zeros_like(mx))
    impl = sum((ops[q] for q in range(3) if implicit[q]), np.zeros_like(mx))
    w = v[sl].copy()
    a_prev = np.zeros_like(w)
    for s in range(3):
        a = expl @ w
        g = impl @ w
        d = dt * (RK_A[s] * a + RK_B[s] * a_prev + RK_G[s] * g)
        for q in range(3):
            if implicit[q]:
                k = CN_FACTOR.get((field, q), 1.0)
                d = np.linalg.solve(np.eye(len(w)) - 0.5 * RK_G[s] * dt * k * ops[q], d)
        w = w + d
        a_prev = a
    fl = st.flow(implicit)
    f = st.zeros()
    f["t"] = u
    fl.load(f)
    for s in range(3):
        fl.predict(s, dt)
    got = fl.get("t")
    fl.close()
    want = st.shape(field, np.pad(w, (1, 1)), my=my, mz=mz)
    return _relmax(got[:, :, sl], want[:, :, sl])


def pressure_update(st, implicit, rk=2, dt=0.07, my=1, mz=1, seed=17):
    """P after update_pressure minus P before equals psi - gamma dt nu sum_d L_d psi over the
    implicit directions, with psi = profile(x) cos(y mode) cos(z mode): L_y psi and L_z psi are
    closed-form eigenvalues, L_x psi is div(grad psi) along x with the Neumann wall values"""
    nx = st.nx
    rng = np.random.default_rng(seed)
    v = np.zeros(nx + 2)
    v[1:nx + 1] = rng.uniform(-1.0, 1.0, nx)
    psi = st.shape("p", v, my=my, mz=mz)
    p0 = rng.uniform(-1.0, 1.0, psi.shape)
    p0[:, :, 0] = p0[:, :, 1]
    p0[:, :, -1] = p0[:, :, -2]
    fl = st.flow(implicit)
    f = st.zeros()
    f["p"] = p0
    f["psi"] = psi
    fl.load(f)
    fl.pressure(rk, dt)
    got = fl.get("p") - p0
    fl.close()
    # x: gradient at the faces i = 1..nx+1 with psi(0) = psi(1), psi(nx+1) = psi(nx), then the
    # divergence at the centres i = 1..nx
    vg = v.copy()
    vg[0], vg[nx + 1] = vg[1], vg[nx]
    grad = (vg[1:] - vg[:-1]) / st.hxxf
    flux = st.jdxf / st.hxxf * grad
    lx = np.zeros(nx + 2)
    lx[1:nx + 1] = (flux[1:] - flux[:-1]) / st.jdxc[1:nx + 1]
    lam_y = -4.0 * np.sin(np.pi * my / st.ny) ** 2 / st.hyxc ** 2
    lam_z = -4.0 * np.sin(np.pi * mz / st.nz) ** 2 / st.hz ** 2
    lpsi = np.zeros_like(psi)
    if implicit[0]:
        lpsi += st.shape("p", lx, my=my, mz=mz)
    if implicit[1]:
        lpsi += lam_y[None, None, :] * psi
    if implicit[2]:
        lpsi += lam_z * psi
    want = psi - RK_G[rk] * dt * st.nu * lpsi
    sl = slice(1, nx + 1)
    return _relmax(got[:, :, sl], want[:, :, sl])


def energy_decay(flow_cls, coefs=(0.8, 0.4, 0.2, 0.1), nsteps_coarse=10):
    """the reference's energy example in small: a quarter cylinder (lx = 1, ly = pi/4,
    lz = 2; exec.sh:9-16) at Re = 1e16 on 32 x 16 x 32 cells, from a smooth solenoidal state,
    advanced over the same time with dt = coef dt_cfl for each coef.  Returns the kinetic and
    scalar energy losses 1 - E/E0 per coef."""
    nx, ny, nz = 32, 16, 32
    lx, ly, lz = 1.0, np.pi / 4, 2.0
    xf, xc = icgen.grid(nx, lx, True, 1.5)
    kw = dict(is_curved=True, lengths=(lx, ly, lz), Re=1e16, Pr=1.0, implicit=(0, 0, 0), scalar=1,
              uy_walls=(1.0, 0.0), t_walls=(0.5, -0.5))
    f = icgen.fields(nx, ny, nz, xf, xc, True, seed=3, amp=0.05, t_amp=0.1, nmodes=3, mode_amp=0.3)
    fl = flow_cls(nx, ny, nz, xf, xc, coef_dt_adv=1.0, coef_dt_dif=0.95, **kw)
    fl.load(f)
    for _ in range(5):          # projected onto solenoidal fields
        fl.step()
    dt0 = fl.decide_dt()
    start = {n: fl.get(n) for n in ("ux", "uy", "uz", "p", "t")}
    fl.close()
    span = nsteps_coarse * coefs[0] * dt0
    out = []
    for c in coefs:
        dt = c * dt0
        n = int(round(span / dt))
        fl = flow_cls(nx, ny, nz, xf, xc, dt_fixed=dt, **kw)
        fl.load(start)
        e0 = fl.energy()
        for _ in range(n):
            fl.step()
        e1 = fl.energy()
        fl.close()
        out.append((1.0 - e1[:3].sum() / e0[:3].sum(), 1.0 - e1[3] / e0[3]))
    return np.array(out)


def decay_orders(losses, coefs=(0.8, 0.4, 0.2, 0.1)):
    """observed orders log(loss_a / loss_b) / log(dt_a / dt_b) between neighbouring dt"""
    c = np.array(coefs)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.log(np.abs(losses[:-1] / losses[1:])) / np.log(c[:-1] / c[1:])[:, None]
