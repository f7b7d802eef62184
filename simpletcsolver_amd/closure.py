"""The learned subgrid closure: data from the solver itself, training on the device, coupling.

No reference counterpart (the reference solves the flow on one grid only).  The pieces:

* ``uniform_grid`` -- a radial grid of equal spacing, so that a coarse cell is exactly the
  union of r x r x r fine cells (a box filter of width r);
* ``box_filter`` -- fine fields -> coarse fields, staggered like the solver's variables:
  a face variable is sampled on the coarse faces (every r-th fine face) and averaged over the
  r x r fine faces that make up the coarse face; cell-centred variables are averaged over the
  r^3 fine cells; the x halos (wall values) are averaged over the wall faces;
* ``network_inputs`` / ``centre_targets`` -- the network's cell-centred channels (u, v, w, T, r)
  and the correction at the cell centres (the x face values averaged to the centres, the
  inverse of the device scatter that averages two centres to a face);
* ``generate`` -- snapshots from a fine run: the target correction of a coarse step of size
  D is (filter(u_f(t + D)) - S(filter(u_f(t)))) / D, where the fine solver advances r steps of
  D / r and S is one step of the coarse solver without closure;
* ``train`` -- Adam on minibatches with the library's forward and backward kernels;
* ``attach`` -- copies trained weights into a batch-1 network of the coarse grid and couples
  it: the solver evaluates it at the start of every step and adds the corrections to the
  explicit right-hand sides of ux, uy, uz and T.
"""
import numpy as np

from . import Cnn, Solver, attach_closure, CNN_SHAPES

CORR_FIELDS = ("ux", "uy", "uz", "t")


def uniform_grid(nx, lx, is_curved):
    """faces xf[0..nx] and centres xc[0..nx+1] (xc[0], xc[nx+1] on the walls)"""
    ri = 1.0 if is_curved else 0.0
    xf = ri + lx * np.arange(nx + 1, dtype=np.float64) / nx
    xf[nx] = ri + lx
    xc = np.empty(nx + 2)
    xc[0] = xf[0]
    xc[1:nx + 1] = 0.5 * xf[:-1] + 0.5 * xf[1:]
    xc[nx + 1] = xf[nx]
    return xf, xc


def clustered_grid(nx, lx, is_curved, stretch, r):
    """a tanh-clustered radial grid (tc_ic_grid's) of nx cells whose every r-th face is kept
    by the coarse grid: the fine faces xf[0..nx], centres xc[0..nx+1] and the coarse faces and
    centres made of the unions of r consecutive fine cells"""
    from . import ic_grid
    xf, xc = ic_grid(nx, lx, is_curved, stretch)
    xfc = xf[::r].copy()
    nc = nx // r
    xcc = np.empty(nc + 2)
    xcc[0] = xfc[0]
    xcc[1:nc + 1] = 0.5 * xfc[:-1] + 0.5 * xfc[1:]
    xcc[nc + 1] = xfc[nc]
    return xf, xc, xfc, xcc


def radial_weights(xf, is_curved):
    """volume weight of each fine cell in x (its width, times its centre radius when curved):
    the box filter's weights along x on a non-uniform radial grid"""
    h = xf[1:] - xf[:-1]
    return h * (0.5 * xf[1:] + 0.5 * xf[:-1]) if is_curved else h


def _xblocks(a, r, w):
    """weighted mean over consecutive blocks of r along the last axis (weights w per cell)"""
    n = a.shape[-1]
    if n % r:
        raise ValueError(f"axis of length {n} is not a multiple of {r}")
    ww = w.reshape(n // r, r)
    b = a.reshape(a.shape[:-1] + (n // r, r))
    return (b * ww).sum(axis=-1) / ww.sum(axis=-1)


def _blocks(a, r, axes):
    """mean over consecutive blocks of r along each of axes"""
    shp = []
    for ax, n in enumerate(a.shape):
        if ax in axes:
            if n % r:
                raise ValueError(f"axis {ax} of length {n} is not a multiple of {r}")
            shp += [n // r, r]
        else:
            shp += [n]
    b = a.reshape(shp)
    red = tuple(i + 1 + sum(1 for x in axes if x < i) for i in sorted(axes))
    return b.mean(axis=red)


def box_filter(f, r, wx=None):
    """fine fields (arrays [z][y][x], x halos included, as Solver.fields) -> coarse fields.
    wx: volume weights of the fine cells along x (radial_weights) for a non-uniform radial
    grid; None: equal weights (uniform grid)"""
    out = {}
    for name, a in f.items():
        if name == "ux":
            out[name] = _blocks(a[:, :, ::r], r, (0, 1))
            continue
        inner = a[:, :, 1:-1]
        walls = a[:, :, [0, -1]]
        if wx is not None:
            inner = _xblocks(inner, r, wx)
            xr = 1
        else:
            xr = r
        if name == "uy":
            ci = _blocks(inner[:, ::r, :], r, (0,)) if xr == 1 else _blocks(inner[:, ::r, :], r, (0, 2))
            cw = _blocks(walls[:, ::r, :], r, (0,))
        elif name == "uz":
            ci = _blocks(inner[::r, :, :], r, (1,)) if xr == 1 else _blocks(inner[::r, :, :], r, (1, 2))
            cw = _blocks(walls[::r, :, :], r, (1,))
        else:
            ci = _blocks(inner, r, (0, 1)) if xr == 1 else _blocks(inner, r, (0, 1, 2))
            cw = _blocks(walls, r, (0, 1))
        out[name] = np.concatenate([cw[:, :, :1], ci, cw[:, :, 1:]], axis=2)
    return out


def network_inputs(f, xc):
    """[nz][ny][nx][5]: u at the centres (mean of the two faces), v, w, T, r -- the channels
    the device gathers from the solver's fields"""
    u = 0.5 * f["ux"][:, :, :-1] + 0.5 * f["ux"][:, :, 1:]
    t = f["t"][:, :, 1:-1] if "t" in f else np.zeros_like(u)
    r = np.broadcast_to(xc[1:-1], u.shape)
    return np.stack([u, f["uy"][:, :, 1:-1], f["uz"][:, :, 1:-1], t, r], axis=-1)


def centre_targets(corr):
    """corrections (fields) -> [nz][ny][nx][4] at the cell centres"""
    cx = 0.5 * corr["ux"][:, :, :-1] + 0.5 * corr["ux"][:, :, 1:]
    ct = corr["t"][:, :, 1:-1] if "t" in corr else np.zeros_like(cx)
    return np.stack([cx, corr["uy"][:, :, 1:-1], corr["uz"][:, :, 1:-1], ct], axis=-1)


class Pair:
    """a fine solver and the coarse solver of the same flow (box filter r in every direction)"""

    def __init__(self, nx, ny, nz, r, lengths, Re, Pr=1.0, is_curved=True, scalar=1,
                 implicit=(1, 0, 0), device=0, stretch=0.0):
        self.r = r
        self.nxc, self.nyc, self.nzc = nx // r, ny // r, nz // r
        if (self.nxc * r, self.nyc * r, self.nzc * r) != (nx, ny, nz):
            raise ValueError("the fine grid must be a multiple of the filter width")
        kw = dict(is_curved=is_curved, lengths=lengths, Re=Re, Pr=Pr, implicit=implicit,
                  scalar=scalar, device=device)
        if stretch > 0.0:
            # radial cells clustered toward both walls; a coarse cell is the union of r fine
            # cells, and the filter weights them by volume
            xf, xc, self.xfc, self.xcc = clustered_grid(nx, lengths[0], is_curved, stretch, r)
            self.wx = radial_weights(xf, is_curved)
        else:
            xf, xc = uniform_grid(nx, lengths[0], is_curved)
            self.xfc, self.xcc = uniform_grid(self.nxc, lengths[0], is_curved)
            self.wx = None
        self.xf, self.xc = xf, xc
        self.fine = Solver(nx, ny, nz, xf, xc, **kw)
        self.coarse = Solver(self.nxc, self.nyc, self.nzc, self.xfc, self.xcc, **kw)
        self.scalar = scalar

    def close(self):
        self.fine.close()
        self.coarse.close()

    def fine_state(self):
        f = self.fine.fields()
        f.pop("psi", None)
        if not self.scalar:
            f.pop("t", None)
        return f

    def coarse_step(self, fc, dt, steps=1, torque=None):
        """the coarse solver from fields fc, `steps` steps of fixed dt; its fields after.
        torque: a list that receives the energy injection (wall torque) of the first step"""
        self.coarse.set_dt_params(0.5, 0.5, dt)
        self.coarse.load_fields(fc)
        for n in range(steps):
            if n == 0 and torque is not None:
                self.coarse.arm_diagnostics(True)
            self.coarse.step()
            if n == 0 and torque is not None:
                torque.append(float(self.coarse.injection()[0]))
                self.coarse.arm_diagnostics(False)
        g = self.coarse.fields()
        g.pop("psi", None)
        if not self.scalar:
            g.pop("t", None)
        return g


def generate(pair, n_samples, spacing, dt_coarse, states=None, log=None, torques=None):
    """n_samples (inputs, targets) pairs from the fine run, `spacing` coarse steps apart.
    The fine solver runs at dt_coarse / r; returns x [n][nz][ny][nx][5] and the target
    corrections y [n][nz][ny][nx][4] (FP32).  states: a list that receives the filtered
    (start, end) fields of every sample; torques: a list that receives, per sample, the
    energy injection (wall torque) of the fine step and of the coarse step from the filtered
    field, the data of a torque calibration of the coarse grid's wall stencil."""
    r = pair.r
    pair.fine.set_dt_params(0.5, 0.5, dt_coarse / r)
    xs, ys = [], []
    for m in range(n_samples):
        if log and m % 8 == 0:
            log(f"sample {m} of {n_samples}")
        f0 = box_filter(pair.fine_state(), r, pair.wx)
        for q in range(r):
            if q == 0 and torques is not None:
                pair.fine.arm_diagnostics(True)
            pair.fine.step()
            if q == 0 and torques is not None:
                tf = float(pair.fine.injection()[0])
                pair.fine.arm_diagnostics(False)
        f1 = box_filter(pair.fine_state(), r, pair.wx)
        tc = [] if torques is not None else None
        s1 = pair.coarse_step(f0, dt_coarse, torque=tc)
        if torques is not None:
            torques.append((tf, tc[0]))
        corr = {n: (f1[n] - s1[n]) / dt_coarse for n in CORR_FIELDS if n in f1}
        if states is not None:
            states.append((f0, f1))
        xs.append(network_inputs(f0, pair.xcc).astype(np.float32))
        ys.append(centre_targets(corr).astype(np.float32))
        for _ in range(r * (spacing - 1)):
            pair.fine.step()
    return np.stack(xs), np.stack(ys)


class Normaliser:
    """per-channel input mean / std and output scale of a training set"""

    def __init__(self, x, y):
        self.in_mean = x.reshape(-1, 5).mean(0).astype(np.float32)
        sd = x.reshape(-1, 5).std(0)
        self.in_std = np.where(sd > 0, sd, 1.0).astype(np.float32)
        sc = np.sqrt((y.reshape(-1, 4) ** 2).mean(0))
        self.out_scale = np.where(sc > 0, sc, 1.0).astype(np.float32)

    def x(self, x):
        return ((x - self.in_mean) / self.in_std).astype(np.float32)

    def y(self, y):
        return (y / self.out_scale).astype(np.float32)


def train(x, y, batch, iters, lr=1e-3, seed=1, device=0, log=None):
    """a network trained on (x, y) with Adam on random minibatches; returns (net, normaliser,
    losses)"""
    norm = Normaliser(x, y)
    xn, yn = norm.x(x), norm.y(y)
    n, nz, ny, nx = x.shape[:4]
    batch = min(batch, n)
    net = Cnn(nx, ny, nz, batch=batch, seed=seed, device=device)
    rng = np.random.default_rng(seed)
    losses = []
    for it in range(iters):
        idx = rng.choice(n, size=batch, replace=False)
        losses.append(net.train_step(xn[idx], yn[idx], lr))
        if log and (it % max(1, iters // 10) == 0 or it == iters - 1):
            log(f"iteration {it}: loss {losses[-1]:.4e}")
    return net, norm, np.array(losses)


def evaluate_loss(net, norm, x, y):
    """mean normalised square error of the network on (x, y), in batches of the network"""
    nb = net.shape[0]
    tot, cnt = 0.0, 0
    for s in range(0, len(x) - nb + 1, nb):
        pred = net.predict(norm.x(x[s:s + nb]))
        d = pred - norm.y(y[s:s + nb])
        tot += float((d.astype(np.float64) ** 2).sum())
        cnt += d.size
    return tot / max(cnt, 1)


def model_power(solver, net1, norm, xf, xc, lengths, is_curved=True):
    """the kinetic-energy input rate of the coupled closure at the solver's current fields:
    sum over the cells of J u . c, with c the corrections the solver adds to du/dt (ux at the
    faces i = 2..nx as the mean of the two neighbouring centres, as the device scatter forms
    them), in the units of the solver's injection and dissipation (energy per unit time, the
    J-weighted sums of src/logging/*.c).  Negative: the closure removes energy, as a subgrid
    dissipation does.  net1: the batch-1 network coupled to the solver; xf, xc: the solver's
    radial faces and centres"""
    f = solver.fields()
    xf, xc = np.asarray(xf), np.asarray(xc)
    nz, ny, nx = f["uy"].shape[0], f["uy"].shape[1], f["uy"].shape[2] - 2
    y = net1.predict(norm.x(network_inputs(f, xc)[None]))[0] * norm.out_scale
    ly, lz = lengths[1], lengths[2]
    hz = lz / nz
    hxxf = xc[1:] - xc[:-1]                        # HXXF(i), i = 1..nx+1
    hxxc = xf[1:] - xf[:-1]                        # HXXC(i), i = 1..nx
    hyxf = xf * ly / ny if is_curved else np.full(nx + 1, ly / ny)
    hyxc = xc[1:-1] * ly / ny if is_curved else np.full(nx, ly / ny)
    jf = hxxf * hyxf * hz                          # JDXF(i), i = 1..nx+1
    jc = hxxc * hyxc * hz                          # JDXC(i), i = 1..nx
    cx = 0.5 * y[:, :, :-1, 0] + 0.5 * y[:, :, 1:, 0]    # faces i = 2..nx
    p = (f["ux"][:, :, 1:nx] * cx * jf[1:nx]).sum()
    p += (f["uy"][:, :, 1:-1] * y[..., 1] * jc).sum()
    p += (f["uz"][:, :, 1:-1] * y[..., 2] * jc).sum()
    return float(p)


def attach(solver, net, norm, device=0):
    """couple a trained network (any batch size) to a solver of the network's grid; returns the
    batch-1 network now coupled (keep it alive while the solver steps)"""
    one = Cnn(solver.nx, solver.ny, solver.nz, batch=1, device=device)
    for l in range(len(CNN_SHAPES)):
        w, b, _, _ = net.layer(l)
        one.set_layer(l, w, b)
    attach_closure(solver, one, norm.in_mean, norm.in_std, norm.out_scale)
    return one


def detach(solver):
    attach_closure(solver, None, np.zeros(5), np.ones(5), np.ones(4))


def step_error(pair, states, dt_coarse):
    """relative L2 error, over ux, uy, uz and T together, of one coarse step from each
    filtered start state against the filtered fine state one step later"""
    num = den = 0.0
    for f0, f1 in states:
        s1 = pair.coarse_step(f0, dt_coarse)
        for n in CORR_FIELDS:
            if n in f1:
                num += float(((s1[n] - f1[n]) ** 2).sum())
                den += float(((f1[n] - f0[n]) ** 2).sum())
    return np.sqrt(num / den)


def _main():
    """time (or profile) the coarse step with an attached network of seeded random weights"""
    import argparse
    import time
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--grid", type=int, nargs=3, default=(32, 64, 64))
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--plain", action="store_true", help="without the network")
    args = ap.parse_args()
    nx, ny, nz = args.grid
    lx = 1.0 / 0.714 - 1.0
    xf, xc = uniform_grid(nx, lx, True)
    s = Solver(nx, ny, nz, xf, xc, is_curved=True, lengths=(lx, np.pi / 3.0, 2.0 * np.pi * lx),
               Re=3000.0, implicit=(1, 0, 0), scalar=1, dt_fixed=2e-3)
    s.ic(seed=1, amp=0.02, t_amp=0.02, nmodes=4, mode_amp=0.1)
    net = None
    if not args.plain:
        net = Cnn(nx, ny, nz, batch=1, seed=1)
        attach_closure(s, net, np.zeros(5), np.ones(5), 1e-3 * np.ones(4))
    s.step_async(20)
    s.synchronize()
    t0 = time.perf_counter()
    s.step_async(args.steps)
    s.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    print(f"{nx}x{ny}x{nz} {'plain' if args.plain else 'with closure'}: {ms:.3f} ms/step", flush=True)
    s.close()
    if net is not None:
        net.close()


if __name__ == "__main__":
    _main()
