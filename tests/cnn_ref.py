"""numpy statement of the closure network (simpletcsolver_amd/csrc/kernels/tck_cnn.hip), the
checker of the device forward and backward passes.

Cells are channels-last arrays [batch, nz, ny, nx, C]; a 3x3x3 convolution reads tap
(dx, dy, dz) = (t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) with zero padding in x and periodic
wrapping in y and z; weights are [27, co, ci].  With bf=True the operands are rounded to
BF16 where the device rounds them (inputs, weights, hidden activations, and the gradients that
enter products); sums are exact (float64), the device's are FP32.
"""
import numpy as np

CI0, C, CO = 16, 32, 16          # padded input, hidden and padded output channels
SHAPES = [(27, C, CI0), (27, C, C), (27, CO, C)]


def bf16(x):
    """round to the nearest BF16 (ties to even), returned as float32"""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def shift(x, t):
    """x at cell + tap t (zero outside in x, periodic in y and z)"""
    dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
    y = np.roll(x, shift=(-dz, -dy), axis=(1, 2))
    out = np.zeros_like(y)
    nx = x.shape[3]
    if dx == 0:
        out[...] = y
    elif dx == 1:
        out[:, :, :, :nx - 1] = y[:, :, :, 1:]
    else:
        out[:, :, :, 1:] = y[:, :, :, :nx - 1]
    return out


def conv(x, w, b=None):
    out = np.zeros(x.shape[:4] + (w.shape[1],), dtype=np.float64)
    for t in range(27):
        out += np.tensordot(shift(x, t).astype(np.float64), w[t].astype(np.float64).T, axes=([4], [0]))
    if b is not None:
        out += b
    return out


def wgrad(g, x):
    """d/dW of sum g . conv(x, W): [27, co, ci]"""
    gf = g.reshape(-1, g.shape[-1]).astype(np.float64)
    return np.stack([gf.T @ shift(x, t).reshape(-1, x.shape[-1]).astype(np.float64) for t in range(27)])


def flip(w):
    """the adjoint kernel: wt[26 - t][ci][co] = w[t][co][ci]"""
    return np.ascontiguousarray(np.transpose(w[::-1], (0, 2, 1)))


def pad_in(x, dtype):
    out = np.zeros(x.shape[:4] + (CI0,), dtype=dtype)
    out[..., :x.shape[-1]] = x
    return out


def _rounder(bf):
    """bf: BF16 rounding of operands (as the device); otherwise float64 throughout"""
    return (lambda a: bf16(np.asarray(a, dtype=np.float32))) if bf else (lambda a: np.asarray(a, dtype=np.float64))


def forward(params, x, bf=True):
    r = _rounder(bf)
    (w1, b1), (w2, b2), (w3, b3) = params
    x0 = r(pad_in(x, np.float32 if bf else np.float64))
    z1 = conv(x0, r(w1), b1)
    a1 = r(np.maximum(z1, 0.0))
    z2 = conv(a1, r(w2), b2)
    a2 = r(np.maximum(z2, 0.0))
    out = conv(a2, r(w3), b3)
    return dict(x0=x0, a1=a1, a2=a2, out=out)


def loss_and_grads(params, x, y, bf=True):
    """mean square over cells and the 4 used outputs; gradients of every weight and bias"""
    r = _rounder(bf)
    f = forward(params, x, bf)
    (w1, b1), (w2, b2), (w3, b3) = params
    n = y.size
    d = f["out"][..., :4] - y
    loss = float((d * d).sum() / n)
    gout = np.zeros_like(f["out"])
    gout[..., :4] = 2.0 * d / n
    g3 = r(gout)
    gw3, gb3 = wgrad(g3, f["a2"]), gout.reshape(-1, CO).sum(0)
    d2 = conv(g3, flip(r(w3)))
    g2f = d2 * (f["a2"] > 0)
    g2 = r(g2f)
    gw2, gb2 = wgrad(g2, f["a1"]), g2f.reshape(-1, C).sum(0)
    d1 = conv(g2, flip(r(w2)))
    g1f = d1 * (f["a1"] > 0)
    g1 = r(g1f)
    gw1, gb1 = wgrad(g1, f["x0"]), g1f.reshape(-1, C).sum(0)
    return loss, [(gw1, gb1), (gw2, gb2), (gw3, gb3)], f


def init_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for l, shp in enumerate(SHAPES):
        w = (rng.standard_normal(shp) * np.sqrt(2.0 / (27 * (5 if l == 0 else C))) * scale).astype(np.float32)
        if l == 0:
            w[:, :, 5:] = 0.0
        if l == 2:
            w[:, 4:, :] = 0.0
        b = (0.1 * rng.standard_normal(shp[1])).astype(np.float32)
        if l == 2:
            b[4:] = 0.0
        out.append((w, b))
    return out
