"""The closure network on the GPU against its numpy statement (tests/cnn_ref.py, pinned by
finite differences in test_cnn_oracle.py): forward outputs and every weight and bias gradient
of the hand-written BF16-MFMA kernels, with the same BF16 roundings, to FP32 accumulation
order.  Training reduces the loss on a fixed batch."""
import numpy as np
import pytest

import cnn_ref as cr

pytestmark = pytest.mark.gpu


def _setup(nb, nz, ny, nx, seed):
    from simpletcsolver_amd import Cnn
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nb, nz, ny, nx, 5)).astype(np.float32)
    y = (0.3 * rng.standard_normal((nb, nz, ny, nx, 4))).astype(np.float32)
    params = cr.init_params(seed)
    net = Cnn(nx, ny, nz, batch=nb, seed=seed)
    for l, (w, b) in enumerate(params):
        net.set_layer(l, w, b)
    return net, params, x, y


@pytest.mark.parametrize("shape", [(2, 4, 6, 16), (1, 3, 8, 21)])
def test_forward_and_gradients_match_numpy(need_gpu, shape):
    nb, nz, ny, nx = shape
    net, params, x, y = _setup(nb, nz, ny, nx, 7)
    out = net.predict(x)
    loss_ref, grads_ref, f = cr.loss_and_grads(params, x, y, bf=True)
    e = np.linalg.norm(out - f["out"][..., :4]) / np.linalg.norm(f["out"][..., :4])
    print(f"forward {shape}: {e:.2e}")
    # the hidden activations are rounded to BF16 from FP32 sums whose order differs from the
    # statement's exact sums: a pre-activation that falls within an FP32 rounding of a BF16
    # rounding boundary lands on the neighbouring BF16 value (2^-8 relative), which moves
    # the outputs at the 1e-5 level
    assert e <= 2e-4
    loss = net.gradients(x, y)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    for l in range(3):
        _, _, gw, gb = net.layer(l)
        rw, rb = grads_ref[l]
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(gb - rb) / max(np.linalg.norm(rb), 1e-30)
        print(f"layer {l}: weight gradient {ew:.2e}, bias gradient {eb:.2e}")
        assert ew <= 2e-3 and eb <= 2e-3
    net.close()


def test_training_reduces_loss(need_gpu):
    net, params, x, y = _setup(2, 4, 8, 16, 11)
    first = net.train_step(x, y, 1e-3)
    for _ in range(60):
        last = net.train_step(x, y, 1e-3)
    print(f"loss {first:.4e} -> {last:.4e}")
    assert last < 0.5 * first
    net.close()
