"""GPU parity of the learned-preconditioner kernels (BF16 MFMA convolutions,
backward data / weight passes, loss, Adam) against the CPU oracle in its
BF16-rounding mode, which rounds at the same points as the GPU and sums
exactly.  The remaining difference is FP32 accumulation on the matrix cores
flipping an occasional BF16 rounding, so the tolerances are tight:

  forward output      relative L2 <= 2e-3
  parameter gradient  relative L2 <= 1e-2 per layer
  loss                relative    <= 1e-4
"""
import numpy as np
import pytest

import oracle as orc
from _pkg import tomo
from test_cnn_oracle import offsets, random_params

pytestmark = pytest.mark.gpu

Y_TOL = 2e-3
G_TOL = 1e-2
L_TOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def inputs(seed, B, H, W):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, H, W, 3)).astype(np.float32)


@pytest.fixture(scope="module")
def net():
    with tomo.Preconditioner(device=0) as n:
        yield n


@pytest.mark.parametrize("shape", [(1, 16, 32), (2, 24, 64), (3, 7, 96)])
def test_forward_matches_oracle(net, shape):
    B, H, W = shape
    print(f"[case] cnn forward {shape}", flush=True)
    p = random_params(100 + H)
    x = inputs(H, B, H, W)
    net.set_params(p)
    y = net.forward(x)
    want = orc.cnn_forward(p, x, mode=1)
    assert rel(y, want) <= Y_TOL, rel(y, want)


def test_forward_closed_form_on_gpu(net):
    """Delta-wired network: y = beta + K (*) relu(x_coverage) (zero padded)."""
    print("[case] cnn closed form", flush=True)
    offs, total = offsets()
    p = np.zeros(total, np.float32)
    p[offs[0][0] + 2 * 9 + 4] = 1.0
    for l in (1, 2, 3):
        p[offs[l][0] + 4] = 1.0
    K = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float32) / 8.0
    p[offs[4][0]:offs[4][0] + 9] = K.ravel()
    p[offs[4][1]] = -0.5
    x = inputs(9, 2, 12, 32)
    x[..., 2] = np.round(x[..., 2] * 8) / 8  # exact in bfloat16
    net.set_params(p)
    y = net.forward(x)
    for b in range(2):
        img = np.pad(np.maximum(x[b, :, :, 2], 0).astype(np.float64), 1)
        want = -0.5 + sum(K[ky, kx] * img[ky:ky + 12, kx:kx + 32] for ky in range(3) for kx in range(3))
        np.testing.assert_allclose(y[b], want, rtol=0, atol=1e-5)


@pytest.mark.parametrize("shape", [(1, 16, 32), (2, 9, 64)])
def test_backward_matches_oracle(net, shape):
    B, H, W = shape
    print(f"[case] cnn backward {shape}", flush=True)
    p = random_params(200 + H)
    x = inputs(300 + H, B, H, W)
    dy = np.random.default_rng(400 + H).standard_normal((B, H, W)).astype(np.float32)
    net.set_params(p)
    g = net.backward(x, dy)
    want = orc.cnn_backward(p, x, dy, mode=1)
    offs, total = offsets()
    bounds = [o[0] for o in offs] + [total]
    for l in range(5):
        sl = slice(bounds[l], bounds[l + 1])
        assert rel(g[sl], want[sl]) <= G_TOL, (l, rel(g[sl], want[sl]))


def test_loss_and_gradient_match_oracle(net):
    print("[case] cnn loss", flush=True)
    B, H, W = 2, 20, 64
    p = random_params(7)
    x = inputs(8, B, H, W)
    t = np.random.default_rng(9).standard_normal((B, H, W)).astype(np.float32)
    mask = [8, 56, 3, 17]
    net.set_params(p)
    L, g = net.loss(x, t, mask)
    Lw, gw = orc.cnn_loss(p, x, t, mask, mode=1)
    assert abs(L - Lw) / Lw <= L_TOL
    assert rel(g, gw) <= G_TOL


def test_backward_is_deterministic(net):
    print("[case] cnn determinism", flush=True)
    p = random_params(3)
    x = inputs(4, 2, 32, 64)
    dy = np.random.default_rng(5).standard_normal((2, 32, 64)).astype(np.float32)
    net.set_params(p)
    g1 = net.backward(x, dy)
    g2 = net.backward(x, dy)
    np.testing.assert_array_equal(g1, g2)


def test_bad_shapes_rejected(net):
    with pytest.raises(tomo.TomoError):
        net.forward(np.zeros((1, 8, 30, 3), np.float32))  # width not a multiple of 32
    with pytest.raises(tomo.TomoError):
        net.loss(np.zeros((1, 8, 32, 3), np.float32), np.zeros((1, 8, 32), np.float32), [0, 40, 0, 8])
