"""Pins of the learned-preconditioner oracle (oracle/cnn_oracle.cpp).

The network has no reference counterpart; the oracle is pinned by
mathematics instead:
- forward: a network wired to pass one input channel through the hidden
  layers and correlate it with a known 3x3 kernel in the last layer equals
  numpy's zero-padded cross-correlation (tap orientation, padding, channel
  order, bias);
- backward: finite differences of the FP64 loss along random directions
  (a planted mirrored-tap mistake in the backward data pass fails this);
- the BF16-rounding mode stays within BF16 precision of the FP64 mode.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
CIN = [3, 32, 32, 32, 32]
COUT = [32, 32, 32, 32, 1]


def offsets():
    off, out = 0, []
    for ci, co in zip(CIN, COUT):
        out.append((off, off + co * ci * 9))  # weights start, biases start
        off += co * ci * 9 + co
    return out, off


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    offs, total = offsets()
    p = np.zeros(total, np.float32)
    for l, (w0, b0) in enumerate(offs):
        sd = np.sqrt(2.0 / (9 * CIN[l])) * scale
        p[w0:b0] = rng.normal(0, sd, b0 - w0)
        p[b0:b0 + COUT[l]] = rng.normal(0, 0.05, COUT[l])
    return p


def test_parameter_count():
    assert orc.cnn_num_params() == 28929 == offsets()[1]


def correlate_same(img, k):
    H, W = img.shape
    pad = np.pad(img, 1)
    out = np.zeros_like(img, dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            out += k[ky, kx] * pad[ky:ky + H, kx:kx + W]
    return out


@pytest.mark.parametrize("channel", [0, 2])
def test_forward_closed_form(channel):
    """Channel `channel` passed through layers 1-4 (centre taps), then a 3x3
    correlation with kernel K and bias beta: y = beta + K * relu(x_c)."""
    offs, total = offsets()
    p = np.zeros(total, np.float32)
    (w0, _b0) = offs[0]
    p[w0 + (0 * CIN[0] + channel) * 9 + 4] = 1.0
    for l in (1, 2, 3):
        w0, _ = offs[l]
        p[w0 + (0 * CIN[l] + 0) * 9 + 4] = 1.0
    K = np.arange(1, 10, dtype=np.float32).reshape(3, 3) / 10.0
    w0, b0 = offs[4]
    p[w0:w0 + 9] = K.ravel()  # o = 0, c = 0
    p[b0] = 0.25
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 7, 9, 3)).astype(np.float32)
    y = orc.cnn_forward(p, x, mode=0)
    for b in range(2):
        want = 0.25 + correlate_same(np.maximum(x[b, :, :, channel], 0.0).astype(np.float64), K)
        np.testing.assert_allclose(y[b], want, rtol=1e-6, atol=1e-6)


def test_backward_matches_finite_differences():
    _fd_check(orc)


def _fd_check(o):
    """Directional finite differences of the FP64 loss against the gradient."""
    p = random_params(11)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2, 6, 10, 3)).astype(np.float32)
    t = rng.standard_normal((2, 6, 10)).astype(np.float32)
    mask = [1, 9, 1, 5]
    _, g = o.cnn_loss(p, x, t, mask, mode=0)
    ratios = []
    for k in range(6):
        v = rng.standard_normal(p.size)
        v *= 1e-5 / np.linalg.norm(v)  # small enough that no ReLU changes side
        pp = (p + v).astype(np.float32)
        pm = (p - v).astype(np.float32)
        dv = pp.astype(np.float64) - pm.astype(np.float64)
        lp = o.cnn_loss(pp, x, t, mask, mode=0, grad=False)
        lm = o.cnn_loss(pm, x, t, mask, mode=0, grad=False)
        ratios.append((lp - lm) / float(g @ dv))
    ratios = np.array(ratios)
    # at most one direction may cross a ReLU kink
    assert np.sum(np.abs(ratios - 1.0) < 1e-5) >= 5, ratios


def test_backward_is_the_loss_gradient():
    """cnn_backward with dy = dL/dy equals the loss gradient."""
    p = random_params(21)
    rng = np.random.default_rng(22)
    x = rng.standard_normal((1, 5, 8, 3)).astype(np.float32)
    t = rng.standard_normal((1, 5, 8)).astype(np.float32)
    mask = [0, 8, 0, 5]
    _, g = orc.cnn_loss(p, x, t, mask, mode=0)
    y = orc.cnn_forward(p, x, mode=0)
    dy = ((y.astype(np.float64) - t) / 40.0).astype(np.float32)
    g2 = orc.cnn_backward(p, x, dy, mode=0)
    assert np.linalg.norm(g - g2) / np.linalg.norm(g) < 1e-5


def test_bf16_mode_within_bf16_precision():
    p = random_params(31)
    rng = np.random.default_rng(32)
    x = rng.standard_normal((2, 8, 16, 3)).astype(np.float32)
    t = rng.standard_normal((2, 8, 16)).astype(np.float32)
    y0 = orc.cnn_forward(p, x, 0)
    y1 = orc.cnn_forward(p, x, 1)
    assert np.linalg.norm(y1 - y0) / np.linalg.norm(y0) < 2e-2
    _, g0 = orc.cnn_loss(p, x, t, [2, 14, 1, 7], 0)
    _, g1 = orc.cnn_loss(p, x, t, [2, 14, 1, 7], 1)
    # four BF16-rounded back-propagation stages and ReLU gates flipped by the
    # rounded activations: about 8 % on this random network
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 0.15


_MUTANT_SCRIPT = r"""
import sys
sys.path[:0] = [{tests!r}, {oracle!r}]
import oracle as o
import test_cnn_oracle as t
try:
    t._fd_check(o)
    print("passed")
except AssertionError:
    print("failed")
"""


@pytest.mark.slow
def test_planted_mirror_mistake_is_caught():
    src = [os.path.join(orc.HERE, f) for f in ("tomo_oracle.cpp", "cnn_oracle.cpp")]
    with tempfile.TemporaryDirectory() as td:
        so = os.path.join(td, "liboracle_mut.so")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                        "-DORC_MUTATION=5", "-o", so, *src], check=True)
        script = _MUTANT_SCRIPT.format(tests=HERE, oracle=orc.HERE)
        env = dict(os.environ, TOMO_ORACLE_LIB=so)
        out = subprocess.run([sys.executable, "-c", script], env=env, check=True, capture_output=True,
                             text=True).stdout
    assert out.strip().splitlines()[-1] == "failed"
