"""Pins of the network oracle (oracle/nn_oracle.py), CPU only:
convolution against scipy.ndimage.correlate (a textbook routine), pooling and
upsampling adjointness, BF16 rounding against known bit patterns, and the full
backward pass against central finite differences of the exact (FP64) network."""
import numpy as np
import pytest
from scipy import ndimage

import nn_oracle as N


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = N.bf16(x)
    # ties to even: 1 + 2^-8 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -7) and r[3] == np.float32(1.0 + 2 ** -6)
    assert r[4] == -2.5
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 5, 6, 7))
    W = rng.normal(size=(4, 3, 27))
    b = rng.normal(size=4)
    y = N.conv(x, W, b)
    for o in range(4):
        ref = np.full(x.shape[1:], b[o])
        for c in range(3):
            ref = ref + ndimage.correlate(x[c], W[o, c].reshape(3, 3, 3), mode="constant", cval=0.0)
        assert np.abs(y[o] - ref).max() < 1e-12


def test_pool_up_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 6, 8))
    y = rng.normal(size=(2, 2, 3, 4))
    assert abs(np.sum(N.pool(x) * y) - np.sum(x * N.pool_backward(y))) < 1e-12
    z = rng.normal(size=(2, 4, 6, 8))
    assert abs(np.sum(N.up(y) * z) - np.sum(y * N.up_backward(z))) < 1e-12


def test_conv_backward_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 4, 5, 6))
    W = rng.normal(size=(2, 3, 27))
    dy = rng.normal(size=(2, 4, 5, 6))
    dx, dW, db = N.conv_backward(x, W, dy)
    # <dy, conv(x)> is linear in x: its x-gradient applied to x gives <dy, conv(x, W, 0)>
    assert abs(np.sum(dx * x) - np.sum(dy * N.conv(x, W, np.zeros(2)))) < 1e-10
    assert abs(np.sum(dW * W) - np.sum(dy * N.conv(x, W, np.zeros(2)))) < 1e-10
    assert np.allclose(db, dy.sum(axis=(1, 2, 3)))


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    c = (2, 3, 4)
    n = 8
    inp = rng.normal(size=(2, n, n, n))
    theta = N.init_params(2, c, seed).astype(np.float64)
    theta[-1 - 0] = 0.05  # a nonzero output bias
    theta = theta + 0.02 * rng.normal(size=theta.shape)  # nonzero biases everywhere
    dout = rng.normal(size=(1, n, n, n))
    out, cache = N.forward(inp, theta, c, round_bf16=False)
    g, gin = N.backward(dout, cache)
    for k in range(4):
        v = rng.normal(size=theta.shape)
        h = 1e-6
        lp = np.sum(dout * N.forward(inp, theta + h * v, c, round_bf16=False)[0])
        lm = np.sum(dout * N.forward(inp, theta - h * v, c, round_bf16=False)[0])
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g @ v)
    v = rng.normal(size=inp.shape)
    h = 1e-6
    fd = (np.sum(dout * N.forward(inp + h * v, theta, c, round_bf16=False)[0]) -
          np.sum(dout * N.forward(inp - h * v, theta, c, round_bf16=False)[0])) / (2 * h)
    assert abs(fd - np.sum(gin * v)) <= 1e-6 * max(1.0, abs(fd))


def test_param_count():
    assert N.num_params(2, (8, 16, 24)) == 27 * 2200 + (8 + 8 + 16 + 16 + 24 + 24 + 16 + 16 + 8 + 1)
