"""The denoiser oracle (numpy U-Net) pinned before it checks the HIP kernels.

* Its convolution equals a direct 3x3 correlation with zero padding (loops over taps).
* Its backward pass equals central differences of its own float64 loss (h = 1e-6) for weights
  and biases of every layer.
* Pooling routes to the first maximum; the BF16 rounding is round-to-nearest-even.
* Its initial parameters equal the HIP denoiser's (same LCG, He bounds), so both start equal.
"""
import numpy as np
import pytest

import denoiser_ref as R


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-39], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 1.5 * 2^-8 rounds up
    np.testing.assert_array_equal(R.bf16(x)[:4], np.array([1.0, 1.0, 1.0078125, 1.0078125], np.float32))
    assert R.bf16(x)[4] == -2.5


def test_conv_equals_direct_correlation():
    rng = np.random.default_rng(0)
    params = rng.standard_normal(R.param_count()) * 0.1
    net = R.Net(params, bf16_mode=False)
    x = rng.standard_normal((2, 8, 12, 9))
    y = net.conv(0, x, relu=False)
    w, b = R.unpack(params)[0]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    ref = np.zeros((2, 8, 12, 32)) + b
    for dy in range(3):
        for dx in range(3):
            ref += np.einsum("nhwc,oc->nhwo", xp[:, dy:dy + 8, dx:dx + 12], w[:, :, dy, dx])
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_pool_backward_first_maximum():
    x = np.array([[1, 3], [3, 0]], np.float64).reshape(1, 2, 2, 1)
    g = R.Net.pool_bwd(x, np.array([[[[5.0]]]]))
    assert g.reshape(2, 2).tolist() == [[0, 5], [0, 0]]


@pytest.mark.parametrize("layer", range(10))
def test_backward_matches_central_differences(layer):
    rng = np.random.default_rng(layer)
    params = R.he_init().astype(np.float64)
    params += rng.standard_normal(params.size) * 0.01  # non-zero biases
    inp = np.concatenate([rng.random((1, 16, 16, 3)) * 2, rng.random((1, 16, 16, 3)), rng.standard_normal((1, 16, 16, 3))], -1)
    tgt = rng.random((1, 16, 16, 3)) * 2
    loss, g, _ = R.Net(params, bf16_mode=False).loss_grad(inp, tgt)
    off = sum(co * ci * 9 + co for ci, co in R.LAYERS[:layer])
    ci, co = R.LAYERS[layer]
    idx = list(rng.choice(co * ci * 9, 4, replace=False) + off) + [off + co * ci * 9 + int(rng.integers(co))]
    h = 1e-6
    for k in idx:
        p, m = params.copy(), params.copy()
        p[k] += h
        m[k] -= h
        fd = (R.Net(p, False).loss_grad(inp, tgt)[0] - R.Net(m, False).loss_grad(inp, tgt)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(abs(g[k]), 1e-3), (layer, k, fd, g[k])


def test_initial_parameters_match_the_hip_layout():
    p = R.he_init()
    assert p.size == R.param_count() == 567_936 + 611  # weights + biases
    w0, b0 = R.unpack(p)[0]
    assert np.abs(w0).max() <= np.sqrt(6 / 81) and np.all(b0 == 0)
