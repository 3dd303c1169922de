"""Learned denoiser on the GPU (BF16 MFMA kernels, through the C ABI) against the numpy oracle
that rounds activations and gradients to BF16 where the kernels store them.  The remaining
differences come from the FP32 summation order (MFMA vs BLAS) and from BF16 values that land on
the other side of a rounding boundary because of it; the bars are stated per check.
"""
import numpy as np
import pytest

import denoiser_ref as R
import protoray_hip as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dn():
    d = P.DenoiserHip(0)
    yield d
    d.close()


def _batch(rng, n, h, w):
    color = rng.gamma(1.5, 0.4, (n, h, w, 3)).astype(np.float32)
    albedo = rng.random((n, h, w, 3)).astype(np.float32)
    normal = rng.standard_normal((n, h, w, 3)).astype(np.float32)
    normal /= np.linalg.norm(normal, axis=-1, keepdims=True)
    target = (color * 0.7 + 0.1).astype(np.float32)
    return np.concatenate([color, albedo, normal], -1), target


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_initial_weights(dn):
    assert dn.n_params == R.param_count()
    assert np.array_equal(dn.getWeights(), R.he_init())
    print(dn.getVersion())


def test_forward_matches_oracle(dn):
    rng = np.random.default_rng(1)
    dn.setWeights(R.he_init() + rng.standard_normal(R.param_count()).astype(np.float32) * 0.01)
    inp, _ = _batch(rng, 1, 48, 64)
    out = dn.denoise(inp[0, ..., :3], inp[0, ..., 3:6], inp[0, ..., 6:9])
    s = R.Net(dn.getWeights(), bf16_mode=True).forward(inp)
    ref = np.expm1(s["pred"][0])
    err = _rel(out, ref)
    print("forward relative L2 error", err)
    assert err < 5e-3
    assert np.mean(np.abs(out - ref) <= 1e-2 * np.abs(ref) + 1e-3) > 0.99


def test_gradient_matches_oracle(dn):
    rng = np.random.default_rng(2)
    dn.setWeights(R.he_init() + rng.standard_normal(R.param_count()).astype(np.float32) * 0.01)
    inp, tgt = _batch(rng, 2, 32, 32)
    loss, g = dn.gradient(inp, tgt)
    rloss, rg, _ = R.Net(dn.getWeights(), bf16_mode=True).loss_grad(inp, tgt)
    assert abs(loss - rloss) <= 1e-3 * rloss
    off = 0
    for i, (ci, co) in enumerate(R.LAYERS):
        nw = co * ci * 9
        ew, eb = _rel(g[off:off + nw], rg[off:off + nw]), _rel(g[off + nw:off + nw + co], rg[off + nw:off + nw + co])
        print(f"layer {i}: weight-gradient relative L2 error {ew:.2e}, bias {eb:.2e}")
        assert ew < 3e-2 and eb < 3e-2
        off += nw + co


def test_training_reduces_the_loss(dn):
    rng = np.random.default_rng(3)
    dn.setWeights(R.he_init())
    inp, tgt = _batch(rng, 4, 64, 64)
    losses = [dn.trainStep(inp, tgt, 2e-3)[0] for _ in range(60)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]
