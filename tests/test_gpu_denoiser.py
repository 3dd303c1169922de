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


@pytest.mark.parametrize("inputs", ["full", "color_albedo", "color"])
def test_trained_network_beats_the_noisy_render(dn, inputs):
    # the weights trained on rendered data for each input set (Denoiser::Params with and without the
    # albedo and normal images; oidn-protoray_amd/weights/unet_sunsky*.npy) on a scene and camera
    # outside the training set: 4 samples, denoised vs not, against 256 samples
    import scenegen as G
    import dataset as D
    import denoise_train as T
    import os
    if not os.path.exists(T.weights_path(inputs)):
        pytest.skip(f"no weights trained for the {inputs} input set")
    dn.setWeights(np.load(T.weights_path(inputs)))
    sc = G.spheres_sunsky(seed=77, n_spheres=300, subdiv=3, palette=64, coated=0.3, width=256, height=192)
    dev = P.DeviceHip(0)
    P.load_scene(dev, sc)
    dev.initCamera(D.look_at((2, 2.5, -1), (-1, 1, 16), 50))
    dev.initFrame(256, 192, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
    dev.initSampler(3)
    dev.clearFrame()
    dev.render(256)
    ref = dev.blitFrame()
    dev.initSampler(0)
    dev.clearFrame()
    dev.render(4)
    noisy, alb, nrm = dev.blitFrame(), dev.blitFrame(P.BUF_ALBEDO1), dev.blitFrame(P.BUF_NORMAL1)
    dev.close()
    out = dn.denoise(noisy, alb if inputs != "color" else None, nrm if inputs == "full" else None)
    e_noisy, e_out = G.relmse(noisy, ref), G.relmse(out, ref)
    print(f"relMSE vs 256 spp: 4 spp {e_noisy:.4f}, denoised 4 spp {e_out:.4f}")
    assert e_out < 0.25 * e_noisy


@pytest.mark.parametrize("inputs", [(True, True), (True, False), (False, False)])
def test_execute_on_the_rendered_frame_equals_host_execute(dn, inputs):
    # prt_denoiser_execute_frame reads the frame's sums and weights in HBM; its output must equal
    # prt_denoiser_execute on the blitted images bit for bit (same means, same edge padding: the
    # height 120 is not a multiple of 16)
    import scenegen as G
    alb_on, nrm_on = inputs
    rng = np.random.default_rng(5)
    dn.setWeights(R.he_init() + rng.standard_normal(R.param_count()).astype(np.float32) * 0.01)
    sc = G.spheres_sunsky(n_spheres=40, subdiv=2, width=224, height=120)
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, sc)
        dev.initFrame(224, 120, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
        dev.initSampler(0)
        dev.clearFrame()
        dev.render(3)
        img, alb, nrm = dev.blitFrame(), dev.blitFrame(P.BUF_ALBEDO1), dev.blitFrame(P.BUF_NORMAL1)
        host = dn.denoise(img, alb if alb_on else None, nrm if nrm_on else None)
        secs = dn.executeFrame(dev, albedo=alb_on, normal=nrm_on)
        out = dn.output()
    finally:
        dev.close()
    assert secs > 0 and dn.kernelMs() > 0
    assert out.shape == host.shape
    assert np.array_equal(out, host)


def test_execute_on_a_batch_frame_equals_host_execute(dn):
    import scenegen as G
    rng = np.random.default_rng(6)
    dn.setWeights(R.he_init() + rng.standard_normal(R.param_count()).astype(np.float32) * 0.01)
    sc = G.spheres_sunsky(n_spheres=40, subdiv=2, width=160, height=96)
    dev = P.DeviceHip(0)
    batch = P.BatchHip(0)
    try:
        P.load_scene(dev, sc)
        batch.addFrame(dev, sc["camera"], 160, 96, 0, 8, snapshots=(4,), aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1)
        batch.render()
        for buf in (P.BATCH_COLOR, P.BATCH_SNAPSHOT):
            img, alb, nrm = batch.get(0, buf), batch.get(0, P.BATCH_ALBEDO1), batch.get(0, P.BATCH_NORMAL1)
            host = dn.denoise(img, alb, nrm)
            dn.executeBatch(batch, 0, buf)
            assert np.array_equal(dn.output(), host)
        with pytest.raises(ValueError):
            dn.executeBatch(batch, 0, P.BATCH_ALBEDO1)
    finally:
        batch.close()
        dev.close()


def test_inference_at_1080p(dn):
    # the learned-denoiser workload's inference size: one 1920 x 1080 frame (padded to 1088 rows),
    # denoised where the renderer left it; equals the host path, and its kernel time is reported
    import scenegen as G
    dn.setWeights(np.load(__import__("denoise_train").weights_path("full")))
    sc = G.spheres_sunsky(n_spheres=64, subdiv=2, width=1920, height=1080)
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, sc)
        dev.initFrame(1920, 1080, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
        dev.initSampler(0)
        dev.clearFrame()
        dev.render(4)
        img, alb, nrm = dev.blitFrame(), dev.blitFrame(P.BUF_ALBEDO1), dev.blitFrame(P.BUF_NORMAL1)
        host = dn.denoise(img, alb, nrm)
        ms = []
        for _ in range(3):
            dn.executeFrame(dev)
            ms.append(dn.kernelMs())
        out = dn.output()
    finally:
        dev.close()
    print(f"1080p inference: {min(ms):.3f} ms of kernels")
    assert np.array_equal(out, host)
    assert np.all(np.isfinite(out))
