"""Volume transmission on the GPU (kShade, and tracePathMega for the batch renderer) against the closed
form of scenegen.glass_slabs and against the oracle (path_renderer_stream.h:652-657,
material/standard.cpp:169-173): a camera ray through slab k returns ambient x (colour_k ^
depth_k) ^ (thickness / cos theta); two slabs of one IOR and different colours get separate media."""
import numpy as np
import pytest

import estimator_closed_form as C
import protoray_hip as P
import scenegen as G
from test_oracle_media import render_slabs, slab_reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def test_slab_absorption_closed_form(dev):
    sc = G.glass_slabs()
    W, H = sc["width"], sc["height"]
    P.load_scene(dev, sc)
    dev.initFrame(W, H)
    dev.setJitter((0.5, 0.5))
    dev.render(4)
    mean = dev.blitFrame().reshape(-1, 3)
    ref, k, keep = slab_reference(sc)
    assert (k[keep] == 0).any() and (k[keep] == 1).any()
    np.testing.assert_allclose(mean[keep], ref[keep], rtol=2e-3)
    np.testing.assert_allclose(mean[keep], render_slabs(sc).T[keep], rtol=1e-4, atol=1e-6)


def test_slab_absorption_batch_renderer(dev):
    # the persistent batch kernel traces the same paths with tracePathMega; batch frames take the
    # sampler's pixel sample (no jitter), so they equal the wavefront renderer without a jitter
    sc = G.glass_slabs()
    W, H = sc["width"], sc["height"]
    P.load_scene(dev, sc)
    dev.disableJitter()
    dev.initFrame(W, H)
    dev.render(4)
    want = dev.blitFrame()
    b = P.BatchHip(0)
    try:
        b.addFrame(dev, sc["camera"], W, H, seed=0, spp=4)
        b.render()
        color = b.get(0)
    finally:
        b.close()
    np.testing.assert_allclose(color, want, rtol=1e-5, atol=1e-7)
    ref, _, keep = slab_reference(sc)
    np.testing.assert_allclose(color.reshape(-1, 3)[keep], ref[keep], rtol=3e-2)
