"""Image textures on the GPU through the C ABI: the textured base colour of a render against
the closed form the oracle tests pin (a ramp texture on a camera-facing quad), the persistent
batch renderer against the wavefront renderer on a textured scene, and the error paths.  Parity
with the oracle on a textured scene is in test_gpu_parity.py."""
import numpy as np
import pytest

import dataset as D
import protoray_hip as P
import scenegen as G
from test_oracle_textures import expected_quad_albedo, textured_quad_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [G.KIND_DIFFUSE, G.KIND_COATED])
def test_textured_base_colour_closed_form(kind):
    scene = textured_quad_scene(kind)
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, scene)
        dev.initFrame(scene["width"], scene["height"], P.AOV_ALBEDO1)
        dev.setJitter((0.5, 0.5))
        dev.render(1)
        got = dev.blitFrame(P.BUF_ALBEDO1).reshape(-1, 3).T
    finally:
        dev.close()
    np.testing.assert_allclose(got, expected_quad_albedo(scene), rtol=0, atol=5e-5)


def test_batch_renderer_on_a_textured_scene():
    sc = G.spheres_textured(seed=3)
    dev = P.DeviceHip(0)
    batch = P.BatchHip(0)
    try:
        P.load_scene(dev, sc)
        cam = D.look_at((0, 2, -2), (1, 1, 15), 50)
        batch.addFrame(dev, cam, 64, 48, 2, 4, snapshots=(1,), aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1)
        batch.render()
        dev.initCamera(cam)
        dev.initFrame(64, 48, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
        dev.initSampler(2)
        dev.clearFrame()
        dev.render(4)
        np.testing.assert_allclose(batch.get(0), dev.blitFrame(), rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(0, P.BATCH_ALBEDO1), dev.blitFrame(P.BUF_ALBEDO1), rtol=1e-5, atol=1e-6)
    finally:
        batch.close()
        dev.close()


def test_texture_errors():
    scene = textured_quad_scene()
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, scene)
        with pytest.raises(ValueError):  # one UV pair per vertex
            dev.initUVs(np.zeros((3, 2), np.float32))
        with pytest.raises(ValueError):  # unknown format
            dev.initTextures([dict(texels=np.zeros((2, 2, 4), np.uint8), format=7)])
        with pytest.raises(ValueError):  # a texture index beyond the set
            dev.setMaterialTexture(1, 5)
        mats = scene["materials"].copy()
        mats[1]["kind"] = G.KIND_CONDUCTOR
        dev.initScene(scene["positions"], scene["indices"], scene["tri_material"], mats)
        with pytest.raises(ValueError):  # base colour textures: diffuse and coated only
            dev.setMaterialTexture(1, 0)
        dev.initScene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"])
        dev.initLights(scene["lights"])
        dev.setMaterialTexture(1, 0)
        dev.initFrame(64, 48)
        with pytest.raises(ValueError):  # the adjoint does not differentiate textures
            dev.adjoint(np.zeros((48, 64, 3), np.float32), 1, 2)
        dev.setMaterialTexture(1, -1)
        dev.render(1)
    finally:
        dev.close()
