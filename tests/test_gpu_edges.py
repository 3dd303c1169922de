"""Edge cases of the render path on the GPU: empty and one-triangle scenes, the frame-size limits
the reference enforces (render/path_renderer_stream.h:109-112: width a multiple of 32, height of
8, each at most 4096), zero-sample renders and empty query batches."""
import numpy as np
import pytest

import oracle as O
import protoray_hip as P
import scenegen as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def _ambient(level):
    lights = np.zeros(1, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_AMBIENT, level, level, level, 0, 0, 0, 0)
    return lights


def _load(dev, pos, idx, tri_mat, mats, lights, max_depth=5):
    dev.initScene(pos, idx, tri_mat, mats)
    dev.initLights(lights, None)
    dev.initRenderer(maxDepth=max_depth)
    dev.initSampler(0)
    dev.initCamera(G._camera())


def test_empty_scene_returns_the_environment(dev):
    # no triangles: every camera ray misses and returns the ambient radiance, exactly
    mats = np.zeros(1, G.MATERIAL_DTYPE)
    _load(dev, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros(0, np.int32), mats, _ambient(0.75))
    dev.initFrame(64, 16, P.AOV_ALBEDO1)
    dev.clearFrame()
    st = dev.render(3)
    img = dev.blitFrame()
    assert st["spp"] == 3
    assert np.all(img == np.float32(0.75))
    assert np.all(dev.blitFrame(P.BUF_ALBEDO1) == 0.0)  # misses write a zero first-hit albedo


def test_one_triangle_scene_matches_oracle(dev):
    # a single triangle in front of the camera: the BVH is a lone leaf
    pos = np.array([[-3, -2, 6], [3, -2, 6], [0, 3, 6]], np.float32)
    idx = np.array([[0, 1, 2]], np.int32)
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1] = G._material(G.KIND_DIFFUSE, (0.6, 0.5, 0.4))
    lights = _ambient(1.0)
    _load(dev, pos, idx, np.array([1], np.int32), mats, lights)
    W, H, spp = 64, 32, 4
    dev.initFrame(W, H)
    dev.clearFrame()
    dev.render(spp)
    img = dev.blitFrame().reshape(-1, 3)
    osc = O.Scene(pos, idx, np.array([1], np.int32), mats)
    osc.set_lights(lights)
    acc, _ = osc.render(G._camera(), W, H, spp_start=0, spp_count=spp, max_depth=5)
    ref = (acc["color"] / acc["weight"]).reshape(3, -1).T
    np.testing.assert_allclose(img, ref, rtol=1e-3, atol=1e-5)
    # an unoccluded two-sided diffuse triangle under a unit ambient light reflects its albedo
    # (one bounce, then the environment again); misses return 1
    hit = np.any(np.abs(ref - 1.0) > 1e-6, axis=1)
    assert hit.any() and (~hit).any()


@pytest.mark.parametrize("W,H", [(60, 32), (64, 30), (4128, 8), (32, 4104), (0, 8), (32, 0)])
def test_frame_size_limits(dev, W, H):
    with pytest.raises(ValueError):
        dev.initFrame(W, H)


def test_largest_frame(dev):
    # 4096 x 4096 is allowed (16.8 M paths, one pass over an ambient-lit empty scene)
    mats = np.zeros(1, G.MATERIAL_DTYPE)
    _load(dev, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros(0, np.int32), mats, _ambient(2.0))
    dev.initFrame(4096, 4096)
    dev.clearFrame()
    dev.render(1)
    img = dev.blitFrame()
    assert img.shape == (4096, 4096, 3) and np.all(img == np.float32(2.0))
    dev.initFrame(64, 16)  # release the large frame


def test_scene_bounds(dev):
    # Device::getSceneBounds: the box of the referenced vertices (an unreferenced far vertex is
    # left out), empty for a scene without triangles
    pos = np.array([[-3, -2, 6], [3, -2, 6], [0, 3, 7], [100, 100, 100]], np.float32)
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1] = G._material(G.KIND_DIFFUSE, (0.5, 0.5, 0.5))
    _load(dev, pos, np.array([[0, 1, 2]], np.int32), np.array([1], np.int32), mats, _ambient(1.0))
    lo, hi = dev.sceneBounds()
    assert np.array_equal(lo, [-3, -2, 6]) and np.array_equal(hi, [3, 3, 7])
    _load(dev, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros(0, np.int32), mats[:1], _ambient(1.0))
    lo, hi = dev.sceneBounds()
    assert np.all(lo > hi)


def test_zero_samples_and_empty_queries(dev):
    scene = G.spheres_ambient(n_spheres=8, subdiv=1)
    P.load_scene(dev, scene)
    dev.initFrame(32, 8)
    dev.clearFrame()
    st = dev.render(0)
    assert st["spp"] == 0
    assert np.all(dev.blitFrame() == 0.0)
    prim, t, u, v = dev.queryRays(np.zeros((0, 8), np.float32))
    assert prim.shape == (0,)
    assert dev.queryOcclusion(np.zeros((0, 8), np.float32)).shape == (0,)


def _shadow_arrays(transparency):
    # floor at y = -1 (diffuse 0.5) under a distant light from straight above, a black quad of
    # the given transparency at y = 1 between the floor and the light (above the camera)
    v = np.array([[-100, -1, -10], [100, -1, -10], [100, -1, 100], [-100, -1, 100],
                  [-100, 1, -10], [100, 1, -10], [100, 1, 100], [-100, 1, 100]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]], np.int32)
    mats = np.zeros(3, G.MATERIAL_DTYPE)
    mats[1] = G._material(G.KIND_DIFFUSE, (0.5, 0.5, 0.5))
    mats[2] = G._material(G.KIND_DIFFUSE, (0, 0, 0), transparency=transparency)
    lights = np.zeros(1, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_DISTANT, 3.0, 3.0, 3.0, 0, 1, 0, np.radians(1.0))
    return v, idx, np.array([1, 1, 2, 2], np.int32), mats, lights


def test_transparent_shadow_scales_direct_light(dev):
    # the transparent-shadow filter in kOccluded: a shadow ray through a surface of transparency t
    # keeps t of the light, so every floor pixel is exactly t times its value under a fully
    # transparent quad (bounce rays leave through the black quad into an empty sky either way)
    W, H, spp = 64, 32, 4

    def render(t):
        v, idx, tm, mats, lights = _shadow_arrays(t)
        _load(dev, v, idx, tm, mats, lights, max_depth=4)
        dev.enableDebug(True)
        dev.initFrame(W, H)
        dev.clearFrame()
        dev.render(spp)
        _, _, first = dev.debugOutputs()
        dev.enableDebug(False)
        return dev.blitFrame().reshape(-1, 3), first

    base, first = render(1.0)
    floor = (first >= 0) & (first < 2)
    assert floor.sum() > 300 and base[floor].min() > 0
    for t in (0.37, 0.8):
        img, _ = render(t)
        np.testing.assert_allclose(img[floor], t * base[floor], rtol=2e-6)
    img, _ = render(0.0)
    assert np.all(img[floor] == 0)


@pytest.mark.parametrize("field,value", [("kind", 5), ("transparency", 1.5), ("transparency", -0.1)])
def test_invalid_material_records(dev, field, value):
    # prt_scene_set_materials rejects an unknown kind and a transparency outside [0, 1]
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1] = G._material(G.KIND_DIFFUSE, (0.5, 0.5, 0.5))
    mats[1][field] = value
    with pytest.raises(ValueError):
        dev.initScene(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.int32), np.array([1], np.int32), mats)


def test_thin_dielectric_needs_an_interface(dev):
    # an IOR equal to the ambient 1 would make it a transparent BTDF (standard.cpp:547-550): rejected
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1] = G._material(G.KIND_THIN_DIELECTRIC, (1, 1, 1), roughness=0.3, ior=1.0)
    with pytest.raises(ValueError):
        dev.initScene(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.int32), np.array([1], np.int32), mats)
