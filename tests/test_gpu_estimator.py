"""HIP renderer (through the C ABI) against the float64 closed forms of tests/estimator_closed_form.py
and against the oracle for the frame outputs of AccumBuffer / FrameBuffer.

* Per-sample estimator on the two-quad scene: every sample whose decisions clear their margins
  equals the closed form to 1e-5 relative.
* Image light: a camera ray that misses everything returns the repeat/clamp bilinear lookup at
  the (u, v) ImageLight::eval assigns its direction; the light CDF equals the closed-form powers.
* Pixel sample: with a jitter every pixel's ray goes through the jitter point; with the triangle
  and Gaussian filters the first hits equal the oracle's bit for bit and the paths match.
* Frame outputs (path_renderer_stream.h:739-874, accum_buffer.h:439-486, 607-649, 752-805): the
  colour mean and variance, the depth, albedo1 / normal1 and albedo / normal AOVs equal the
  oracle's on pixels whose paths matched in every pass; the half buffer equals the mean of the
  class-B passes' samples; getFrameError equals the reference's relative-L1 formula against the
  checkpoint taken at the last power-of-two sample count; the accumulators of the debug build
  equal those of the benchmarked (debug-off) build bit for bit.
"""
import math

import numpy as np
import pytest

import estimator_closed_form as C
import oracle as O
import protoray_hip as P
import scenegen as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def test_per_sample_estimator_closed_form(dev):
    sc = C.two_quad_scene(G)
    W, H = 64, 32
    dev.enableDebug(True)
    P.load_scene(dev, sc)
    dev.initFrame(W, H)
    total = 0
    for si in range(6):
        dev.render(1)
        L, _, first = dev.debugOutputs()
        ref, ok = C.two_quad_samples(sc, W, H, si, O.sampler)
        assert ok.mean() > 0.85
        np.testing.assert_allclose(L[:, ok].astype(np.float64), ref[:, ok], rtol=1e-5, atol=1e-7)
        assert (first[ok] == 0).any() and (first[ok] == 2).any()
        total += int(ok.sum())
    assert total > 0.85 * 6 * W * H
    dev.enableDebug(False)


def _pixel_dirs(cam, W, H, jx, jy):
    U, V, N = (np.asarray(cam[k], np.float64) for k in ("U", "V", "N"))
    hh = math.tan(0.5 * float(cam["fov"]))
    hw = hh * W / H
    ys, xs = np.divmod(np.arange(W * H), W)
    d = (-hw * U + hh * V - N)[None] + ((xs + jx) / W)[:, None] * (2 * hw * U)[None] \
        + ((ys + jy) / H)[:, None] * (-2 * hh * V)[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("view", ["+z", "-x", "up"])
def test_image_light_lookup_on_miss_rays(dev, view):
    Wimg, Himg = 32, 16
    ys, xs = np.meshgrid(np.linspace(0, 1, Himg), np.linspace(0, 1, Wimg), indexing="ij")
    env = np.stack([1 + np.sin(6 * xs) * ys, 2 + np.cos(4 * ys), 0.5 + xs * ys], -1).astype(np.float32)
    v = np.array([[-1, -1, -60], [1, -1, -60], [0, 1, -60]], np.float32)  # out of every view below
    mats = np.zeros(2, P.MATERIAL_DTYPE)
    lights = np.zeros(1, P.LIGHT_DTYPE)
    scale = (1.5, 0.5, 2.0)
    lights[0] = (G.LIGHT_IMAGE, *scale, 0, 0, 0, 0)
    cam = G._camera(fov_deg=90.0)
    if view == "-x":
        cam["U"], cam["V"], cam["N"] = (0, 0, 1), (0, 1, 0), (1, 0, 0)
    elif view == "up":
        cam["U"], cam["V"], cam["N"] = (1, 0, 0), (0, 0, 1), (0, -1, 0)
    scene = dict(positions=v, indices=np.array([[0, 1, 2]], np.int32), tri_material=np.ones(1, np.int32),
                 materials=mats, lights=lights, env=env, camera=cam, max_depth=2)
    W, H = 64, 32
    dev.enableDebug(True)
    P.load_scene(dev, scene)
    dev.initFrame(W, H)
    dev.setJitter((0.5, 0.5))
    assert dev.hasJitter() and dev.getJitter() == (0.5, 0.5)
    dev.render(1)
    L, _, first = dev.debugOutputs()
    dev.enableDebug(False)
    assert (first == -1).all()
    d = _pixel_dirs(cam, W, H, 0.5, 0.5)
    u, vv = C.image_light_uv(d)
    want = (C.texture_bilinear(env, u, vv) * np.array(scale)).T
    keep = np.abs(d[:, 1]) < 0.999  # u is ill-conditioned at the poles
    np.testing.assert_allclose(L[:, keep], want[:, keep], rtol=3e-4, atol=3e-4)


def test_light_cdf_closed_form(dev):
    sc = G.spheres_ambient(n_spheres=6, subdiv=1)
    Wimg, Himg = 16, 8
    env = np.random.default_rng(2).uniform(0.1, 3.0, size=(Himg, Wimg, 3)).astype(np.float32)
    lights = np.zeros(3, P.LIGHT_DTYPE)
    d = np.array([0.2, 0.9, -0.3]) / np.linalg.norm([0.2, 0.9, -0.3])
    lights[0] = (G.LIGHT_AMBIENT, 0.3, 0.4, 0.5, 0, 0, 0, 0)
    lights[1] = (G.LIGHT_DISTANT, 50.0, 40.0, 30.0, *d, 0.05)
    lights[2] = (G.LIGHT_IMAGE, 0.7, 0.8, 0.9, 0, 0, 0, 0)
    dev.initScene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    dev.initLights(lights, env)
    r = C.scene_radius(sc["positions"], sc["indices"])
    np.testing.assert_allclose(dev.lightCdf(), C.light_cdf(C.light_powers(lights, r, env)), atol=2e-6)


def _run_pair(dev, scene, W, H, spp, aovs_mask=0, oracle_aovs=(), setup=None):
    """Renders spp passes on the GPU (debug on) and in the oracle; returns per-pass GPU samples,
    the oracle accumulators and the mask of pixels whose paths matched in every pass."""
    dev.enableDebug(True)
    P.load_scene(dev, scene)
    dev.initFrame(W, H, aovs_mask)
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"], scene.get("normals"))
    osc.set_lights(scene["lights"], scene.get("env"))
    if setup:
        setup(dev, osc)
    acc = None
    same_all = np.ones(W * H, bool)
    samples, accums = [], []
    for s in range(spp):
        dev.render(1)
        L, hsh, first = dev.debugOutputs()
        acc, _ = osc.render(scene["camera"], W, H, spp_start=s, spp_count=1, seed=0, max_depth=scene["max_depth"],
                            rr_depth=scene.get("rr_depth", 8), aovs=oracle_aovs, accum=acc, record=True)
        assert np.array_equal(first, acc["_first"]), f"pass {s}: first hits differ"
        same = hsh == acc["_hash"]
        assert same.mean() >= 0.998, f"pass {s}: only {same.mean():.4f} identical paths"
        np.testing.assert_allclose(L[:, same], acc["_sampleL"][:, same], rtol=1e-3, atol=1e-5)
        same_all &= same
        samples.append(L.copy())
        accums.append(dev.getAccum())
    dev.enableDebug(False)
    return samples, accums, acc, same_all


@pytest.mark.parametrize("ftype", [P.FILTER_TRIANGLE, P.FILTER_GAUSSIAN, P.FILTER_BLACKMAN_HARRIS])
def test_pixel_filter_parity(dev, ftype):
    scene = G.spheres_ambient(n_spheres=24, subdiv=2)

    def setup(d, o):
        d.initFilter(ftype)
        o.set_pixel_sample(filter_type=ftype)
        assert d.getFilterWidth() == (2.0 if ftype == P.FILTER_TRIANGLE else 3.0)

    _run_pair(dev, scene, 64, 64, 3, setup=setup)


def test_jitter_parity(dev):
    scene = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24)

    def setup(d, o):
        d.setJitter((0.25, 0.625))
        o.set_pixel_sample(jitter=(0.25, 0.625))

    _run_pair(dev, scene, 96, 64, 2, setup=setup)
    dev.initFrame(96, 64)  # a new frame has no jitter (FrameBuffer ctor)
    assert not dev.hasJitter() and dev.getFilterWidth() == 0.0


def test_frame_outputs_against_oracle(dev):
    scene = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24)
    W, H, spp = 96, 64, 6
    mask = P.AOV_ALBEDO | P.AOV_ALBEDO1 | P.AOV_NORMAL | P.AOV_NORMAL1 | P.AOV_DEPTH | P.AOV_VAR | P.AOV_HALF
    samples, accums, acc, same = _run_pair(dev, scene, W, H, spp, mask,
                                           ("albedo", "albedo1", "normal", "normal1", "depth", "color2"))
    assert same.mean() > 0.98
    w = acc["weight"]
    r = np.where(w == 0, np.float32(3.402823466e38), np.float32(1) / w)
    # colour mean and sample variance (accum_buffer.h:607-631) on pixels whose every path matched
    img, var = dev.blitFrame(variance=True)
    img, var = img.reshape(-1, 3).T, var.reshape(-1, 3).T
    mean_o = acc["color"] * r
    var_o = np.maximum(acc["color2"] - acc["color"] ** 2 * r, 0) / np.maximum(w - 1, 1)
    np.testing.assert_allclose(img[:, same], mean_o[:, same], rtol=1e-3, atol=1e-5)
    # the variance is a difference of moments: its error scales with the second moment E[c^2]
    m2 = acc["color2"] * r
    err = np.abs(var - var_o)[:, same]
    tol = 2e-3 * np.abs(var_o[:, same]) + 2e-3 * m2[:, same] + 1e-6
    assert (err <= tol).all(), f"variance: {(err > tol).sum()} elements off, worst {np.max(err - tol)}"
    assert (var_o[:, same] > 0).mean() > 0.5
    # first-hit AOVs: every pixel (first hits are bit-exact)
    for buf, name in ((P.BUF_ALBEDO1, "albedo1"), (P.BUF_NORMAL1, "normal1")):
        g = dev.blitFrame(buf).reshape(-1, 3).T
        o = acc[name] / np.maximum(acc[name + "W"], 1)
        assert np.isfinite(g).all()
        np.testing.assert_allclose(g, o, rtol=1e-4, atol=1e-5, err_msg=name)
    depth = dev.blitFrame(P.BUF_DEPTH).ravel()
    depth_o = (acc["depth"] / np.maximum(acc["depthW"], 1)).ravel()
    np.testing.assert_allclose(depth, depth_o, rtol=1e-5)
    assert np.isinf(depth).any() and np.isfinite(depth).mean() > 0.3
    # AOVs at the first non-delta vertex (:822-874): on the matched pixels
    for buf, name in ((P.BUF_ALBEDO, "albedo"), (P.BUF_NORMAL, "normal")):
        g = dev.blitFrame(buf).reshape(-1, 3).T
        o = acc[name] / np.maximum(acc[name + "W"], 1)
        np.testing.assert_allclose(g[:, same], o[:, same], rtol=1e-4, atol=1e-5, err_msg=name)
    # the half buffer holds the class-B passes (count with an odd number of bits in count & 0xaaaaaaaa)
    classB = [bin(k & 0xaaaaaaaa).count("1") & 1 for k in range(spp)]
    assert classB == [0, 0, 1, 1, 0, 0]
    sB = np.zeros((3, W * H), np.float32)
    for k in range(spp):
        if classB[k]:
            sB += samples[k]
    half = dev.blitHalf().reshape(-1, 3).T
    np.testing.assert_array_equal(half, sB * (np.float32(1) / np.float32(sum(classB))))
    # getFrameError (accum_buffer.h:752-805) against the checkpoint copied when the count was 4
    # (FrameBuffer::begin at power-of-two counts, accum_buffer.h:252-260)
    s6, w6 = accums[-1]
    s4, w4 = accums[3]
    I = s6 * (np.float32(1) / w6)
    B = s4 * (np.float32(1) / w4)
    norm = I.sum(axis=0, dtype=np.float32)
    norm = np.where(norm < 1, np.sqrt(norm), norm).astype(np.float32)
    err = (np.abs(I - B).sum(axis=0, dtype=np.float32) / (norm + np.float32(1e-4))).astype(np.float64).sum() / (W * H)
    assert abs(dev.getFrameError() - err) <= 1e-5 * err
    assert err > 0


def test_debug_build_accumulates_like_the_benchmarked_build(dev):
    scene = G.spheres_lobes()
    W, H = 96, 64
    out = []
    for debug in (False, True):
        dev.enableDebug(debug)
        P.load_scene(dev, scene)
        dev.initFrame(W, H, P.AOV_ALBEDO | P.AOV_NORMAL1 | P.AOV_VAR)
        dev.render(4)
        out.append((dev.getAccum(), dev.blitFrame(P.BUF_ALBEDO), dev.blitFrame(P.BUF_NORMAL1),
                    dev.blitFrame(variance=True)[1]))
    dev.enableDebug(False)
    (a0, al0, n0, v0), (a1, al1, n1, v1) = out
    for x, y in ((a0[0], a1[0]), (a0[1], a1[1]), (al0, al1), (n0, n1), (v0, v1)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
