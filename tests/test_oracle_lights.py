"""Pins the oracle's environment lights, light powers, per-sample estimator and pixel sample
against float64 restatements written from the reference's source text (tests/estimator_closed_form.py).
The reference's light and integrator headers include the absent TBB headers, so these closed forms
replace golden vectors for them.

* Image light (light/image_light.h:66-147): with one bright texel, the direction through that
  texel's centre returns the texel and the pdf W H / (2 pi^2 sin theta); zenith, nadir and the
  four horizon axes land on the texels the reference's lat-long convention assigns; on a smooth
  image, evaluation equals the repeat/clamp bilinear lookup at the (u, v) of the direction.
* Light powers (ambient_light.h:53-57, distant_light.h:73-78, image_light.h:136-148): the light
  CDF equals the closed-form powers for ambient + distant + image lights.
* Per-sample estimator (path_renderer_stream.h:285-937): on a two-quad diffuse scene every
  sample's radiance equals the float64 restatement (power-heuristic MIS, the RR-adjusted NEE pdf,
  roulette only after the AOVs are written, the light pick from the third light-sample
  dimension) to 1e-5; each of five plausible slips in restating it disagrees.
* Pixel sample (image/pixel_filter.h, path_renderer_stream.h:228-236): the jitter replaces the
  sampler's value; the triangle filter's warp equals its analytic inverse CDF, the Gaussian and
  Blackman-Harris warps equal the float64 inverse of their integrated profiles.
"""
import math

import numpy as np
import pytest

import estimator_closed_form as C
import oracle as O
import scenegen as G


def _one_triangle_scene():
    v = np.array([[-1, -1, 20], [1, -1, 20], [0, 1, 20]], np.float32)
    mats = np.zeros(2, O.MATERIAL_DTYPE)
    return O.Scene(v, np.array([[0, 1, 2]], np.int32), np.ones(1, np.int32), mats)


def _image_light(scale=(1.0, 1.0, 1.0)):
    lights = np.zeros(1, O.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_IMAGE, *scale, 0, 0, 0, 0)
    return lights


def test_image_light_single_texel_mapping():
    W, H = 16, 8
    s = _one_triangle_scene()
    for y in range(H):
        for x in range(0, W, 3):
            env = np.zeros((H, W, 3), np.float32)
            env[y, x] = (2.0, 3.0, 4.0)
            s.set_lights(_image_light(), env)
            d = C.image_light_dir((x + 0.5) / W, (y + 0.5) / H)
            out = s.env_eval(d[None])[0]
            np.testing.assert_allclose(out[:3], env[y, x], rtol=1e-4)
            np.testing.assert_allclose(out[3], C.image_light_pdf_single_texel(W, H, y), rtol=1e-4)
            # the neighbouring texel centres see nothing of it but their bilinear share (zero at the centre)
            d2 = C.image_light_dir((x + 1.5) / W, (y + 0.5) / H)
            assert np.abs(s.env_eval(d2[None])[0, :3]).max() < 1e-3


# (direction, expected (x, y, seam) of the bilinear lookup): worked by hand from ImageLight::eval,
# theta = acos(y), phi = atan2(x, -z), u = 0.5 + phi / 2 pi, v = 1 - theta / pi, and the texture's
# texel centres at (i + 0.5) / size, u repeating and v clamped, rows in texture order (row 0 is
# the nadir, image/image_texture.h:396); the image holds x, y and a marker of column 0
AXES = [((0, 1, 0), (None, 7.0, None)), ((0, -1, 0), (None, 0.0, None)), ((0, 0, -1), (7.5, 3.5, 0.0)),
        ((1, 0, 0), (11.5, 3.5, 0.0)), ((-1, 0, 0), (3.5, 3.5, 0.0)), ((0, 0, 1), (7.5, 3.5, 0.5))]


@pytest.mark.parametrize("d,want", AXES)
def test_image_light_axes(d, want):
    W, H = 16, 8
    s = _one_triangle_scene()
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    env = np.stack([xs, ys, (xs == 0)], -1).astype(np.float32)
    s.set_lights(_image_light(), env)
    out = s.env_eval(np.array([d], np.float32))[0]
    for k in range(3):
        if want[k] is not None:
            assert abs(out[k] - want[k]) < 1e-4, (d, k, out[:3])


def test_image_light_matches_bilinear_lookup():
    W, H = 32, 16
    rng = np.random.default_rng(5)
    ys, xs = np.meshgrid(np.linspace(0, 1, H), np.linspace(0, 1, W), indexing="ij")
    env = np.stack([1 + np.sin(6 * xs) * ys, 2 + np.cos(4 * ys), 0.5 + xs * ys], -1).astype(np.float32)
    s = _one_triangle_scene()
    s.set_lights(_image_light((1.5, 0.5, 2.0)), env)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = s.env_eval(d.astype(np.float32))
    u, v = C.image_light_uv(d)
    want = C.texture_bilinear(env, u, v) * np.array([1.5, 0.5, 2.0])
    np.testing.assert_allclose(out[:, :3], want, rtol=2e-4, atol=2e-4)


def test_light_powers_closed_form():
    sc = G.spheres_ambient(n_spheres=6, subdiv=1)
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    W, H = 16, 8
    rng = np.random.default_rng(2)
    env = rng.uniform(0.1, 3.0, size=(H, W, 3)).astype(np.float32)
    lights = np.zeros(3, O.LIGHT_DTYPE)
    d = np.array([0.2, 0.9, -0.3]) / np.linalg.norm([0.2, 0.9, -0.3])
    lights[0] = (G.LIGHT_AMBIENT, 0.3, 0.4, 0.5, 0, 0, 0, 0)
    lights[1] = (G.LIGHT_DISTANT, 50.0, 40.0, 30.0, *d, 0.05)
    lights[2] = (G.LIGHT_IMAGE, 0.7, 0.8, 0.9, 0, 0, 0, 0)
    s.set_lights(lights, env)
    r = C.scene_radius(sc["positions"], sc["indices"])
    assert abs(s.radius - r) < 1e-5 * r
    want = C.light_cdf(C.light_powers(lights, r, env))
    np.testing.assert_allclose(s.light_cdf(), want, atol=2e-6)


def _two_quads():
    sc = C.two_quad_scene(G)
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    s.set_lights(sc["lights"])
    return sc, s


def test_per_sample_estimator_closed_form():
    sc, s = _two_quads()
    W, H = 32, 16
    for si in range(4):
        acc, _ = s.render(sc["camera"], W, H, spp_start=si, spp_count=1, seed=0, max_depth=sc["max_depth"],
                          rr_depth=sc["rr_depth"], record=True)
        ref, ok = C.two_quad_samples(sc, W, H, si, O.sampler)
        assert ok.mean() > 0.85
        L = acc["_sampleL"].astype(np.float64)
        np.testing.assert_allclose(L[:, ok], ref[:, ok], rtol=1e-5, atol=1e-7)
        # the comparison covers every kind of sample: direct views of the wall and the floor, paths
        # that escape to the sky and paths ended by roulette
        assert (acc["_first"][ok] == 0).any() and (acc["_first"][ok] == 2).any()
        assert (ref[:, ok] > 0).all(axis=0).mean() > 0.5


@pytest.mark.parametrize("variant", C.VARIANTS)
def test_estimator_slips_are_detected(variant):
    sc, s = _two_quads()
    W, H = 32, 16
    bad = tot = 0
    for si in range(2):
        acc, _ = s.render(sc["camera"], W, H, spp_start=si, spp_count=1, seed=0, max_depth=sc["max_depth"],
                          rr_depth=sc["rr_depth"], record=True)
        ref, ok = C.two_quad_samples(sc, W, H, si, O.sampler, variant=variant)
        L = acc["_sampleL"].astype(np.float64)
        rel = np.abs(L - ref) / np.maximum(np.abs(ref), 1e-5)
        bad += int((rel[:, ok].max(axis=0) > 1e-4).sum())
        tot += int(ok.sum())
    assert bad > 0.05 * tot, f"{variant}: only {bad} of {tot} samples differ"


def _filter_inverse_cdf(profile, radius, u):
    x = np.linspace(-radius, radius, 200001)
    f = np.maximum(profile(x), 0)
    c = np.concatenate([[0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(x))])
    return np.interp(u, c / c[-1], x)


@pytest.mark.parametrize("ftype,width", [(1, 2.0), (1, 3.0), (2, 3.0), (3, 3.0)])
def test_pixel_filter_warp(ftype, width):
    s = _one_triangle_scene()
    s.set_pixel_sample(filter_type=ftype, width=width)
    u = np.linspace(0.001, 0.999, 501)
    out = s.pixel_sample(np.stack([u, u[::-1]], 1))
    r = width / 2
    if ftype == 1:  # triangle: density (r - |x|) / r^2, inverse CDF in closed form
        want = np.where(u < 0.5, -r + r * np.sqrt(2 * u), r - r * np.sqrt(2 * (1 - u)))
    elif ftype == 2:  # Gaussian, sigma = width / 6, minus its value at the radius
        sig = width / 6
        want = _filter_inverse_cdf(lambda x: np.exp(-x * x / (2 * sig * sig)) - math.exp(-r * r / (2 * sig * sig)), r, u)
    else:
        want = _filter_inverse_cdf(lambda x: 0.35875 - 0.48829 * np.cos((x / r + 1) * math.pi)
                                   + 0.14128 * np.cos(2 * (x / r + 1) * math.pi)
                                   - 0.01168 * np.cos(3 * (x / r + 1) * math.pi), r, u)
    # the table has 4097 entries over the width: its piecewise-linear inverse is within ~1e-3 of the width
    np.testing.assert_allclose(out[:, 0] - 0.5, want, atol=2e-3 * width)
    np.testing.assert_allclose(out[:, 1] - 0.5, want[::-1], atol=2e-3 * width)


def test_jitter_replaces_the_pixel_sample():
    s = _one_triangle_scene()
    s.set_pixel_sample(jitter=(0.25, 0.75))
    out = s.pixel_sample(np.random.default_rng(0).uniform(size=(10, 2)))
    assert (out == np.array([0.25, 0.75], np.float32)).all()
    s.set_pixel_sample()
    sv = np.random.default_rng(0).uniform(size=(10, 2)).astype(np.float32)
    assert (s.pixel_sample(sv) == sv).all()
