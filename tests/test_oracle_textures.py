"""Image textures in the oracle: the texel decode and the UV interpolation against vectors the
reference's own code wrote (image/pixel.h, geometry/triangle_mesh.cpp), and the bilinear lookup
(image/image_texture.h:186-227) and the textured base colour of a render against closed forms.
The reference's lookup itself cannot be compiled here (image/image.h includes sys/tasking.h,
whose TBB headers are absent), so the lookup is pinned by what bilinear filtering fixes: texel
centres return texels, midpoints return means, linear ramps come back linear, repeat is
periodic and clamp holds the edge texel.
"""
import numpy as np
import pytest

import oracle as O
import scenegen as G


def _scene_with_texture(tex):
    # a one-triangle scene to hold textures for lookups
    pos = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
    sc = O.Scene(pos, np.array([[0, 1, 2]], np.int32), np.array([1], np.int32), np.zeros(2, O.MATERIAL_DTYPE))
    k = sc.add_texture(tex)
    return sc, k


def test_texel_decode_matches_reference(golden):
    # every byte value through a 256 x 1 texture, looked up at the texel centres (clamp addressing)
    ramp = np.zeros((1, 256, 4), np.uint8)
    ramp[0, :, 0] = np.arange(256)
    ramp[0, :, 1] = np.arange(256)[::-1]
    ramp[0, :, 2] = np.arange(256)
    uv = np.stack([(np.arange(256) + 0.5) / 256.0, np.full(256, 0.5)], axis=1)
    for fmt, name in ((0, "pixel_srgb8"), (1, "pixel_linear8")):
        sc, k = _scene_with_texture(dict(texels=ramp, format=fmt, address=(1, 1)))
        got = sc.texture_lookup(k, uv)
        ref = golden[name][:, 0]
        assert np.array_equal(got[:, 0], ref), name
        assert np.array_equal(got[:, 1], ref[::-1]), name


def test_uv_at_hits_matches_reference_post_intersect(golden):
    vuv = golden["texcoord_vertex_uv"]
    n = vuv.shape[0]
    pos = np.random.default_rng(0).random((n, 3)).astype(np.float32)
    idx = np.arange(n, dtype=np.int32).reshape(-1, 3)
    sc = O.Scene(pos, idx, np.ones(idx.shape[0], np.int32), np.zeros(2, O.MATERIAL_DTYPE))
    prim, huv = golden["texcoord_prim"], golden["texcoord_hit_uv"]
    # without texcoords the hit's barycentric (u, v) are the UVs (triangle_mesh.cpp:234-237)
    assert np.array_equal(sc.hit_uv(prim, huv), golden["texcoord_out_none"])
    sc.set_textures({"uvs": vuv})
    # bit-exact: the reference's build contracts b0 uv0 + b1 uv1 + b2 uv2 into two fused
    # multiply-adds (GCC's default for GNU C++), which the oracle and the kernels write out
    assert np.array_equal(sc.hit_uv(prim, huv), golden["texcoord_out"])


def _rgb32(h, w, fn):
    t = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        for x in range(w):
            t[y, x] = fn(x, y)
    return t


def test_bilinear_lookup_closed_forms():
    rng = np.random.default_rng(3)
    H, W = 3, 5
    img = rng.random((H, W, 3)).astype(np.float32)
    for addr in ((0, 0), (1, 1), (0, 1)):
        sc, k = _scene_with_texture(dict(texels=img, format=2, address=addr))
        # texel centres: texel row y (from the bottom) is image row H - 1 - y
        for y in range(H):
            for x in range(W):
                got = sc.texture_lookup(k, [[(x + 0.5) / W, (y + 0.5) / H]])[0]
                np.testing.assert_allclose(got, img[H - 1 - y, x], rtol=0, atol=1e-6)
        # midpoint between two horizontal neighbours (not across the wrap)
        got = sc.texture_lookup(k, [[2.0 / W, 0.5 / H]])[0]
        np.testing.assert_allclose(got, 0.5 * (img[H - 1, 1] + img[H - 1, 2]), rtol=0, atol=1e-6)
    # repeat: periodic in u and v; the wrap blends the last texel with the first
    sc, k = _scene_with_texture(dict(texels=img, format=2, address=(0, 0)))
    uv = rng.random((64, 2)).astype(np.float32) * 0.8 + 0.1
    np.testing.assert_allclose(sc.texture_lookup(k, uv + np.float32(3.0)), sc.texture_lookup(k, uv), rtol=0, atol=2e-6)
    got = sc.texture_lookup(k, [[0.0, 0.5 / H]])[0]
    np.testing.assert_allclose(got, 0.5 * (img[H - 1, W - 1] + img[H - 1, 0]), rtol=0, atol=1e-6)
    # clamp: beyond the edges the edge texels
    sc, k = _scene_with_texture(dict(texels=img, format=2, address=(1, 1)))
    np.testing.assert_allclose(sc.texture_lookup(k, [[-4.0, -2.0]])[0], img[H - 1, 0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(sc.texture_lookup(k, [[7.0, 9.0]])[0], img[0, W - 1], rtol=0, atol=1e-7)


def test_bilinear_lookup_reproduces_linear_ramps():
    # R = x / (W - 1), G = y / (H - 1) over texel indices (y from the bottom): with clamp
    # addressing the lookup is the linear function clamp(u W - 0.5, 0, W - 1) / (W - 1)
    H, W = 6, 9
    img = _rgb32(H, W, lambda x, y: (x / (W - 1), (H - 1 - y) / (H - 1), 0.25))
    sc, k = _scene_with_texture(dict(texels=img, format=2, address=(1, 1)))
    uv = np.random.default_rng(4).uniform(-0.2, 1.2, (500, 2)).astype(np.float32)
    got = sc.texture_lookup(k, uv)
    er = np.clip(uv[:, 0] * W - 0.5, 0, W - 1) / (W - 1)
    eg = np.clip(uv[:, 1] * H - 0.5, 0, H - 1) / (H - 1)
    np.testing.assert_allclose(got[:, 0], er, rtol=0, atol=2e-6)
    np.testing.assert_allclose(got[:, 1], eg, rtol=0, atol=2e-6)
    np.testing.assert_allclose(got[:, 2], 0.25, rtol=0, atol=1e-7)


def test_uv_scale_and_offset():
    img = np.random.default_rng(5).random((4, 7, 3)).astype(np.float32)
    sc, k0 = _scene_with_texture(dict(texels=img, format=2, address=(0, 1)))
    k1 = sc.add_texture(dict(texels=img, format=2, address=(0, 1), uv_scale=(2.5, 0.75), uv_offset=(0.125, -0.25)))
    uv = np.random.default_rng(6).random((100, 2)).astype(np.float32)
    moved = uv * np.array([2.5, 0.75], np.float32) + np.array([0.125, -0.25], np.float32)
    np.testing.assert_allclose(sc.texture_lookup(k1, uv), sc.texture_lookup(k0, moved), rtol=0, atol=1e-7)


def textured_quad_scene(kind=G.KIND_DIFFUSE, scale=(0.9, 1.3, 0.6), a=7.0, z0=8.0, W=64, H=48):
    """A quad at z = z0 spanning [-a, a]^2 in front of the pinhole camera of the generators, UVs
    (x + a) / 2a, (y + a) / 2a, its base colour a ramp texture (R = u ramp, G = v ramp) times
    `scale`, lit by a constant ambient light; one bounce."""
    pos = np.array([[-a, -a, z0], [a, -a, z0], [a, a, z0], [-a, a, z0]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    uvs = (pos[:, :2] + a) / (2 * a)
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1]["kind"] = kind
    mats[1]["r"], mats[1]["g"], mats[1]["b"] = scale
    if kind == G.KIND_COATED:
        mats[1]["er"], mats[1]["eg"], mats[1]["eb"] = (0.8, 0.9, 1.0)
        mats[1]["roughness"], mats[1]["ior"], mats[1]["thickness"], mats[1]["weight"] = 0.3, 1.5, 1.0, 0.7
    TW, TH = 17, 11
    tex = _rgb32(TH, TW, lambda x, y: (x / (TW - 1), (TH - 1 - y) / (TH - 1), 0.8))
    lights = np.zeros(1, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_AMBIENT, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    return dict(name="textured_quad", positions=pos, indices=idx, tri_material=np.ones(2, np.int32), materials=mats,
                uvs=uvs, textures=[dict(texels=tex, format=2, address=(1, 1))], material_textures=[(1, 0)],
                lights=lights, env=None, camera=G._camera(fov_deg=60.0), width=W, height=H, spp=1, max_depth=1,
                seed=0, a=a, z0=z0, tex_size=(TW, TH), scale=scale)


def expected_quad_albedo(scene):
    """First-hit albedo of the textured quad per pixel centre, in closed form: the pixel-centre
    camera ray meets the plane z = z0 at (x, y); uv = ((x + a) / 2a, (y + a) / 2a); the ramps
    give R, G; clamp(ramp * scale); a coated base reports its albedo lerp(c, c t^d, w)."""
    W, H, a, z0 = scene["width"], scene["height"], scene["a"], scene["z0"]
    TW, TH = scene["tex_size"]
    cam = scene["camera"]
    cfg = np.array([W, H, cam["fov"], cam["nearClip"], 0.0, 1.0, 0, 0, 0], np.float32)
    py, px = np.mgrid[0:H, 0:W]
    s = np.zeros((W * H, 4), np.float32)
    s[:, 0] = ((px.ravel() + 0.5) / W).astype(np.float32)
    s[:, 1] = ((py.ravel() + 0.5) / H).astype(np.float32)
    d = O.camera_rays(cfg, s)[:, 0, 3:6].astype(np.float64)
    t = z0 / d[:, 2]
    x, y = d[:, 0] * t, d[:, 1] * t
    u, v = (x + a) / (2 * a), (y + a) / (2 * a)
    r = np.clip(u * TW - 0.5, 0, TW - 1) / (TW - 1)
    g = np.clip(v * TH - 0.5, 0, TH - 1) / (TH - 1)
    c = np.clip(np.stack([r, g, np.full_like(r, 0.8)]) * np.array(scene["scale"])[:, None], 0, 1)
    m = scene["materials"][1]
    if m["kind"] == G.KIND_COATED:  # bsdf/rough_coating.h:567-571
        tr = np.array([m["er"], m["eg"], m["eb"]], np.float64)[:, None] ** float(m["thickness"])
        c = c + float(m["weight"]) * (c * tr - c)
    return c


@pytest.mark.parametrize("kind", [G.KIND_DIFFUSE, G.KIND_COATED])
def test_textured_base_colour_in_a_render(kind):
    scene = textured_quad_scene(kind)
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"])
    osc.set_lights(scene["lights"])
    osc.set_textures(scene)
    osc.set_pixel_sample(jitter=(0.5, 0.5))
    acc, _ = osc.render(scene["camera"], scene["width"], scene["height"], 0, 1, seed=0, max_depth=1, aovs=("albedo1",))
    got = acc["albedo1"] / np.maximum(acc["albedo1W"], 1)
    np.testing.assert_allclose(got, expected_quad_albedo(scene), rtol=0, atol=5e-5)
    # and the colour follows it: a diffuse plane under a unit ambient light and one bounce
    # returns emission 0 + albedo x (ambient seen through the hemisphere) = albedo in expectation;
    # per sample it is the estimator's value, so only the albedo is pinned exactly here
    assert np.isfinite(acc["color"]).all()
