"""Closed forms and invariants that pin the oracle's lights and integrator (their reference
headers cannot be compiled here: they need TBB headers).

* A ray that misses everything under a constant ambient light returns exactly its radiance.
* A diffuse plane of albedo rho under a constant ambient light of radiance 1 reflects exactly
  rho (the plane sees only the upper hemisphere and cannot see itself).
* A white rough conductor with energy compensation reflects ~1 under the same light
  (Kulla-Conty: the multiple-scattering lobe restores the lost energy).
* Light sampling: the returned weight equals eval / pdf at the sampled direction, and the pdf
  integrates to 1 over the sphere.
* BSDF sampling: for the non-delta lobes the sample weight equals eval / pdf.
* Coated diffuse (rough_coating.h): with weight 0 the coat vanishes and the lobe is exactly the
  diffuse base; a white base under a clear coat conserves energy (Kelemen-style compensation
  makes the directional albedo Ro + (1 - Ro) = 1 up to LUT interpolation), so the furnace
  returns ~1; a tinted coat absorbs more along longer paths, so thicker coats are darker.
* Area light (light/triangle_light.h): a camera ray that hits an emitter returns its emission
  exactly (no MIS at depth 0); a diffuse floor under an emitting ceiling that fills its
  hemisphere reflects rho * Le (light sampling and BSDF sampling combined by MIS stay unbiased
  only if both pdfs are right); the light pick is proportional to the lights' powers.
"""
import numpy as np
import pytest

import oracle as O
import scenegen as G


def plane_scene(kind, rgb=(0.6, 0.6, 0.6), edge=(1, 1, 1), roughness=0.5, ior=1.5, thickness=1.0, weight=1.0):
    # 2 x 2 km quad at y = 0, facing +y
    v = np.array([[-1000, 0, -1000], [1000, 0, -1000], [1000, 0, 1000], [-1000, 0, 1000]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    mats = np.zeros(2, O.MATERIAL_DTYPE)
    m = mats[1]
    m["kind"] = kind
    m["r"], m["g"], m["b"] = rgb
    m["er"], m["eg"], m["eb"] = edge
    m["roughness"], m["ior"], m["thickness"], m["weight"] = roughness, ior, thickness, weight
    s = O.Scene(v, idx, np.ones(2, np.int32), mats)
    lights = np.zeros(1, O.LIGHT_DTYPE)
    lights[0] = (0, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    s.set_lights(lights)
    cam = np.zeros((), O.CAMERA_DTYPE)
    cam["pos"] = (0, 5, 0)
    # looking straight down: view direction -N = (0, -1, 0)
    cam["U"], cam["V"], cam["N"] = (1, 0, 0), (0, 0, 1), (0, 1, 0)
    cam["fov"] = np.radians(40)
    cam["nearClip"] = 0.01
    return s, cam


def test_miss_returns_ambient_exactly():
    sc = G.spheres_ambient(n_spheres=8, subdiv=1)
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    s.set_lights(sc["lights"])
    acc, _ = s.render(sc["camera"], 64, 32, 0, 1, max_depth=5, record=True)
    miss = acc["_first"] == -1
    assert miss.sum() > 100
    assert np.all(acc["_sampleL"][:, miss] == 1.0)


def test_diffuse_plane_reflects_albedo():
    rho = 0.6
    s, cam = plane_scene(0, rgb=(rho, rho, rho))
    acc, _ = s.render(cam, 32, 32, 0, 256, max_depth=5)
    img = acc["color"] / acc["weight"]
    assert abs(img.mean() - rho) < 0.01 * rho


@pytest.mark.parametrize("roughness", [0.2, 0.7])
def test_white_conductor_furnace(roughness):
    s, cam = plane_scene(1, rgb=(0.99, 0.99, 0.99), edge=(1, 1, 1), roughness=roughness)
    acc, _ = s.render(cam, 32, 32, 0, 256, max_depth=5)
    img = acc["color"] / acc["weight"]
    assert 0.93 < img.mean() < 1.01


def sunsky_scene(lights=None):
    sc = G.spheres_sunsky(n_spheres=4, subdiv=1, palette=4)
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    s.set_lights(sc["lights"] if lights is None else sc["lights"][lights], sc["env"])
    return s


def test_light_sample_matches_eval_over_pdf():
    s = sunsky_scene()
    rng = np.random.default_rng(0)
    smp = s.light_sample(rng.random((4096, 3)).astype(np.float32))
    wi, weight, pdf = smp[:, :3], smp[:, 3:6], smp[:, 6]
    ev = s.env_eval(wi)
    ok = (pdf > 0) & (ev[:, 3] > 0)
    assert ok.mean() > 0.95
    # eval sums all lights; the sampled light's contribution dominates except where a
    # direction falls inside two lights at once, so compare the MIS-free one-sample estimator
    np.testing.assert_allclose((weight * pdf[:, None])[ok].sum(0), ev[ok, :3].sum(0), rtol=0.05)


def test_image_light_pdf_integrates_to_one():
    s = sunsky_scene(lights=[2])  # the sky image light alone
    rng = np.random.default_rng(1)
    n = 400000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ev = s.env_eval(d.astype(np.float32))
    assert abs(ev[:, 3].mean() * 4 * np.pi - 1.0) < 0.01


@pytest.mark.parametrize("which", [0, 1])
def test_cone_light_pdf_times_solid_angle_is_one(which):
    s = sunsky_scene(lights=[which])
    l = s.lights[0]
    axis = np.array([[l["dx"], l["dy"], l["dz"]]], np.float32)
    axis /= np.linalg.norm(axis)
    pdf = s.env_eval(axis)[0, 3]
    solid = 2 * np.pi * (1 - np.cos(np.float64(l["angle"])))
    assert abs(pdf * solid - 1.0) < 1e-3
    # outside the cone the light contributes nothing
    assert s.env_eval(-axis)[0, 3] == 0


def test_bsdf_sample_weight_is_eval_over_pdf(golden):
    flat, _ = O.compute_luts()
    p, geo, smp = golden["bsdf_params"], golden["bsdf_geo"].copy(), golden["bsdf_s"]
    so, st, _ = O.bsdf(flat, p, geo, smp)
    valid = np.abs(so[:, :3]).max(1) > 0
    geo[:, 15:18] = so[:, 3:6]  # evaluate at the sampled direction
    _, _, ev = O.bsdf(flat, p, geo, smp)
    rough = p[:, 7] > 0.05
    m = valid & rough
    np.testing.assert_allclose(so[m, 6], ev[m, 3], rtol=2e-3)
    np.testing.assert_allclose(so[m, :3], ev[m, :3] / ev[m, 3:4], rtol=2e-3, atol=1e-6)


def _coat_rows(golden, weight=None):
    cp = golden["coat_params"].copy()
    if weight is not None:
        cp[:, 9] = weight
    return np.concatenate([np.full((cp.shape[0], 1), 3, np.float32), cp], 1)


def test_coat_with_zero_weight_is_the_diffuse_base(golden):
    flat, _ = O.compute_luts()
    geo, smp = golden["coat_geo"], golden["coat_s"]
    pc = _coat_rows(golden, weight=0.0)
    pd = np.zeros((pc.shape[0], 9), np.float32)
    pd[:, 1:4] = pc[:, 1:4]
    so_c, st_c, ev_c = O.bsdf(flat, pc, geo, smp)
    so_d, st_d, ev_d = O.bsdf(flat, pd, geo, smp)
    # eval: same value and pdf where wo is above the shading normal and wi on the same side of
    # the shading and geometric normals (the coat's validity test, rough_coating.h:309-315);
    # zero elsewhere
    dot = lambda a, b: np.einsum("ij,ij->i", a, b)
    wo, wi, N, Ng = geo[:, 12:15], geo[:, 15:18], geo[:, 6:9], geo[:, 9:12]
    ok_eval = (dot(wo, N) > 0) & (dot(wi, N) * dot(wi, Ng) > 0)
    assert ok_eval.mean() > 0.6
    np.testing.assert_allclose(ev_c[ok_eval, :4], ev_d[ok_eval, :4], rtol=1e-5, atol=1e-7)
    assert (ev_c[~ok_eval, :4] == 0).all()
    np.testing.assert_allclose(ev_c[:, 4:7], ev_d[:, 4:7], rtol=1e-6)
    ok = (np.abs(so_c[:, :3]).max(1) > 0) & (np.abs(so_d[:, :3]).max(1) > 0)
    assert ok.mean() > 0.8
    np.testing.assert_allclose(so_c[ok], so_d[ok], rtol=1e-5, atol=1e-6)
    assert (st_c[ok] == 1).all()


def test_coated_sample_weight_is_eval_over_pdf(golden):
    flat, _ = O.compute_luts()
    p, geo, smp = _coat_rows(golden), golden["coat_geo"].copy(), golden["coat_s"]
    so, st, _ = O.bsdf(flat, p, geo, smp)
    valid = np.abs(so[:, :3]).max(1) > 0
    geo[:, 15:18] = so[:, 3:6]
    _, _, ev = O.bsdf(flat, p, geo, smp)
    np.testing.assert_allclose(so[valid, 6], ev[valid, 3], rtol=2e-3)
    np.testing.assert_allclose(so[valid, :3], ev[valid, :3] / ev[valid, 3:4], rtol=2e-3, atol=1e-6)


@pytest.mark.parametrize("roughness", [0.2, 0.6])
def test_white_coated_diffuse_furnace(roughness):
    s, cam = plane_scene(3, rgb=(1.0, 1.0, 1.0), edge=(1, 1, 1), roughness=roughness)
    acc, _ = s.render(cam, 32, 32, 0, 256, max_depth=6)
    img = acc["color"] / acc["weight"]
    # a diffuse plane of albedo 1 seen through a clear coat: energy conserved within ~2 %
    assert 0.97 < img.mean() < 1.02


def test_tinted_coat_absorbs_with_thickness():
    means = []
    for thickness in (0.0, 0.5, 2.0):
        s, cam = plane_scene(3, rgb=(0.8, 0.8, 0.8), edge=(0.5, 0.9, 0.9), roughness=0.3, thickness=thickness)
        acc, _ = s.render(cam, 32, 32, 0, 64, max_depth=4)
        means.append((acc["color"] / acc["weight"]).reshape(3, -1).mean(1))
    r = [m[0] for m in means]
    assert r[0] > r[1] > r[2]
    # the green/blue channels (coat color 0.9) lose less than red (0.5)
    assert means[2][1] / means[0][1] > means[2][0] / means[0][0]


def _slab(rho, Le, ceiling_y=1.0, half=1000.0):
    """A diffuse floor (y = 0, facing up) under an emitting ceiling (y = ceiling_y, emitting down)."""
    h = half
    v = np.array([[-h, 0, -h], [h, 0, -h], [h, 0, h], [-h, 0, h],
                  [-h, ceiling_y, -h], [h, ceiling_y, -h], [h, ceiling_y, h], [-h, ceiling_y, h]], np.float32)
    # floor normal +y: cross(v2 - v0, v1 - v0) ordering as in plane_scene; ceiling normal -y
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]], np.int32)
    mats = np.zeros(3, O.MATERIAL_DTYPE)
    mats[1]["kind"] = 0
    mats[1]["r"] = mats[1]["g"] = mats[1]["b"] = rho
    mats[2]["kind"] = 0
    mats[2]["emission"] = Le
    s = O.Scene(v, idx, np.array([1, 1, 2, 2], np.int32), mats)
    s.set_lights(np.zeros(0, O.LIGHT_DTYPE))
    return s


def test_area_light_geometry_orientation():
    v = np.array([[-1, 1, -1], [1, 1, -1], [1, 1, 1]], np.float32)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    assert n[1] < 0  # the ceiling winding emits downward


def test_direct_view_of_an_emitter_is_exact():
    Le = (2.0, 3.0, 0.5)
    s = _slab(0.5, Le, ceiling_y=1.0)
    cam = np.zeros((), O.CAMERA_DTYPE)
    cam["pos"] = (0, 0.5, 0)
    cam["U"], cam["V"], cam["N"] = (1, 0, 0), (0, 0, 1), (0, -1, 0)  # looking up (+y)
    cam["fov"] = np.radians(30)
    cam["nearClip"] = 0.01
    acc, _ = s.render(cam, 32, 32, 0, 1, max_depth=3, record=True)
    assert s.light_count() == 1
    np.testing.assert_array_equal(acc["_sampleL"], np.array(Le, np.float32)[:, None].repeat(32 * 32, 1))


@pytest.mark.parametrize("rho", [0.3, 0.8])
def test_floor_under_emitting_ceiling_reflects_rho_le(rho):
    Le = 2.0
    s = _slab(rho, (Le, Le, Le), ceiling_y=1.0)
    cam = np.zeros((), O.CAMERA_DTYPE)
    cam["pos"] = (0, 0.5, 0)
    cam["U"], cam["V"], cam["N"] = (1, 0, 0), (0, 0, 1), (0, 1, 0)  # looking down
    cam["fov"] = np.radians(20)
    cam["nearClip"] = 0.01
    acc, _ = s.render(cam, 32, 32, 0, 128, max_depth=1)
    img = acc["color"] / acc["weight"]
    # the 2 km ceiling at height 1 covers the floor's hemisphere up to ~1e-6 of its cosine mass
    assert abs(img.mean() / (rho * Le) - 1.0) < 0.01


def test_light_pick_proportional_to_power():
    Le = 2.0
    s = _slab(0.5, (Le, Le, Le), half=10.0)
    amb = np.zeros(1, O.LIGHT_DTYPE)
    amb[0] = (0, 0.7, 0.7, 0.7, 0, 0, 0, 0)
    s.set_lights(amb)
    cdf = s.light_cdf()
    assert s.light_count() == 2 and cdf.shape == (3,)
    # TriangleLight::getPower = pi * sum |cross(e1, e2)| * luminance = pi * 2 * area * Le
    area_power = np.pi * 2 * (20.0 * 20.0) * Le
    amb_power = 4 * np.pi ** 2 * s.radius ** 2 * 0.7
    np.testing.assert_allclose(cdf[1], area_power / (area_power + amb_power), rtol=1e-5)


def _shadow_scene(transparency, thin=False):
    """Floor at y = -1 (diffuse 0.5) under a distant light from straight above, with a black
    quad at y = 1 between the floor and the light (the camera at the origin looks along +z, so
    its rays to the floor do not cross the quad)."""
    v = np.array([[-100, -1, -10], [100, -1, -10], [100, -1, 100], [-100, -1, 100],
                  [-100, 1, -10], [100, 1, -10], [100, 1, 100], [-100, 1, 100]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]], np.int32)
    mats = np.zeros(3, O.MATERIAL_DTYPE)
    mats[1]["kind"] = G.KIND_DIFFUSE
    mats[1]["r"] = mats[1]["g"] = mats[1]["b"] = 0.5
    if thin:
        mats[2]["kind"] = G.KIND_THIN_DIELECTRIC
        mats[2]["r"], mats[2]["g"], mats[2]["b"] = 0.9, 0.6, 0.3
        mats[2]["roughness"], mats[2]["ior"], mats[2]["thickness"] = 0.3, 1.5, 0.5
    else:
        mats[2]["kind"] = G.KIND_DIFFUSE  # black: its lobe has no importance, only the transparent BTDF is sampled
        mats[2]["transparency"] = transparency
    s = O.Scene(v, idx, np.array([1, 1, 2, 2], np.int32), mats)
    lights = np.zeros(1, O.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_DISTANT, 3.0, 3.0, 3.0, 0, 1, 0, np.radians(1.0))
    s.set_lights(lights)
    return s


def test_transparent_shadow_scales_direct_light():
    # render/transparent_filter.h:42-93: a shadow ray crossing a transparent (non-thin) surface
    # keeps 1 - opacity of the light.  The floor's bounce rays leave through the black quad into
    # an empty sky either way, so every floor pixel is exactly t times its unshadowed value.
    W, H, spp = 64, 32, 4
    cam = G._camera()
    open_sky = _shadow_scene(0.0)
    ref, _ = _shadow_scene(1.0).render(cam, W, H, 0, spp, max_depth=4, record=True)  # quad fully transparent
    for t in (0.37, 0.8):
        acc, _ = _shadow_scene(t).render(cam, W, H, 0, spp, max_depth=4, record=True)
        floor = (acc["_first"] >= 0) & (acc["_first"] < 2)
        assert floor.sum() > 300
        img, base = acc["color"][:, floor], ref["color"][:, floor]
        assert base.min() > 0
        np.testing.assert_allclose(img, t * base, rtol=2e-6)
    # an opaque quad (transparency 0) blocks the direct light completely
    acc, _ = open_sky.render(cam, W, H, 0, spp, max_depth=4, record=True)
    floor = (acc["_first"] >= 0) & (acc["_first"] < 2)
    assert np.all(acc["color"][:, floor] == 0)


def test_thin_dielectric_shadow_transmittance():
    # a thin rough dielectric between the floor and a light from straight above: the shadow ray
    # keeps (1 - R(cos 1)) pow(color, thickness / cos_t) (thin_rough_dielectric.h:307-321), the
    # same value the golden-pinned transparency gives at normal incidence
    W, H, spp = 64, 32, 4
    cam = G._camera()
    ref, _ = _shadow_scene(1.0).render(cam, W, H, 0, spp, max_depth=1, record=True)
    acc, _ = _shadow_scene(0.0, thin=True).render(cam, W, H, 0, spp, max_depth=1, record=True)
    floor = (acc["_first"] >= 0) & (acc["_first"] < 2)
    ratio = acc["color"][:, floor] / ref["color"][:, floor]
    # the light's cone is 1 degree: the ratio varies by well under 1e-3 around its mean
    expect = ratio.mean(axis=1, keepdims=True)
    np.testing.assert_allclose(ratio, np.broadcast_to(expect, ratio.shape), rtol=2e-3)
    flat, _ = O.compute_luts()
    m = np.zeros(1, O.MATERIAL_DTYPE)
    m[0] = _thin_record()
    geo = np.array([[1, 0, 0, 0, 0, 1, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, -1, 0]], np.float32)  # N = Ng = wo = +y
    _, _, _, aux = O.bsdf_mat(flat, m, geo, np.zeros((1, 3), np.float32))
    np.testing.assert_allclose(expect[:, 0], aux[0, 6:9], rtol=2e-3)


def _thin_record():
    r = np.zeros((), O.MATERIAL_DTYPE)
    r["kind"] = G.KIND_THIN_DIELECTRIC
    r["r"], r["g"], r["b"] = 0.9, 0.6, 0.3
    r["roughness"], r["ior"], r["thickness"] = 0.3, 1.5, 0.5
    return r
