"""HIP path (through the C ABI) against the oracle and the reference's golden vectors.

Bars: integer work (sampler values, closest-hit primitive IDs and barycentrics) bit-exact;
per-pixel FP32 radiance within 1e-3 relative (1e-5 absolute floor) on identical paths, i.e.
samples whose hit / lobe event hash matches the oracle's.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
import protoray_hip as P
import scenegen as G
from conftest import results_dir

pytestmark = pytest.mark.gpu

RTOL = 1e-3
ATOL = 1e-5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def test_info(dev):
    info = dev.getInfo()
    print(info)
    assert "gfx950" in info


def test_sampler_bit_exact(dev, golden):
    o1, o2, o3 = dev.sampler(golden["sampler_in"], 64)
    assert np.array_equal(o1.view(np.uint32), golden["sampler_1d"].view(np.uint32))
    assert np.array_equal(o2.view(np.uint32), golden["sampler_2d"].view(np.uint32))
    assert np.array_equal(o3.view(np.uint32), golden["sampler_3d"].view(np.uint32))


def test_luts(dev, golden):
    lut = dev.luts()
    _, orc = O.compute_luts()
    off = 0
    for name, n in O.LUT_SIZES:
        np.testing.assert_allclose(lut[off:off + n], orc[name], atol=1e-6, err_msg=name)
        off += n
    np.testing.assert_allclose(lut[:1024], golden["lut_ggx_dir"].ravel(), atol=5e-6)
    np.testing.assert_allclose(lut[-256:], golden["lut_sheen_dir"].ravel(), atol=5e-6)


def _random_scene(seed, ntri=3000):
    rng = np.random.default_rng(seed)
    c = rng.uniform([-10, -2, 0], [10, 8, 30], size=(ntri, 3))
    v = (c[:, None, :] + rng.normal(size=(ntri, 3, 3)) * 0.8).astype(np.float32).reshape(-1, 3)
    idx = np.arange(ntri * 3, dtype=np.int32).reshape(-1, 3)
    # a shared-edge fan so exact-t ties occur on shared vertices
    fan_c = np.array([0, 3, 15], np.float32)
    fan = [fan_c]
    for k in range(8):
        a = 2 * np.pi * k / 8
        fan.append(fan_c + np.array([np.cos(a), np.sin(a), 0.3], np.float32))
    base = v.shape[0]
    v = np.concatenate([v, np.array(fan, np.float32)])
    fidx = np.array([[base, base + 1 + k, base + 1 + (k + 1) % 8] for k in range(8)], np.int32)
    idx = np.concatenate([idx, fidx])
    return v, idx


def _random_rays(seed, n, toward=None):
    rng = np.random.default_rng(seed)
    o = rng.uniform([-12, -3, -5], [12, 9, 5], size=(n, 3))
    if toward is None:
        d = rng.normal(size=(n, 3))
    else:
        d = toward[rng.integers(0, len(toward), n)] - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    far = np.full(n, 3.4028234663852886e38)
    far[: n // 4] = rng.uniform(1, 30, n // 4)  # bounded segments
    return np.concatenate([o, d, np.zeros((n, 1)), far[:, None]], axis=1).astype(np.float32)


def test_closest_hit_matches_brute_force(dev):
    v, idx = _random_scene(1)
    tri_mat = np.ones(idx.shape[0], np.int32)
    mats = np.zeros(2, P.MATERIAL_DTYPE)
    dev.initScene(v, idx, tri_mat, mats)
    stats = dev.bvhStats()
    print(stats)
    assert stats["triangles"] == idx.shape[0]
    scene = O.Scene(v, idx, tri_mat, np.zeros(2, O.MATERIAL_DTYPE))
    centroids = v[idx].mean(axis=1)
    rays = _random_rays(2, 20000, toward=centroids)
    # rays through the fan's shared centre vertex hit several triangles at the same t
    fan_rays = _random_rays(3, 64)
    fan_rays[:, 3:6] = np.array([0, 3, 15], np.float32) - fan_rays[:, :3]
    fan_rays[:, 3:6] /= np.linalg.norm(fan_rays[:, 3:6], axis=1, keepdims=True)
    rays = np.concatenate([rays, fan_rays])
    prim, t, u, w = dev.queryRays(rays)
    bprim, bt, bu, bw = scene.intersect(rays, mode=1)
    assert (prim >= 0).sum() > 10000
    assert np.array_equal(prim, bprim)
    hit = prim >= 0
    assert np.array_equal(t[hit].view(np.uint32), bt[hit].view(np.uint32))
    assert np.array_equal(u[hit].view(np.uint32), bu[hit].view(np.uint32))
    assert np.array_equal(w[hit].view(np.uint32), bw[hit].view(np.uint32))
    occ = dev.queryOcclusion(rays)
    bocc, _, _, _ = scene.intersect(rays, mode=3)
    assert np.array_equal(occ, (bocc >= 0).astype(np.int32))


def _record_fraction(name, spp, frac):
    out = results_dir()
    if out:
        with open(os.path.join(out, "parity_fractions.txt"), "a") as f:
            f.write(f"{name} pass {spp}: {frac:.5f}\n")


def _parity(dev, scene, W, H, spp, min_match):
    dev.enableDebug(True)
    P.load_scene(dev, scene)
    dev.initFrame(W, H, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"], scene.get("normals"))
    osc.set_textures(scene)  # UVs, textures and instances (before the lights: the radius)
    osc.set_lights(scene["lights"], scene.get("env"))
    acc = None
    for s in range(spp):
        dev.render(1)
        L, hsh, first = dev.debugOutputs()
        acc, _ = osc.render(scene["camera"], W, H, spp_start=s, spp_count=1, seed=0, max_depth=scene["max_depth"],
                            rr_depth=scene.get("rr_depth", 8), aovs=("albedo1", "normal1"), accum=acc, record=True)
        ref_first = acc["_first"]
        # first hits are pure intersection work: bit-exact
        assert np.array_equal(first, ref_first), f"spp {s}: {np.sum(first != ref_first)} first-hit mismatches"
        same = hsh == acc["_hash"]
        frac = same.mean()
        print(f"spp {s}: identical paths {frac:.4f}")
        _record_fraction(scene["name"], s, frac)
        assert frac >= min_match
        ref = acc["_sampleL"]
        bad = same & ~np.isclose(L, ref, rtol=RTOL, atol=ATOL).all(axis=0)
        if bad.any():
            out = results_dir()
            if out:
                np.savez(os.path.join(out, f"parity_fail_{scene['name']}_{s}.npz"), L=L, hsh=hsh,
                         first=first, refL=ref, refHash=acc["_hash"], refFirst=ref_first, bad=np.flatnonzero(bad))
            print("mismatched pixels", np.flatnonzero(bad), L[:, bad].T, ref[:, bad].T)
        np.testing.assert_allclose(L[:, same], ref[:, same], rtol=RTOL, atol=ATOL)
    gs, gw = dev.getAccum()
    assert np.array_equal(gw, acc["weight"])
    img = dev.blitFrame()
    assert np.isfinite(img).all()
    for buf, name in ((P.BUF_ALBEDO1, "albedo1"), (P.BUF_NORMAL1, "normal1")):
        got = dev.blitFrame(buf).reshape(-1, 3).T
        assert np.isfinite(got).all(), name
        ref = acc[name] / np.maximum(acc[name + "W"], 1)
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5, err_msg=name)
    dev.enableDebug(False)


def test_render_parity_ambient_diffuse(dev):
    scene = G.spheres_ambient(n_spheres=24, subdiv=2)
    _parity(dev, scene, 64, 64, 4, 0.998)


def test_render_parity_sunsky_mixed(dev):
    scene = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24)
    _parity(dev, scene, 96, 64, 3, 0.998)


def test_render_parity_sunsky_coated(dev):
    # every diffuse palette entry coated (kind 3, RoughCoatingBsdf over a diffuse base)
    scene = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24, coated=1.0)
    assert (scene["materials"]["kind"] == G.KIND_COATED).sum() >= 5
    _parity(dev, scene, 96, 64, 3, 0.998)


def test_render_parity_default_depth(dev):
    # the reference's default maxDepth of 100 (render/path_renderer_stream.h:114): Russian roulette from
    # depth 8 on, depths beyond 9 enqueued only while paths survive, and sampler dimensions up to
    # 4 + 7 * 100
    scene = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24)
    scene["max_depth"] = 100
    _parity(dev, scene, 96, 64, 3, 0.998)
    st = dev.render(1)
    assert st["spp"] >= 1


def test_render_parity_remaining_lobes(dev):
    # Oren-Nayar bases (alone and under a coat), thin rough dielectrics, opacity mixes with the
    # transparent BTDF (one material fully transparent), sheen layers over every kind, and shadow
    # rays through the thin and transparent surfaces (the transparent-shadow filter)
    scene = G.spheres_lobes()
    kinds = scene["materials"]["kind"]
    assert (kinds == G.KIND_THIN_DIELECTRIC).sum() >= 3 and (scene["materials"]["transparency"] > 0).sum() >= 9
    assert (scene["materials"]["sheen_weight"] > 0).sum() >= 10
    _parity(dev, scene, 96, 64, 4, 0.998)


def test_render_parity_standard_materials(dev):
    # kind-5 StandardMaterials (material/standard.cpp:409-627): the opaque, translucent and metal
    # bases mixed by weight (partial metalness, partial transmission, base / specular weights below
    # 1), coats over metal, glass and a diffuse-only base, coloured solid glass whose interior
    # medium absorbs (kind 5 and kind 2), thin-walled partial transmission under a coat (the
    # transparent-shadow filter), opacity and sheen over the mix
    scene = G.spheres_standard()
    kinds = scene["materials"]["kind"]
    assert (kinds == G.KIND_STANDARD).sum() >= 20 and (kinds == G.KIND_DIELECTRIC).sum() >= 3
    _parity(dev, scene, 96, 64, 4, 0.998)


def test_render_parity_textured_materials(dev):
    # image textures on base colours (diffuse with Lambert and Oren-Nayar bases, coated diffuse):
    # per-vertex UVs, an sRGB RGBA8 checker tiled with repeat addressing, an FP32 gradient with
    # clamp addressing, a linear RGBA8 texture with mixed addressing, UV scales and offsets
    scene = G.spheres_textured()
    assert len(scene["material_textures"]) >= 10
    _parity(dev, scene, 96, 64, 4, 0.998)


def test_render_parity_instanced_scene(dev):
    # a two-level scene: six meshes in object space (smooth normals, a textured Oren-Nayar base,
    # coated, conductor, dielectric media interface, thin dielectric seen by shadow rays through
    # the transparent-shadow filter) placed by 41 instances with rotations, non-uniform scales and
    # translations
    scene = G.spheres_instanced()
    _parity(dev, scene, 96, 64, 4, 0.998)
    dev.initInstances(None, [])  # back to one level for the other tests on this context


@pytest.mark.parametrize("ambient", [0.5, None])
def test_render_parity_area_lights(dev, ambient):
    # emissive spheres (one TriangleLight over their triangles) with and without an ambient light
    scene = G.spheres_ambient(n_spheres=24, subdiv=2, emissive=6, ambient=ambient)
    _parity(dev, scene, 64, 64, 4, 0.998)


@pytest.mark.parametrize("instanced", [False, True])
def test_render_parity_scene_file(dev, tmp_path, instanced):
    # a glTF file (smooth-normal spheres instanced by a node tree, a ground quad, an emissive quad,
    # Standard / metal / glass / emissive materials) through prt_scene_file_load, rendered with
    # interpolated vertex normals, against the oracle on the same arrays
    import scene_files as S
    w = S.GltfWriter()
    w.doc["materials"] = [
        {"name": "plastic", "pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.3, 0.2, 1], "metallicFactor": 0.0,
                                                     "roughnessFactor": 0.4}},
        {"name": "metal", "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.8, 0.5, 1], "metallicFactor": 1.0,
                                                   "roughnessFactor": 0.3}},
        {"name": "glass", "pbrMetallicRoughness": {"metallicFactor": 0.0, "roughnessFactor": 0.15},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0},
                        "KHR_materials_volume": {"thicknessFactor": 1.0}}},
        {"name": "lamp", "emissiveFactor": [1.0, 0.9, 0.8],
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 6.0}}},
        {"name": "floor", "pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.5, 0.5, 1], "metallicFactor": 0.0,
                                                   "roughnessFactor": 0.8},
         "extensions": {"KHR_materials_specular": {"specularFactor": 0.0}}},
    ]
    sp, si = S.icosphere(2)
    pa, na, ia = w.accessor(sp, 5126, "VEC3"), w.accessor(sp, 5126, "VEC3"), w.accessor(si.ravel(), 5125, "SCALAR")
    spheres = [w.mesh([{"attributes": {"POSITION": pa, "NORMAL": na}, "indices": ia, "material": m}]) for m in range(3)]
    quad = np.array([[-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 0, 1]], np.float32)
    qa, qi = w.accessor(quad, 5126, "VEC3"), w.accessor(np.array([0, 2, 1, 1, 2, 3]), 5125, "SCALAR")
    floor = w.mesh([{"attributes": {"POSITION": qa}, "indices": qi, "material": 4}])
    lamp = w.mesh([{"attributes": {"POSITION": qa}, "indices": qi, "material": 3}])
    rng = np.random.default_rng(3)
    kids = [w.node(mesh=spheres[i % 3], translation=[float(rng.uniform(-4, 4)), float(rng.uniform(0, 2.5)),
                                                     float(rng.uniform(6, 14))],
                   scale=[float(rng.uniform(0.4, 1.0))] * 3) for i in range(12)]
    w.node(children=kids, root=True, rotation=S.quaternion((0, 1, 0), 0.2))
    w.node(mesh=floor, root=True, translation=[0, -1, 10], scale=[20, 1, 20])
    w.node(mesh=lamp, root=True, translation=[0, 6, 10], rotation=S.quaternion((1, 0, 0), math.pi), scale=[2, 1, 2])
    path = tmp_path / "parity.gltf"
    w.save(str(path))
    sc = P.load_scene_file(path, instanced=instanced)
    if instanced:  # the spheres' three meshes and the floor in two levels, the emissive lamp flattened
        assert len(sc["instances"]) == 14 and len(sc["mesh_begin"]) == 6
    else:
        assert sc["info"]["instance_count"] == 14
    assert sc["normals"] is not None
    lights = np.zeros(2, G.LIGHT_DTYPE)
    d = np.array([-0.4, 0.8, -0.45])
    d /= np.linalg.norm(d)
    lights[0] = (G.LIGHT_AMBIENT, 0.3, 0.35, 0.4, 0, 0, 0, 0)
    lights[1] = (G.LIGHT_DISTANT, 20.0, 19.0, 17.0, d[0], d[1], d[2], math.radians(2.0))
    sc.update(name="gltf_file_two_level" if instanced else "gltf_file", lights=lights, env=None, camera=G._camera(),
              max_depth=5, seed=0)
    _parity(dev, sc, 64, 64, 3, 0.998)
    if instanced:
        dev.initInstances(None, [])


def test_render_stats_and_error(dev):
    scene = G.spheres_ambient(n_spheres=16, subdiv=1)
    P.load_scene(dev, scene)
    dev.initFrame(64, 32)
    st = dev.render(4)
    assert st["spp"] == 4 and st["rays"] > 64 * 32 * 4
    err = dev.getFrameError()
    assert 0.0 <= err < 1.0
    with pytest.raises(ValueError):
        dev.initFrame(60, 32)  # width not a multiple of 32 (path_renderer_stream.h:109-110)


def _band_parity(dev, scene, rows, spp_check, spp_total, min_match):
    """Full-size frame on the GPU; the oracle renders a band of rows of the same passes."""
    W, H = scene["width"], scene["height"]
    dev.enableDebug(True)
    P.load_scene(dev, scene)
    dev.initFrame(W, H)
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"])
    osc.set_lights(scene["lights"], scene.get("env"))
    band = np.zeros(W * H, bool)
    band[rows[0] * W:rows[1] * W] = True
    acc = None
    for s in range(spp_check):
        dev.render(1)
        L, hsh, first = dev.debugOutputs()
        acc, _ = osc.render(scene["camera"], W, H, spp_start=s, spp_count=1, max_depth=scene["max_depth"], rows=rows,
                            accum=acc, record=True)
        assert np.array_equal(first[band], acc["_first"][band])
        same = band & (hsh == acc["_hash"])
        frac = same.sum() / band.sum()
        print(f"{scene['name']} spp {s}: identical paths on the band {frac:.4f}")
        _record_fraction(scene["name"] + "_band", s, frac)
        assert frac >= min_match
        np.testing.assert_allclose(L[:, same], acc["_sampleL"][:, same], rtol=RTOL, atol=ATOL)
    dev.enableDebug(False)
    if spp_total > spp_check:
        dev.render(spp_total - spp_check)
    _, w = dev.getAccum()
    assert (w == spp_total).all()  # exactly one sample per pixel per pass
    img = dev.blitFrame()
    assert np.isfinite(img).all() and (img >= 0).all()


def test_full_size_spheres_ambient_512(dev):
    # 512x512, 16 spp, depth 5, 327,682 triangles
    _band_parity(dev, G.spheres_ambient(), (240, 272), 2, 16, 0.998)


def test_full_size_spheres_sunsky_1080p(dev):
    # 1920x1080, depth 8, 20,971,522 triangles; one pass checked on a band, 4 passes rendered
    _band_parity(dev, G.spheres_sunsky(), (536, 544), 1, 4, 0.998)
