"""The oracle (CPU restatement) against golden vectors produced by the reference's own code.

Fixtures: tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py from the
harness oracle/ref_golden.cpp compiled against /root/reference.  Integer work is compared
bit for bit; float work to tolerances that cover the reference's rcp/rsqrt approximations
(math/simd/vfloat8_avx2.h:227-232,327-342) and ISPC transcendentals (math/simd.h:150-657).
"""
import numpy as np
import pytest

import oracle as O


def test_sampler_bit_exact(golden):
    # sampling/sobol_burley_sampler_simd.h:36-85 via sobol_burley.h:222-297
    o1, o2, o3 = O.sampler(golden["sampler_in"], 64)
    assert np.array_equal(o1.view(np.uint32), golden["sampler_1d"].view(np.uint32))
    assert np.array_equal(o2.view(np.uint32), golden["sampler_2d"].view(np.uint32))
    assert np.array_equal(o3.view(np.uint32), golden["sampler_3d"].view(np.uint32))


def test_sampler_closed_forms():
    # Dimension 0 of a 1D sample is Owen-scrambled van der Corput: over 2^k consecutive
    # indices the samples are stratified, one per interval of width 2^-k.
    seeds = np.array([[7, 3 | (5 << 16), i] for i in range(256)], np.uint32)
    o1, o2, _ = O.sampler(seeds, 8)
    for d in range(8):
        assert sorted(np.floor(o1[:, d] * 256).astype(int)) == list(range(256))
    # A 2D set is a (0,8,2)-net: every 16x16 elementary box holds exactly one point
    for d in range(8):
        cells = np.floor(o2[:, d, 0] * 16).astype(int) * 16 + np.floor(o2[:, d, 1] * 16).astype(int)
        assert len(np.unique(cells)) == 256
        cells = np.floor(o2[:, d, 0] * 256).astype(int) * 1 + np.floor(o2[:, d, 1] * 1).astype(int)
        assert len(np.unique(cells)) == 256
    assert (o1 >= 0).all() and (o1 < 1).all()


def test_morton_order(golden):
    # image/morton.h:32-39 for 32-pixel-wide tiles
    assert np.array_equal(O.morton(1024, 32), golden["morton32"])


def test_camera_rays(golden):
    # camera/pinhole_camera.h:128-133, camera/thin_lens_camera.h:175-182
    for c in range(golden["camera_cfg"].shape[0]):
        out = O.camera_rays(golden["camera_cfg"][c], golden["camera_s"][c])
        ref = golden["camera_rays"][c]
        np.testing.assert_allclose(out[..., :6], ref[..., :6], atol=1e-6, rtol=0)
        np.testing.assert_allclose(out[..., 6], ref[..., 6], rtol=2e-5)


def test_shape_warps(golden):
    c, sp, d, cn = O.warps(golden["warp_s"])
    np.testing.assert_allclose(c, golden["warp_cosine"], atol=1e-6)
    np.testing.assert_allclose(sp, golden["warp_sphere"], atol=1e-6)
    np.testing.assert_allclose(d, golden["warp_disk"], atol=1e-6)
    ref_cone = golden["warp_cone"].transpose(0, 2, 1, 3).reshape(-1, 2, 4)
    np.testing.assert_allclose(cn[..., :3], ref_cone[..., :3], atol=5e-6)
    np.testing.assert_allclose(cn[..., 3], ref_cone[..., 3], rtol=1e-5)


def test_distributions(golden):
    idx, pdf_sr, pdf_i = O.dist1d(golden["dist1_w"], golden["dist1_s"])
    assert np.array_equal(idx, golden["dist1_idx"])
    np.testing.assert_allclose(pdf_sr, golden["dist1_pdf_sr"], atol=1e-6)
    np.testing.assert_allclose(pdf_i, golden["dist1_pdf_i"], atol=1e-7)
    idx, pdf_sr, cells = O.dist2d(golden["dist2_w"], golden["dist2_s"])
    assert np.array_equal(idx, golden["dist2_idx"])
    np.testing.assert_allclose(pdf_sr[:, 0], golden["dist2_pdf_sr"][:, 0], rtol=1e-5)
    np.testing.assert_allclose(pdf_sr[:, 1:], golden["dist2_pdf_sr"][:, 1:], atol=2e-5)
    np.testing.assert_allclose(cells, golden["dist2_pdf_cells"], atol=1e-7)


def test_energy_compensation_luts(golden):
    # bsdf/albedo_tables.cpp:173-186.  The column cosThetaO = 0 at eta = 1 is a singular limit
    # (grazing VNDF sampling at alpha = 0.001) where both sides amplify rounding differences; no
    # supported material reads it (IOR is clamped to [1/3, 3] and eta = 1 means no interface).
    _, L = O.compute_luts()
    np.testing.assert_allclose(L["ggx_dir"], golden["lut_ggx_dir"].ravel(), atol=5e-6)
    np.testing.assert_allclose(L["ggx_avg"], golden["lut_ggx_avg"].ravel(), atol=5e-6)
    eta_dir = L["diel_eta_dir"].reshape(16, 16, 16).copy()
    ref_eta_dir = golden["lut_diel_eta_dir"].copy()
    eta_dir[:, 15, 0] = ref_eta_dir[:, 15, 0] = 0
    np.testing.assert_allclose(eta_dir, ref_eta_dir, atol=2e-5)
    rcp_dir = L["diel_rcpeta_dir"].reshape(16, 16, 16).copy()
    ref_rcp_dir = golden["lut_diel_rcpeta_dir"].copy()
    np.testing.assert_allclose(rcp_dir[:, 0, 0], ref_rcp_dir[:, 0, 0], atol=2e-3)
    rcp_dir[:, 0, 0] = ref_rcp_dir[:, 0, 0] = 0
    np.testing.assert_allclose(rcp_dir, ref_rcp_dir, atol=2e-5)
    np.testing.assert_allclose(L["diel_rcpeta_avg"], golden["lut_diel_rcpeta_avg"].ravel(), atol=2e-5)
    # the eta average integrates the singular column with weight ~0: still within 1e-3
    np.testing.assert_allclose(L["diel_eta_avg"], golden["lut_diel_eta_avg"].ravel(), atol=1e-3)
    # MicrofacetDielectricReflectionAlbedoTable (albedo_tables.cpp:109-116), read by the coating
    refl_dir = L["diel_refl_dir"].reshape(16, 16, 16).copy()
    ref_refl_dir = golden["lut_diel_refl_dir"].copy()
    refl_dir[:, 15, 0] = ref_refl_dir[:, 15, 0] = 0
    np.testing.assert_allclose(refl_dir, ref_refl_dir, atol=2e-5)
    np.testing.assert_allclose(L["diel_refl_avg"], golden["lut_diel_refl_avg"].ravel(), atol=1e-3)
    # MicrofacetSheenAlbedoTable (albedo_tables.cpp:155-170)
    np.testing.assert_allclose(L["sheen_dir"], golden["lut_sheen_dir"].ravel(), atol=5e-6)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bsdf_sample_eval(golden, kind):
    # bsdf/diffuse.h:55-68, rough_conductor.h:132-202, rough_dielectric.h:296-526 wrapped in
    # MultiBsdfSimd (multi_bsdf.h:166-210), evaluated on the same frames, directions and samples
    flat, _ = O.compute_luts()
    p, geo, s = golden["bsdf_params"], golden["bsdf_geo"], golden["bsdf_s"]
    so, st, ev = O.bsdf(flat, p, geo, s)
    m = p[:, 0] == kind
    ref_s, ref_t, ref_e = golden["bsdf_sample"][m], golden["bsdf_sample_type"][m], golden["bsdf_eval"][m]
    so, st, ev = so[m], st[m], ev[m]
    valid_ref = np.abs(ref_s[:, :3]).max(1) > 0
    valid = np.abs(so[:, :3]).max(1) > 0
    assert np.array_equal(valid, valid_ref)
    assert np.array_equal(st[valid], ref_t[valid])
    np.testing.assert_allclose(so[valid, :3], ref_s[valid, :3], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(so[valid, 3:6], ref_s[valid, 3:6], atol=2e-5)
    rough0 = p[m][valid, 7] == 0
    np.testing.assert_allclose(so[valid][~rough0, 6], ref_s[valid][~rough0, 6], rtol=1e-3)
    np.testing.assert_allclose(so[valid][rough0, 6], ref_s[valid][rough0, 6], rtol=1e-2)
    np.testing.assert_allclose(ev[:, :3], ref_e[:, :3], rtol=5e-4, atol=1e-6)
    erough0 = p[m][:, 7] == 0
    np.testing.assert_allclose(ev[~erough0, 3], ref_e[~erough0, 3], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(ev[erough0, 3], ref_e[erough0, 3], rtol=1e-2, atol=1e-6)


def test_coated_diffuse_sample_eval(golden):
    # bsdf/rough_coating.h:297-593 over bsdf/diffuse.h:55-68 in MultiBsdfSimd: absorption through
    # the coat, Kelemen energy conservation, coat/base pick by importance; albedo :567-571
    flat, _ = O.compute_luts()
    cp = golden["coat_params"]
    p = np.concatenate([np.full((cp.shape[0], 1), 3, np.float32), cp], 1)
    so, st, ev = O.bsdf(flat, p, golden["coat_geo"], golden["coat_s"])
    ref_s, ref_t, ref_e = golden["coat_sample"], golden["coat_sample_type"], golden["coat_eval"]
    valid_ref = np.abs(ref_s[:, :3]).max(1) > 0
    valid = np.abs(so[:, :3]).max(1) > 0
    assert valid.sum() > 0.8 * len(valid)
    assert np.array_equal(valid, valid_ref)
    assert np.array_equal(st[valid], ref_t[valid])
    assert set(np.unique(st[valid])) == {1, 2}  # both the base and the coat get sampled
    np.testing.assert_allclose(so[valid, :3], ref_s[valid, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(so[valid, 3:6], ref_s[valid, 3:6], atol=2e-5)
    np.testing.assert_allclose(so[valid, 6], ref_s[valid, 6], rtol=1e-3)
    assert (np.abs(ev[:, :3]).max(1) > 0).mean() > 0.6
    np.testing.assert_allclose(ev[:, :3], ref_e[:, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(ev[:, 3], ref_e[:, 3], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(ev[:, 4:7], ref_e[:, 4:7], rtol=1e-5)


def lobe_materials(golden):
    """Material records of the golden lobe rows (MATERIAL_DTYPE)."""
    p = golden["lobe_params"]
    m = np.zeros(p.shape[0], O.MATERIAL_DTYPE)
    m["kind"] = p[:, 0].astype(np.int32)
    for k, f in enumerate(("r", "g", "b", "er", "eg", "eb", "roughness", "ior", "thickness", "weight", "transparency",
                           "diffuse_roughness")):
        m[f] = p[:, 1 + k]
    return m


@pytest.mark.parametrize("variant", range(6))
def test_remaining_lobes_sample_eval(golden, variant):
    # bsdf/rough_diffuse.h (Oren-Nayar, Fujii), a coating over it, bsdf/thin_rough_dielectric.h,
    # and the opacity mix with bsdf/transparent.h in MultiBsdfSimd (multi_bsdf.h:151-459), composed
    # as material/standard.cpp:409-627 composes them; variants: 0 rough diffuse, 1 coated rough
    # diffuse, 2 thin dielectric, 3 rough diffuse + transparency, 4 thin dielectric +
    # transparency, 5 conductor + transparency (every 4th batch fully transparent)
    flat, _ = O.compute_luts()
    mats, geo, s = lobe_materials(golden), golden["lobe_geo"], golden["lobe_s"]
    so, st, ev, aux = O.bsdf_mat(flat, mats, geo, s)
    sel = (np.arange(len(mats)) // 8) % 6 == variant
    ref_s, ref_t, ref_e = golden["lobe_sample"][sel], golden["lobe_sample_type"][sel], golden["lobe_eval"][sel]
    so, st, ev, aux = so[sel], st[sel], ev[sel], aux[sel]
    valid_ref = np.abs(ref_s[:, :3]).max(1) > 0
    valid = np.abs(so[:, :3]).max(1) > 0
    assert valid.mean() > 0.5
    assert np.array_equal(valid, valid_ref)
    assert np.array_equal(st, ref_t)
    # the reference's SIMD code leaves the reflected direction in rejected lanes; a zero weight
    # ends the path either way, so directions are compared where the weight is non-zero
    np.testing.assert_allclose(so[valid, :3], ref_s[valid, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(so[valid, 3:6], ref_s[valid, 3:6], atol=2e-5)
    np.testing.assert_allclose(so[:, 6], ref_s[:, 6], rtol=1e-3, atol=1e-7)
    np.testing.assert_allclose(ev[:, :3], ref_e[:, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(ev[:, 3], ref_e[:, 3], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(aux[:, 0:3], golden["lobe_albedo"][sel], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(aux[:, 3:6], golden["lobe_normal"][sel], atol=1e-6)
    if variant in (2, 4):  # thin-walled: what a shadow ray keeps through the surface
        np.testing.assert_allclose(aux[:, 6:9], golden["lobe_transparency"][sel], rtol=5e-4, atol=1e-6)
    if variant >= 3:
        types = set(np.unique(st[valid]))
        assert 128 in types  # the transparent BTDF gets sampled (BsdfType::Transparency)
        assert len(types) > 1 or variant == 5


def sheen_materials(golden):
    """Material records of the golden sheen rows (MATERIAL_DTYPE)."""
    p = golden["sheen_params"]
    m = np.zeros(p.shape[0], O.MATERIAL_DTYPE)
    m["kind"] = p[:, 0].astype(np.int32)
    for k, f in enumerate(("r", "g", "b", "er", "eg", "eb", "roughness", "ior", "thickness", "weight", "transparency",
                           "diffuse_roughness", "sheen_weight")):
        m[f] = p[:, 1 + k]
    m["sheen_color"] = p[:, 14:17]
    m["sheen_roughness"] = p[:, 17]
    return m


@pytest.mark.parametrize("variant", range(6))
def test_sheen_layer_sample_eval(golden, variant):
    # bsdf/sheen.h:199-451 (Conty-Kulla microfiber distribution, sheen_distribution.h) over the
    # lobe of each kind, as material/standard.cpp:599-611 layers it; variants: 0 diffuse,
    # 1 conductor, 2 coated diffuse, 3 thin dielectric, 4 rough diffuse + transparency,
    # 5 rough dielectric
    flat, _ = O.compute_luts()
    mats = sheen_materials(golden)
    so, st, ev, aux = O.bsdf_mat(flat, mats, golden["sheen_geo"], golden["sheen_s"])
    sel = (np.arange(len(mats)) // 8) % 6 == variant
    ref_s, ref_t, ref_e = golden["sheen_sample"][sel], golden["sheen_sample_type"][sel], golden["sheen_eval"][sel]
    so, st, ev, aux = so[sel], st[sel], ev[sel], aux[sel]
    valid_ref = np.abs(ref_s[:, :3]).max(1) > 0
    valid = np.abs(so[:, :3]).max(1) > 0
    assert valid.mean() > 0.5
    assert np.array_equal(valid, valid_ref)
    assert np.array_equal(st, ref_t)
    assert (st[valid] == 1).any()  # the sheen reflection gets sampled (BsdfType::DiffuseReflection)
    np.testing.assert_allclose(so[valid, :3], ref_s[valid, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(so[valid, 3:6], ref_s[valid, 3:6], atol=2e-5)
    np.testing.assert_allclose(so[:, 6], ref_s[:, 6], rtol=1e-3, atol=1e-7)
    np.testing.assert_allclose(ev[:, :3], ref_e[:, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(ev[:, 3], ref_e[:, 3], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(aux[:, 0:3], golden["sheen_albedo"][sel], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(aux[:, 3:6], golden["sheen_normal"][sel], atol=1e-6)
    if variant == 3:
        np.testing.assert_allclose(aux[:, 6:9], golden["sheen_transparency"][sel], rtol=5e-4, atol=1e-6)


@pytest.mark.parametrize("suffix", ["", "_n"])
def test_post_intersect_and_frame(golden, suffix):
    # geometry/triangle_mesh.cpp:152-257, core/ray_simd.h:63-68, core/shading_context.h:144-191
    normals = golden["pi_normals_n"] if suffix else None
    ctx, fr = O.post_intersect(golden["pi_verts" + suffix], normals, golden["pi_indices" + suffix],
                               golden["pi_rays" + suffix], golden["pi_prim" + suffix])
    ref = golden["pi_ctx" + suffix]
    np.testing.assert_allclose(ctx[:, :3], ref[:, :3], atol=1e-5)
    np.testing.assert_allclose(ctx[:, 3], ref[:, 3], rtol=1e-6)
    np.testing.assert_allclose(ctx[:, 4:], ref[:, 4:], atol=1e-5)
    np.testing.assert_allclose(fr, golden["pi_frame" + suffix], atol=1e-5)


STD_FIELDS = ("base_weight", "r", "g", "b", "metalness", "diffuse_roughness", "specular_weight", "er", "eg", "eb",
              "roughness", "ior", "transmission_weight", "transmission_color", "transmission_depth", "thin_walled",
              "thickness", "coat_weight", "coat_color", "coat_roughness", "coat_ior", "coat_thickness", "sheen_weight",
              "sheen_color", "sheen_roughness")


def std_materials(golden):
    """Kind-5 material records of the golden StandardMaterial rows, and their backfacing flags."""
    p = golden["std_params"]
    m = np.zeros(p.shape[0], O.MATERIAL_DTYPE)
    m["kind"] = 5
    c = 0
    for f in STD_FIELDS:
        width = 3 if f in ("transmission_color", "coat_color", "sheen_color") else 1
        m[f] = p[:, c:c + width] if width == 3 else (p[:, c].astype(np.int32) if f == "thin_walled" else p[:, c])
        c += width
    m["transparency"] = 1.0 - p[:, 31]
    return m, p[:, 32].astype(np.int32)


@pytest.mark.parametrize("variant", range(8))
def test_standard_material_sample_eval(golden, variant):
    # material/standard.cpp:409-627 for uniform parameters: the opaque, translucent and metal bases
    # weighted in one MultiBsdfSimd (multi_bsdf.h:151-459), the coat over any base
    # (rough_coating.h:270-593 with a generic base), sheen and opacity; variants: 0 partial
    # metalness, 1 partial transmission (outside / inside), 2 coat over metal, 3 coat over glass
    # (outside / inside), 4 every lobe with sheen, opacity and a rough diffuse base, 5 thin-walled
    # partial transmission under a coat (the transparent BTDF at the ambient IOR), 6 base and
    # specular weights below 1, 7 a coat over a diffuse-only base (Kelemen's energy term)
    flat, _ = O.compute_luts()
    mats, inside = std_materials(golden)
    so, st, ev, aux = O.bsdf_mat(flat, mats, golden["std_geo"], golden["std_s"], inside)
    sel = (np.arange(len(mats)) // 8) % 8 == variant
    ref_s, ref_t, ref_e = golden["std_sample"][sel], golden["std_sample_type"][sel], golden["std_eval"][sel]
    so, st, ev, aux = so[sel], st[sel], ev[sel], aux[sel]
    valid_ref = np.abs(ref_s[:, :3]).max(1) > 0
    valid = np.abs(so[:, :3]).max(1) > 0
    assert valid.mean() > 0.5
    assert np.array_equal(valid, valid_ref)
    assert np.array_equal(st[valid], ref_t[valid])
    np.testing.assert_allclose(so[valid, :3], ref_s[valid, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(so[valid, 3:6], ref_s[valid, 3:6], atol=2e-5)
    np.testing.assert_allclose(so[valid, 6], ref_s[valid, 6], rtol=1e-3, atol=1e-7)
    assert (np.abs(ev[:, :3]).max(1) > 0).mean() > 0.3
    np.testing.assert_allclose(ev[:, :3], ref_e[:, :3], rtol=5e-4, atol=1e-6)
    np.testing.assert_allclose(ev[:, 3], ref_e[:, 3], rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(aux[:, 0:3], golden["std_albedo"][sel], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(aux[:, 3:6], golden["std_normal"][sel], atol=1e-6)
    transparent = (mats["transparency"][sel] != 0) | ((mats["transmission_weight"][sel] > 0) & (mats["thin_walled"][sel] != 0))
    if transparent.any():  # what a shadow ray keeps through the surface
        np.testing.assert_allclose(aux[transparent, 6:9], golden["std_transparency"][sel][transparent], rtol=5e-4, atol=1e-6)
    types = set(np.unique(st[valid]).tolist())
    assert len(types) > 1 or variant == 2  # more than one lobe gets sampled
