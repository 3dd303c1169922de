"""The differentiable-rendering checker: the oracle's FP64 gradient of mean((I - target)^2).

Pins, in order:
* the FP64 BSDF value of the derivative code equals the float BSDF evaluation that the golden
  vectors pin (diffuse and coated rows);
* its albedo and alpha derivatives equal central differences of that FP64 value;
* the whole estimator (suffix radiance times df/f at every vertex, next-event terms, the
  per-pixel adjoint) equals central differences of the oracle's own rendered loss for diffuse
  albedos, whose sampling does not depend on the albedo (cosine sampling, light picks and MIS
  weights do not see it, and Russian roulette starts beyond the path depth), so the rendered
  loss is a smooth function of the albedo along fixed samples.
"""
import numpy as np
import pytest

import oracle as O
import scenegen as G


def _loss(scene, mats, W, H, spp, target, max_depth=4):
    s = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], mats)
    s.set_lights(scene["lights"], scene.get("env"))
    acc, _ = s.render(scene["camera"], W, H, 0, spp, max_depth=max_depth)
    img = (acc["color"] / acc["weight"]).reshape(3, -1).T.astype(np.float64)
    return float(np.mean((img - target.reshape(-1, 3)) ** 2)), img


def _coat_rows(golden):
    cp = golden["coat_params"]
    return np.concatenate([np.full((cp.shape[0], 1), 3, np.float32), cp], 1)


def _kind_rows(golden, kind):
    sel = golden["bsdf_params"][:, 0] == kind
    return golden["bsdf_params"][sel], golden["bsdf_geo"][sel], golden["bsdf_s"][sel]


def test_grad_value_matches_bsdf_eval(golden):
    # the FP64 value of the derivative code equals the pinned float BSDF evaluation, every lobe
    flat, _ = O.compute_luts()
    for p, geo, s in ((_coat_rows(golden), golden["coat_geo"], golden["coat_s"]),
                      _kind_rows(golden, 0), _kind_rows(golden, 1), _kind_rows(golden, 2)):
        g = O.bsdf_grad(flat, p, geo)
        _, _, ev = O.bsdf(flat, p, geo, s)
        assert (np.abs(ev[:, :3]).max(1) > 0).mean() > 0.3
        # FP64 against FP32: relative 2e-4, or 1e-5 of the row's largest channel for tiny values.
        # Through the refraction half vector (dielectric), whose denominator cosOH + cosIH / eta
        # cancels near grazing, FP32 carries 1e-3 relative and 1e-7 absolute.  Where alpha sits at
        # its 0.001 clamp, FP32 D (up to ~1e5 at the peak) carries about 1e-2 relative and 2e-7
        # absolute off the peak.
        scale = np.abs(ev[:, :3]).max(1, keepdims=True)
        sharp = (p[:, 7:8].astype(np.float64) ** 2) <= 1e-3
        diel = p[0, 0] == 2
        rtol = np.where(sharp, 2e-2, 1e-3 if diel else 2e-4)
        atol = np.where(sharp, 2e-7, 1e-7 if diel else 1e-9)
        assert np.all(np.abs(g[:, :3] - ev[:, :3]) <= rtol * np.abs(ev[:, :3]) + 1e-5 * scale + atol)


def test_grad_matches_central_differences(golden):
    flat, _ = O.compute_luts()
    p, geo = _coat_rows(golden).astype(np.float64), golden["coat_geo"]
    g = O.bsdf_grad(flat, p, geo)
    live = np.abs(g[:, :3]).max(1) > 1e-6
    # albedo: the value is linear in it, so a difference of +-1e-3 is exact up to rounding
    h = 1e-3
    for c in range(3):
        pp, pm = p.copy(), p.copy()
        pp[:, 1 + c] += h
        pm[:, 1 + c] -= h
        fd = (O.bsdf_grad(flat, pp, geo)[:, c] - O.bsdf_grad(flat, pm, geo)[:, c]) / (2 * h)
        np.testing.assert_allclose(g[live, 3 + c], fd[live], rtol=1e-4, atol=1e-9)
    # alpha = roughness^2: perturb alpha by +-h; rows whose roughness step crosses a row of the
    # albedo tables (the value is piecewise linear in roughness there) are left out
    r = p[:, 7]
    alpha = r * r
    rp, rm = np.sqrt(alpha + h), np.sqrt(alpha - h)
    same_cell = (np.floor(rp * 15) == np.floor(r * 15)) & (np.floor(rm * 15) == np.floor(r * 15))
    pp, pm = p.copy(), p.copy()
    pp[:, 7], pm[:, 7] = rp, rm
    vp, vm = O.bsdf_grad(flat, pp, geo), O.bsdf_grad(flat, pm, geo)
    # the float roughness of a row is what the value sees: difference over the realised alpha step
    ap = np.float32(rp).astype(np.float64) ** 2
    am = np.float32(rm).astype(np.float64) ** 2
    m = live & same_cell
    assert m.sum() > 500
    for c in range(3):
        fd = (vp[:, c] - vm[:, c]) / (ap - am)
        scale = np.abs(g[m, 6 + c]).max()
        np.testing.assert_allclose(g[m, 6 + c], fd[m], rtol=2e-3, atol=2e-4 * scale)


@pytest.mark.parametrize("kind", [1, 2])
def test_conductor_dielectric_grad_matches_central_differences(golden, kind):
    # rough conductor: reflectivity and alpha; rough dielectric: alpha (it has no albedo)
    flat, _ = O.compute_luts()
    p0, geo, _ = _kind_rows(golden, kind)
    p = np.zeros((p0.shape[0], 11))
    p[:, :9] = p0
    g = O.bsdf_grad(flat, p, geo)
    live = np.abs(g[:, :3]).max(1) > 1e-5
    h = 1e-3
    if kind == 1:
        for c in range(3):
            r = p[:, 1 + c]
            ok = live & (r > 2 * h) & (r < 0.99 - 2 * h)
            pp, pm = p.copy(), p.copy()
            pp[:, 1 + c] += h
            pm[:, 1 + c] -= h
            fd = (O.bsdf_grad(flat, pp, geo)[:, c] - O.bsdf_grad(flat, pm, geo)[:, c]) / (2 * h)
            assert ok.sum() > 200
            scale = np.abs(g[ok, 3 + c]).max()
            np.testing.assert_allclose(g[ok, 3 + c], fd[ok], rtol=2e-3, atol=1e-4 * scale)
    else:
        assert np.all(g[:, 3:6] == 0)
    # alpha: rows whose roughness step stays inside one row of the albedo tables (32 rows for the
    # conductor's GGX tables, 16 for the dielectric's)
    n = 31 if kind == 1 else 15
    r = p[:, 7]
    alpha = r * r
    rp, rm = np.sqrt(alpha + h), np.sqrt(np.maximum(alpha - h, 0))
    same_cell = (np.floor(rp * n) == np.floor(r * n)) & (np.floor(rm * n) == np.floor(r * n)) & (alpha - h > 0.001)
    pp, pm = p.copy(), p.copy()
    pp[:, 7], pm[:, 7] = rp, rm
    vp, vm = O.bsdf_grad(flat, pp, geo), O.bsdf_grad(flat, pm, geo)
    ap = np.float32(rp).astype(np.float64) ** 2
    am = np.float32(rm).astype(np.float64) ** 2
    m = live & same_cell
    assert m.sum() > 100
    for c in range(3):
        fd = (vp[:, c] - vm[:, c]) / (ap - am)
        scale = np.abs(g[m, 6 + c]).max()
        np.testing.assert_allclose(g[m, 6 + c], fd[m], rtol=5e-3, atol=5e-4 * scale)


def test_adjoint_matches_central_differences_diffuse_albedo():
    W, H, spp = 64, 64, 4
    sc = G.spheres_ambient(n_spheres=24, subdiv=2, width=W, height=H)
    mats = sc["materials"]
    tgt_mats = mats.copy()
    rng = np.random.default_rng(3)
    for i in range(2, mats.shape[0]):
        for ch in ("r", "g", "b"):
            tgt_mats[i][ch] = np.clip(mats[i][ch] + rng.choice([-0.1, 0.1]), 0.05, 0.95)
    _, target = _loss(sc, tgt_mats, W, H, spp, np.zeros((H * W, 3)))
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], mats)
    s.set_lights(sc["lights"])
    loss, grad = s.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    loss0, _ = _loss(sc, mats, W, H, spp, target)
    assert abs(loss - loss0) <= 1e-9 * max(loss0, 1e-12) + 1e-15
    assert np.all(grad[:, 3] == 0)  # diffuse materials have no alpha
    h = 1e-3
    big = np.argsort(-np.abs(grad[2:, :3]).max(1))[:4] + 2  # the four materials with the largest gradients
    for mi in list(big) + [1]:  # and the ground
        for c, ch in enumerate(("r", "g", "b")):
            mp, mm = mats.copy(), mats.copy()
            mp[mi][ch] += h
            mm[mi][ch] -= h
            fd = (_loss(sc, mp, W, H, spp, target)[0] - _loss(sc, mm, W, H, spp, target)[0]) / (2 * h)
            assert abs(fd - grad[mi, c]) <= 2e-3 * abs(grad[mi, c]) + 1e-4 * np.abs(grad[:, :3]).max(), (mi, c, fd, grad[mi, c])


def test_adjoint_matches_central_differences_conductor_reflectivity():
    # rough conductors sample the visible normals (independent of the reflectivity) and their
    # sampling pdf does not see it either, so the rendered loss is smooth in the reflectivity along
    # fixed samples: the whole estimator through the conductor's Fresnel and multiple-scattering
    # terms equals central differences.  Dielectrics in the same scene have no albedo gradient.
    W, H, spp = 64, 64, 4
    sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=9, width=W, height=H, mixed=True)
    mats = sc["materials"]
    _, target = _loss(sc, G.perturbed_materials(mats), W, H, spp, np.zeros((H * W, 3)))
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], mats)
    s.set_lights(sc["lights"])
    loss, grad = s.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    assert abs(loss - _loss(sc, mats, W, H, spp, target)[0]) <= 1e-9 * loss
    kinds = mats["kind"]
    assert np.all(grad[kinds == G.KIND_DIELECTRIC, :3] == 0) and np.all(grad[kinds == G.KIND_DIELECTRIC, 3] != 0)
    h = 1e-3
    for mi in np.where(kinds == G.KIND_CONDUCTOR)[0]:
        assert grad[mi, 3] != 0
        for c, ch in enumerate(("r", "g", "b")):
            mp, mm = mats.copy(), mats.copy()
            mp[mi][ch] += h
            mm[mi][ch] -= h
            step = float(mp[mi][ch]) - float(mm[mi][ch])
            fd = (_loss(sc, mp, W, H, spp, target)[0] - _loss(sc, mm, W, H, spp, target)[0]) / step
            assert abs(fd - grad[mi, c]) <= 2e-4 * abs(grad[mi, c]) + 1e-6 * np.abs(grad[:, :3]).max(), (mi, c, fd, grad[mi, c])
