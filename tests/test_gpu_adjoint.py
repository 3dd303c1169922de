"""Path-replay adjoint on the GPU (through the C ABI) against the oracle's FP64 gradient and
central differences of the rendered loss.

* The replay follows the primal paths: its per-sample radiance equals the wavefront renderer's.
* Its gradient equals the oracle's FP64 gradient of the same estimator (FP32 sums on the GPU).
* For diffuse albedos (sampling independent of the parameter, Russian roulette beyond the path
  depth) it equals central differences of the GPU-rendered loss with h = 1e-3.
* The differentiable-rendering configuration (512x512, 32 spp, depth 4, 64 coated materials,
  327,682 triangles) is run in full; the printed errors are the ones the design notes state.
"""
import json
import os
import time

import numpy as np
import pytest

import oracle as O
import protoray_hip as P
import scenegen as G
from conftest import results_dir

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def _load(dev, scene, mats=None, max_depth=4):
    m = scene["materials"] if mats is None else mats
    dev.initScene(scene["positions"], scene["indices"], scene["tri_material"], m)
    dev.initLights(scene["lights"], scene.get("env"))
    dev.initRenderer(maxDepth=max_depth)
    dev.initSampler(0)
    dev.initCamera(scene["camera"])


def _render(dev, W, H, spp):
    dev.initFrame(W, H)
    dev.initSampler(0)
    dev.clearFrame()
    dev.render(spp)
    return dev.blitFrame().reshape(-1, 3).astype(np.float64)


@pytest.mark.parametrize("which", ["adjoint", "sunsky"])
def test_replay_radiance_equals_primal_sample(dev, which):
    if which == "adjoint":
        sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=8, width=64, height=64)
        W, H, depth = 64, 64, 4
    else:
        sc = G.spheres_sunsky(n_spheres=48, subdiv=2, palette=24, coated=0.5, width=96, height=64)
        W, H, depth = 96, 64, 8
    dev.enableDebug(True)
    _load(dev, sc, max_depth=depth)
    dev.initFrame(W, H)
    for s in range(3):
        dev.render(1)
        L, _, _ = dev.debugOutputs()
        R = dev.adjointReplay(s)
        assert np.array_equal(L > 0, R > 0)
        np.testing.assert_allclose(R, L, rtol=1e-5, atol=1e-7)
    dev.enableDebug(False)


def test_gradient_matches_fp64_oracle(dev):
    W, H, spp = 64, 64, 4
    sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=8, width=W, height=H)
    target_mats = G.perturbed_materials(sc["materials"])
    _load(dev, sc, target_mats)
    target = _render(dev, W, H, spp)
    _load(dev, sc)
    loss, grad, ms = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    osc.set_lights(sc["lights"])
    oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    print("loss", loss, oloss, "ms", ms)
    assert abs(loss - oloss) <= 1e-4 * oloss
    scale = np.abs(ograd).max(axis=0)
    err = np.abs(grad - ograd)
    print("max abs err / max |g| per (r, g, b, alpha):", err.max(axis=0) / scale)
    assert np.all(err <= 2e-3 * np.abs(ograd) + 1e-3 * scale)


def open_room(albedo=(0.9, 0.85, 0.8)):
    """A box [-2, 2] x [-2, 2] x [1, 9] open towards the pinhole camera at the origin (its face at
    z = 1 missing), coated walls (material 1) and floor, ceiling and back (material 2) of high
    albedo, lit by a unit ambient light seen only through the opening: most paths bounce many
    times before they leave."""
    x0, x1, y0, y1, z0, z1 = -2.0, 2.0, -2.0, 2.0, 1.0, 9.0
    quads = [[(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)],   # floor
             [(x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)],   # ceiling
             [(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)],   # back
             [(x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)],   # left
             [(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)]]   # right
    pos, idx, tm = [], [], []
    for qi, q in enumerate(quads):
        b = len(pos)
        pos += q
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        tm += [2 if qi < 3 else 1] * 2
    mats = np.zeros(3, G.MATERIAL_DTYPE)
    for i, rough in ((1, 0.45), (2, 0.3)):
        mats[i]["kind"] = G.KIND_COATED
        mats[i]["r"], mats[i]["g"], mats[i]["b"] = albedo
        mats[i]["er"], mats[i]["eg"], mats[i]["eb"] = (1.0, 1.0, 1.0)
        mats[i]["roughness"], mats[i]["ior"], mats[i]["thickness"], mats[i]["weight"] = rough, 1.5, 1.0, 1.0
    lights = np.zeros(1, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_AMBIENT, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    return dict(positions=np.array(pos, np.float32), indices=np.array(idx, np.int32),
                tri_material=np.array(tm, np.int32), materials=mats, lights=lights, env=None,
                camera=G._camera(fov_deg=60.0))


def test_gradient_beyond_eight_vertices_matches_fp64_oracle(dev):
    # the kernel records path vertices eight depths at a time, so at depth 12 the vertices at
    # depths 8..11 come from a second hook replay and a second block reduction; in the open room
    # most paths are still bouncing at depth 8
    W, H, spp = 64, 64, 4
    sc = open_room()
    target_mats = G.perturbed_materials(sc["materials"])
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    osc.set_lights(sc["lights"])
    res = {}
    for depth in (8, 12):
        _load(dev, sc, target_mats, max_depth=depth)
        target = _render(dev, W, H, spp)
        _load(dev, sc, max_depth=depth)
        loss, grad, _ = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
        oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=depth)
        print("depth", depth, "loss", loss, oloss)
        assert abs(loss - oloss) <= 1e-4 * oloss
        scale = np.abs(ograd).max(axis=0)
        assert np.all(np.abs(grad - ograd) <= 2e-3 * np.abs(ograd) + 1e-3 * scale)
        res[depth] = (grad, ograd)
    # the deeper vertices change the gradient, and the change is the oracle's change
    dg, dog = res[12][0] - res[8][0], res[12][1] - res[8][1]
    print("depth 12 - depth 8, oracle: max |change| per (r, g, b, alpha):", np.abs(dog).max(axis=0))
    print("GPU change - oracle change:", np.abs(dg - dog).max(axis=0))
    print("scale:", np.abs(res[12][1]).max(axis=0))
    assert np.all(np.abs(dog).max(axis=0) > 0.02 * np.abs(res[12][1]).max(axis=0))
    assert np.all(np.abs(dg - dog) <= 0.05 * np.abs(dog).max(axis=0))


def test_gradient_two_level_scene_matches_fp64_oracle(dev):
    # 40 instances of six sphere meshes (diffuse, Oren-Nayar, coated, conductor, dielectric, thin
    # dielectric) over a ground quad: the replays trace through the top-level tree and the mesh
    # trees and shade in the instances' frames; base-colour textures dropped (not differentiated)
    W, H, spp = 64, 64, 4
    sc = {k: v for k, v in G.spheres_instanced(seed=3).items() if k not in ("uvs", "textures", "material_textures")}
    target_mats = G.perturbed_materials(sc["materials"])
    _load_any = lambda mats: (P.load_scene(dev, dict(sc, materials=mats)), dev.initRenderer(maxDepth=4),
                              dev.initSampler(0), dev.initCamera(sc["camera"]))
    _load_any(target_mats)
    target = _render(dev, W, H, spp)
    _load_any(sc["materials"])
    loss, grad, _ = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"], sc.get("normals"))
    osc.set_textures(sc)
    osc.set_lights(sc["lights"], sc.get("env"))
    oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    print("loss", loss, oloss)
    assert abs(loss - oloss) <= 1e-4 * oloss
    scale = np.abs(ograd).max(axis=0)
    err = np.abs(grad - ograd)
    print("max abs err / max |g| per (r, g, b, alpha):", err.max(axis=0) / scale)
    assert np.abs(ograd).max() > 0
    assert np.all(err <= 2e-3 * np.abs(ograd) + 1e-3 * scale)


def test_gradient_matches_fp64_oracle_conductor_dielectric(dev):
    # coated, rough conductor and rough dielectric materials in turn: reflectivity and alpha of
    # the conductors, alpha of the dielectrics (refraction into the spheres and out again)
    W, H, spp = 64, 64, 4
    sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=9, width=W, height=H, mixed=True)
    _load(dev, sc, G.perturbed_materials(sc["materials"]))
    target = _render(dev, W, H, spp)
    _load(dev, sc)
    loss, grad, ms = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    osc.set_lights(sc["lights"])
    oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    kinds = sc["materials"]["kind"]
    print("loss", loss, oloss, "ms", ms)
    assert abs(loss - oloss) <= 1e-4 * oloss
    assert np.all(grad[kinds == G.KIND_DIELECTRIC, :3] == 0)
    scale = np.abs(ograd).max(axis=0)
    err = np.abs(grad - ograd)
    for k in (G.KIND_CONDUCTOR, G.KIND_DIELECTRIC, G.KIND_COATED):
        sel = kinds == k
        print("kind", k, "max abs err / max |g| per (r, g, b, alpha):", err[sel].max(axis=0) / scale)
        assert np.all(np.abs(ograd[sel, 3]) > 0)
    assert np.all(err <= 2e-3 * np.abs(ograd) + 1e-3 * scale)


def test_gradient_matches_fp64_oracle_remaining_lobes(dev):
    # Oren-Nayar bases, thin dielectrics (not differentiated), opacity mixes (continuations carry
    # no gradient) and shadow rays through transparent surfaces
    W, H, spp = 64, 64, 4
    sc = G.spheres_lobes(width=W, height=H)
    mats = sc["materials"]
    target_mats = mats.copy()
    for i in range(2, len(mats)):
        for ch in ("r", "g", "b"):
            target_mats[i][ch] = min(max(mats[i][ch] + (0.1 if (i + ord(ch)) % 2 else -0.1), 0.02), 0.98)
    dev.initScene(sc["positions"], sc["indices"], sc["tri_material"], target_mats)
    dev.initLights(sc["lights"], sc["env"])
    dev.initRenderer(maxDepth=4)
    dev.initSampler(0)
    dev.initCamera(sc["camera"])
    target = _render(dev, W, H, spp)
    dev.initScene(sc["positions"], sc["indices"], sc["tri_material"], mats)
    dev.initLights(sc["lights"], sc["env"])
    loss, grad, _ = dev.adjoint(target.reshape(H, W, 3), spp, mats.shape[0])
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], mats)
    osc.set_lights(sc["lights"], sc["env"])
    oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=4)
    print("loss", loss, oloss)
    out = results_dir()
    if out:
        scale_ = np.abs(ograd).max(axis=0)
        with open(os.path.join(out, "adjoint_remaining_lobes.json"), "w") as f:
            json.dump({"loss_rel_err": abs(loss - oloss) / oloss,
                       "grad_err_over_scale": (np.abs(grad - ograd).max(axis=0) / scale_).tolist()}, f)
    # camera rays, hit points and frames round alike on both sides (no FP contraction in either),
    # so shadow rays that graze their own transparent surface decide alike too
    assert abs(loss - oloss) <= 1e-4 * oloss
    assert np.all(grad[mats["kind"] == G.KIND_THIN_DIELECTRIC] == 0)
    scale = np.abs(ograd).max(axis=0)
    err = np.abs(grad - ograd)
    print("max abs err / max |g| per (r, g, b, alpha):", err.max(axis=0) / scale)
    assert np.all(err <= 2e-3 * np.abs(ograd) + 1e-3 * scale)


def test_conductor_reflectivity_matches_central_differences(dev):
    # the conductor's sampling does not depend on its reflectivity: the adjoint equals central
    # differences of the GPU-rendered loss
    W, H, spp = 64, 64, 8
    sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=9, width=W, height=H, mixed=True)
    _load(dev, sc, G.perturbed_materials(sc["materials"]))
    target = _render(dev, W, H, spp)
    _load(dev, sc)
    _, grad, _ = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
    h = 1e-3
    for mi in np.where(sc["materials"]["kind"] == G.KIND_CONDUCTOR)[0]:
        for c, ch in enumerate(("r", "g", "b")):
            mp, mm = sc["materials"].copy(), sc["materials"].copy()
            mp[mi][ch] += h
            mm[mi][ch] -= h
            step = float(mp[mi][ch]) - float(mm[mi][ch])
            _load(dev, sc, mp)
            lp = np.mean((_render(dev, W, H, spp) - target) ** 2)
            _load(dev, sc, mm)
            lm = np.mean((_render(dev, W, H, spp) - target) ** 2)
            fd = (lp - lm) / step
            print(f"conductor {mi} {ch}: adjoint {grad[mi, c]:.6e} central difference {fd:.6e}")
            assert abs(fd - grad[mi, c]) <= 1e-2 * abs(grad[mi, c]) + 1e-7


def test_ground_albedo_matches_central_differences(dev):
    W, H, spp = 64, 64, 8
    sc = G.spheres_adjoint(n_spheres=24, subdiv=2, n_materials=8, width=W, height=H)
    target_mats = G.perturbed_materials(sc["materials"])
    target_mats[1]["r"] = target_mats[1]["g"] = target_mats[1]["b"] = 0.6  # ground 0.5 -> 0.6
    _load(dev, sc, target_mats)
    target = _render(dev, W, H, spp)
    _load(dev, sc)
    loss, grad, _ = dev.adjoint(target.reshape(H, W, 3), spp, sc["materials"].shape[0])
    h = 1e-3
    for c, ch in enumerate(("r", "g", "b")):
        mp, mm = sc["materials"].copy(), sc["materials"].copy()
        mp[1][ch] += h
        mm[1][ch] -= h
        _load(dev, sc, mp)
        lp = np.mean((_render(dev, W, H, spp) - target) ** 2)
        _load(dev, sc, mm)
        lm = np.mean((_render(dev, W, H, spp) - target) ** 2)
        fd = (lp - lm) / (2 * h)
        print(f"ground {ch}: adjoint {grad[1, c]:.6e} central difference {fd:.6e}")
        assert abs(fd - grad[1, c]) <= 5e-3 * abs(grad[1, c]) + 1e-7


def test_differentiable_rendering_configuration(dev):
    # 512x512, 32 spp, depth 4, 64 coated materials over 256 icospheres (327,682 triangles); the
    # target moves every albedo channel by +-0.1 and every alpha by +-0.05
    sc = G.spheres_adjoint()
    W, H, spp = sc["width"], sc["height"], sc["spp"]
    nmat = sc["materials"].shape[0]
    target_mats = G.perturbed_materials(sc["materials"])
    _load(dev, sc, target_mats)
    target = _render(dev, W, H, spp)
    _load(dev, sc)
    t0 = time.perf_counter()
    loss, grad, ms = dev.adjoint(target.reshape(H, W, 3), spp, nmat)
    wall = time.perf_counter() - t0
    # FP64 oracle gradient of the same estimator
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    osc.set_lights(sc["lights"])
    t1 = time.perf_counter()
    oloss, ograd = osc.adjoint(sc["camera"], W, H, target, spp, max_depth=sc["max_depth"])
    cpu_s = time.perf_counter() - t1
    scale = np.abs(ograd[2:]).max(axis=0)
    rel = np.abs(grad - ograd)[2:] / scale
    # central differences of the GPU-rendered loss (h = 1e-3) for the parameters with the
    # largest gradients: albedo (coated: the coat/base pick depends on the albedo) and alpha
    # (the microfacet sampling depends on alpha), so these differentiate the sampling too
    fd_rows = []
    order = np.argsort(-np.abs(ograd[2:, 3]))[:3] + 2
    for mi in order:
        for k, field in ((0, "r"), (3, "roughness")):
            mp, mm = sc["materials"].copy(), sc["materials"].copy()
            if k == 3:
                a = float(sc["materials"][mi]["roughness"]) ** 2
                mp[mi]["roughness"], mm[mi]["roughness"] = np.sqrt(a + 1e-3), np.sqrt(a - 1e-3)
                step = float(np.float32(np.sqrt(a + 1e-3))) ** 2 - float(np.float32(np.sqrt(a - 1e-3))) ** 2
            else:
                mp[mi][field] += 1e-3
                mm[mi][field] -= 1e-3
                step = float(mp[mi][field]) - float(mm[mi][field])
            _load(dev, sc, mp)
            lp = np.mean((_render(dev, W, H, spp) - target) ** 2)
            _load(dev, sc, mm)
            lm = np.mean((_render(dev, W, H, spp) - target) ** 2)
            fd = (lp - lm) / step
            fd_rows.append({"material": int(mi), "param": "alpha" if k == 3 else "albedo_r", "adjoint": float(grad[mi, k]),
                            "central_difference": float(fd), "rel_err": float(abs(fd - grad[mi, k]) / abs(grad[mi, k]))})
    res = {"config": f"spheres_adjoint {W}x{H}, {spp} spp, depth {sc['max_depth']}, {nmat - 2} coated materials, "
                     f"{sc['indices'].shape[0]} triangles",
           "loss_gpu": loss, "loss_fp64": oloss, "adjoint_kernel_ms": ms, "adjoint_call_s": wall, "oracle_fp64_s": cpu_s,
           "max_err_vs_fp64_over_max_grad": {"albedo_r": float(rel[:, 0].max()), "albedo_g": float(rel[:, 1].max()),
                                             "albedo_b": float(rel[:, 2].max()), "alpha": float(rel[:, 3].max())},
           "central_differences": fd_rows}
    print(json.dumps(res, indent=1))
    out = results_dir()
    if out:
        with open(os.path.join(out, "adjoint_config.json"), "w") as f:
            json.dump(res, f, indent=1)
    assert abs(loss - oloss) <= 1e-4 * oloss
    assert rel.max() < 5e-3
    for r in fd_rows:  # sampling-dependent parameters: the detached gradient agrees in expectation
        assert np.sign(r["central_difference"]) == np.sign(r["adjoint"]) and r["rel_err"] < 0.5
