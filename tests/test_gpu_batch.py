"""Persistent batch renderer (one launch, captured in a HIP graph) against the wavefront renderer.

Each batch frame must equal prt_render of the same scene, camera, seed and sample count: final
mean, the snapshot means after 1 and 4 samples, and the first-hit albedo / normal AOVs (the
batch lane traces the same paths with the same code and sums the samples in the same order).
"""
import json
import os

import numpy as np
import pytest

import dataset as D
import protoray_hip as P
import scenegen as G
from conftest import results_dir

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _scene_dev(sc, max_depth=8):
    d = P.DeviceHip(0)
    P.load_scene(d, sc)
    d.initRenderer(maxDepth=max_depth)
    return d


def _wavefront(dev, cam, W, H, seed, spp):
    dev.initCamera(cam)
    dev.initFrame(W, H, P.AOV_ALBEDO1 | P.AOV_NORMAL1 | P.AOV_DEPTH)
    dev.initSampler(seed)
    dev.clearFrame()
    out = {}
    for s in range(1, spp + 1):
        dev.render(1)
        if s in (1, 4, spp):
            out[s] = dev.blitFrame()
    return out, dev.blitFrame(P.BUF_ALBEDO1), dev.blitFrame(P.BUF_NORMAL1), dev.blitFrame(P.BUF_DEPTH)


def test_batch_equals_wavefront_renderer_standard_materials():
    # kind-5 materials: the batch kernel built with them (and, in the same launch, a frame of a
    # scene without them) equals the wavefront renderer's level-2 shading kernel
    scS = G.spheres_standard(seed=5)
    scA = G.spheres_sunsky(seed=1, n_spheres=32, subdiv=2, palette=16)
    dS, dA = _scene_dev(scS), _scene_dev(scA)
    cam = D.look_at((0, 2, -2), (1, 1, 15), 50)
    frames = [(dS, 64, 48, 3, 4), (dA, 32, 32, 1, 4)]
    batch = P.BatchHip(0)
    for d, W, H, seed, spp in frames:
        batch.addFrame(d, cam, W, H, seed, spp, snapshots=(1, 4), aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1 | P.AOV_DEPTH)
    batch.render()
    for i, (d, W, H, seed, spp) in enumerate(frames):
        ref, alb, nrm, dep = _wavefront(d, cam, W, H, seed, spp)
        np.testing.assert_allclose(batch.get(i), ref[spp], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_ALBEDO1), alb, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(batch.get(i, P.BATCH_NORMAL1), nrm, rtol=1e-5, atol=1e-6)
    batch.close()
    dS.close()
    dA.close()


def test_batch_equals_wavefront_renderer_two_level_scene():
    # a two-level scene (41 instances of 7 meshes, a textured and a transparent material) and a
    # one-level scene in the same launch: the kernel built for two-level scenes takes each frame's
    # traversal from its own scene
    scI = G.spheres_instanced(seed=2)
    scA = G.spheres_sunsky(seed=1, n_spheres=32, subdiv=2, palette=16)
    dI, dA = _scene_dev(scI), _scene_dev(scA)
    frames = [(dI, scI["camera"], 64, 48, 3, 4), (dA, D.look_at((0, 2, -2), (1, 1, 15), 50), 32, 32, 1, 4)]
    batch = P.BatchHip(0)
    for d, cam, W, H, seed, spp in frames:
        batch.addFrame(d, cam, W, H, seed, spp, snapshots=(1, 4), aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1 | P.AOV_DEPTH)
    batch.render()
    for i, (d, cam, W, H, seed, spp) in enumerate(frames):
        ref, alb, nrm, dep = _wavefront(d, cam, W, H, seed, spp)
        assert ref[spp].max() > 0
        np.testing.assert_allclose(batch.get(i), ref[spp], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_SNAPSHOT), ref[1], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_ALBEDO1), alb, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(batch.get(i, P.BATCH_NORMAL1), nrm, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(batch.get(i, P.BATCH_DEPTH), dep, rtol=1e-5, atol=1e-5)
    batch.close()
    dI.close()
    dA.close()


def test_batch_equals_wavefront_renderer():
    scA = G.spheres_sunsky(seed=1, n_spheres=64, subdiv=2, palette=32, coated=0.5)
    scB = G.spheres_ambient(seed=2, n_spheres=48, subdiv=2, emissive=4, ambient=0.3)
    scC = G.spheres_lobes(seed=3)  # thin dielectrics, transparency, Oren-Nayar bases
    dA, dB, dC = _scene_dev(scA), _scene_dev(scB, max_depth=5), _scene_dev(scC)
    cams = [D.look_at((0, 2, -2), (1, 1, 15), 50), D.look_at((-3, 1, 0), (0, 2, 20), 40, lens_radius=0.05, focal=15)]
    frames = [(dA, cams[0], 96, 64, 3, 6), (dB, cams[1], 64, 48, 11, 5), (dA, cams[1], 128, 32, 0, 4),
              (dC, cams[0], 64, 64, 5, 4)]
    batch = P.BatchHip(0)
    for d, cam, W, H, seed, spp in frames:
        batch.addFrame(d, cam, W, H, seed, spp, snapshots=(1, 4), aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1 | P.AOV_DEPTH)
    r1 = batch.render()
    assert r1["samples"] == sum(W * H * spp for _, _, W, H, _, spp in frames)
    first = [batch.get(i) for i in range(len(frames))]
    for i, (d, cam, W, H, seed, spp) in enumerate(frames):
        ref, alb, nrm, dep = _wavefront(d, cam, W, H, seed, spp)
        np.testing.assert_allclose(batch.get(i), ref[spp], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_SNAPSHOT), ref[1], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_SNAPSHOT + 1), ref[4], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(batch.get(i, P.BATCH_ALBEDO1), alb, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(batch.get(i, P.BATCH_NORMAL1), nrm, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(batch.get(i, P.BATCH_DEPTH), dep, rtol=1e-5)
    # the captured graph replays: same work, same result
    r2 = batch.render()
    assert r2["samples"] == r1["samples"]
    for i in range(len(frames)):
        assert np.array_equal(batch.get(i), first[i])
    # the autotuner picks a launch shape; the result does not depend on it
    t = batch.tune(1)
    assert t["tiles_per_fetch"] in (1, 2, 4, 8) and 1 <= t["blocks_per_cu"] <= 4
    batch.render()
    for i in range(len(frames)):
        assert np.array_equal(batch.get(i), first[i])
    batch.close()
    dA.close()
    dB.close()
    dC.close()


def test_batch_follows_a_changed_scene():
    # geometry, materials and BVH of a frame's scene context replaced after the batch rendered:
    # the next render rebuilds the frame's kernel arguments (the old device buffers are released)
    # and equals a fresh batch of the changed scene
    scA = G.spheres_ambient(seed=4, n_spheres=24, subdiv=2)
    scB = G.spheres_ambient(seed=5, n_spheres=40, subdiv=2)
    d = _scene_dev(scA, max_depth=5)
    cam = D.look_at((0, 2, -2), (1, 1, 15), 50)
    batch = P.BatchHip(0)
    batch.addFrame(d, cam, 64, 64, 1, 3)
    batch.render()
    before = batch.get(0)
    mats = scB["materials"].copy()
    mats["r"] = 1.0 - mats["r"]
    d.initScene(scB["positions"], scB["indices"], scB["tri_material"], mats)
    with pytest.raises(ValueError):
        batch.render()  # new geometry needs its lights set again (the area light depends on it)
    d.initLights(scB["lights"])
    batch.render()
    after = batch.get(0)
    assert not np.array_equal(after, before)
    fresh = P.BatchHip(0)
    fresh.addFrame(d, cam, 64, 64, 1, 3)
    fresh.render()
    assert np.array_equal(after, fresh.get(0))
    fresh.close()
    batch.close()
    d.close()


def test_dataset_configuration_batch(tmp_path):
    # the 64-frame, 8-scene configuration (100 k .. 4 M triangles, 256x256 .. 2048x1152, depth 8),
    # with a 64-sample reference instead of 4096 to keep the test short; two frames are written
    s = D.render_dataset(outdir=None, ref_spp=64)
    print(json.dumps(s))
    out = results_dir()
    if out:
        with open(os.path.join(out, "batch_config.json"), "w") as f:
            json.dump(s, f, indent=1)
    assert s["samples"] == s["pixels"] * 64
    assert min(s["triangles"]) >= 90_000 and max(s["triangles"]) <= 4_200_000
    D.render_dataset(outdir=str(tmp_path), ref_spp=4, n_frames=2, n_scenes=2, max_tris=200_000, tune=False)
    img, names = D.read_exr(tmp_path / "frame00_ref.hdr.exr")
    assert names == ["B", "G", "R"] and np.isfinite(img).all() and img.max() > 0
    assert (tmp_path / "frame01_0004.hdr.exr").exists() and (tmp_path / "frame01.json").exists()
