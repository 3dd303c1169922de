"""Two-level scenes on the GPU through the C ABI: closest hits (triangle, instance, t, u, v) bit
for bit against the oracle's brute force, the world normals of a scaled and rotated instance
against the closed form, and the calls that take one-level scenes only.  Render parity on an
instanced scene is in test_gpu_parity.py."""
import math

import numpy as np
import pytest

import oracle as O
import protoray_hip as P
import scenegen as G
from test_oracle_instances import _oracle, _rays_towards, cube_closed_form, cube_scene, rot_y

pytestmark = pytest.mark.gpu


def test_instanced_hits_bit_exact():
    sc = G.spheres_instanced()
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, sc)
        rays = _rays_towards(20000, (0, 2, -2), ((-10, -1, 5), (10, 8, 30)), 5)
        got = dev.queryRaysInst(rays)
        plain = dev.queryRays(rays)
    finally:
        dev.close()
    ref = _oracle(sc).intersect_inst(rays, 1)
    assert (got[0] >= 0).sum() > 5000
    for g, r in zip(got, ref):
        assert np.array_equal(np.asarray(g).view(np.int32), np.asarray(r).view(np.int32))
    for g, r in zip(plain, got[:4]):
        assert np.array_equal(g, r)


def test_instance_normals_closed_form():
    sc = cube_scene(rot_y(-25.0) @ np.array([[1, 0, 0], [0, math.cos(-0.5), -math.sin(-0.5)], [0, math.sin(-0.5), math.cos(-0.5)]]),
                    (1.5, 0.7, 2.2), (-0.4, 0.2, 10.0))
    W, H = sc["width"], sc["height"]
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, sc)
        dev.initFrame(W, H, P.AOV_NORMAL1)
        dev.setJitter((0.5, 0.5))
        dev.render(1)
        got = dev.blitFrame(P.BUF_NORMAL1).reshape(-1, 3)
        lo, hi = dev.sceneBounds()
    finally:
        dev.close()
    cam = sc["camera"]
    cfg = np.array([W, H, cam["fov"], cam["nearClip"], 0.0, 1.0, 0, 0, 0], np.float32)
    py, px = np.mgrid[0:H, 0:W]
    s2 = np.zeros((W * H, 4), np.float32)
    s2[:, 0], s2[:, 1] = (px.ravel() + 0.5) / W, (py.ravel() + 0.5) / H
    cr = O.camera_rays(cfg, s2)[:, 0, :]
    prays = np.concatenate([cr[:, :6], np.zeros((W * H, 1)), np.full((W * H, 1), 1e30)], 1).astype(np.float32)
    phit, _, n_w = cube_closed_form(sc, prays)
    ok = phit & (np.abs(np.einsum("ij,ij->i", got, n_w) - 1) < 1e-4)
    assert phit.sum() > 200 and ok.sum() >= 0.98 * phit.sum()
    np.testing.assert_allclose(got[ok], n_w[ok], rtol=0, atol=2e-6)
    # Instance::getBounds: the corners of the cube through the transform
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]) @ sc["L"].T + sc["t"]
    np.testing.assert_allclose(lo, corners.min(axis=0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(hi, corners.max(axis=0), rtol=1e-6, atol=1e-6)


def test_two_level_limits():
    sc = G.spheres_instanced()
    dev = P.DeviceHip(0)
    try:
        P.load_scene(dev, sc)
        nt = sc["indices"].shape[0]
        eye = np.hstack([np.eye(3), np.zeros((3, 1))])
        with pytest.raises(ValueError):  # meshes must cover the triangles
            dev.initInstances([0, nt - 1], [(0, eye)])
        with pytest.raises(ValueError):  # singular transform
            dev.initInstances([0, nt], [(0, np.zeros((3, 4)))])
        with pytest.raises(ValueError):  # unknown mesh
            dev.initInstances([0, nt], [(1, eye)])
        dev.initInstances(sc["mesh_begin"], sc["instances"])
        dev.initLights(sc["lights"], sc["env"])
        dev.initFrame(64, 48)
        with pytest.raises(ValueError):
            dev.adjoint(np.zeros((48, 64, 3), np.float32), 1, sc["materials"].shape[0])
        # emissive geometry stays one-level
        mats = sc["materials"].copy()
        mats[2]["emission"] = (1.0, 1.0, 1.0)
        dev.initScene(sc["positions"], sc["indices"], sc["tri_material"], mats, sc["normals"])
        dev.initInstances(sc["mesh_begin"], sc["instances"])
        with pytest.raises(ValueError):  # emissive triangles in a mesh placed several times
            dev.initLights(sc["lights"], sc["env"])
    finally:
        dev.close()


def test_queries_on_identity_instance_equal_the_one_level_scene():
    # a scene with transparent materials, uploaded flat and as one mesh placed once by the
    # identity: the transparent-shadow occlusion query and the pixel query (hit list with
    # post-intersection frames and transparencies) agree; the identity still renormalises the
    # frame, so vectors agree to rounding
    sc = G.spheres_lobes()
    eye = np.hstack([np.eye(3), np.zeros((3, 1))])
    two = dict(sc, mesh_begin=[0, sc["indices"].shape[0]], instances=[(0, eye)])
    d1, d2 = P.DeviceHip(0), P.DeviceHip(0)
    try:
        P.load_scene(d1, sc)
        P.load_scene(d2, two)
        W, H = 96, 64
        d1.initFrame(W, H)
        d2.initFrame(W, H)
        rng = np.random.default_rng(5)
        n = 4096
        o = np.tile(np.asarray(sc["camera"]["pos"], np.float32), (n, 1))
        dirs = rng.normal(size=(n, 3))
        dirs[:, 2] = np.abs(dirs[:, 2]) + 0.5
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        rays = np.concatenate([o, dirs, np.full((n, 1), 1e-3), np.full((n, 1), 3.4e38)], 1).astype(np.float32)
        occ1, occ2 = d1.queryOcclusion(rays), d2.queryOcclusion(rays)
        assert occ1.sum() > n // 4 and np.array_equal(occ1, occ2)
        compared = multi = 0
        for y in range(0, H, 4):
            for x in range(0, W, 4):
                h1, h2 = d1.queryPixel(x, y), d2.queryPixel(x, y)
                assert len(h1) == len(h2)
                if len(h1) == 0:
                    continue
                compared += 1
                multi += len(h1) > 1
                for f in ("prim", "material", "material_type", "dist", "depth", "uv"):
                    assert np.array_equal(h1[f], h2[f]), f
                np.testing.assert_allclose(h2["p"], h1["p"], rtol=1e-6, atol=1e-6)
                for f in ("Ng", "N", "U", "V", "T"):
                    np.testing.assert_allclose(h2[f], h1[f], rtol=0, atol=2e-6)
        assert compared > 100 and multi > 0
    finally:
        d1.close()
        d2.close()
