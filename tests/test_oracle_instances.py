"""Two-level scenes in the oracle (geometry/instance.h, accel/embree_intersector.cpp:100-129):
the per-instance traversal against brute force, identity instances against the one-level scene,
and a scaled, rotated, translated cube against closed forms for the hit distance and the world
normal (the inverse transpose of the linear part).  Embree's instance traversal cannot run here
(Embree is absent), so parity of the two-level path rests on these and on the GPU comparing
bit for bit with this oracle."""
import math

import numpy as np

import oracle as O
import scenegen as G


def _rays_towards(n, origin, target_box, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(target_box[0]), np.asarray(target_box[1])
    tgt = lo + rng.random((n, 3)) * (hi - lo)
    d = tgt - np.asarray(origin)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.tile(np.asarray(origin, np.float32), (n, 1))
    return np.concatenate([o, d, np.zeros((n, 1)), np.full((n, 1), 1e30)], 1).astype(np.float32)


def _oracle(scene):
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"], scene.get("normals"))
    osc.set_textures(scene)
    osc.set_lights(scene["lights"], scene.get("env"))
    return osc


def test_instanced_traversal_equals_brute_force():
    sc = G.spheres_instanced()
    osc = _oracle(sc)
    rays = _rays_towards(6000, (0, 2, -2), ((-10, -1, 5), (10, 8, 30)), 1)
    a, b = osc.intersect_inst(rays, 0), osc.intersect_inst(rays, 1)
    assert (a[0] >= 0).sum() > 1500 and len(set(a[4][a[0] >= 0])) > 20
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_identity_instances_equal_the_one_level_scene():
    sc = G.spheres_sunsky(n_spheres=24, subdiv=2, palette=12, width=64, height=48)
    one = _oracle(sc)
    two_scene = dict(sc, mesh_begin=[0, sc["indices"].shape[0]], instances=[(0, np.hstack([np.eye(3), np.zeros((3, 1))]))])
    two = _oracle(two_scene)
    rays = _rays_towards(4000, (0, 2, -2), ((-10, -1, 5), (10, 8, 30)), 2)
    p1, t1, u1, v1 = one.intersect(rays)
    p2, t2, u2, v2, i2 = two.intersect_inst(rays)
    assert np.array_equal(p1, p2) and np.array_equal(t1, t2) and np.array_equal(u1, u2) and np.array_equal(v1, v2)
    assert np.all(i2[p2 >= 0] == 0) and np.all(i2[p2 < 0] == -1)
    assert abs(one.radius - two.radius) <= 1e-6 * one.radius
    acc1, _ = one.render(sc["camera"], 64, 48, 0, 2, seed=0, max_depth=8, aovs=("albedo1", "normal1"))
    acc2, _ = two.render(sc["camera"], 64, 48, 0, 2, seed=0, max_depth=8, aovs=("albedo1", "normal1"))
    # the identity still renormalises the frame (an ulp here and there): equal up to rounding
    np.testing.assert_allclose(acc2["albedo1"], acc1["albedo1"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(acc2["normal1"], acc1["normal1"], rtol=0, atol=2e-6)
    assert np.mean(np.isclose(acc2["color"], acc1["color"], rtol=1e-4, atol=1e-6)) > 0.99


def cube_scene(R, S, t, W=64, H=48):
    """The cube [-1, 1]^3 (12 triangles, no vertex normals) placed by x' = R S x + t in front of
    the generators' pinhole camera, diffuse, under a unit ambient light, one bounce."""
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = []
    for ax in range(3):
        for sgn in (-1, 1):
            quad = [i for i in range(8) if v[i, ax] == sgn]
            a, b, c, d = sorted(quad, key=lambda i: math.atan2(*[v[i, k] for k in range(3) if k != ax][::-1]))
            faces += [[a, b, c], [a, c, d]]
    mats = np.zeros(2, G.MATERIAL_DTYPE)
    mats[1]["kind"] = G.KIND_DIFFUSE
    mats[1]["r"], mats[1]["g"], mats[1]["b"] = 0.5, 0.5, 0.5
    lights = np.zeros(1, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_AMBIENT, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    L = np.asarray(R) @ np.diag(S)
    return dict(name="cube", positions=v, indices=np.array(faces, np.int32), tri_material=np.ones(12, np.int32),
                materials=mats, mesh_begin=[0, 12], instances=[(0, np.hstack([L, np.asarray(t)[:, None]]))],
                lights=lights, env=None, camera=G._camera(fov_deg=50.0), width=W, height=H, spp=1, max_depth=1, seed=0,
                L=L, t=np.asarray(t, np.float64))


def rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def cube_closed_form(scene, rays):
    """Slab test of each ray against the cube in its object space (float64): hit distance and
    the world normal of the face entered, normalize(L^-T n) turned towards the ray."""
    L, t = scene["L"], scene["t"]
    Li = np.linalg.inv(L)
    o = (rays[:, :3].astype(np.float64) - t) @ Li.T
    d = rays[:, 3:6].astype(np.float64) @ Li.T
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (-1 - o) / d, (1 - o) / d
    tn, tf = np.minimum(t0, t1), np.maximum(t0, t1)
    tin, tout = tn.max(axis=1), tf.min(axis=1)
    hit = (tin <= tout) & (tout > 0) & (tin > 0)
    ax = tn.argmax(axis=1)
    n_obj = np.zeros_like(o)
    n_obj[np.arange(len(o)), ax] = -np.sign(d[np.arange(len(o)), ax])
    n_w = n_obj @ Li  # (L^-T n) as rows
    n_w /= np.linalg.norm(n_w, axis=1, keepdims=True)
    return hit, tin, n_w


def test_transformed_cube_closed_form():
    sc = cube_scene(rot_y(35.0) @ np.array([[1, 0, 0], [0, math.cos(0.4), -math.sin(0.4)], [0, math.sin(0.4), math.cos(0.4)]]),
                    (2.0, 0.6, 1.3), (0.5, -0.3, 9.0))
    osc = _oracle(sc)
    rays = _rays_towards(3000, (0.3, 0.5, 0.0), ((-2, -2, 7), (3, 1.5, 11)), 3)
    prim, t, u, v, inst = osc.intersect_inst(rays)
    hit, tc, _ = cube_closed_form(sc, rays)
    assert hit.sum() > 1000
    assert np.array_equal(prim >= 0, hit)
    np.testing.assert_allclose(t[hit], tc[hit], rtol=2e-6)
    assert np.all(inst[hit] == 0)
    # world normals through the first-hit normal AOV at pixel centres
    osc.set_pixel_sample(jitter=(0.5, 0.5))
    W, H = sc["width"], sc["height"]
    acc, _ = osc.render(sc["camera"], W, H, 0, 1, seed=0, max_depth=1, aovs=("normal1",))
    got = (acc["normal1"] / np.maximum(acc["normal1W"], 1)).T
    cam = sc["camera"]
    cfg = np.array([W, H, cam["fov"], cam["nearClip"], 0.0, 1.0, 0, 0, 0], np.float32)
    py, px = np.mgrid[0:H, 0:W]
    s2 = np.zeros((W * H, 4), np.float32)
    s2[:, 0], s2[:, 1] = (px.ravel() + 0.5) / W, (py.ravel() + 0.5) / H
    cr = O.camera_rays(cfg, s2)[:, 0, :]
    prays = np.concatenate([cr[:, :6], np.zeros((W * H, 1)), np.full((W * H, 1), 1e30)], 1).astype(np.float32)
    phit, _, n_w = cube_closed_form(sc, prays)
    # pixels near a cube edge may see another face after rounding: compare where the closed form
    # is robust (the entry face clearly first)
    assert phit.sum() > 200
    ok = phit & (np.abs(np.einsum("ij,ij->i", got, n_w) - 1) < 1e-4)
    assert ok.sum() >= 0.98 * phit.sum()
    np.testing.assert_allclose(got[ok], n_w[ok], rtol=0, atol=2e-6)
    assert np.all(got[~phit] == 0)
