"""Seeded synthetic scenes for the stream path tracer's workloads.

Two generators, both deterministic for a seed (PCG32, the generator of math/random.h:17-60):

* ``spheres_ambient`` -- a 40x40 ground quad at y = -1 plus N icospheres with centres uniform in
  [-10,10]x[0,8]x[5,30] and radii U[0.2,1.0], all diffuse with albedo U[0.1,0.9]^3, lit by a
  constant ambient light of radiance 1.  512x512, 16 spp, max depth 5, pinhole camera
  (vertical fov 45 deg) at the origin looking +z.  Defaults: 256 spheres at subdivision 3.
* ``spheres_sunsky`` -- the same generator with 4096 icospheres at subdivision 4
  (20,971,522 triangles), materials drawn per sphere from a seeded palette (40 % diffuse,
  30 % rough GGX conductor with gold-like eta/k, 30 % rough GGX dielectric of IOR 1.5), lit by a
  sun/sky (elevation 35 deg, turbidity 3) plus one distant light; 1920x1080, 64 spp, max
  depth 8, thin-lens camera (lens radius 0.05, focus distance 15).

Choices the workload description leaves open, fixed here:
  * the ground quad spans x in [-20,20], z in [-2.5,37.5] (centred under the spheres), faces +y,
    diffuse albedo 0.5;
  * "GGX alpha a" maps to roughness sqrt(a) because the reference uses alpha = roughness^2
    (bsdf/microfacet.h:10-15);
  * gold eta/k (RGB) are drawn within +-10 % of (0.143, 0.374, 1.442) / (3.983, 2.385, 1.603) and
    converted to the artistic reflectivity / edge tint the reference's conductor takes by
    inverting Gulbrandsen's mapping (bsdf/optics.h:214-226);
  * the material palette has 1024 entries (the reference's stream sort holds at most 2048
    material IDs, render/renderer_stream.h:110);
  * the distant light has radiance (300, 280, 250), direction normalize(-0.6, 0.7, -0.4) and an
    angular diameter of 2 deg; the sun azimuth is 30 deg (reference convention atan2(x, -z));
  * lens radius = 0.05 (the "aperture").
Material ID 0 is reserved for the environment (render/scene.cpp:204-207); triangles use IDs >= 1.
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

MATERIAL_DTYPE = np.dtype([("kind", "<i4"), ("r", "<f4"), ("g", "<f4"), ("b", "<f4"),
                           ("er", "<f4"), ("eg", "<f4"), ("eb", "<f4"), ("roughness", "<f4"),
                           ("ior", "<f4"), ("thickness", "<f4"), ("weight", "<f4"),
                           ("emission", "<f4", 3), ("transparency", "<f4"),
                           ("diffuse_roughness", "<f4"), ("sheen_weight", "<f4"), ("sheen_color", "<f4", 3),
                           ("sheen_roughness", "<f4"),
                           ("transmission_color", "<f4", 3), ("transmission_depth", "<f4"),
                           ("base_weight", "<f4"), ("metalness", "<f4"), ("specular_weight", "<f4"),
                           ("transmission_weight", "<f4"), ("coat_weight", "<f4"), ("coat_color", "<f4", 3),
                           ("coat_roughness", "<f4"), ("coat_ior", "<f4"), ("coat_thickness", "<f4"),
                           ("thin_walled", "<i4")])
LIGHT_DTYPE = np.dtype([("kind", "<i4"), ("r", "<f4"), ("g", "<f4"), ("b", "<f4"),
                        ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("angle", "<f4")])
CAMERA_DTYPE = np.dtype([("pos", "<f4", 3), ("U", "<f4", 3), ("V", "<f4", 3), ("N", "<f4", 3),
                         ("fov", "<f4"), ("nearClip", "<f4"), ("lensRadius", "<f4"),
                         ("focalDistance", "<f4")])
KIND_DIFFUSE, KIND_CONDUCTOR, KIND_DIELECTRIC, KIND_COATED, KIND_THIN_DIELECTRIC = 0, 1, 2, 3, 4
LIGHT_AMBIENT, LIGHT_DISTANT, LIGHT_IMAGE = 0, 1, 2


class Pcg32:
    """PCG-XSH-RR 64/32 (O'Neill 2014), seeded like math/random.h:29-36."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed=0xcafef00dd15ea5e5, seq=1442695040888963407 >> 1):
        self.inc = ((seq << 1) | 1) & self.MASK
        self.state = (seed + self.inc) & self.MASK
        self.state = (self.state * self.MULT + self.inc) & self.MASK

    def u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xffffffff
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xffffffff

    def f(self, a=0.0, b=1.0):
        return a + (self.u32() * 2.0 ** -32) * (b - a)


def icosphere(subdiv):
    """Unit icosphere (float64): 10*4^s + 2 vertices, 20*4^s outward-wound triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces, np.int64)
    # make every triangle wind counter-clockwise seen from outside (cross product points out)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(axis=1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def gold_artistic(rng):
    """Gold-like eta/k (RGB, +-10 %) -> artistic reflectivity r and edge tint g."""
    eta0, k0 = (0.143, 0.374, 1.442), (3.983, 2.385, 1.603)
    r, g = [], []
    for c in range(3):
        n = eta0[c] * rng.f(0.9, 1.1)
        k = k0[c] * rng.f(0.9, 1.1)
        rc = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)
        rc = min(rc, 0.99)
        nmin = (1 - rc) / (1 + rc)
        nmax = (1 + math.sqrt(rc)) / (1 - math.sqrt(rc))
        gc = (nmax - n) / (nmax - nmin)
        r.append(rc)
        g.append(min(max(gc, 0.0), 1.0))
    return r, g


def _material(kind, rgb=(0, 0, 0), edge=(0, 0, 0), roughness=0.0, ior=1.5, thickness=0.0, weight=0.0,
              transparency=0.0, diffuse_roughness=0.0):
    m = np.zeros((), MATERIAL_DTYPE)
    m["transparency"] = transparency
    m["diffuse_roughness"] = diffuse_roughness
    m["kind"] = kind
    m["r"], m["g"], m["b"] = rgb
    m["er"], m["eg"], m["eb"] = edge
    m["roughness"] = roughness
    m["ior"] = ior
    m["thickness"] = thickness
    m["weight"] = weight
    return m


def _camera(fov_deg=45.0, lens_radius=0.0, focal=1.0, near=0.01):
    c = np.zeros((), CAMERA_DTYPE)
    c["pos"] = (0, 0, 0)
    # looking along +z: the view direction is -N (camera/pinhole_camera.h:86-89)
    c["U"], c["V"], c["N"] = (-1, 0, 0), (0, 1, 0), (0, 0, -1)
    c["fov"] = math.radians(fov_deg)
    c["nearClip"] = near
    c["lensRadius"] = lens_radius
    c["focalDistance"] = focal
    return c


def _assemble(rng, n_spheres, subdiv, sphere_material):
    Vs, Fs = icosphere(subdiv)
    # ground quad y = -1, x in [-20, 20], z in [-2.5, 37.5], normal +y
    gv = np.array([[-20, -1, -2.5], [20, -1, -2.5], [20, -1, 37.5], [-20, -1, 37.5]], np.float64)
    gf = np.array([[0, 2, 1], [0, 3, 2]], np.int64)
    centers = np.zeros((n_spheres, 3))
    radii = np.zeros(n_spheres)
    mat_ids = np.zeros(n_spheres, np.int64)
    for i in range(n_spheres):
        centers[i] = (rng.f(-10, 10), rng.f(0, 8), rng.f(5, 30))
        radii[i] = rng.f(0.2, 1.0)
        mat_ids[i] = sphere_material(i)
    pos = np.concatenate([gv, (Vs[None] * radii[:, None, None] + centers[:, None, :]).reshape(-1, 3)])
    idx = np.concatenate([gf, (Fs[None] + 4 + Vs.shape[0] * np.arange(n_spheres)[:, None, None]).reshape(-1, 3)])
    tri_mat = np.concatenate([np.ones(2, np.int64), np.repeat(mat_ids, Fs.shape[0])])
    return pos.astype(np.float32), idx.astype(np.int32), tri_mat.astype(np.int32)


def spheres_ambient(seed=0, n_spheres=256, subdiv=3, width=512, height=512, spp=16, max_depth=5, emissive=0,
                    emission=4.0, ambient=1.0):
    """emissive: the first `emissive` spheres also emit emission x a random color (an area light
    over their triangles, drawn from a second PCG32 stream so the geometry is unchanged);
    ambient: radiance of the ambient light, None for no environment light at all."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5))]
    sphere_mats = []

    def pick(i):
        sphere_mats.append(_material(KIND_DIFFUSE, (rng.f(0.1, 0.9), rng.f(0.1, 0.9), rng.f(0.1, 0.9))))
        return 2 + i

    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, pick)
    erng = Pcg32(seed + 7919)
    for i in range(min(emissive, n_spheres)):
        sphere_mats[i]["emission"] = (emission * erng.f(0.3, 1.0), emission * erng.f(0.3, 1.0), emission * erng.f(0.3, 1.0))
    lights = np.zeros(0 if ambient is None else 1, LIGHT_DTYPE)
    if ambient is not None:
        lights[0] = (LIGHT_AMBIENT, ambient, ambient, ambient, 0, 0, 0, 0)
    return dict(name="spheres_ambient", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats + sphere_mats, MATERIAL_DTYPE), lights=lights, env=None,
                camera=_camera(), width=width, height=height, spp=spp, max_depth=max_depth, seed=0)


def spheres_adjoint(seed=0, n_spheres=256, subdiv=3, width=512, height=512, spp=32, max_depth=4, n_materials=64,
                    mixed=False):
    """Differentiable-rendering scene (config-0 style): the ground (diffuse 0.5) plus n_spheres
    icospheres sharing n_materials coated-diffuse materials (material IDs 2 .. n_materials+1;
    albedo U[0.2,0.8]^3, coat IOR 1.5, coat alpha U[0.05,0.4], i.e. roughness sqrt(alpha)),
    under the ambient light and a distant light.  mixed=True makes every third material a rough
    conductor (reflectivity = the albedo, edge tint 0.5) and every third, starting at the second,
    a rough dielectric (IOR 1.5), with the same alpha draw."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5))]
    for i in range(n_materials):
        rgb = (rng.f(0.2, 0.8), rng.f(0.2, 0.8), rng.f(0.2, 0.8))
        alpha = rng.f(0.05, 0.4)
        if mixed and i % 3 == 1:
            mats.append(_material(KIND_CONDUCTOR, rgb, (0.5, 0.5, 0.5), math.sqrt(alpha)))
        elif mixed and i % 3 == 2:
            mats.append(_material(KIND_DIELECTRIC, (1, 1, 1), (0, 0, 0), math.sqrt(alpha), 1.5))
        else:
            mats.append(_material(KIND_COATED, rgb, (1, 1, 1), math.sqrt(alpha), 1.5, thickness=1.0, weight=1.0))
    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, lambda i: 2 + i % n_materials)
    lights = np.zeros(2, LIGHT_DTYPE)
    d = np.array([-0.5, 0.8, -0.3])
    d /= np.linalg.norm(d)
    lights[0] = (LIGHT_AMBIENT, 0.6, 0.6, 0.6, 0, 0, 0, 0)
    lights[1] = (LIGHT_DISTANT, 40.0, 38.0, 35.0, d[0], d[1], d[2], math.radians(3.0))
    return dict(name="spheres_adjoint", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats, MATERIAL_DTYPE), lights=lights, env=None,
                camera=_camera(), width=width, height=height, spp=spp, max_depth=max_depth, seed=0)


def perturbed_materials(materials, seed=1, d_albedo=0.1, d_alpha=0.05):
    """Target parameters of the differentiable-rendering configuration: every coated material's
    albedo moved by +-d_albedo per channel and its alpha by +-d_alpha (signs from PCG32 seed);
    conductors likewise (reflectivity and alpha), dielectrics in alpha only."""
    rng = Pcg32(seed)
    m = materials.copy()
    for i in range(m.shape[0]):
        if m[i]["kind"] not in (KIND_COATED, KIND_CONDUCTOR, KIND_DIELECTRIC):
            continue
        for ch in ("r", "g", "b") if m[i]["kind"] != KIND_DIELECTRIC else ():
            m[i][ch] = min(max(m[i][ch] + (d_albedo if rng.f() < 0.5 else -d_albedo), 0.02), 0.98)
        alpha = float(m[i]["roughness"]) ** 2 + (d_alpha if rng.f() < 0.5 else -d_alpha)
        m[i]["roughness"] = math.sqrt(min(max(alpha, 0.01), 0.9))
    return m


def load_sunsky(path=os.path.join(HERE, "sunsky_e35_t3_az30.npz")):
    z = np.load(path)
    # the bake is stored zenith-first; the image light reads texture rows bottom-up
    # (image/image_texture.h:396), so flip to nadir-first
    return dict(sky=np.ascontiguousarray(z["sky"][::-1]), sun_radiance=z["sun_radiance"],
                sun_radius=float(z["sun_radius"][0]), sun_dir=z["sun_dir"])


def spheres_sunsky(seed=0, n_spheres=4096, subdiv=4, width=1920, height=1080, spp=64, max_depth=8, palette=1024,
                   coated=0.0):
    """coated: fraction of the diffuse palette entries given a rough coat (kind 3: IOR U[1.3,1.8],
    roughness U[0.1,0.7], tinted coat color and thickness for half of them, the StandardMaterial
    specular layer - white, thickness 1, weight 1 - for the other half).  0 keeps the random
    stream, and so the scene, of the plain generator."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5))]
    for _ in range(palette):
        u = rng.f()
        if u < 0.4:
            rgb = (rng.f(0.1, 0.9), rng.f(0.1, 0.9), rng.f(0.1, 0.9))
            if coated > 0 and rng.f() < coated:
                ior, rough = rng.f(1.3, 1.8), rng.f(0.1, 0.7)
                if rng.f() < 0.5:
                    mats.append(_material(KIND_COATED, rgb, (1, 1, 1), rough, ior, thickness=1.0, weight=1.0))
                else:
                    tint = (rng.f(0.3, 1.0), rng.f(0.3, 1.0), rng.f(0.3, 1.0))
                    mats.append(_material(KIND_COATED, rgb, tint, rough, ior, thickness=rng.f(0.2, 2.0),
                                          weight=rng.f(0.3, 1.0)))
            else:
                mats.append(_material(KIND_DIFFUSE, rgb))
        elif u < 0.7:
            alpha = rng.f(0.05, 0.5)
            r, g = gold_artistic(rng)
            mats.append(_material(KIND_CONDUCTOR, r, g, math.sqrt(alpha)))
        else:
            alpha = rng.f(0.01, 0.3)
            mats.append(_material(KIND_DIELECTRIC, roughness=math.sqrt(alpha), ior=1.5))
    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, lambda i: 2 + rng.u32() % palette)
    ss = load_sunsky()
    d = np.array([-0.6, 0.7, -0.4])
    d /= np.linalg.norm(d)
    lights = np.zeros(3, LIGHT_DTYPE)
    # render/scene.cpp:375-431 order: distant light, then the sun/sky pair (sun, sky)
    lights[0] = (LIGHT_DISTANT, 300.0, 280.0, 250.0, d[0], d[1], d[2], 0.5 * math.radians(2.0))
    sr = ss["sun_radiance"]
    sd = ss["sun_dir"]
    lights[1] = (LIGHT_DISTANT, sr[0], sr[1], sr[2], sd[0], sd[1], sd[2], ss["sun_radius"])
    lights[2] = (LIGHT_IMAGE, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    return dict(name="spheres_sunsky", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats, MATERIAL_DTYPE), lights=lights, env=ss["sky"],
                camera=_camera(lens_radius=0.05, focal=15.0), width=width, height=height, spp=spp,
                max_depth=max_depth, seed=0)


def spheres_textured(seed=0, n_spheres=48, subdiv=2, width=96, height=64, spp=4, max_depth=8, palette=24):
    """The sun/sky spheres with image textures on base colours: per-vertex UVs (the ground's
    planar x/z map, a latitude/longitude map on each sphere), three textures (an 8x8 sRGB RGBA8
    checker with repeat addressing and a 4x tiling on the ground, a 16x4 RGB FP32 gradient with
    clamp addressing, a 5x3 linear RGBA8 texture repeating in u and clamped in v with an offset),
    bound to the ground, to diffuse spheres (some with Oren-Nayar bases) and to coated spheres;
    conductors and dielectrics stay uniform."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.9, 0.9, 0.9))]
    bind = [(1, 0)]
    for i in range(palette):
        v = i % 4
        rgb = (rng.f(0.5, 1.2), rng.f(0.5, 1.2), rng.f(0.5, 1.2))  # scales of the texture (clamped product)
        if v == 0:
            mats.append(_material(KIND_DIFFUSE, rgb, diffuse_roughness=0.6 if i % 8 == 0 else 0.0))
            bind.append((len(mats) - 1, 1 + (i // 4) % 2))
        elif v == 1:
            mats.append(_material(KIND_COATED, rgb, (1, 1, 1), math.sqrt(rng.f(0.02, 0.4)), rng.f(1.3, 1.8), thickness=1.0,
                                  weight=1.0))
            bind.append((len(mats) - 1, (i // 4) % 3))
        elif v == 2:
            r, g = gold_artistic(rng)
            mats.append(_material(KIND_CONDUCTOR, r, g, math.sqrt(rng.f(0.05, 0.5))))
        else:
            mats.append(_material(KIND_DIELECTRIC, roughness=math.sqrt(rng.f(0.01, 0.3)), ior=1.5))
    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, lambda i: 2 + rng.u32() % palette)
    # UVs: ground x/z over [0, 1]^2; spheres by latitude / longitude of the unit icosphere vertex
    Vs, _ = icosphere(subdiv)
    ground = pos[:4]
    uv_g = np.stack([(ground[:, 0] + 20.0) / 40.0, (ground[:, 2] + 2.5) / 40.0], axis=1)
    uv_s = np.stack([0.5 + np.arctan2(Vs[:, 2], Vs[:, 0]) / (2 * math.pi), 0.5 + np.arcsin(np.clip(Vs[:, 1], -1, 1)) / math.pi],
                    axis=1)
    uvs = np.concatenate([uv_g, np.tile(uv_s, (n_spheres, 1))]).astype(np.float32)
    trng = Pcg32(seed + 101)
    checker = np.zeros((8, 8, 4), np.uint8)
    ca = [int(trng.f(40, 255)) for _ in range(3)]
    cb = [int(trng.f(0, 120)) for _ in range(3)]
    for y in range(8):
        for x in range(8):
            checker[y, x, :3] = ca if (x + y) % 2 == 0 else cb
            checker[y, x, 3] = 255
    gradient = np.zeros((4, 16, 3), np.float32)
    for y in range(4):
        for x in range(16):
            gradient[y, x] = (x / 15.0, 0.2 + 0.2 * y, 1.0 - x / 15.0)
    small = np.array([[[int(trng.f(0, 255)) for _ in range(3)] + [255] for _ in range(5)] for _ in range(3)], np.uint8)
    textures = [
        dict(texels=checker, format=0, address=(0, 0), uv_scale=(4.0, 4.0), uv_offset=(0.0, 0.0)),
        dict(texels=gradient, format=2, address=(1, 1), uv_scale=(1.0, 1.0), uv_offset=(0.0, 0.0)),
        dict(texels=small, format=1, address=(0, 1), uv_scale=(2.0, 1.5), uv_offset=(0.25, -0.2)),
    ]
    sc = spheres_sunsky(n_spheres=0, width=width, height=height, spp=spp, max_depth=max_depth, palette=1)
    return dict(name="spheres_textured", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats, MATERIAL_DTYPE), lights=sc["lights"], env=sc["env"], uvs=uvs,
                textures=textures, material_textures=bind,
                camera=_camera(lens_radius=0.05, focal=15.0), width=width, height=height, spp=spp,
                max_depth=max_depth, seed=0)


def _rotation(rng):
    """A random rotation (unit quaternion from three uniforms, Shoemake 1992) as a 3 x 3 matrix."""
    u1, u2, u3 = rng.f(), rng.f(), rng.f()
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x, y, z, w = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def spheres_instanced(seed=0, n_instances=40, subdiv=2, width=96, height=64, spp=4, max_depth=8):
    """A two-level scene: meshes in object space (the ground quad and six unit icospheres with
    smooth vertex normals, each with its own material: diffuse, Oren-Nayar diffuse with a base
    colour texture, coated, rough conductor, rough dielectric (a media interface), thin rough
    dielectric (transparent to shadow rays)), placed by instances: the ground once (identity),
    the spheres n_instances times with random rotations, non-uniform scales (0.3 to 1.2 per axis)
    and translations over the sun/sky spheres' volume.  Lit by the sun/sky pair and a distant
    light."""
    rng = Pcg32(seed)
    Vs, Fs = icosphere(subdiv)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5))]
    r, g = gold_artistic(rng)
    mats += [_material(KIND_DIFFUSE, (0.7, 0.3, 0.2)),
             _material(KIND_DIFFUSE, (1.0, 1.0, 1.0), diffuse_roughness=0.5),
             _material(KIND_COATED, (0.2, 0.5, 0.7), (1, 1, 1), 0.3, 1.5, thickness=1.0, weight=1.0),
             _material(KIND_CONDUCTOR, r, g, 0.35),
             _material(KIND_DIELECTRIC, roughness=0.2, ior=1.5),
             _material(KIND_THIN_DIELECTRIC, (0.8, 0.9, 0.7), roughness=0.25, ior=1.4, thickness=0.5)]
    gv = np.array([[-20, -1, -2.5], [20, -1, -2.5], [20, -1, 37.5], [-20, -1, 37.5]], np.float64)
    pos = [gv] + [Vs] * 6
    nrm = [np.tile([0.0, 1.0, 0.0], (4, 1))] + [Vs] * 6
    idx, tm, begin = [np.array([[0, 2, 1], [0, 3, 2]])], [np.ones(2, np.int64)], [0, 2]
    off = 4
    for k in range(6):
        idx.append(Fs + off)
        tm.append(np.full(Fs.shape[0], 2 + k))
        off += Vs.shape[0]
        begin.append(begin[-1] + Fs.shape[0])
    uv_s = np.stack([0.5 + np.arctan2(Vs[:, 2], Vs[:, 0]) / (2 * math.pi), 0.5 + np.arcsin(np.clip(Vs[:, 1], -1, 1)) / math.pi], axis=1)
    uvs = np.concatenate([(gv[:, [0, 2]] + [20.0, 2.5]) / 40.0] + [uv_s] * 6).astype(np.float32)
    instances = [(0, np.hstack([np.eye(3), np.zeros((3, 1))]))]
    for _ in range(n_instances):
        mesh = 1 + rng.u32() % 6
        R = _rotation(rng)
        S = np.diag([rng.f(0.3, 1.2), rng.f(0.3, 1.2), rng.f(0.3, 1.2)])
        t = np.array([rng.f(-8, 8), rng.f(0, 6), rng.f(6, 28)])
        instances.append((mesh, np.hstack([R @ S, t[:, None]])))
    checker = np.zeros((4, 4, 4), np.uint8)
    for y in range(4):
        for x in range(4):
            checker[y, x] = (230, 200, 60, 255) if (x + y) % 2 == 0 else (40, 90, 200, 255)
    sc = spheres_sunsky(n_spheres=0, width=width, height=height, spp=spp, max_depth=max_depth, palette=1)
    return dict(name="spheres_instanced", positions=np.concatenate(pos).astype(np.float32),
                normals=np.concatenate(nrm).astype(np.float32), indices=np.concatenate(idx).astype(np.int32),
                tri_material=np.concatenate(tm).astype(np.int32), materials=np.array(mats, MATERIAL_DTYPE),
                uvs=uvs, textures=[dict(texels=checker, format=0, address=(0, 0), uv_scale=(3.0, 2.0))],
                material_textures=[(3, 0)], mesh_begin=begin, instances=instances, lights=sc["lights"], env=sc["env"],
                camera=_camera(lens_radius=0.05, focal=15.0), width=width, height=height, spp=spp, max_depth=max_depth, seed=0)


def spheres_lobes(seed=0, n_spheres=48, subdiv=2, width=96, height=64, spp=4, max_depth=8, palette=24):
    """The sun/sky scene of spheres_sunsky with the remaining StandardMaterial lobes: rough
    diffuse (Oren-Nayar) bases, coats over them, thin rough dielectrics (thin-walled bubbles with
    tinted transmission), opacity < 1 on diffuse, conductor and coated materials (one fully
    transparent), and a sheen layer on every other material (rough dielectrics among them);
    shadow rays pass through the thin and transparent ones."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5), diffuse_roughness=0.5)]
    for i in range(palette):
        v = i % 6
        rgb = (rng.f(0.1, 0.9), rng.f(0.1, 0.9), rng.f(0.1, 0.9))
        rough = math.sqrt(rng.f(0.02, 0.4))
        if v == 0:
            mats.append(_material(KIND_DIFFUSE, rgb, diffuse_roughness=rng.f(0.2, 1.0)))
        elif v == 1:
            mats.append(_material(KIND_COATED, rgb, (1, 1, 1), rough, rng.f(1.3, 1.8), thickness=1.0, weight=1.0,
                                  diffuse_roughness=rng.f(0.2, 1.0)))
        elif v == 2:
            tint = (rng.f(0.4, 1.0), rng.f(0.4, 1.0), rng.f(0.4, 1.0))
            mats.append(_material(KIND_THIN_DIELECTRIC, tint, roughness=rough, ior=rng.f(1.3, 1.8),
                                  thickness=rng.f(0.2, 2.0)))
        elif v == 3:
            mats.append(_material(KIND_DIFFUSE, rgb, transparency=rng.f(0.2, 0.8)))
        elif v == 4:
            r, g = gold_artistic(rng)
            mats.append(_material(KIND_CONDUCTOR, r, g, rough, transparency=1.0 if i == 4 else rng.f(0.2, 0.8)))
        else:
            mats.append(_material(KIND_COATED, rgb, (1, 1, 1), rough, 1.5, thickness=1.0, weight=1.0,
                                  transparency=rng.f(0.2, 0.8)))
    for i in range(palette // 4):  # a few rough dielectrics (media interfaces) with sheen
        mats.append(_material(KIND_DIELECTRIC, roughness=math.sqrt(rng.f(0.02, 0.3)), ior=1.5))
    srng = Pcg32(seed + 101)
    for i in range(2, len(mats), 2):
        mats[i]["sheen_weight"] = srng.f(0.3, 1.0)
        mats[i]["sheen_color"] = (srng.f(0.3, 1.0), srng.f(0.3, 1.0), srng.f(0.3, 1.0))
        mats[i]["sheen_roughness"] = srng.f(0.1, 0.9)
    palette = len(mats) - 2
    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, lambda i: 2 + rng.u32() % palette)
    ss = load_sunsky()
    d = np.array([-0.6, 0.7, -0.4])
    d /= np.linalg.norm(d)
    lights = np.zeros(3, LIGHT_DTYPE)
    lights[0] = (LIGHT_DISTANT, 300.0, 280.0, 250.0, d[0], d[1], d[2], 0.5 * math.radians(2.0))
    sr, sd = ss["sun_radiance"], ss["sun_dir"]
    lights[1] = (LIGHT_DISTANT, sr[0], sr[1], sr[2], sd[0], sd[1], sd[2], ss["sun_radius"])
    lights[2] = (LIGHT_IMAGE, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    return dict(name="spheres_lobes", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats, MATERIAL_DTYPE), lights=lights, env=ss["sky"],
                camera=_camera(), width=width, height=height, spp=spp, max_depth=max_depth, seed=0)


def relmse(x, ref):
    """mean over pixels and channels of (x - y)^2 / (y^2 + 1e-2)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(ref, np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


def glass_slabs(slabs=(((0.25, 0.5, 0.8), 0.5), ((0.5, 0.5, 0.5), 2.0)), thickness=2.0, ior=1.0001, z0=5.0,
                ambient=1.0, width=64, height=32, max_depth=6):
    """Media absorption with a closed form: side by side (slab k covers x in [xs[k], xs[k + 1]], xs
    evenly from -40 to 40), slabs of one IOR but their own transmission colour and depth, each two
    facing quads at z = z0 and
    z0 + thickness of a smooth dielectric so close to the ambient IOR that the rays pass almost
    straight and unreflected (Fresnel 2.5e-9 at normal incidence), under an ambient light; the
    camera at the origin looks along +z.  A camera ray through slab k returns ambient x
    (colour_k ^ depth_k) ^ (thickness / cos theta), theta its angle to the z axis
    (path_renderer_stream.h:652-657, material/standard.cpp:169-173).  slabs: ((colour, depth), ...)
    in order of increasing x."""
    n = len(slabs)
    pos, idx, tri_mat = [], [], []
    xs = np.linspace(-40.0, 40.0, n + 1)
    for k in range(n):
        x0, x1 = xs[k + 1], xs[k]  # the camera's U is -x: slab 0 fills the left of the image
        for z in (z0, z0 + thickness):
            b = len(pos)
            pos += [[x0, -40, z], [x1, -40, z], [x1, 40, z], [x0, 40, z]]
            idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
            tri_mat += [1 + k, 1 + k]
    mats = np.zeros(1 + n, MATERIAL_DTYPE)
    for k, (col, depth) in enumerate(slabs):
        m = mats[1 + k]
        m["kind"] = KIND_DIELECTRIC
        m["ior"] = ior
        m["roughness"] = 0.0
        m["transmission_color"] = col
        m["transmission_depth"] = depth
    lights = np.zeros(1, LIGHT_DTYPE)
    lights[0] = (LIGHT_AMBIENT, ambient, ambient, ambient, 0, 0, 0, 0)
    return dict(name="glass_slabs", positions=np.array(pos, np.float32), indices=np.array(idx, np.int32),
                tri_material=np.array(tri_mat, np.int32), materials=mats, lights=lights, env=None,
                camera=_camera(fov_deg=60.0), width=width, height=height, spp=1, max_depth=max_depth, seed=0,
                rr_depth=100, slab_thickness=thickness, slab_z0=z0,
                slab_media=[(np.array(c, np.float64), d) for c, d in slabs])


def glass_slab_radiance(scene, dirs):
    """The closed form of glass_slabs for camera directions dirs (n x 3): the ambient radiance times
    the slab's medium colour to the power of the path length through it, for the slab the ray
    enters (returned as k per ray)."""
    d = np.asarray(dirs, np.float64)
    amb = float(scene["lights"][0]["r"])
    cos = d[:, 2] / np.linalg.norm(d, axis=1)
    media = scene["slab_media"]
    n = len(media)
    x_hit = d[:, 0] / d[:, 2] * scene["slab_z0"]
    xs = np.linspace(-40.0, 40.0, n + 1)
    k = np.clip(np.searchsorted(xs, x_hit) - 1, 0, n - 1)
    out = np.zeros((d.shape[0], 3))
    for j, (col, depth) in enumerate(media):
        sel = k == j
        out[sel] = amb * (col[None] ** depth) ** (scene["slab_thickness"] / cos[sel, None])
    return out, k


KIND_STANDARD = 5


def standard_material(base_color=(0.8, 0.8, 0.8), base_weight=1.0, metalness=0.0, diffuse_roughness=0.0,
                      specular_weight=1.0, specular_color=(1, 1, 1), specular_roughness=0.3, specular_ior=1.5,
                      transmission_weight=0.0, transmission_color=(1, 1, 1), transmission_depth=1.0, thin_walled=0,
                      thickness=1.0, coat_weight=0.0, coat_color=(1, 1, 1), coat_roughness=0.1, coat_ior=1.6,
                      coat_thickness=0.5, sheen_weight=0.0, sheen_color=(1, 1, 1), sheen_roughness=0.3, opacity=1.0,
                      emission=(0, 0, 0)):
    """A kind-5 record: StandardMaterial's uniform parameters with the reference's defaults
    (material/standard.cpp:27-81; coat roughness 0.1 instead of 0, the rough coat being the one
    modelled)."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["kind"] = KIND_STANDARD
    m["r"], m["g"], m["b"] = base_color
    m["base_weight"] = base_weight
    m["metalness"] = metalness
    m["diffuse_roughness"] = diffuse_roughness
    m["specular_weight"] = specular_weight
    m["er"], m["eg"], m["eb"] = specular_color
    m["roughness"] = specular_roughness
    m["ior"] = specular_ior
    m["transmission_weight"] = transmission_weight
    m["transmission_color"] = transmission_color
    m["transmission_depth"] = transmission_depth
    m["thin_walled"] = thin_walled
    m["thickness"] = thickness
    m["coat_weight"] = coat_weight
    m["coat_color"] = coat_color
    m["coat_roughness"] = coat_roughness
    m["coat_ior"] = coat_ior
    m["coat_thickness"] = coat_thickness
    m["sheen_weight"] = sheen_weight
    m["sheen_color"] = sheen_color
    m["sheen_roughness"] = sheen_roughness
    m["transparency"] = 1.0 - opacity
    m["emission"] = emission
    return m


def spheres_standard(seed=0, n_spheres=48, subdiv=2, width=96, height=64, spp=4, max_depth=8, palette=24):
    """The sun/sky scene of spheres_lobes with kind-5 StandardMaterials: partial metalness, partial
    transmission (a rough dielectric beside the coated diffuse base), coloured solid glass (the
    interior medium absorbs), coats over metal and over glass, thin-walled partial transmission
    under a coat (shadow rays pass through), opacity and sheen over mixes, base and specular weights
    below 1, and a coat over a diffuse-only base; plus coloured kind-2 glass."""
    rng = Pcg32(seed)
    mats = [_material(KIND_DIFFUSE, (0.8, 0.8, 0.8)), _material(KIND_DIFFUSE, (0.5, 0.5, 0.5))]
    for i in range(palette):
        v = i % 8
        rgb = (rng.f(0.1, 0.9), rng.f(0.1, 0.9), rng.f(0.1, 0.9))
        rough = math.sqrt(rng.f(0.02, 0.4))
        tint = (rng.f(0.3, 1.0), rng.f(0.3, 1.0), rng.f(0.3, 1.0))
        if v == 0:
            m = standard_material(rgb, metalness=rng.f(0.2, 0.8), specular_roughness=rough)
        elif v == 1:
            m = standard_material(rgb, transmission_weight=rng.f(0.2, 0.8), specular_roughness=rough,
                                  transmission_color=tint, transmission_depth=rng.f(0.5, 2.0))
        elif v == 2:
            m = standard_material(rgb, metalness=1.0, specular_roughness=rough, coat_weight=rng.f(0.4, 1.0),
                                  coat_color=tint, coat_roughness=math.sqrt(rng.f(0.01, 0.2)))
        elif v == 3:
            m = standard_material(rgb, transmission_weight=1.0, specular_roughness=rough, transmission_color=tint,
                                  transmission_depth=rng.f(0.3, 1.5), coat_weight=rng.f(0.3, 1.0),
                                  coat_roughness=math.sqrt(rng.f(0.01, 0.2)))
        elif v == 4:
            m = standard_material(rgb, metalness=rng.f(0.2, 0.5), transmission_weight=rng.f(0.2, 0.6),
                                  diffuse_roughness=rng.f(0.1, 0.9), specular_roughness=rough, coat_weight=rng.f(0.3, 0.9),
                                  sheen_weight=rng.f(0.2, 0.9), sheen_color=tint, opacity=rng.f(0.5, 0.95))
        elif v == 5:
            m = standard_material(rgb, transmission_weight=rng.f(0.3, 0.9), thin_walled=1, thickness=rng.f(0.2, 1.5),
                                  transmission_color=tint, specular_roughness=rough, coat_weight=rng.f(0.3, 0.9),
                                  specular_ior=1.0 if i == 5 else 1.5)
        elif v == 6:
            m = standard_material(rgb, base_weight=rng.f(0.3, 0.9), specular_weight=rng.f(0.2, 0.9),
                                  metalness=rng.f(0.2, 0.8), specular_roughness=rough)
        else:
            m = standard_material(rgb, specular_weight=0.0, coat_weight=rng.f(0.3, 1.0),
                                  metalness=rng.f(0.2, 0.6) if (i // 8) % 2 else 0.0, specular_roughness=rough)
        mats.append(m)
    for i in range(palette // 6):  # coloured kind-2 glass
        g = _material(KIND_DIELECTRIC, roughness=math.sqrt(rng.f(0.02, 0.3)), ior=1.5)
        g["transmission_color"] = (rng.f(0.3, 1.0), rng.f(0.3, 1.0), rng.f(0.3, 1.0))
        g["transmission_depth"] = rng.f(0.5, 2.0)
        mats.append(g)
    palette = len(mats) - 2
    pos, idx, tri_mat = _assemble(rng, n_spheres, subdiv, lambda i: 2 + rng.u32() % palette)
    ss = load_sunsky()
    d = np.array([-0.6, 0.7, -0.4])
    d /= np.linalg.norm(d)
    lights = np.zeros(3, LIGHT_DTYPE)
    lights[0] = (LIGHT_DISTANT, 300.0, 280.0, 250.0, d[0], d[1], d[2], 0.5 * math.radians(2.0))
    sr, sd = ss["sun_radiance"], ss["sun_dir"]
    lights[1] = (LIGHT_DISTANT, sr[0], sr[1], sr[2], sd[0], sd[1], sd[2], ss["sun_radius"])
    lights[2] = (LIGHT_IMAGE, 1.0, 1.0, 1.0, 0, 0, 0, 0)
    return dict(name="spheres_standard", positions=pos, indices=idx, tri_material=tri_mat,
                materials=np.array(mats, MATERIAL_DTYPE), lights=lights, env=ss["sky"],
                camera=_camera(), width=width, height=height, spp=spp, max_depth=max_depth, seed=0)
