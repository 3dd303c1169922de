"""ctypes binding of the oracle library (oracle/liboracle.so).  TEST INFRASTRUCTURE ONLY.

The oracle is a clean-room CPU restatement of ProtoRay's stream path tracer
(render/path_renderer_stream.h) and the pieces it calls.  Only tests/, the smoke check and
bench.py's CPU-baseline leg use it, always as the checker or the reported baseline, never as
the product path.

Parity status of each oracle function (what pins it):
  sampler, morton order          bit-exact vs reference golden vectors (tests/golden)
  camera rays, warps, 1D/2D dist reference golden vectors (few-ulp tolerance)
  LUT precompute                 reference golden vectors (albedo_tables.cpp output)
  diffuse / conductor / dielectric BSDF sample+eval (MultiBsdf-wrapped)
                                 reference golden vectors + invariants (sample = eval/pdf,
                                 white-furnace bounds)
  postIntersect + makeFrame      reference golden vectors
  closest / any hit              brute force over all triangles (Embree is absent: its own
                                 hit arithmetic is unpinned)
  lights (ambient/distant/image) and the integrator
                                 parity unpinned against the reference (their headers need TBB
                                 headers that are absent); pinned by invariants and closed forms
                                 (pdf normalisation, sample/eval consistency, constant-ambient
                                 and furnace images) in tests/test_oracle_integrator.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

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
LUT_SIZES = [("ggx_dir", 32 * 32), ("ggx_avg", 32), ("diel_eta_dir", 16 ** 3), ("diel_eta_avg", 16 ** 2),
             ("diel_rcpeta_dir", 16 ** 3), ("diel_rcpeta_avg", 16 ** 2), ("diel_refl_dir", 16 ** 3),
             ("diel_refl_avg", 16 ** 2), ("sheen_dir", 16 ** 2)]
LUT_FLOATS = sum(n for _, n in LUT_SIZES)


def _p(a, t=ctypes.c_float):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


_lib_name = "liboracle.so"


def select_library(name):
    """Chooses the build to load before the first call: liboracle.so (the parity checker, FP
    contraction off) or liboracle_bench.so (the timed CPU baseline, AVX2 + FMA)."""
    global _lib_name
    if _lib is not None and name != _lib_name:
        raise RuntimeError("the oracle library is already loaded")
    _lib_name = name


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, _lib_name)
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp, i, f, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
        L.orc_scene_create.restype = vp
        L.orc_scene_create.argtypes = [i, vp, vp, i, vp, vp, i, vp]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_set_uvs.argtypes = [vp, vp, i]
        L.orc_scene_add_texture.restype = i
        L.orc_scene_add_texture.argtypes = [vp, i, i, i, i, i, vp, vp]
        L.orc_material_set_texture.argtypes = [vp, i, i]
        L.orc_texture_lookup.argtypes = [vp, i, i, vp, vp]
        L.orc_hit_uv.argtypes = [vp, i, vp, vp, vp]
        L.orc_scene_set_instances.argtypes = [vp, i, vp, i, vp]
        L.orc_intersect_inst.argtypes = [vp, i, vp, i, vp, vp, vp, vp, vp]
        L.orc_scene_radius.restype = f
        L.orc_scene_radius.argtypes = [vp]
        L.orc_scene_set_lights.restype = i
        L.orc_scene_set_lights.argtypes = [vp, i, vp, i, i, vp, i]
        L.orc_get_light_cdf.argtypes = [vp, vp]
        L.orc_get_luts.argtypes = [vp, vp]
        L.orc_compute_luts.argtypes = [vp]
        L.orc_intersect.argtypes = [vp, i, vp, i, vp, vp, vp, vp]
        L.orc_render.restype = ctypes.c_longlong
        L.orc_render.argtypes = [vp, vp, i, i, i, i, i, i, u, i, i, f, f, i, vp, vp, vp]
        L.orc_sampler.argtypes = [i, vp, i, vp, vp, vp]
        L.orc_morton.argtypes = [i, i, vp]
        L.orc_camera_rays.argtypes = [vp, i, vp, vp]
        L.orc_warps.argtypes = [i, vp, vp, vp, vp, vp]
        L.orc_dist1d.argtypes = [vp, i, i, vp, vp, vp, vp]
        L.orc_dist2d.argtypes = [vp, i, i, i, vp, vp, vp, vp]
        L.orc_bsdf.argtypes = [i, vp, vp, vp, vp, vp, vp, vp]
        L.orc_post_intersect.argtypes = [i, vp, vp, vp, i, vp, vp, vp, vp]
        L.orc_env_eval.argtypes = [vp, i, vp, vp]
        L.orc_light_sample.argtypes = [vp, i, vp, vp, vp]
        L.orc_light_count.argtypes = [vp]
        L.orc_adjoint.restype = ctypes.c_double
        L.orc_adjoint.argtypes = [vp, vp, i, i, i, u, i, i, f, vp, i, vp]
        L.orc_bsdf_grad.argtypes = [i, vp, vp, vp, vp]
        L.orc_bsdf_mat.argtypes = [i, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_set_pixel_sample.argtypes = [vp, i, f, f, i, f, f, i]
        L.orc_pixel_sample.argtypes = [vp, i, vp, vp]
        _lib = L
    return _lib


def c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def sampler(seed_pid_idx, ndim):
    q = np.ascontiguousarray(seed_pid_idx, dtype=np.uint32)
    n = q.shape[0]
    o1 = np.zeros((n, ndim), np.float32)
    o2 = np.zeros((n, ndim, 2), np.float32)
    o3 = np.zeros((n, ndim, 3), np.float32)
    lib().orc_sampler(n, q.ctypes.data, ndim, o1.ctypes.data, o2.ctypes.data, o3.ctypes.data)
    return o1, o2, o3


def morton(n, tile):
    xy = np.zeros((n, 2), np.int32)
    lib().orc_morton(n, tile, xy.ctypes.data)
    return xy


def camera_rays(cfg, s):
    cfg, s = c32(cfg), c32(s)
    out = np.zeros((s.shape[0], 2, 7), np.float32)
    lib().orc_camera_rays(cfg.ctypes.data, s.shape[0], s.ctypes.data, out.ctypes.data)
    return out


def warps(s):
    s = c32(s)
    n = s.shape[0]
    c, sp, d, cn = (np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32), np.zeros((n, 2), np.float32),
                    np.zeros((n, 2, 4), np.float32))
    lib().orc_warps(n, s.ctypes.data, c.ctypes.data, sp.ctypes.data, d.ctypes.data, cn.ctypes.data)
    return c, sp, d, cn


def dist1d(w, s):
    w, s = c32(w), c32(s)
    n = s.shape[0]
    idx = np.zeros(n, np.int32)
    pdf_sr = np.zeros((n, 2), np.float32)
    pdf_i = np.zeros(w.shape[0] + 2, np.float32)
    lib().orc_dist1d(w.ctypes.data, w.shape[0], n, s.ctypes.data, idx.ctypes.data, pdf_sr.ctypes.data, pdf_i.ctypes.data)
    return idx, pdf_sr, pdf_i


def dist2d(w, s):
    w, s = c32(w), c32(s)
    H, W = w.shape
    n = s.shape[0]
    idx = np.zeros((n, 2), np.int32)
    pdf_sr = np.zeros((n, 3), np.float32)
    cells = np.zeros((H, W), np.float32)
    lib().orc_dist2d(w.ctypes.data, W, H, n, s.ctypes.data, idx.ctypes.data, pdf_sr.ctypes.data, cells.ctypes.data)
    return idx, pdf_sr, cells


def compute_luts():
    out = np.zeros(LUT_FLOATS, np.float32)
    lib().orc_compute_luts(out.ctypes.data)
    res, off = {}, 0
    for name, n in LUT_SIZES:
        res[name] = out[off:off + n].copy()
        off += n
    return out, res


def bsdf(luts_flat, params, geo, s):
    """params: n x [kind r g b er eg eb roughness eta] (+ [thickness weight] for coated rows)."""
    params, geo, s = c32(params), c32(geo), c32(s)
    if params.shape[1] == 9:
        params = np.ascontiguousarray(np.concatenate([params, np.zeros((params.shape[0], 2), np.float32)], 1))
    n = params.shape[0]
    so = np.zeros((n, 7), np.float32)
    st = np.zeros(n, np.int32)
    ev = np.zeros((n, 7), np.float32)  # value rgb, pdf, albedo rgb
    lib().orc_bsdf(n, c32(luts_flat).ctypes.data, params.ctypes.data, geo.ctypes.data, s.ctypes.data,
                   so.ctypes.data, st.ctypes.data, ev.ctypes.data)
    return so, st, ev


def bsdf_mat(luts_flat, mats, geo, s, inside=None):
    """The renderer's whole BSDF for material records (MATERIAL_DTYPE) on supplied frames:
    (sample n x 7 [weight wi pdf], type n, eval n x 4 [value pdf], aux n x 9 [albedo, normal,
    transparency of a transparent material]).  inside (n, optional): the path travels in the
    material's interior medium, so a media interface is seen from its back."""
    mats = np.ascontiguousarray(mats, MATERIAL_DTYPE)
    geo, s = c32(geo), c32(s)
    n = mats.shape[0]
    so = np.zeros((n, 7), np.float32)
    st = np.zeros(n, np.int32)
    ev = np.zeros((n, 4), np.float32)
    aux = np.zeros((n, 9), np.float32)
    ins = None if inside is None else i32(inside)
    lib().orc_bsdf_mat(n, c32(luts_flat).ctypes.data, mats.ctypes.data, geo.ctypes.data, s.ctypes.data,
                       so.ctypes.data, st.ctypes.data, ev.ctypes.data, aux.ctypes.data,
                       None if ins is None else ins.ctypes.data)
    return so, st, ev, aux


def bsdf_grad(luts_flat, params, geo):
    """FP64 value and parameter derivatives of the BSDF: n x [value rgb, d/d albedo rgb, d/d alpha rgb].
    params rows as bsdf() (11 columns; 9-column rows are padded)."""
    params, geo = c32(params), c32(geo)
    if params.shape[1] == 9:
        params = np.ascontiguousarray(np.concatenate([params, np.zeros((params.shape[0], 2), np.float32)], 1))
    out = np.zeros((params.shape[0], 9), np.float64)
    lib().orc_bsdf_grad(params.shape[0], c32(luts_flat).ctypes.data, params.ctypes.data, geo.ctypes.data, out.ctypes.data)
    return out


def post_intersect(verts, normals, indices, rays, prim):
    verts, rays = c32(verts), c32(rays)
    indices, prim = i32(indices), i32(prim)
    nrm = None if normals is None else c32(normals)
    n = rays.shape[0]
    ctx = np.zeros((n, 16), np.float32)
    fr = np.zeros((n, 9), np.float32)
    lib().orc_post_intersect(verts.shape[0], verts.ctypes.data, None if nrm is None else nrm.ctypes.data,
                             indices.ctypes.data, n, rays.ctypes.data, prim.ctypes.data, ctx.ctypes.data, fr.ctypes.data)
    return ctx, fr


class Scene:
    """Oracle scene: geometry + materials + lights, with its own CPU BVH."""

    def __init__(self, positions, indices, tri_material, materials, normals=None):
        self.positions = c32(positions).reshape(-1, 3)
        self.indices = i32(indices).reshape(-1, 3)
        self.tri_material = i32(tri_material)
        self.materials = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        self.normals = None if normals is None else c32(normals).reshape(-1, 3)
        self.h = lib().orc_scene_create(self.positions.shape[0], self.positions.ctypes.data,
                                        None if self.normals is None else self.normals.ctypes.data,
                                        self.indices.shape[0], self.indices.ctypes.data, self.tri_material.ctypes.data,
                                        self.materials.shape[0], self.materials.ctypes.data)

    def set_lights(self, lights, env_rgb=None, lut_max=128):
        self.lights = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        env = None if env_rgb is None else c32(env_rgb)
        H, W = (0, 0) if env is None else env.shape[:2]
        self.env = env
        rc = lib().orc_scene_set_lights(self.h, self.lights.shape[0], self.lights.ctypes.data, W, H,
                                        None if env is None else env.ctypes.data, lut_max)
        if rc != 0:
            raise ValueError("invalid light list")

    def set_textures(self, scene):
        """UV set 0, textures and baseColor bindings of a scene dict (keys "uvs", "textures",
        "material_textures"; see scenes/scenegen.py), as the renderer gets them."""
        uv = scene.get("uvs")
        if uv is not None:
            self._uv = c32(uv)
            lib().orc_scene_set_uvs(self.h, self._uv.ctypes.data, self._uv.shape[0])
        for t in scene.get("textures", ()):
            self.add_texture(t)
        for mat, tex in scene.get("material_textures", ()):
            lib().orc_material_set_texture(self.h, int(mat), int(tex))
        if scene.get("instances"):
            self.set_instances(scene["mesh_begin"], scene["instances"])

    def set_instances(self, mesh_begin, instances):
        """Two-level scene: meshes as triangle ranges, instances [(mesh, 3 x 4 [L | t]), ...]
        (empty: one level).  Set the lights afterwards (their powers use the scene radius)."""
        mb = i32(mesh_begin)
        rec = np.zeros(len(instances), np.dtype([("mesh", "<i4"), ("xfm", "<f4", 12)]))
        for k, (mesh, xfm) in enumerate(instances):
            rec[k]["mesh"] = mesh
            rec[k]["xfm"] = np.asarray(xfm, np.float32).reshape(12)
        self._inst = (mb, rec)
        lib().orc_scene_set_instances(self.h, max(mb.shape[0] - 1, 0), mb.ctypes.data, rec.shape[0], rec.ctypes.data)

    def intersect_inst(self, rays, mode=0):
        """Closest hits with their instances: (prim, t, u, v, inst); mode 1 = brute force."""
        rays = c32(rays)
        n = rays.shape[0]
        prim, inst = np.zeros(n, np.int32), np.zeros(n, np.int32)
        t, u, v = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        lib().orc_intersect_inst(self.h, n, rays.ctypes.data, mode, prim.ctypes.data, t.ctypes.data, u.ctypes.data,
                                 v.ctypes.data, inst.ctypes.data)
        return prim, t, u, v, inst

    def add_texture(self, t):
        texels = np.ascontiguousarray(t["texels"], np.float32 if t["format"] == 2 else np.uint8)
        H, W = texels.shape[:2]
        au, av = t.get("address", (0, 0))
        prm = c32(list(t.get("uv_scale", (1.0, 1.0))) + list(t.get("uv_offset", (0.0, 0.0))))
        return lib().orc_scene_add_texture(self.h, W, H, int(t["format"]), int(au), int(av), prm.ctypes.data, texels.ctypes.data)

    def hit_uv(self, prim, uv):
        """UV set 0 at hits (prim, barycentric u, v)."""
        prim, uv = i32(prim), c32(uv).reshape(-1, 2)
        out = np.zeros_like(uv)
        lib().orc_hit_uv(self.h, prim.shape[0], prim.ctypes.data, uv.ctypes.data, out.ctypes.data)
        return out

    def texture_lookup(self, tex, uv):
        """The bilinear lookup of texture `tex` at n (u, v) pairs (n x 3)."""
        uv = c32(uv).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), np.float32)
        lib().orc_texture_lookup(self.h, tex, uv.shape[0], uv.ctypes.data, out.ctypes.data)
        return out

    def set_pixel_sample(self, jitter=None, filter_type=0, width=0.0, sigma=0.0, bias=1):
        """Pixel sample of later renders: a fixed jitter (jx, jy), or the pixel filter (0 box,
        1 triangle, 2 Gaussian, 3 Blackman-Harris) applied to the sampler's value."""
        jx, jy = (0.0, 0.0) if jitter is None else jitter
        lib().orc_set_pixel_sample(self.h, 0 if jitter is None else 1, jx, jy, filter_type, width, sigma, bias)

    def pixel_sample(self, s2):
        """Pixel-sample positions (n x 2) for sampler values s2 (n x 2) under the jitter / filter."""
        s2 = c32(s2)
        out = np.zeros_like(s2)
        lib().orc_pixel_sample(self.h, s2.shape[0], s2.ctypes.data, out.ctypes.data)
        return out

    @property
    def radius(self):
        return lib().orc_scene_radius(self.h)

    def light_count(self):
        """Lights in the sampling list: the area light (emissive triangles) first, if any."""
        return lib().orc_light_count(self.h)

    def light_cdf(self):
        out = np.zeros(self.light_count() + 1, np.float32)
        lib().orc_get_light_cdf(self.h, out.ctypes.data)
        return out

    def intersect(self, rays, mode=0):
        rays = c32(rays)
        n = rays.shape[0]
        prim = np.zeros(n, np.int32)
        t, u, v = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        lib().orc_intersect(self.h, n, rays.ctypes.data, mode, prim.ctypes.data, t.ctypes.data, u.ctypes.data, v.ctypes.data)
        return prim, t, u, v

    def env_eval(self, dirs):
        dirs = c32(dirs)
        out = np.zeros((dirs.shape[0], 4), np.float32)
        lib().orc_env_eval(self.h, dirs.shape[0], dirs.ctypes.data, out.ctypes.data)
        return out

    def light_sample(self, s3, org=(0.0, 0.0, 0.0)):
        """n x [wi3 weight3 pdf dist] for light samples taken from the point org."""
        s3 = c32(s3)
        o = c32(org)
        out = np.zeros((s3.shape[0], 8), np.float32)
        lib().orc_light_sample(self.h, s3.shape[0], s3.ctypes.data, o.ctypes.data, out.ctypes.data)
        return out

    def adjoint(self, camera, W, H, target, spp, seed=0, max_depth=4, rr_depth=8, min_weight=1e-4, threads=None):
        """FP64 gradient of mean((I - target)^2) w.r.t. (albedo r, g, b, alpha) per material: (loss, grad[nMat, 4])."""
        cam = np.ascontiguousarray(camera, CAMERA_DTYPE)
        t = c32(target)
        nmat = self.materials.shape[0]
        grad = np.zeros((nmat, 4), np.float64)
        threads = threads or min(os.cpu_count() or 1, 16)
        loss = lib().orc_adjoint(self.h, cam.ctypes.data, W, H, spp, seed, max_depth, rr_depth, min_weight, t.ctypes.data,
                                 threads, grad.ctypes.data)
        return loss, grad

    def render(self, camera, W, H, spp_start=0, spp_count=1, seed=0, max_depth=5, rr_depth=8, min_weight=1e-4,
               max_radiance=3.4028234663852886e38, threads=None, rows=None, aovs=(), accum=None, record=False):
        """Accumulates spp_count sample passes into `accum` (dict of FP32 sums); returns (accum, rays)."""
        npix = W * H
        if accum is None:
            accum = {"color": np.zeros((3, npix), np.float32), "weight": np.zeros(npix, np.float32)}
            for a in aovs:
                if a == "color2":  # sums of squares of the color samples (AccumFlag::SampleVar)
                    accum[a] = np.zeros((3, npix), np.float32)
                    continue
                accum[a] = np.zeros((1, npix) if a == "depth" else (3, npix), np.float32)
                accum[a + "W"] = np.zeros(npix, np.float32)
        names = ["color", "weight", "albedo", "albedoW", "albedo1", "albedo1W", "normal", "normalW", "normal1",
                 "normal1W", "depth", "depthW"]
        bufs = (ctypes.c_void_p * 14)()
        for k, nm in enumerate(names):
            bufs[k] = accum[nm].ctypes.data if nm in accum else None
        bufs[13] = accum["color2"].ctypes.data if "color2" in accum else None
        sampleL = hashes = first = None
        if record:
            sampleL = np.zeros((3, npix), np.float32)
            hashes = np.zeros(npix, np.uint32)
            first = np.full(npix, -3, np.int32)
            bufs[12] = sampleL.ctypes.data
        cam = np.ascontiguousarray(camera, dtype=CAMERA_DTYPE)
        r0, r1 = (0, H) if rows is None else rows
        rays = lib().orc_render(self.h, cam.ctypes.data, W, H, r0, r1, spp_start, spp_count, seed, max_depth, rr_depth,
                                min_weight, max_radiance, threads or os.cpu_count(), bufs,
                                None if hashes is None else hashes.ctypes.data,
                                None if first is None else first.ctypes.data)
        if record:
            accum["_sampleL"], accum["_hash"], accum["_first"] = sampleL, hashes, first
        return accum, rays

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None
