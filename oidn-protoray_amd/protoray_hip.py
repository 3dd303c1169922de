"""Python binding of the MI355X stream path tracer (libprt_hip.so, C ABI in include/prt_hip.h).

``DeviceHip`` mirrors the reference's ``Device`` interface (render/device.h:56-114) for the
stream path renderer: initScene / initLights / initRenderer / initSampler / initCamera /
initFrame / clearFrame / render / blitFrame / getFrameError / queryRay.  Arguments keep the
reference's meaning (Props names in the docstrings); errors raise like the reference's C++
exceptions (ValueError for std::invalid_argument / std::logic_error, RuntimeError otherwise).

The library is loaded from this directory; there is no fallback: without the built HIP
library (``make -C oidn-protoray_amd``) every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprt_hip.so")

PRT_OK, PRT_ERR_INVALID, PRT_ERR_HIP, PRT_ERR_NOMEM, PRT_ERR_STATE = 0, -1, -2, -3, -4
AOV_ALBEDO, AOV_ALBEDO1, AOV_NORMAL, AOV_NORMAL1, AOV_DEPTH, AOV_VAR, AOV_HALF = 1, 2, 4, 8, 16, 32, 64
FILTER_BOX, FILTER_TRIANGLE, FILTER_GAUSSIAN, FILTER_BLACKMAN_HARRIS = 0, 1, 2, 3
BUF_COLOR, BUF_ALBEDO, BUF_ALBEDO1, BUF_NORMAL, BUF_NORMAL1, BUF_DEPTH = range(6)
BATCH_COLOR, BATCH_ALBEDO1, BATCH_NORMAL1, BATCH_DEPTH, BATCH_SNAPSHOT = 0, 1, 2, 3, 16
DENOISE_ALBEDO, DENOISE_NORMAL = 1, 2
INSTANCE_DTYPE = np.dtype([("mesh", "<i4"), ("xfm", "<f4", 12)])  # prt_instance
TEX_RGBA8_SRGB, TEX_RGBA8_LINEAR, TEX_RGB32F = 0, 1, 2
TEX_REPEAT, TEX_CLAMP = 0, 1


class Texture(ctypes.Structure):  # prt_texture (image/image_texture.h:148-300)
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("format", ctypes.c_int32),
                ("address_u", ctypes.c_int32), ("address_v", ctypes.c_int32),
                ("uv_scale", ctypes.c_float * 2), ("uv_offset", ctypes.c_float * 2), ("texels", ctypes.c_void_p)]
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
PIXEL_HIT_DTYPE = np.dtype([("prim", "<i4"), ("material", "<i4"), ("material_type", "<i4"), ("p", "<f4", 3),
                            ("Ng", "<f4", 3), ("N", "<f4", 3), ("U", "<f4", 3), ("V", "<f4", 3), ("uv", "<f4", 2),
                            ("T", "<f4", 3), ("eps", "<f4"), ("dist", "<f4"), ("depth", "<f4")])
LIGHT_DTYPE = np.dtype([("kind", "<i4"), ("r", "<f4"), ("g", "<f4"), ("b", "<f4"),
                        ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("angle", "<f4")])
CAMERA_DTYPE = np.dtype([("pos", "<f4", 3), ("U", "<f4", 3), ("V", "<f4", 3), ("N", "<f4", 3),
                         ("fov", "<f4"), ("nearClip", "<f4"), ("lensRadius", "<f4"),
                         ("focalDistance", "<f4")])


class PrtStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_double), ("ms", ctypes.c_double), ("mray", ctypes.c_double),
                ("spp", ctypes.c_int32)]


class RendererParams(ctypes.Structure):
    _fields_ = [("maxDepth", ctypes.c_int32), ("rrDepth", ctypes.c_int32), ("minWeight", ctypes.c_float),
                ("maxRadiance", ctypes.c_float)]


_lib = None

# (name, argtypes) of every entry point in include/prt_hip.h
_vp, _i, _u, _f = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float
SIGNATURES = {
    "prt_context_create": [_i, ctypes.POINTER(ctypes.c_void_p)],
    "prt_context_destroy": [_vp],
    "prt_last_error": [],
    "prt_debug_copy": [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)],
    "prt_device_synchronize": [_vp],
    "prt_get_info": [_vp, ctypes.c_char_p, _i],
    "prt_scene_set_geometry": [_vp, _i, _vp, _vp, _i, _vp, _vp],
    "prt_scene_set_materials": [_vp, _i, _vp],
    "prt_scene_set_lights": [_vp, _i, _vp, _i, _i, _vp, _i],
    "prt_bvh_build": [_vp, ctypes.POINTER(ctypes.c_double)],
    "prt_renderer_init": [_vp, _vp],
    "prt_sampler_init": [_vp, _u],
    "prt_set_camera": [_vp, _vp],
    "prt_frame_init": [_vp, _i, _i, _u],
    "prt_frame_clear": [_vp],
    "prt_render": [_vp, _i, ctypes.POINTER(PrtStats)],
    "prt_blit": [_vp, _i, _vp, _vp],
    "prt_get_accum": [_vp, _vp, _vp],
    "prt_blit_half": [_vp, _vp],
    "prt_set_jitter": [_vp, _i, _f, _f],
    "prt_get_jitter": [_vp, _vp, _vp],
    "prt_set_pixel_filter": [_vp, _i, _f, _f, _i],
    "prt_get_filter_width": [_vp, ctypes.POINTER(ctypes.c_float)],
    "prt_frame_error": [_vp, ctypes.POINTER(ctypes.c_float)],
    "prt_query_rays": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "prt_query_occlusion": [_vp, _i, _vp, _vp],
    "prt_query_pixel": [_vp, ctypes.c_float, ctypes.c_float, _i, _vp, _vp],
    "prt_query_ray_hits": [_vp, _vp, _i, _vp, _vp],
    "prt_frame_error_exposure": [_vp, _f, ctypes.POINTER(ctypes.c_float)],
    "prt_camera_matrices": [_vp, _i, _i, _vp, _vp, _vp],
    "prt_scene_bounds": [_vp, _vp, _vp],
    "prt_query_rays_device": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "prt_query_rays_device_noalloc": [_vp, _i, _vp, _vp, _vp, _vp, _vp],
    "prt_debug_enable": [_vp, _i],
    "prt_debug_get": [_vp, _vp, _vp, _vp],
    "prt_debug_sampler": [_vp, _i, _vp, _i, _vp, _vp, _vp],
    "prt_debug_luts": [_vp, _vp, ctypes.c_int32, _vp],
    "prt_debug_light_cdf": [_vp, _vp, ctypes.c_int32, _vp],
    "prt_debug_trace_steps": [_vp, _vp],
    "prt_debug_count_steps": [_vp, _i],
    "prt_adjoint": [_vp, _vp, _i, _vp, _vp, _i, _vp],
    "prt_denoiser_create": [_i, _vp],
    "prt_denoiser_destroy": [_vp],
    "prt_denoiser_init": [_vp, _vp],
    "prt_denoiser_execute": [_vp, _vp],
    "prt_denoiser_get_version": [_vp, ctypes.c_char_p, _i],
    "prt_denoiser_kernel_ms": [_vp, _vp],
    "prt_denoiser_param_count": [_vp, _vp],
    "prt_denoiser_get_weights": [_vp, _vp, ctypes.c_int64],
    "prt_denoiser_set_weights": [_vp, _vp, ctypes.c_int64],
    "prt_denoiser_train_step": [_vp, _vp, _vp, _i, _i, _i, _f, _vp, _vp],
    "prt_denoiser_gradient": [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, ctypes.c_int64],
    "prt_denoiser_execute_frame": [_vp, _vp, _i, _f, _vp],
    "prt_scene_set_uvs": [_vp, _vp, _i],
    "prt_scene_set_instances": [_vp, _i, _vp, _i, _vp],
    "prt_query_rays_inst": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "prt_scene_file_load_instanced": [ctypes.c_char_p, _vp, _vp],
    "prt_scene_file_get_instances": [_vp, _vp, _vp, _vp, _vp],
    "prt_scene_set_textures": [_vp, _i, _vp],
    "prt_material_set_texture": [_vp, _i, _i, _i],
    "prt_denoiser_execute_batch": [_vp, _vp, _i, _i, _i, _f, _vp],
    "prt_denoiser_get_output": [_vp, _vp, _i, _i],
    "prt_batch_create": [_i, _vp],
    "prt_batch_destroy": [_vp],
    "prt_batch_add_frame": [_vp, _vp, _vp, _i, _i, _u, _i, _vp, _i, _u, _vp],
    "prt_batch_tune": [_vp, _i, _vp, _vp, _vp],
    "prt_batch_render": [_vp, _vp, _vp],
    "prt_batch_get": [_vp, _i, _i, _vp],
    "prt_debug_adjoint_replay": [_vp, _u, _vp],
    "prt_debug_bvh_stats": [_vp, _vp],
    "prt_debug_kernel_ms": [_vp, _vp, _vp, _vp],
    "prt_scene_file_load": [ctypes.c_char_p, _vp, _vp],
    "prt_scene_file_get": [_vp, _vp, _vp, _vp, _vp, _vp],
    "prt_scene_file_material_name": [_vp, _i, ctypes.c_char_p, _i],
    "prt_scene_file_destroy": [_vp],
}


def load_library(path=LIB_PATH):
    """Loads libprt_hip.so and declares every entry point; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libprt_hip.so is not built (run `make -C oidn-protoray_amd`); "
                               "there is no CPU fallback")
        L = ctypes.CDLL(path)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "prt_last_error" else ctypes.c_int32
        _lib = L
    return _lib


def _check(rc):
    if rc != PRT_OK:
        msg = _lib.prt_last_error().decode()
        if rc in (PRT_ERR_INVALID, PRT_ERR_STATE):
            raise ValueError(msg)
        raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else a.ctypes.data


def camera_matrices(camera, width, height):
    """(worldToViewD3D, viewToClipD3D, worldToRaster) of a CAMERA_DTYPE record for an image of
    width x height, each 4 x 4 in row-major notation (camera/pinhole_camera.h:41-68).  Host only."""
    cam = np.ascontiguousarray(camera, CAMERA_DTYPE)
    out = np.zeros((3, 16), np.float32)
    _check(load_library().prt_camera_matrices(_ptr(cam), int(width), int(height), _ptr(out[0]), _ptr(out[1]),
                                              _ptr(out[2])))
    return tuple(m.reshape(4, 4) for m in out)


class DeviceHip:
    """Device for the stream path renderer on one HIP device (cf. DeviceCpu)."""

    def __init__(self, device=0):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.prt_context_create(device, ctypes.byref(h)))
        self.h = h
        self.size = None
        self.aov_mask = 0

    def close(self):
        if getattr(self, "h", None):
            _lib.prt_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def getInfo(self):
        buf = ctypes.create_string_buffer(256)
        _check(_lib.prt_get_info(self.h, buf, 256))
        return buf.value.decode()

    # --- scene (Device::initScene, Scene ctor render/scene.cpp:65-134) -------------------
    def initScene(self, positions, indices, tri_material, materials, normals=None):
        self._pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self._idx = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        self._mat = np.ascontiguousarray(tri_material, np.int32)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        _check(_lib.prt_scene_set_geometry(self.h, self._pos.shape[0], _ptr(self._pos), _ptr(nrm),
                                           self._idx.shape[0], _ptr(self._idx), _ptr(self._mat)))
        m = np.ascontiguousarray(materials, MATERIAL_DTYPE)
        _check(_lib.prt_scene_set_materials(self.h, m.shape[0], _ptr(m)))
        ms = ctypes.c_double(0)
        _check(_lib.prt_bvh_build(self.h, ctypes.byref(ms)))
        self.build_ms = ms.value
        return self.build_ms

    def initLights(self, lights, env_rgb=None, lut_size=128):
        """lights: LIGHT_DTYPE records in reference order; env_rgb: H x W x 3, row 0 = nadir."""
        l = np.ascontiguousarray(lights, LIGHT_DTYPE)
        env = None if env_rgb is None else np.ascontiguousarray(env_rgb, np.float32)
        H, W = (0, 0) if env is None else env.shape[:2]
        _check(_lib.prt_scene_set_lights(self.h, l.shape[0], _ptr(l), W, H, _ptr(env), lut_size))

    # --- renderer (RendererFactoryStream::make "pathStream") ------------------------------
    def initRenderer(self, maxDepth=100, rrDepth=8, minWeight=1e-4, maxRadiance=3.4028234663852886e38):
        p = RendererParams(maxDepth, rrDepth, minWeight, maxRadiance)
        _check(_lib.prt_renderer_init(self.h, ctypes.byref(p)))

    def initSampler(self, seed):
        _check(_lib.prt_sampler_init(self.h, seed))

    def initCamera(self, camera):
        c = np.ascontiguousarray(camera, CAMERA_DTYPE)
        _check(_lib.prt_set_camera(self.h, _ptr(c)))

    def initInstances(self, mesh_begin, instances):
        """Two-level scene: meshes as triangle ranges (mesh_begin, n_meshes + 1 entries) placed by
        instances [(mesh, 3 x 4 [L | t]), ...]; rebuilds the trees (set the lights afterwards)."""
        if not instances:
            _check(_lib.prt_scene_set_instances(self.h, 0, None, 0, None))
        else:
            mb = np.ascontiguousarray(mesh_begin, np.int32)
            rec = np.zeros(len(instances), INSTANCE_DTYPE)
            for k, (mesh, xfm) in enumerate(instances):
                rec[k]["mesh"] = mesh
                rec[k]["xfm"] = np.asarray(xfm, np.float32).reshape(12)
            self._inst = (mb, rec)
            _check(_lib.prt_scene_set_instances(self.h, mb.shape[0] - 1, _ptr(mb), rec.shape[0], _ptr(rec)))
        ms = ctypes.c_double(0)
        _check(_lib.prt_bvh_build(self.h, ctypes.byref(ms)))
        self.build_ms = ms.value
        return self.build_ms

    def queryRaysInst(self, rays):
        """Closest hits with their instances: (prim, t, u, v, inst)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim, inst = np.zeros(n, np.int32), np.zeros(n, np.int32)
        t, u, v = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(_lib.prt_query_rays_inst(self.h, n, _ptr(rays), _ptr(prim), _ptr(t), _ptr(u), _ptr(v), _ptr(inst)))
        return prim, t, u, v, inst

    def initUVs(self, uv):
        """Texture coordinate set 0 (nverts x 2), or None to drop it."""
        if uv is None:
            _check(_lib.prt_scene_set_uvs(self.h, None, 0))
            return
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        _check(_lib.prt_scene_set_uvs(self.h, _ptr(uv), uv.shape[0]))

    def initTextures(self, textures):
        """Textures as dicts: texels (H x W x 4 uint8 for formats 0 / 1, H x W x 3 float32 for 2;
        image rows top to bottom), format, address (u, v), uv_scale, uv_offset."""
        recs = (Texture * max(len(textures), 1))()
        keep = []
        for k, t in enumerate(textures):
            tx = np.ascontiguousarray(t["texels"], np.float32 if t["format"] == TEX_RGB32F else np.uint8)
            keep.append(tx)
            au, av = t.get("address", (TEX_REPEAT, TEX_REPEAT))
            su, sv = t.get("uv_scale", (1.0, 1.0))
            ou, ov = t.get("uv_offset", (0.0, 0.0))
            recs[k] = Texture(tx.shape[1], tx.shape[0], int(t["format"]), int(au), int(av), (su, sv), (ou, ov), tx.ctypes.data)
        _check(_lib.prt_scene_set_textures(self.h, len(textures), recs))

    def setMaterialTexture(self, material, texture, parameter=0):
        """Binds texture index `texture` (-1: none) to the material's base colour."""
        _check(_lib.prt_material_set_texture(self.h, material, parameter, texture))

    def initFrame(self, width, height, aov_mask=0):
        _check(_lib.prt_frame_init(self.h, width, height, aov_mask))
        self.size = (width, height)
        self.aov_mask = aov_mask

    def clearFrame(self):
        _check(_lib.prt_frame_clear(self.h))

    def render(self, spp=1):
        """Renders `spp` sample passes; returns the stats dict (rays, ms, mray, spp)."""
        st = PrtStats()
        _check(_lib.prt_render(self.h, spp, ctypes.byref(st)))
        return {"rays": st.rays, "ms": st.ms, "mray": st.mray, "spp": st.spp}

    def blitFrame(self, buffer=BUF_COLOR, variance=False):
        W, H = self.size
        ch = 1 if buffer == BUF_DEPTH else 3
        out = np.zeros((H, W, ch) if ch == 3 else (H, W), np.float32)
        var = np.zeros((H, W, 3), np.float32) if variance else None
        _check(_lib.prt_blit(self.h, buffer, _ptr(out), _ptr(var)))
        return (out, var) if variance else out

    def blitHalf(self):
        """Mean of the class-B samples (AccumBuffer3f::blit destHalf); needs AOV_HALF."""
        W, H = self.size
        out = np.zeros((H, W, 3), np.float32)
        _check(_lib.prt_blit_half(self.h, _ptr(out)))
        return out

    def setJitter(self, jitter):
        """Device::setJitter: the sub-pixel position (x, y) in [0, 1]^2 of every pixel's sample."""
        _check(_lib.prt_set_jitter(self.h, 1, float(jitter[0]), float(jitter[1])))

    def disableJitter(self):
        _check(_lib.prt_set_jitter(self.h, 0, 0.0, 0.0))

    def hasJitter(self):
        has = ctypes.c_int32(0)
        _check(_lib.prt_get_jitter(self.h, ctypes.byref(has), None))
        return bool(has.value)

    def getJitter(self):
        has = ctypes.c_int32(0)
        j = np.zeros(2, np.float32)
        _check(_lib.prt_get_jitter(self.h, ctypes.byref(has), _ptr(j)))
        if not has.value:
            raise ValueError("no jitter set")
        return float(j[0]), float(j[1])

    def initFilter(self, filter_type, width=0.0, sigma=0.0, bias=True):
        """Device::initFilter (sample mode): FILTER_BOX / TRIANGLE / GAUSSIAN / BLACKMAN_HARRIS."""
        _check(_lib.prt_set_pixel_filter(self.h, int(filter_type), float(width), float(sigma), 1 if bias else 0))

    def getFilterWidth(self):
        w = ctypes.c_float(0)
        _check(_lib.prt_get_filter_width(self.h, ctypes.byref(w)))
        return w.value

    def getAccum(self):
        W, H = self.size
        s = np.zeros((3, W * H), np.float32)
        w = np.zeros(W * H, np.float32)
        _check(_lib.prt_get_accum(self.h, _ptr(s), _ptr(w)))
        return s, w

    def getFrameError(self, exposure=1.0):
        e = ctypes.c_float(0)
        _check(_lib.prt_frame_error_exposure(self.h, float(exposure), ctypes.byref(e)))
        return e.value

    def queryRayHits(self, ray, max_hits=1):
        """Hit properties along one ray [ox oy oz dx dy dz near far] (Renderer::queryRay)."""
        r = np.ascontiguousarray(ray, np.float32)
        hits = np.zeros(max_hits, PIXEL_HIT_DTYPE)
        n = ctypes.c_int32(0)
        _check(_lib.prt_query_ray_hits(self.h, _ptr(r), max_hits, _ptr(hits), ctypes.byref(n)))
        return hits[:n.value]

    # --- queries (Renderer::queryRay / queryOcclusionRay) ---------------------------------
    def queryRays(self, rays):
        """rays: n x 8 [o d near far] -> (prim, t, u, v)."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        prim = np.zeros(n, np.int32)
        t, u, v = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(_lib.prt_query_rays(self.h, n, _ptr(r), _ptr(prim), _ptr(t), _ptr(u), _ptr(v)))
        return prim, t, u, v

    def queryRaysDevice(self, n, rays_ptr, prim_ptr, t_ptr, u_ptr, v_ptr, noalloc=False):
        """Closest hits of n rays already in device memory (raw device addresses), asynchronous
        on the context stream; noalloc selects prt_query_rays_device_noalloc."""
        fn = _lib.prt_query_rays_device_noalloc if noalloc else _lib.prt_query_rays_device
        _check(fn(self.h, int(n), rays_ptr, prim_ptr, t_ptr, u_ptr, v_ptr))

    def queryOcclusion(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(r.shape[0], np.int32)
        _check(_lib.prt_query_occlusion(self.h, r.shape[0], _ptr(r), _ptr(occ)))
        return occ

    def sceneBounds(self):
        """Device::getSceneBounds: (lo, hi) of the vertices the triangles reference."""
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        _check(_lib.prt_scene_bounds(self.h, _ptr(lo), _ptr(hi)))
        return lo, hi

    def queryPixel(self, x, y, max_hits=32):
        """Renderer::queryPixel (render/renderer.h:82-121): the hits of the ray through the centre
        of pixel (x, y) with the lens at its centre, stepping past transparent ones; a
        PIXEL_HIT_DTYPE array (prim, material, material_type, p, Ng, N, U, V, uv, T, eps, dist,
        depth)."""
        W, H = self.size
        hits = np.zeros(max_hits, PIXEL_HIT_DTYPE)
        n = ctypes.c_int32(0)
        _check(_lib.prt_query_pixel(self.h, (x + 0.5) / W, (y + 0.5) / H, max_hits, _ptr(hits), ctypes.byref(n)))
        return hits[:n.value]

    # --- debug / parity hooks ---------------------------------------------------------------
    def enableDebug(self, enable=True):
        _check(_lib.prt_debug_enable(self.h, 1 if enable else 0))

    def debugOutputs(self):
        W, H = self.size
        L = np.zeros((3, W * H), np.float32)
        hsh = np.zeros(W * H, np.uint32)
        first = np.zeros(W * H, np.int32)
        _check(_lib.prt_debug_get(self.h, _ptr(L), _ptr(hsh), _ptr(first)))
        return L, hsh, first

    def sampler(self, seed_pid_idx, ndim):
        q = np.ascontiguousarray(seed_pid_idx, np.uint32)
        n = q.shape[0]
        o1, o2, o3 = (np.zeros((n, ndim), np.float32), np.zeros((n, ndim, 2), np.float32),
                      np.zeros((n, ndim, 3), np.float32))
        _check(_lib.prt_debug_sampler(self.h, n, _ptr(q), ndim, _ptr(o1), _ptr(o2), _ptr(o3)))
        return o1, o2, o3

    # --- differentiable rendering ----------------------------------------------------------
    def adjoint(self, target, spp, n_materials):
        """Path-replay adjoint: (loss, grad [n_materials x 4] = dLoss/d(r, g, b, alpha), device ms)."""
        t = np.ascontiguousarray(target, np.float32)
        grad = np.zeros((n_materials, 4), np.float32)
        loss, ms = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(_lib.prt_adjoint(self.h, _ptr(t), int(spp), ctypes.byref(loss), _ptr(grad), grad.size, ctypes.byref(ms)))
        return loss.value, grad, ms.value

    def adjointReplay(self, sample_index):
        W, H = self.size
        out = np.zeros((3, W * H), np.float32)
        _check(_lib.prt_debug_adjoint_replay(self.h, int(sample_index), _ptr(out)))
        return out

    def luts(self):
        n = ctypes.c_int32(0)
        _check(_lib.prt_debug_luts(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(_lib.prt_debug_luts(self.h, _ptr(out), out.size, ctypes.byref(n)))
        return out

    def countSteps(self, enable=True):
        """Renders from now on use the traversal kernels that count their loop iterations."""
        _check(_lib.prt_debug_count_steps(self.h, 1 if enable else 0))

    def traceSteps(self):
        """Traversal loop iterations of the last render: lane sums and wave sums (closest, shadow)."""
        s = np.zeros(4, np.float64)
        _check(_lib.prt_debug_trace_steps(self.h, _ptr(s)))
        return {"closest_lane_steps": s[0], "closest_wave_steps": s[1], "shadow_lane_steps": s[2],
                "shadow_wave_steps": s[3]}

    def lightCdf(self):
        n = ctypes.c_int32(0)
        _check(_lib.prt_debug_light_cdf(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(_lib.prt_debug_light_cdf(self.h, _ptr(out), out.size, ctypes.byref(n)))
        return out

    def bvhStats(self):
        s = np.zeros(4, np.int64)
        _check(_lib.prt_debug_bvh_stats(self.h, _ptr(s)))
        return {"inner_nodes": int(s[0]), "triangles": int(s[1]), "max_depth": int(s[2])}

    def copyCalibration(self, nbytes):
        """Runs one device-to-device copy of nbytes (a copy kernel of known traffic); returns ms."""
        ms = ctypes.c_double(0)
        _check(_lib.prt_debug_copy(self.h, int(nbytes), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(_lib.prt_device_synchronize(self.h))

    def kernelTimes(self):
        """Per-kernel-class device time of the last render call (HIP events) and the traversal
        kernels' counters (rays, inner-node visits, triangle tests)."""
        ms = np.zeros(4, np.float64)
        cn = np.zeros(6, np.float64)
        ln = np.zeros(2, np.float64)
        _check(_lib.prt_debug_kernel_ms(self.h, _ptr(ms), _ptr(cn), _ptr(ln)))
        return {"generate_ms": ms[0], "intersect_ms": ms[1], "shade_ms": ms[2], "occluded_ms": ms[3],
                "rays": cn[0], "node_visits": cn[1], "tri_tests": cn[2],
                "shadow_rays": cn[3], "shadow_node_visits": cn[4], "shadow_tri_tests": cn[5],
                "intersect_launches": ln[0], "occluded_launches": ln[1]}


class SceneFileInfo(ctypes.Structure):  # prt_scene_file_info
    _fields_ = [("vertex_count", ctypes.c_int64), ("triangle_count", ctypes.c_int64), ("material_count", ctypes.c_int32),
                ("has_normals", ctypes.c_int32), ("has_sun_sky", ctypes.c_int32), ("sun_direction", ctypes.c_float * 3),
                ("approximated_materials", ctypes.c_int32), ("instance_count", ctypes.c_int32)]


def load_scene_file(path, instanced=False):
    """Scene::Scene(path) for .rkscene / .gltf / .mesh files (render/scene.cpp:64-372): the
    flattened world-space mesh, per-triangle material IDs and material records as a scene dict
    (positions, normals or None, indices, tri_material, materials, material_names, info).
    instanced: keep instanced objects in two levels (meshes in object space, "mesh_begin" and
    "instances" for DeviceHip.initInstances).  Host-side parsing only; lights, camera and
    renderer settings are the caller's."""
    L = load_library()
    h = ctypes.c_void_p()
    info = SceneFileInfo()
    loader = L.prt_scene_file_load_instanced if instanced else L.prt_scene_file_load
    _check(loader(os.fsencode(str(path)), ctypes.byref(h), ctypes.byref(info)))
    try:
        nv, nt, nm = info.vertex_count, info.triangle_count, info.material_count
        pos = np.zeros((nv, 3), np.float32)
        nrm = np.zeros((nv, 3), np.float32) if info.has_normals else None
        idx = np.zeros((nt, 3), np.int32)
        tm = np.zeros(nt, np.int32)
        mats = np.zeros(nm, MATERIAL_DTYPE)
        _check(L.prt_scene_file_get(h, _ptr(pos), _ptr(nrm), _ptr(idx), _ptr(tm), _ptr(mats)))
        names = []
        buf = ctypes.create_string_buffer(1024)
        for i in range(nm):
            _check(L.prt_scene_file_material_name(h, i, buf, 1024))
            names.append(buf.value.decode(errors="replace"))
        two = {}
        if instanced:
            nmesh, ninst = ctypes.c_int32(0), ctypes.c_int32(0)
            _check(L.prt_scene_file_get_instances(h, ctypes.byref(nmesh), None, ctypes.byref(ninst), None))
            mb = np.zeros(nmesh.value + 1, np.int32)
            rec = np.zeros(ninst.value, INSTANCE_DTYPE)
            _check(L.prt_scene_file_get_instances(h, ctypes.byref(nmesh), _ptr(mb), ctypes.byref(ninst), _ptr(rec)))
            two = {"mesh_begin": mb, "instances": [(int(r["mesh"]), r["xfm"].reshape(3, 4)) for r in rec]}
    finally:
        L.prt_scene_file_destroy(h)
    return {"positions": pos, "normals": nrm, "indices": idx, "tri_material": tm, "materials": mats,
            "material_names": names, **two,
            "info": {"approximated_materials": info.approximated_materials, "instance_count": info.instance_count,
                     "has_sun_sky": bool(info.has_sun_sky), "sun_direction": tuple(info.sun_direction)}}


def load_scene(dev, scene):
    """Uploads a scene dict from scenes/scenegen.py and configures renderer, camera and frame."""
    dev.initScene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"], scene.get("normals"))
    if scene.get("uvs") is not None:
        dev.initUVs(scene["uvs"])
    if scene.get("textures"):
        dev.initTextures(scene["textures"])
    for mat, tex in scene.get("material_textures", ()):
        dev.setMaterialTexture(mat, tex)
    if scene.get("instances"):
        dev.initInstances(scene["mesh_begin"], scene["instances"])
    dev.initLights(scene["lights"], scene.get("env"))
    dev.initRenderer(maxDepth=scene["max_depth"], rrDepth=scene.get("rr_depth", 8))
    dev.initSampler(scene.get("seed", 0))
    dev.initCamera(scene["camera"])


class BatchHip:
    """Persistent batch renderer: frames of uneven size from several scenes in one launch."""

    def __init__(self, device=0):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.prt_batch_create(device, ctypes.byref(h)))
        self.h = h
        self.frames = []
        self.scenes = []  # keeps the scene devices alive while the batch exists

    def close(self):
        if getattr(self, "h", None):
            _lib.prt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def addFrame(self, scene_dev, camera, width, height, seed, spp, snapshots=(), aov_mask=0):
        cam = np.ascontiguousarray(camera, CAMERA_DTYPE)
        snap = np.ascontiguousarray(snapshots, np.int32)
        fid = ctypes.c_int32(0)
        _check(_lib.prt_batch_add_frame(self.h, scene_dev.h, _ptr(cam), width, height, seed, spp,
                                        _ptr(snap) if snap.size else None, int(snap.size), aov_mask, ctypes.byref(fid)))
        self.frames.append((width, height, tuple(int(x) for x in snapshots), aov_mask))
        self.scenes.append(scene_dev)
        return fid.value

    def tune(self, calibration_spp=1):
        f, b, ms = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_double(0.0)
        _check(_lib.prt_batch_tune(self.h, calibration_spp, ctypes.byref(f), ctypes.byref(b), ctypes.byref(ms)))
        return {"tiles_per_fetch": f.value, "blocks_per_cu": b.value, "ms": ms.value}

    def render(self):
        ms, n = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(_lib.prt_batch_render(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return {"ms": ms.value, "samples": n.value}

    def get(self, frame, buffer=BATCH_COLOR):
        W, H = self.frames[frame][:2]
        out = np.zeros((H, W) if buffer == BATCH_DEPTH else (H, W, 3), np.float32)
        _check(_lib.prt_batch_get(self.h, frame, buffer, _ptr(out)))
        return out


class DenoiserParams(ctypes.Structure):  # prt_denoiser_params = Denoiser::Params (image/denoiser.h:14-29)
    _fields_ = [("color", ctypes.c_void_p), ("albedo", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("output", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("hdr", ctypes.c_int32), ("hdrScale", ctypes.c_float)]


class DenoiserHip:
    """The learned denoiser behind the Denoiser interface (init / execute / getVersion)."""

    def __init__(self, device=0):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.prt_denoiser_create(device, ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64(0)
        _check(_lib.prt_denoiser_param_count(self.h, ctypes.byref(n)))
        self.n_params = n.value
        self._keep = None
        self.initialWeights = self.getWeights()  # the initializer's, before any training or loading

    def close(self):
        if getattr(self, "h", None):
            _lib.prt_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def getVersion(self):
        buf = ctypes.create_string_buffer(128)
        _check(_lib.prt_denoiser_get_version(self.h, buf, 128))
        return buf.value.decode()

    def init(self, color, albedo, normal, output, hdr=True, hdrScale=float("nan")):
        """Denoiser::init; albedo and normal may be None (colour-only or colour + albedo input)."""
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (color, albedo, normal)]
        if not (output.dtype == np.float32 and output.flags["C_CONTIGUOUS"]):
            raise ValueError("output must be a C-contiguous float32 array")
        H, W = arrs[0].shape[:2]
        self._keep = arrs + [output]
        ptr = lambda a: None if a is None else a.ctypes.data
        p = DenoiserParams(ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), output.ctypes.data, W, H, int(hdr), hdrScale)
        _check(_lib.prt_denoiser_init(self.h, ctypes.byref(p)))

    def execute(self):
        t = ctypes.c_double(0.0)
        _check(_lib.prt_denoiser_execute(self.h, ctypes.byref(t)))
        return t.value

    def kernelMs(self):
        t = ctypes.c_double(0.0)
        _check(_lib.prt_denoiser_kernel_ms(self.h, ctypes.byref(t)))
        return t.value

    def executeFrame(self, dev, albedo=True, normal=True, hdrScale=float("nan")):
        """Denoises dev's frame on the device (colour + first-hit albedo / normal AOVs); the
        result stays on the device until output().  Returns the call's wall time in seconds."""
        t = ctypes.c_double(0.0)
        inputs = (DENOISE_ALBEDO if albedo else 0) | (DENOISE_NORMAL if normal else 0)
        _check(_lib.prt_denoiser_execute_frame(self.h, dev.h, inputs, hdrScale, ctypes.byref(t)))
        self._outShape = dev.size[::-1]
        return t.value

    def executeBatch(self, batch, frame, buffer=0, albedo=True, normal=True, hdrScale=float("nan")):
        """Denoises a rendered batch frame (its colour or a snapshot) on the device."""
        t = ctypes.c_double(0.0)
        inputs = (DENOISE_ALBEDO if albedo else 0) | (DENOISE_NORMAL if normal else 0)
        _check(_lib.prt_denoiser_execute_batch(self.h, batch.h, frame, buffer, inputs, hdrScale, ctypes.byref(t)))
        self._outShape = batch.frames[frame][1::-1]
        return t.value

    def output(self):
        """The last executeFrame / executeBatch result as an H x W x 3 array."""
        H, W = self._outShape
        out = np.zeros((H, W, 3), np.float32)
        _check(_lib.prt_denoiser_get_output(self.h, _ptr(out), W, H))
        return out

    def denoise(self, color, albedo, normal):
        out = np.zeros(np.shape(color), np.float32)
        self.init(color, albedo, normal, out)
        self.execute()
        return out

    def getWeights(self):
        w = np.zeros(self.n_params, np.float32)
        _check(_lib.prt_denoiser_get_weights(self.h, _ptr(w), w.size))
        return w

    def setWeights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(_lib.prt_denoiser_set_weights(self.h, _ptr(w), w.size))

    def trainStep(self, inp, target, lr):
        """inp: n x h x w x 9 (colour, albedo, normal); target: n x h x w x 3.  Returns (loss, device ms)."""
        inp = np.ascontiguousarray(inp, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w = inp.shape[:3]
        loss, ms = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(_lib.prt_denoiser_train_step(self.h, _ptr(inp), _ptr(target), n, h, w, lr, ctypes.byref(loss), ctypes.byref(ms)))
        return loss.value, ms.value

    def gradient(self, inp, target):
        inp = np.ascontiguousarray(inp, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w = inp.shape[:3]
        loss = ctypes.c_double(0.0)
        g = np.zeros(self.n_params, np.float32)
        _check(_lib.prt_denoiser_gradient(self.h, _ptr(inp), _ptr(target), n, h, w, ctypes.byref(loss), _ptr(g), g.size))
        return loss.value, g
