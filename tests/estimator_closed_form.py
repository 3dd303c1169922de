"""Float64 restatements, written from the reference's source text, of the parts of the stream path
integrator whose reference headers cannot be compiled here (they include the absent TBB headers):
the environment lights' direction mapping and powers, and the per-sample estimator of
render/path_renderer_stream.h on a two-quad diffuse scene.  Test infrastructure: the tests compare
the oracle and the HIP renderer against these, sample by sample.

The scene is chosen so that every quantity has a closed form: two axis-aligned diffuse quads (a
floor and a wall in front of the camera) under an ambient light plus a distant light, so a ray's
hit, its shading frame and every shadow test are plane intersections.  Decisions that sit within
a small margin of a boundary (a quad edge, the light-pick CDF, the distant light's cone, the
Russian-roulette threshold, a direction grazing a surface) are reported per sample, so a test can
leave those few samples out instead of loosening its tolerance.
"""
import math

import numpy as np

PI = math.pi

# --- lights ------------------------------------------------------------------------------------


def luminance(c):
    """image/color.h:17-22"""
    return 0.212671 * c[..., 0] + 0.715160 * c[..., 1] + 0.072169 * c[..., 2]


def scene_radius(positions, indices):
    """render/scene.cpp:365-367: half the diagonal of the box of the referenced vertices."""
    p = np.asarray(positions, np.float64)[np.asarray(indices).ravel()]
    return 0.5 * float(np.linalg.norm(p.max(0) - p.min(0)))


def light_powers(lights, radius, env=None):
    """Light::getPower of each environment light (light/ambient_light.h:53-57,
    light/distant_light.h:73-78 with angle = the half angle, light/image_light.h:136-148 with the
    mean over every texel of the scaled image).  `lights` rows: (kind, r, g, b, dx, dy, dz, angle)."""
    out = []
    for l in lights:
        Ke = np.array([l["r"], l["g"], l["b"]], np.float64)
        if l["kind"] == 0:
            out.append(4 * PI ** 2 * radius ** 2 * luminance(Ke))
        elif l["kind"] == 1:
            out.append(4 * PI ** 2 * math.sin(0.5 * float(l["angle"])) ** 2 * radius ** 2 * luminance(Ke))
        else:
            img = np.asarray(env, np.float64) * Ke
            out.append(4 * PI ** 2 * radius ** 2 * luminance(img).mean())
    return np.array(out)


def light_cdf(powers):
    """Distribution1D::init over the powers (sampling/distribution1d.cpp:20-35)."""
    c = np.concatenate([[0.0], np.cumsum(powers)])
    c = c / c[-1] if c[-1] > 0 else c * 0
    c[-1] = 1.0
    return c


def image_light_dir(u, v):
    """ImageLight::sample's (u, v) -> direction (light/image_light.h:112-126): v = 0 is the
    nadir (texture row 0), u = 0.5 looks along -z."""
    theta = (1.0 - v) * PI
    phi = (u - 0.5) * 2 * PI
    return np.stack([np.sin(phi) * np.sin(theta), np.cos(theta), -np.cos(phi) * np.sin(theta)], -1)


def image_light_uv(d):
    """ImageLight::eval's direction -> (u, v) (light/image_light.h:79-86)."""
    d = np.asarray(d, np.float64)
    theta = np.arccos(np.clip(d[..., 1], -1, 1))
    phi = np.arctan2(d[..., 0], -d[..., 2])
    return 0.5 + phi / (2 * PI), 1.0 - theta / PI


def texture_bilinear(tex, u, v):
    """ImageTexture::get with TextureAddress::Repeat in u and Clamp in v (image/image_texture.h:
    196-225), texel centres at (i + 0.5) / size; tex rows in texture order (row 0 = v = 0)."""
    tex = np.asarray(tex, np.float64)
    H, W = tex.shape[:2]
    x = u - 0.5 / W
    y = v - 0.5 / H
    xr = (x - np.floor(x)) * W
    yr = np.clip(y * H, 0.0, H - 1)
    sx, sy = xr - np.floor(xr), yr - np.floor(yr)
    x0 = np.clip(xr.astype(np.int64), 0, W - 1)
    y0 = np.clip(yr.astype(np.int64), 0, H - 1)
    x1 = np.where(x0 + 1 == W, 0, x0 + 1)
    y1 = np.minimum(y0 + 1, H - 1)
    c0 = tex[y0, x0] * (1 - sx)[..., None] + tex[y0, x1] * sx[..., None]
    c1 = tex[y1, x0] * (1 - sx)[..., None] + tex[y1, x1] * sx[..., None]
    return c0 * (1 - sy)[..., None] + c1 * sy[..., None]


def image_light_pdf_single_texel(W, H, y):
    """The image light's pdf in solid angle where the whole distribution sits in one texel of row
    y: the cell's probability 1 times W H over 2 pi^2 sin(theta) (light/image_light.h:85-86)."""
    theta = (1.0 - (y + 0.5) / H) * PI
    return W * H / (2 * PI * PI * math.sin(theta))


# --- the two-quad scene ------------------------------------------------------------------------

FLOOR = dict(axis=1, value=0.0, lo=(-20.0, -4.0), hi=(20.0, 40.0))   # y = 0, (x, z) extent
WALL = dict(axis=2, value=12.0, lo=(-20.0, 0.0), hi=(20.0, 16.0))    # z = 12, (x, y) extent
RHO = (np.array([0.7, 0.5, 0.3]), np.array([0.3, 0.6, 0.8]))
AMBIENT = np.array([0.4, 0.45, 0.5])
SUN = np.array([30.0, 28.0, 25.0])
SUN_DIR = np.array([0.3, 0.7, 0.2]) / np.linalg.norm([0.3, 0.7, 0.2])
SUN_HALF_ANGLE = 0.1
CAM_POS = np.array([0.0, 3.0, 0.0])
CAM_FOV = math.radians(60.0)


def two_quad_scene(scenegen):
    """The scene as renderer inputs (scenegen supplies the record dtypes): a floor (+y) and a wall
    facing the camera (-z), diffuse, under the ambient and the distant light; pinhole camera at
    (0, 3, 0) looking along +z."""
    G = scenegen
    v = np.array([[-20, 0, -4], [20, 0, -4], [20, 0, 40], [-20, 0, 40],
                  [-20, 0, 12], [20, 0, 12], [20, 16, 12], [-20, 16, 12]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]], np.int32)
    tri_mat = np.array([1, 1, 2, 2], np.int32)
    mats = np.zeros(3, G.MATERIAL_DTYPE)
    for k in (1, 2):
        mats[k]["kind"] = G.KIND_DIFFUSE
        mats[k]["r"], mats[k]["g"], mats[k]["b"] = RHO[k - 1]
    lights = np.zeros(2, G.LIGHT_DTYPE)
    lights[0] = (G.LIGHT_AMBIENT, *AMBIENT, 0, 0, 0, 0)
    lights[1] = (G.LIGHT_DISTANT, *SUN, *SUN_DIR, SUN_HALF_ANGLE)
    cam = np.zeros((), G.CAMERA_DTYPE)
    cam["pos"] = CAM_POS
    cam["U"], cam["V"], cam["N"] = (-1, 0, 0), (0, 1, 0), (0, 0, -1)
    cam["fov"] = CAM_FOV
    cam["nearClip"] = 0.01
    return dict(name="two_quads", positions=v, indices=idx, tri_material=tri_mat, materials=mats, lights=lights,
                env=None, camera=cam, max_depth=2, rr_depth=0, seed=0)


def _geometric_normal(quad, positions, indices, tri):
    p = np.asarray(positions, np.float64)[np.asarray(indices)[tri]]
    n = np.cross(p[1] - p[0], p[2] - p[0])
    return n / np.linalg.norm(n)


def _intersect(o, d, tnear, quad):
    """Plane hit of quad; returns (t, margin): margin = distance of the hit to the quad's edge
    (negative outside), inf when the ray is parallel or the plane is behind."""
    a = quad["axis"]
    if d[a] == 0.0:
        return math.inf, -math.inf
    t = (quad["value"] - o[a]) / d[a]
    if not t >= tnear:
        return math.inf, -math.inf
    p = o + t * d
    other = [k for k in range(3) if k != a]
    m = min(min(p[k] - quad["lo"][i], quad["hi"][i] - p[k]) for i, k in enumerate(other))
    return (t if m >= 0 else math.inf), m


def _make_frame(N):
    """math/basis3.h:173-180"""
    U0 = np.array([0.0, N[2], -N[1]])
    U1 = np.array([-N[2], 0.0, N[0]])
    U = U0 if U0 @ U0 > U1 @ U1 else U1
    U = U / np.linalg.norm(U)
    V = np.cross(N, U)
    return U, V / np.linalg.norm(V)


def _mis(a, b):
    """render/path_renderer_stream.h:1077-1080 (power heuristic)"""
    return a * a / (a * a + b * b) if (a * a + b * b) > 0 else 0.0


# Deliberate deviations two_quad_samples can apply, each a plausible slip in restating the
# estimator; a test checks that every one of them disagrees with the renderer under test.
VARIANTS = ("balance_heuristic", "nee_pdf_without_rr", "light_dims_of_bsdf", "distant_power_full_angle",
            "rr_before_aovs")


def two_quad_samples(scene, W, H, sample_index, sampler_fn, margin_eps=1e-4, graze_eps=0.05, variant=None):
    """Per-pixel sample radiance (3, W*H) of pass sample_index and a bool mask of samples whose
    decisions all clear margin_eps and whose rays leave their surface at a cosine of at least
    graze_eps (a float32 hit point lies about an ulp of its coordinates off the plane, so a ray
    leaving at a grazing angle can re-cross its own surface beyond the offset of 0x1.fp-18 times
    the point's magnitude).  sampler_fn(seed_pid_idx n x 3, ndim) -> (1D, 2D, 3D) values
    of the Sobol-Burley sampler (sampling/sobol_burley_sampler_simd.h:27-85)."""
    pos, idx = scene["positions"], scene["indices"]
    quads = [FLOOR, WALL]
    ng_raw = [_geometric_normal(FLOOR, pos, idx, 0), _geometric_normal(WALL, pos, idx, 2)]
    radius = scene_radius(pos, idx)
    lights = scene["lights"].copy()
    if variant == "distant_power_full_angle":
        lights["angle"] = np.where(lights["kind"] == 1, 2 * lights["angle"], lights["angle"])
    powers = light_powers(lights, radius)
    mis = (lambda a, b: a / (a + b) if a + b > 0 else 0.0) if variant == "balance_heuristic" else _mis
    cdf = light_cdf(powers)
    cos_sun = math.cos(SUN_HALF_ANGLE)
    sun_pdf = 1.0 / (4 * PI * math.sin(0.5 * SUN_HALF_ANGLE) ** 2)
    sunU, sunV = _make_frame(SUN_DIR)
    max_depth, rr_depth, min_weight = scene["max_depth"], scene["rr_depth"], 1e-4
    ys, xs = np.divmod(np.arange(W * H), W)
    q = np.stack([np.full(W * H, scene["seed"]), xs | (ys << 16), np.full(W * H, sample_index)], 1).astype(np.uint32)
    ndim = 4 + 7 * (max_depth + 1)
    s1, s2, s3 = (a.astype(np.float64) for a in sampler_fn(q, ndim))
    hh = math.tan(0.5 * CAM_FOV)
    hw = hh * W / H
    U, V, N = np.array([-1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, -1.0])
    imageO, imageDx, imageDy = -hw * U + hh * V - N, 2 * hw * U, -2 * hh * V
    out = np.zeros((3, W * H))
    ok = np.ones(W * H, bool)

    def env(d):
        Le = AMBIENT.copy()
        pdf = (1 / (4 * PI)) * (cdf[1] - cdf[0])
        c = d @ SUN_DIR
        if c >= cos_sun:
            Le = Le + SUN
            pdf += sun_pdf * (cdf[2] - cdf[1])
        return Le, pdf, abs(c - cos_sun)

    def trace(o, d, tnear):
        best, which, m = math.inf, -1, math.inf
        for k, qd in enumerate(quads):
            t, mk = _intersect(o, d, tnear, qd)
            if mk != -math.inf:
                m = min(m, abs(mk))
            if t < best:
                best, which = t, k
        return best, which, m

    for i in range(W * H):
        margin = math.inf
        x, y = xs[i], ys[i]
        psx, psy = s2[i, 0]
        d = imageO + ((x + psx) / W) * imageDx + ((y + psy) / H) * imageDy
        ln = np.linalg.norm(d)
        d = d / ln
        o, tnear = CAM_POS.copy(), 0.01 * ln
        T, L = np.ones(3), np.zeros(3)
        last_pdf, last_smooth = 1.0, False
        aov_done = False
        graze = False
        for depth in range(max_depth + 1):
            dim = 4 + 7 * depth
            t, k, m = trace(o, d, tnear)
            margin = min(margin, m)
            if k < 0:  # miss (:285-415)
                Le, lpdf, mc = env(d)
                margin = min(margin, mc)
                w = mis(last_pdf, lpdf) if last_smooth else 1.0
                L = L + T * w * Le
                break
            if depth >= max_depth:
                break
            p = o + t * d
            eps = max(t, np.abs(p).max()) * 1.9375 * 2.0 ** -18  # core/ray_simd.h:63-68 (0x1.fp-18)
            ngr = ng_raw[k]
            fU, fV = _make_frame(ngr)
            Nf = -ngr if ngr @ d > 0 else ngr
            # both AOV flags are set after the first (diffuse, non-delta) vertex
            m_rr = aov_done or variant == "rr_before_aovs"
            rho = RHO[k]
            # next-event estimation (:659-718)
            a0, a1, a2 = s3[i, dim + (0 if variant == "light_dims_of_bsdf" else 3)]
            c = a2
            li = int(np.searchsorted(cdf[1:-1], c, side="right"))
            margin = min(margin, np.min(np.abs(cdf[1:-1] - c)))
            pick = cdf[li + 1] - cdf[li]
            if li == 0:  # ambient: uniform sphere (shape_sampler.h:23-40)
                ct = 1 - 2 * a0
                st = 2 * math.sqrt(a0 * (1 - a0))
                wiL = np.array([math.cos(2 * PI * a1) * st, math.sin(2 * PI * a1) * st, ct])
                lpdf, lw = 1 / (4 * PI), AMBIENT * 4 * PI
            else:  # distant: uniform cone in makeFrame(dir) (distant_light.h:66-71, shape_sampler.h:100-117)
                ct = 1 - a0 * (1 - cos_sun)
                st = math.sqrt(max(1 - ct * ct, 0.0))
                wiL = math.cos(2 * PI * a1) * st * sunU + math.sin(2 * PI * a1) * st * sunV + ct * SUN_DIR
                lpdf, lw = sun_pdf, SUN / sun_pdf
            lpdf *= pick
            lw = lw / pick
            cosL = wiL @ Nf
            graze = graze or abs(cosL) < graze_eps
            if cosL > 0:
                f = rho / PI * cosL
                bpdf = cosL / PI
                dirT = T * f
                if depth >= rr_depth and variant != "nee_pdf_without_rr":
                    cont = min(luminance(dirT / bpdf), 0.95) if m_rr else 1.0
                    bpdf *= cont
                dirL = dirT * mis(lpdf, bpdf) * lw
                ts, ks, ms = trace(p, wiL, eps)
                margin = min(margin, ms)
                if ks < 0:
                    L = L + dirL
            # BSDF sample: cosine hemisphere in the frame (diffuse.h:61-67, shape_sampler.h:69-86)
            b0, b1, _ = s3[i, dim]
            ctb, stb = math.sqrt(b0), math.sqrt(1 - b0)
            wi = math.cos(2 * PI * b1) * stb * fU + math.sin(2 * PI * b1) * stb * fV + ctb * Nf
            pdf = ctb / PI
            gn = wi @ Nf
            graze = graze or abs(gn) < graze_eps
            ext = gn > 0
            aov_done = True
            if not ext:
                break
            T = T * rho
            ext = T.max() >= min_weight
            if depth >= rr_depth:  # Russian roulette (:880-887)
                cont = min(luminance(T), 0.95) if m_rr else 1.0
                u = s1[i, dim + 6]
                margin = min(margin, abs(u - cont))
                ext = ext and u <= cont
                T = T / cont
                pdf *= cont
            if not ext:
                break
            o, d, tnear = p, wi, eps
            last_pdf, last_smooth = pdf, True
        out[:, i] = np.maximum(L, 0.0)
        ok[i] = margin >= margin_eps and not graze
    return out, ok


def pinhole_dirs(cam, W, H, jx, jy):
    """Unit directions of the pinhole camera's rays through (x + jx, y + jy) of every pixel, row
    by row (camera/pinhole_camera.h:86-117; the view looks along -N, image rows go down V)."""
    U, V, N = (np.asarray(cam[k], np.float64) for k in ("U", "V", "N"))
    hh = math.tan(0.5 * float(cam["fov"]))
    hw = hh * W / H
    ys, xs = np.divmod(np.arange(W * H), W)
    d = (-hw * U + hh * V - N)[None] + ((xs + jx) / W)[:, None] * (2 * hw * U)[None] \
        + ((ys + jy) / H)[:, None] * (-2 * hh * V)[None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)
