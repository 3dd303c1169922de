/*
 * prt_hip.h -- C ABI of the MI355X (gfx950) wavefront path tracer that stands in for
 * ProtoRay's stream renderer (RendererFactoryStream -> PathRendererStream behind the
 * Device / Renderer interfaces).
 *
 * Conventions: every call returns PRT_OK (0) or a negative PRT_ERR_* code; on error
 * prt_last_error() returns a message (thread-local).  No C++ exception crosses this ABI (the
 * reference throws std::invalid_argument / std::logic_error / std::runtime_error, e.g.
 * render/path_renderer_stream.h:109-112,153-154; those conditions map to PRT_ERR_INVALID).
 * All pointers are host pointers unless the name says "device".  A context owns one HIP
 * device and one stream; calls on one context must not overlap.
 */
#ifndef PRT_HIP_H
#define PRT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_OK 0
#define PRT_ERR_INVALID -1   /* bad argument / state (reference: std::invalid_argument, std::logic_error) */
#define PRT_ERR_HIP -2       /* HIP runtime failure */
#define PRT_ERR_NOMEM -3     /* device allocation failed */
#define PRT_ERR_STATE -4     /* call order violated (e.g. render before a scene) */

typedef struct prt_context prt_context;

/* Material record (148 bytes; a zero field is the reference's default except where noted).  Uniform StandardMaterial parameter sets of the stream path
 * (material/standard.cpp:21-183 ctor, :409-597 SIMD lobes), each a MultiBsdf with one lobe:
 *   kind 0 diffuse    : baseColor = (r,g,b), specularWeight 0          -> DiffuseBrdf
 *   kind 1 conductor  : metalness 1, baseColor = (r,g,b), specularColor = (er,eg,eb),
 *                       specularRoughness = roughness                  -> RoughConductorBrdf
 *                                                                         <FresnelConductorArtistic>
 *   kind 2 dielectric : transmissionWeight 1, specularIOR = ior,
 *                       specularRoughness = roughness                  -> RoughDielectricBsdf;
 *                       its interior medium absorbs by transmission_color ^ transmission_depth
 *                       per unit length (see the fields below)
 *   kind 3 coated     : a rough dielectric layer over a diffuse base   -> RoughCoatingBsdf
 *                       (rough_coating.h:270-593) over DiffuseBrdf.  The StandardMaterial
 *                       default (metalness 0, specularWeight w > 0, specularRoughness > 0) is
 *                       baseColor = (r,g,b), coat color (er,eg,eb) = 1, thickness = 1,
 *                       weight = w, ior = specularIOR (standard.cpp:512-518); a coat layer
 *                       over a diffuse base (specularWeight 0, coatWeight w) takes coatColor,
 *                       coatThickness, coatRoughness and coatIOR (standard.cpp:580-597).
 *   kind 4 thin dielectric : thinWalled 1, transmissionWeight 1, specularIOR = ior,
 *                       specularRoughness = roughness, transmissionColor = (r,g,b),
 *                       thickness / transmissionDepth = thickness -> ThinRoughDielectricBsdf
 *                       (thin_rough_dielectric.h:170-345): glossy reflection plus an unbent
 *                       transmission (a continuation lobe: the ray goes straight through).
 *   kind 5 standard : any uniform StandardMaterial parameter set (material/standard.cpp:409-611
 *                       for ambientIOR 1): the opaque base (diffuse or Oren-Nayar, under a
 *                       specular coating of weight specular_weight when specularIOR differs from
 *                       1), the translucent base (rough dielectric, thin rough dielectric when
 *                       thin_walled, the transparent BTDF at IOR 1) and the metal base (rough
 *                       conductor, artistic Fresnel of baseColor * base_weight and specularColor)
 *                       in one MultiBsdf weighted (1 - metalness)(1 - transmission),
 *                       (1 - metalness) transmission specular and metalness specular; a rough
 *                       coat (coat_*) over that mix; then sheen and opacity as for every kind.
 *                       Reads the kind-5 fields below; a transmissive, non-thin-walled one is a
 *                       media interface like kind 2.
 * thickness is read by kinds 3, 4 and 5 (thin-walled transmission), weight by kind 3 only.
 * Zero roughness is modelled by the rough lobes at the minimum GGX alpha (0.001), where the
 * reference switches to its smooth lobes.
 * transparency = 1 - opacity (standard.cpp:152-157, :412-420, :613-627), any kind: a value
 * above 1e-5 mixes the lobe (scale opacity) with a transparent BTDF (scale transparency,
 * transparent.h) in a MultiBsdf.  A material with transparency > 0, and every kind 4, is
 * transparent: shadow rays pass through it, attenuated by its transparency
 * (render/transparent_filter.h:42-93).
 * diffuse_roughness (kinds 0 and 3): above 0 the diffuse base is the Oren-Nayar BRDF of
 * rough_diffuse.h (Fujii's form) instead of the Lambertian one (standard.cpp:446-453).
 * sheen_weight, sheen_color, sheen_roughness: the sheen layer over any kind (standard.cpp:
 * 599-611, bsdf/sheen.h), active above a weight of 1e-5; the reference's default sheen
 * roughness is 0.3 (a zero here means roughness 0, alpha 0.001).
 * emission = emissionColor * emissionLuminance (standard.cpp:71-74, :663-666), any kind: a
 * non-zero emission makes the material emissive, its triangles form the scene's area light
 * (light/triangle_light.h, render/scene.cpp:83-116) and a path that hits one ends there.
 * Material ID 0 is reserved for the environment (render/scene.cpp:204-207). */
#define PRT_KIND_DIFFUSE 0
#define PRT_KIND_CONDUCTOR 1
#define PRT_KIND_DIELECTRIC 2
#define PRT_KIND_COATED 3
#define PRT_KIND_THIN_DIELECTRIC 4
#define PRT_KIND_STANDARD 5
typedef struct prt_material {
  int32_t kind;
  float r, g, b;
  float er, eg, eb;
  float roughness;
  float ior;
  float thickness;
  float weight;
  float emission[3];
  float transparency;
  float diffuse_roughness;
  float sheen_weight;
  float sheen_color[3];
  float sheen_roughness;
  /* kinds 2 and 5: the medium inside a transmissive, non-thin-walled surface, transmission_color ^
   * transmission_depth (material/standard.cpp:169-173); a path segment that ends at a hit inside
   * it attenuates its throughput by that colour ^ hit distance (path_renderer_stream.h:652-657).
   * The reference's default is depth 1 with colour 1 (clear); depth 0 (a zeroed record) is clear
   * as well, since colour ^ 0 = 1. */
  float transmission_color[3];
  float transmission_depth;
  /* kind 5 only: the StandardMaterial parameters read by no single-lobe kind (zero is NOT the
   * default here: set every field).  The fields above keep their meaning: (r,g,b) baseColor,
   * (er,eg,eb) specularColor, roughness specularRoughness, ior specularIOR, thickness (thin-walled
   * transmission), emission, transparency, diffuse_roughness, sheen_*; weight is unused. */
  float base_weight;          /* baseWeight, default 1 */
  float metalness;            /* default 0 */
  float specular_weight;      /* specularWeight, default 1 */
  float transmission_weight;  /* default 0 */
  float coat_weight;          /* default 0 */
  float coat_color[3];        /* default 1 */
  float coat_roughness;       /* default 0 (> 0 required where the coat is present) */
  float coat_ior;             /* coatIOR, default 1.6 */
  float coat_thickness;       /* default 0.5 (StandardMaterial's fixed coat thickness) */
  int32_t thin_walled;        /* default 0 */
} prt_material;

/* Environment light record (32 bytes): light/ambient_light.h (kind 0, Ke = rgb),
 * light/distant_light.h (kind 1, Ke = rgb, direction, half angle in radians),
 * light/image_light.h (kind 2, the lat-long image given to prt_scene_set_lights, scaled by rgb).
 * A sun/sky light (light/sun_sky_light.cpp) is passed as its baked pair: a distant sun plus
 * an image sky, in that order. */
#define PRT_LIGHT_AMBIENT 0
#define PRT_LIGHT_DISTANT 1
#define PRT_LIGHT_IMAGE 2
typedef struct prt_light {
  int32_t kind;
  float r, g, b;
  float dx, dy, dz;
  float angle;
} prt_light;

/* Camera (camera/pinhole_camera.h:72-117, camera/thin_lens_camera.h:155-161): position,
 * basis U, V, N (the view looks along -N), vertical fov in radians, near clip, lens radius
 * (0 selects the pinhole camera, render/device_cpu.cpp:117-128) and focal distance. */
typedef struct prt_camera {
  float pos[3];
  float U[3], V[3], N[3];
  float fov, nearClip, lensRadius, focalDistance;
} prt_camera;

/* Renderer properties (render/path_renderer_stream.h:114-116, render/integrator.h:58-62). */
typedef struct prt_renderer_params {
  int32_t maxDepth;   /* default 100 */
  int32_t rrDepth;    /* default 8 */
  float minWeight;    /* default 1e-4 */
  float maxRadiance;  /* default FLT_MAX (no clamping) */
} prt_renderer_params;

/* Per-call statistics (render/path_renderer_stream.h:167-178): rays traced (primary,
 * extension and shadow), render time, Mray/s, samples per pixel rendered. */
typedef struct prt_stats {
  double rays;
  double ms;
  double mray;
  int32_t spp;
} prt_stats;

/* AOV selection bits for prt_frame_init (render/device_cpu.cpp:178-232 aov* props). */
#define PRT_AOV_ALBEDO   (1u << 0)  /* aovAlbedo  : albedo at the first non-delta hit */
#define PRT_AOV_ALBEDO1  (1u << 1)  /* aovAlbedo1 : first-hit albedo */
#define PRT_AOV_NORMAL   (1u << 2)  /* aovNormal  */
#define PRT_AOV_NORMAL1  (1u << 3)  /* aovNormal1 */
#define PRT_AOV_DEPTH    (1u << 4)  /* aovDepth   : linear depth of the first hit */
#define PRT_AOV_VAR      (1u << 5)  /* aovHdrVar  : color sample variance (sum of squares) */
#define PRT_AOV_HALF     (1u << 6)  /* AccumFlag::HalfBuffer of the color buffer: the class-B samples
                                       (image/accum_buffer.h:36-47, 455-470), read by prt_blit_half */
#define PRT_AOV_ALL      0x7fu

/* Buffer selectors for prt_blit. */
#define PRT_BUF_COLOR   0
#define PRT_BUF_ALBEDO  1
#define PRT_BUF_ALBEDO1 2
#define PRT_BUF_NORMAL  3
#define PRT_BUF_NORMAL1 4
#define PRT_BUF_DEPTH   5

/* --- context ------------------------------------------------------------------------- */
/* Replaces DeviceCpu construction (render/device_cpu.cpp:20-23, incl. the albedo LUT
 * precompute, bsdf/albedo_tables.cpp:173-186). */
int prt_context_create(int device_ordinal, prt_context** out);
int prt_context_destroy(prt_context* ctx);
const char* prt_last_error(void);
/* Copies `bytes` bytes device-to-device with hipMemcpyAsync (a plain copy kernel of known
 * size, used to calibrate the HBM FETCH_SIZE / WRITE_SIZE counters); ms receives its time. */
int prt_debug_copy(prt_context* ctx, int64_t bytes, double* ms);
/* Waits for all work on the context's device (hipDeviceSynchronize); brackets timed regions. */
int prt_device_synchronize(prt_context* ctx);
/* Device name, CU count and HBM size (Device::getInfo, render/device.h:60). */
int prt_get_info(prt_context* ctx, char* buf, int len);

/* --- scene (Device::initScene / Scene ctor, render/scene.cpp:65-134) ------------------ */
/* positions: nverts x 3; normals: nverts x 3 or NULL; indices: ntris x 3;
 * tri_material: ntris material IDs (>= 1).  Uploads to HBM. */
int prt_scene_set_geometry(prt_context* ctx, int32_t nverts, const float* positions, const float* normals,
                           int32_t ntris, const int32_t* indices, const int32_t* tri_material);
int prt_scene_set_materials(prt_context* ctx, int32_t n, const prt_material* materials);
/* Image textures on material parameters (TriangleMesh texcoords, geometry/triangle_mesh.cpp:
 * 83-90, 131; ImageTexture, image/image_texture.h:148-300; MaterialParam3f,
 * material/material_param.h:14-60).  set_uvs: texture coordinate set 0, nverts x 2 floats for the
 * vertices of the current geometry (interpolated barycentrically at a hit; NULL or no call: the
 * hit's barycentric (u, v) are the coordinates, as for a mesh without texcoords; new geometry
 * drops them).  set_textures: the scene's textures, each width x height texels given as image
 * rows top to bottom (the reference stores them bottom to top, image_texture.h:352-353, so
 * (u, v) = (0, 0) is the lower-left corner of the image), decoded when uploaded: 8-bit RGBA
 * through the reference's 2.2-gamma table (sRGB, image/pixel.cpp:21) or /255 (linear), or RGB
 * FP32; the lookup is the reference's bilinear filter with half-texel offset and repeat or clamp
 * addressing per axis, at uv * uv_scale + uv_offset (ImageTextureParams).  material_set_texture:
 * parameter PRT_PARAM_BASE_COLOR of a diffuse or coated material (kinds 0, 3) becomes
 * clamp(texture(uv) x (r, g, b)), the value a textured baseColor parameter gives
 * (material/standard.cpp:424); texture -1 unbinds.  prt_scene_set_materials drops the bindings.
 * The adjoint rejects scenes with textured materials. */
#define PRT_TEX_RGBA8_SRGB 0
#define PRT_TEX_RGBA8_LINEAR 1
#define PRT_TEX_RGB32F 2
#define PRT_TEX_REPEAT 0
#define PRT_TEX_CLAMP 1
#define PRT_PARAM_BASE_COLOR 0
typedef struct prt_texture {
  int32_t width, height;
  int32_t format;                /* PRT_TEX_* */
  int32_t address_u, address_v;  /* PRT_TEX_REPEAT or PRT_TEX_CLAMP */
  float uv_scale[2], uv_offset[2];
  const void* texels;            /* width x height: 4 bytes (R, G, B, A) or 3 floats each */
} prt_texture;
int prt_scene_set_uvs(prt_context* ctx, const float* uv, int32_t nverts);
int prt_scene_set_textures(prt_context* ctx, int32_t n, const prt_texture* textures);
int prt_material_set_texture(prt_context* ctx, int32_t material, int32_t parameter, int32_t texture);
/* Device::initLights (render/scene.cpp:365-450): n (0..8) environment lights in reference
 * order, optional lat-long environment image env_rgb (env_h x env_w x 3, texture row order:
 * row 0 = nadir, image/image_texture.h:396), lut_max as the image light's lutSize prop.  Call
 * after the geometry and materials: the area light (emissive triangles) heads the light list
 * and is built here; with no lights at all a default ambient light (Ke = 1) is added. */
int prt_scene_set_lights(prt_context* ctx, int32_t n, const prt_light* lights, int32_t env_w, int32_t env_h,
                         const float* env_rgb, int32_t lut_max);
/* Two-level scenes (Instance, geometry/instance.h:12-60; Embree instances,
 * accel/embree_intersector.cpp:100-129).  The geometry's triangles fall into n_meshes meshes,
 * mesh m being triangles [mesh_begin[m], mesh_begin[m + 1]) (mesh_begin[0] = 0, mesh_begin[n_meshes]
 * = the triangle count); each of the n instances places one mesh in the world by the affine map
 * x_world = L x + t given as the rows of [L | t].  Positions, normals and UVs of the geometry are
 * in the meshes' object space; only instanced meshes are in the world.  prt_bvh_build then builds
 * one object-space tree per mesh and a top-level tree over the instances; rays enter an instance
 * through the inverse map (t keeps its meaning), and the hit frame leaves it by the linear part
 * (tangents) and its inverse transpose (normals), renormalised.  Hits report the scene triangle
 * (mesh_begin[m] + local index); prt_query_rays_inst also reports the instance (ties in t go to
 * the smaller instance, then the smaller triangle).  n = 0 returns to a one-level scene; new
 * geometry drops the instances.  Two-level scenes render with prt_render and the batch
 * renderer, differentiate with prt_adjoint and answer prt_query_rays / prt_query_rays_device /
 * prt_query_occlusion (transparent shadows included) and the pixel and hit queries; emissive
 * triangles must lie in a mesh placed exactly once by the identity (the reference flattens
 * emissive objects, render/scene.cpp:343-363; prt_scene_set_lights checks it), and the
 * allocation-free query takes one-level scenes. */
typedef struct prt_instance {
  int32_t mesh;
  float xfm[12];  /* rows of [L | t] */
} prt_instance;
int prt_scene_set_instances(prt_context* ctx, int32_t n_meshes, const int32_t* mesh_begin, int32_t n,
                            const prt_instance* instances);
/* BVH build on the device (replaces the Embree build, accel/embree_intersector.cpp:11-90).
 * build_ms receives the build time (Embree's buildMs stat). */
int prt_bvh_build(prt_context* ctx, double* build_ms);
/* Device::getSceneBounds (render/device.h:65): the box of the vertices the triangles reference
 * (empty: lo > hi). */
int prt_scene_bounds(prt_context* ctx, float lo[3], float hi[3]);

/* --- renderer (Device::initRenderer -> RendererFactoryStream::make "pathStream",
 *     render/renderer_factory_stream.cpp:17-76) --------------------------------------- */
int prt_renderer_init(prt_context* ctx, const prt_renderer_params* params);
/* Renderer::initSampler (render/path_renderer_stream.h:144-148): seed, sampleIndex = 0. */
int prt_sampler_init(prt_context* ctx, uint32_t seed);
/* Device::initCamera (render/device_cpu.cpp:130-134). */
int prt_set_camera(prt_context* ctx, const prt_camera* camera);
/* The camera's matrices for an image of width x height (camera/pinhole_camera.h:41-68, the same
 * for the thin lens; Device::getWorldToViewD3D / getViewToClipD3D / getWorldToRaster,
 * render/device.h:85-87), each 16 floats row-major: worldToViewD3D (row-vector convention, the
 * translation in the last row), viewToClipD3D (reverse infinite perspective), worldToRaster
 * (column-vector convention: raster = M [p 1], divided by its w).  Host only: no device is
 * touched.  Any output may be null. */
int prt_camera_matrices(const prt_camera* camera, int32_t width, int32_t height, float* world_to_view_d3d,
                        float* view_to_clip_d3d, float* world_to_raster);
/* Device::initFrame / clearFrame (render/device_cpu.cpp:178-232, FrameBuffer::clear).
 * Width must be a multiple of 32 and height of 8, each <= 4096
 * (render/path_renderer_stream.h:109-112). */
int prt_frame_init(prt_context* ctx, int32_t width, int32_t height, uint32_t aov_mask);
int prt_frame_clear(prt_context* ctx);
/* Renderer::render (render/path_renderer_stream.h:150-187) called n_spp times: each pass
 * renders one sample per pixel with the current sampleIndex, accumulates it in sample
 * order, and increments sampleIndex.  Synchronous. */
int prt_render(prt_context* ctx, int32_t n_spp, prt_stats* stats);
/* Device::blitFrame (render/device_cpu.cpp:316-376, AccumBuffer3f::blit accum_buffer.h:607-649):
 * mean = sum * rcpSafe(weight), interleaved RGB (depth: one float).  var may be NULL. */
int prt_blit(prt_context* ctx, int32_t buffer, float* dst, float* var);
/* The half buffer of AccumBuffer3f::blit (destHalf, accum_buffer.h:633-643): the mean of the
 * class-B samples (sample counts with an odd number of bits set in count & 0xaaaaaaaa),
 * interleaved RGB.  Needs PRT_AOV_HALF in prt_frame_init. */
int prt_blit_half(prt_context* ctx, float* dst);
/* Device::setJitter / disableJitter / hasJitter / getJitter (render/device.h:90-94,
 * image/frame_buffer.h:75-78): with enable != 0 every pixel's sample position is (jx, jy) in
 * [0, 1]^2 instead of the sampler's dimension-set 0 value (path_renderer_stream.h:228-231);
 * enable = 0 disables it.  The jitter belongs to the frame: prt_frame_init resets it. */
int prt_set_jitter(prt_context* ctx, int32_t enable, float jx, float jy);
int prt_get_jitter(prt_context* ctx, int32_t* has, float* jxy);
/* Device::initFilter (render/device_cpu.cpp:235-282) in PixelFilterMode::Sample: the pixel
 * sample is warped through the filter's tabulated distribution (image/pixel_filter.h:30-66).
 * width <= 0 takes the reference default (triangle 2, Gaussian and Blackman-Harris 3); sigma <= 0
 * the Gaussian default width / 6; bias: subtract the Gaussian's value at the radius.  Without
 * a frame the call does nothing (as DeviceCpu::initFilter); prt_frame_init resets the filter to
 * the box.  prt_get_filter_width: Device::getFilterWidth (0 for the box). */
#define PRT_FILTER_BOX 0
#define PRT_FILTER_TRIANGLE 1
#define PRT_FILTER_GAUSSIAN 2
#define PRT_FILTER_BLACKMAN_HARRIS 3
int prt_set_pixel_filter(prt_context* ctx, int32_t type, float width, float sigma, int32_t bias);
int prt_get_filter_width(prt_context* ctx, float* width);
/* Raw FP32 accumulator sums (planar, 3 x W*H) and weights (W*H) of the color buffer. */
int prt_get_accum(prt_context* ctx, float* sum, float* weight);
/* Device::getFrameError (AccumBuffer3f::getError, accum_buffer.h:752-805) against the
 * checkpoint copy taken at the last power-of-two sample count. */
int prt_frame_error(prt_context* ctx, float* error);
/* getError(toneMapper) with a tone mapper: both images scaled by its exposure first
 * (accum_buffer.h:787-792). */
int prt_frame_error_exposure(prt_context* ctx, float exposure, float* error);

/* --- queries (Renderer::queryRay / queryOcclusionRay, render/renderer.h:47-128) ------- */
/* rays: n x [ox oy oz dx dy dz near far]; closest hit per ray: prim (-1 = miss), t, u, v. */
int prt_query_rays(prt_context* ctx, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v);
/* The same with the hit's instance (-1 for a miss or a one-level scene): the instID of an
 * Embree hit in a two-level scene. */
int prt_query_rays_inst(prt_context* ctx, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v,
                        int32_t* inst);
/* any hit in [near, far]: occluded[i] = 1 / 0.  In a scene with transparent materials the ray
 * passes them as a shadow ray does (renderer_stream.h:64-97 with the transparent-shadow
 * filter): occluded once the transmittance falls below 1e-5. */
int prt_query_occlusion(prt_context* ctx, int32_t n, const float* rays, int32_t* occluded);
/* One hit of a pixel query (Renderer::queryPixel, render/renderer.h:50-121, with the hit
 * properties of RendererStream::queryRay, render/renderer_stream.h:16-62). */
typedef struct prt_pixel_hit {
  int32_t prim, material;
  int32_t material_type;  /* MaterialType bits (material/material.h:17-25): 1 transparent,
                             4 transmissive, 8 media interface, 16 emissive */
  float p[3], Ng[3], N[3], U[3], V[3];  /* hit point, oriented geometric normal, shading frame */
  float uv[2];            /* the hit's barycentric (u, v): the uv the reference reports without
                             texture coordinates (geometry/triangle_mesh.cpp:131) */
  float T[3];             /* transparency along the ray (0 for an opaque material) */
  float eps, dist, depth; /* hit offset, distance along the ray, distance along the view axis */
} prt_pixel_hit;
/* The ray through image point (px, py) in [0, 1]^2 (x right, y down; (x + 0.5) / W,
 * (y + 0.5) / H for pixel (x, y)) of the current camera with the lens at its centre: up to
 * max_hits (<= 32) hits, stepping past each transparent one (T >= 1e-5) by 16 x its offset.
 * Needs the scene, BVH, camera and frame size. */
int prt_query_pixel(prt_context* ctx, float px, float py, int32_t max_hits, prt_pixel_hit* hits, int32_t* count);
/* Renderer::queryRay with the hit properties of RendererStream::queryRay (render/renderer_stream.h:
 * 16-62) for one ray [ox oy oz dx dy dz near far]: up to max_hits (<= 32) hits along it, stepping
 * past each transparent one as prt_query_pixel does (max_hits = 1: the closest hit only); depth
 * is the distance along the ray.  *count = 0 for a miss. */
int prt_query_ray_hits(prt_context* ctx, const float* ray, int32_t max_hits, prt_pixel_hit* hits, int32_t* count);
/* Same two queries on device buffers already in HBM, on the context stream (asynchronous). */
int prt_query_rays_device(prt_context* ctx, int32_t n, const float* rays_device, int32_t* prim_device,
                          float* t_device, float* u_device, float* v_device);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops): prt_query_rays_device with the same results, whose traversal
 * stack overflows into a fixed 16 MiB static device array (128 blocks x 256 lanes x 64 entries)
 * instead of the context's buffer; it calls no hipMalloc / hipFree.  Trees of more than 24 levels
 * are rejected (PRT_ERR_INVALID).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other contexts or streams) race; calls on one context are ordered by its
 * stream.  Asynchronous on the context stream. */
int prt_query_rays_device_noalloc(prt_context* ctx, int32_t n, const float* rays_device, int32_t* prim_device,
                                  float* t_device, float* u_device, float* v_device);

/* --- differentiable rendering ----------------------------------------------------------
 * Path-replay adjoint.  No counterpart in the reference (it renders forward only); added for
 * inverse rendering of material parameters.  Clears the frame and renders spp sample passes
 * (indices 0 .. spp-1 of the sampler seed); forms the loss mean over pixels and channels of
 * (I - target)^2 with target an H x W x 3 image in prt_blit layout; returns *loss and
 * dLoss / d(albedo r, g, b, GGX alpha) per material in grad[4 * material + k] (grad_capacity
 * floats, >= 4 x the material count).  Differentiated: the diffuse colour of kinds 0 and 3, the
 * reflectivity (baseColor) of kind 1, and alpha (= roughness^2) of kinds 1, 2 and 3 (the
 * microfacet lobe, through D, G2 and the albedo tables of the multiple-scattering term); kind 0
 * has no alpha, kind 2 no albedo and kind 4 is not differentiated, so those entries are 0.  A
 * sheen layer attenuates the derivatives of the lobe under it; opacity scales value and
 * derivatives alike.  The gradient is that of the
 * primal estimator with its samples fixed (detached sampling); paths are replayed from the
 * seed, not stored; a path deeper than 8 vertices is replayed once more per further 8 depths
 * (the records of its vertices are reduced 8 depths at a time).  *ms: device time of the replay
 * kernels. */
int prt_adjoint(prt_context* ctx, const float* target, int32_t spp, double* loss, float* grad, int32_t grad_capacity,
                double* ms);
/* Replays sample sample_index of every pixel as the adjoint does and returns its radiance
 * (planar [3][H*W]); equal to the primal sample's radiance (the replay follows the same path). */
int prt_debug_adjoint_replay(prt_context* ctx, uint32_t sample_index, float* radiance);

/* --- persistent batch rendering (dataset generation) -------------------------------------
 * No counterpart in the reference's renderer API (its batch mode, main/render/render_window.cpp:
 * 630-698, renders preset after preset through Device::render); added to render many frames of
 * uneven resolution from several scenes in one persistent launch, captured in a HIP graph.
 * Each frame names a scene context (prt_context with geometry, BVH, materials, lights and
 * renderer parameters set; it must outlive the batch), a camera, a size (multiples of 8, at most
 * 4096), a Sobol-Burley seed and a sample count; sample indices 0 .. spp-1 are rendered in
 * order, as prt_render would, and the mean is kept, plus the means after snapshot_spp[k]
 * samples (ascending, at most 4) and, when aov_mask has PRT_AOV_ALBEDO1 or PRT_AOV_NORMAL1,
 * the first-hit albedo and normal means, with PRT_AOV_DEPTH the first-hit depth mean (distance
 * along the view axis, +inf where a sample missed).  prt_batch_tune times every (tiles per fetch, blocks
 * per CU) candidate on the batch with calibration_spp samples per frame and keeps the fastest.
 * prt_batch_render: *ms = device time of the launch, *samples = path samples rendered.
 * prt_batch_get copies one buffer of a frame as H x W x 3 floats (prt_blit layout; the depth as
 * H x W).  A scene
 * context changed after its frames were prepared (geometry, materials, lights, BVH or renderer
 * parameters) makes the next prt_batch_render / prt_batch_tune rebuild every frame's kernel
 * arguments and the captured graph, so they never read the context's released buffers; the
 * outputs of earlier renders are released with them. */
#define PRT_BATCH_COLOR 0
#define PRT_BATCH_ALBEDO1 1
#define PRT_BATCH_NORMAL1 2
#define PRT_BATCH_DEPTH 3
#define PRT_BATCH_SNAPSHOT 16 /* + k: the mean after snapshot_spp[k] samples */
typedef struct prt_batch prt_batch;
int prt_batch_create(int32_t device_ordinal, prt_batch** out);
int prt_batch_destroy(prt_batch* batch);
int prt_batch_add_frame(prt_batch* batch, prt_context* scene, const prt_camera* cam, int32_t width, int32_t height,
                        uint32_t seed, int32_t spp, const int32_t* snapshot_spp, int32_t n_snapshots, uint32_t aov_mask,
                        int32_t* frame_id);
int prt_batch_tune(prt_batch* batch, int32_t calibration_spp, int32_t* tiles_per_fetch, int32_t* blocks_per_cu,
                   double* best_ms);
int prt_batch_render(prt_batch* batch, double* ms, double* samples);
int prt_batch_get(prt_batch* batch, int32_t frame, int32_t buffer, float* dst);

/* --- learned denoiser ---------------------------------------------------------------------
 * The Denoiser plug-in (image/denoiser.h:10-41): init with the colour / albedo / normal /
 * output images (H x W x 3 floats each; albedo and normal may be NULL, as Params allows, but a
 * normal needs an albedo: missing images enter the network as zero channels, so set the weights
 * trained for that input set), execute (returns the denoising time in seconds), version.  The filter is a U-Net (4 levels, 3x3 convolutions, 32 / 48 / 64 / 96 / 128 channels,
 * ReLU, max pooling, nearest upsampling with skips) on BF16 matrix cores with FP32 accumulation.
 * Weights: param_count FP32 values (per layer: weights [cout][cin][3][3] then bias [cout]; the
 * layers enc0..enc4, dec3..dec0, out).  Training (no counterpart in the reference, which calls
 * OIDN's pretrained filters): train_step runs forward, backward and one Adam step on a batch of
 * n images h x w (multiples of 16) given as input (n*h*w*9: colour, albedo, normal) and target
 * (n*h*w*3), with the L1 loss of log(1 + colour); gradient returns the loss and the FP32
 * parameter gradient of a batch without updating. */
typedef struct prt_denoiser prt_denoiser;
typedef struct prt_denoiser_params {   /* Denoiser::Params (image/denoiser.h:14-29) */
  const float* color;
  const float* albedo;
  const float* normal;
  float* output;
  int32_t width, height;
  int32_t hdr;
  float hdrScale;                      /* <= 0 or NaN: 1 */
} prt_denoiser_params;
int prt_denoiser_create(int32_t device_ordinal, prt_denoiser** out);
int prt_denoiser_destroy(prt_denoiser* dn);
int prt_denoiser_init(prt_denoiser* dn, const prt_denoiser_params* params);
int prt_denoiser_execute(prt_denoiser* dn, double* seconds);
int prt_denoiser_get_version(prt_denoiser* dn, char* buf, int32_t len);
int prt_denoiser_kernel_ms(prt_denoiser* dn, double* ms);
int prt_denoiser_param_count(prt_denoiser* dn, int64_t* n);
int prt_denoiser_get_weights(prt_denoiser* dn, float* weights, int64_t n);
int prt_denoiser_set_weights(prt_denoiser* dn, const float* weights, int64_t n);
int prt_denoiser_train_step(prt_denoiser* dn, const float* input, const float* target, int32_t n, int32_t h, int32_t w,
                            float lr, double* loss, double* ms);
int prt_denoiser_gradient(prt_denoiser* dn, const float* input, const float* target, int32_t n, int32_t h, int32_t w,
                          double* loss, float* grad, int64_t grad_count);
/* Denoising a frame where rendering left it, in HBM, with no host copy in between: the
 * reference's render window blits the accumulation buffer into host images and calls
 * Denoiser::execute on them (main/render/render_window.cpp:485-530 init, :1034-1064
 * updateFrame then execute; image/denoiser.h:10-41);
 * here the network reads the frame's planar sums and weights directly.  execute_frame takes the
 * context's frame colour (sum / weight, as prt_blit forms it) and, selected by inputs, its
 * first-hit albedo and normal (PRT_AOV_ALBEDO1 / PRT_AOV_NORMAL1; a normal needs an albedo);
 * execute_batch takes a rendered batch frame's colour (buffer PRT_BATCH_COLOR) or a snapshot
 * (PRT_BATCH_SNAPSHOT + k) with that frame's first-hit AOVs.  The call waits for the kernels
 * enqueued on the producer's stream, runs the network on the denoiser's stream, and returns
 * when the result is in the denoiser's device buffer (seconds: the call's wall time;
 * prt_denoiser_kernel_ms: its kernels).  get_output copies that W x H x 3 result to the host.
 * The output equals prt_denoiser_execute on the blitted images. */
#define PRT_DENOISE_ALBEDO 1
#define PRT_DENOISE_NORMAL 2
int prt_denoiser_execute_frame(prt_denoiser* dn, prt_context* ctx, int32_t inputs, float hdr_scale, double* seconds);
int prt_denoiser_execute_batch(prt_denoiser* dn, prt_batch* batch, int32_t frame, int32_t buffer, int32_t inputs,
                               float hdr_scale, double* seconds);
int prt_denoiser_get_output(prt_denoiser* dn, float* dst, int32_t width, int32_t height);

/* --- scene files ------------------------------------------------------------------------
 * The loaders of Scene::Scene(path, props) (render/scene.cpp:64-372): .rkscene (JSON + .rkbin,
 * rkscene/rkscene.h:847-903), .gltf (rkscene/rkscene_gltf.h; also node matrices and data-URI
 * buffers) and .mesh (TriangleMesh blob, geometry/triangle_mesh.cpp:362-380, sys/blob.h:13-37,
 * with the .mtl of the same base name, material/mtl_loader.cpp).  The file becomes one
 * world-space triangle mesh: instanced objects are flattened (each instance's copy
 * transformed; normals by the inverse transpose), as the reference does for emissive objects
 * or with its `flatten` prop (render/scene.cpp:343-363).  Material IDs follow the reference:
 * 0 = environment, emissive materials first, the default material last (.mesh files carry
 * their IDs).  Standard material parameters map to the records above; a material using
 * textures, sheen, subsurface, thin walls, opacity, partial metalness/transmission, smooth
 * (roughness 0) lobes or a coat over a specular layer is mapped to its nearest record and
 * counted in approximated_materials.  Host only: no device is touched.  prt_scene_file_get
 * copies the arrays (sizes from the info; normals only when has_normals). */
typedef struct prt_scene_file prt_scene_file;
typedef struct prt_scene_file_info {
  int64_t vertex_count, triangle_count;
  int32_t material_count;          /* including ID 0 */
  int32_t has_normals;
  int32_t has_sun_sky;             /* an rkscene sunSky light: its direction below */
  float sun_direction[3];
  int32_t approximated_materials;
  int32_t instance_count;          /* instances flattened */
} prt_scene_file_info;
int prt_scene_file_load(const char* path, prt_scene_file** out, prt_scene_file_info* info);
/* The same file kept in two levels (render/scene.cpp:325-363 without the `flatten` prop): each
 * instanced object becomes a mesh in object space placed by one prt_instance per transform;
 * objects without instances and emissive objects (which the reference always flattens) form one
 * world-space mesh placed by the identity.  get_instances returns the meshes' triangle ranges
 * (n_meshes + 1 entries) and the instances, for prt_scene_set_instances; null arrays return
 * the counts only.  info.instance_count counts the instance records. */
int prt_scene_file_load_instanced(const char* path, prt_scene_file** out, prt_scene_file_info* info);
int prt_scene_file_get_instances(prt_scene_file* f, int32_t* n_meshes, int32_t* mesh_begin, int32_t* n_instances,
                                 prt_instance* instances);
int prt_scene_file_get(prt_scene_file* f, float* positions, float* normals, int32_t* indices, int32_t* tri_material,
                       prt_material* materials);
int prt_scene_file_material_name(prt_scene_file* f, int32_t id, char* buf, int32_t len);
int prt_scene_file_destroy(prt_scene_file* f);

/* --- per-sample debug outputs of the last pass (parity checks) ------------------------- */
/* When enabled, each pass also stores per pixel: the clamped sample radiance (planar
 * 3 x W*H), a hash of the path's hit / lobe sequence and the first-hit primitive ID. */
int prt_debug_enable(prt_context* ctx, int32_t enable);
int prt_debug_get(prt_context* ctx, float* sample_l, uint32_t* path_hash, int32_t* first_prim);
/* Sobol-Burley sampler on the device: seed_pid_idx n x 3, outputs n x ndim (1D), n x ndim x 2,
 * n x ndim x 3 (sampling/sobol_burley_sampler_simd.h:36-85). */
int prt_debug_sampler(prt_context* ctx, int32_t n, const uint32_t* seed_pid_idx, int32_t ndim, float* out1,
                      float* out2, float* out3);
/* The energy-compensation LUTs computed at context creation (14368 floats: GGX dir 32x32,
 * GGX avg 32, dielectric eta dir 16^3, eta avg 16^2, rcp-eta dir 16^3, rcp-eta avg 16^2,
 * dielectric reflection dir 16^3, avg 16^2, sheen dir 16^2; bsdf/albedo_tables.h:14-300).  *count receives the
 * number of floats; out (capacity floats) may be NULL to query it.  PRT_ERR_INVALID if
 * capacity is too small. */
int prt_debug_luts(prt_context* ctx, float* out, int32_t capacity, int32_t* count);
/* The light-power distribution the renderer picks lights from (Scene::lightDistribution,
 * render/scene.cpp:440-445): *count = lights + 1 CDF entries (the area light first, if any, then
 * the environment lights in order); out (capacity floats) may be NULL to query the count. */
int prt_debug_light_cdf(prt_context* ctx, float* out, int32_t capacity, int32_t* count);
/* BVH statistics: inner node count, triangle count, max depth, scene triangle count. */
int prt_debug_bvh_stats(prt_context* ctx, int64_t* stats4);
/* Device-side kernel timing of the last prt_render call, per kernel class
 * (generate, intersect, shade, occluded) in ms, from HIP events on the context stream, and the
 * kernels' own counters (the analogue of RayStats::nodeCount / primCount, core/ray_stats.h:10-45):
 * counters6 = closest-hit rays, inner-node visits, triangle tests, then the same three for the
 * shadow (any-hit) rays; launches2 = number of intersect and occluded launches. */
int prt_debug_kernel_ms(prt_context* ctx, double* ms4, double* counters6, double* launches2);
/* Renders with enable != 0 use builds of the traversal kernels that also count their loop
 * iterations (prt_debug_trace_steps); the default builds keep their registers and occupancy. */
int prt_debug_count_steps(prt_context* ctx, int32_t enable);
/* SIMD utilisation of the traversal kernels in the last prt_render call (counted when enabled by
 * prt_debug_count_steps, zero otherwise): steps4 = traversal loop
 * iterations (inner-node and leaf visits) summed over lanes, then summed over waves (per ray
 * batch, the longest lane's count), for closest-hit then shadow rays.  Active-lane fraction =
 * lane steps / (64 x wave steps) (the analogue of RayStats' active-lane counts,
 * core/ray_stats.h:10-45). */
int prt_debug_trace_steps(prt_context* ctx, double* steps4);

#ifdef __cplusplus
}
#endif

#endif /* PRT_HIP_H */
