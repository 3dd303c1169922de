// DeviceHip: the stream path renderer on an MI355X behind ProtoRay's Device interface
// (render/device.h:56-114); see device_hip.h.  Every method names the DeviceCpu code it stands in
// for (render/device_cpu.cpp) and the C ABI entry point it calls.
#include "device_hip.h"

#include <cstring>
#include <stdexcept>

#include "image/linear_tone_mapper.h"
#include "image/reinhard_tone_mapper.h"
#include "image/aces_tone_mapper.h"
#include "image/filmic_tone_mapper.h"
#include "image/log_tone_mapper.h"
#include "image/pixel.h"

namespace prt {

static_assert(sizeof(Vec3f) == 3 * sizeof(float), "Vec3f images are read as packed RGB floats");

namespace {

// No C++ exception crosses the C ABI; its status codes become the exceptions the reference throws
// for the same conditions (invalid arguments, call order, runtime failures).
void check(int rc)
{
  if (rc == PRT_OK) return;
  if (rc == PRT_ERR_INVALID) throw std::invalid_argument(prt_last_error());
  if (rc == PRT_ERR_STATE) throw std::logic_error(prt_last_error());
  throw std::runtime_error(prt_last_error());
}

Vec3f vec3(const float* v) { return Vec3f(v[0], v[1], v[2]); }

Mat4f mat4(const float* m)  // row-major values, as the ABI returns them
{
  return Mat4f(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], m[12], m[13], m[14], m[15]);
}

// AOV props of DeviceCpu::initFrame (render/device_cpu.cpp:178-232) this device does not produce
const char* const kUnsupportedAovs[] = {
  "aovHdrUnclamp", "aovHdrUnclampVar", "aovAlbedoVar", "aovDiffuseAlbedo", "aovDiffuseAlbedo1", "aovSpecularAlbedo",
  "aovSpecularAlbedo1", "aovNormalVar", "aovDepthVar", "aovHWDepth", "aovRoughness", "aovRoughness1", "aovAlpha",
  "aovSh", "aovMotionBack", "aovMotionFore"};

} // namespace

DeviceHip::DeviceHip(int deviceOrdinal)
{
  check(prt_context_create(deviceOrdinal, &ctx));  // DeviceCpu ctor: albedo-table precompute
}

DeviceHip::~DeviceHip()
{
  prt_context_destroy(ctx);
}

std::string DeviceHip::getInfo()
{
  char buf[256];
  check(prt_get_info(ctx, buf, sizeof(buf)));
  return buf;
}

// ---- scene ----------------------------------------------------------------------------

void DeviceHip::initScene(const std::string& path, const Props& props)
{
  // Scene::Scene(path, props) (render/scene.cpp:64-372) through the library's loaders; the file
  // becomes one world-space mesh with the reference's material numbering
  prt_scene_file* file = nullptr;
  prt_scene_file_info info{};
  check(prt_scene_file_load(path.c_str(), &file, &info));
  struct Closer { prt_scene_file* f; ~Closer() { prt_scene_file_destroy(f); } } closer{file};
  std::vector<float> pos(size_t(info.vertex_count) * 3), nrm(info.has_normals ? size_t(info.vertex_count) * 3 : 0);
  std::vector<int32_t> idx(size_t(info.triangle_count) * 3), triMat(size_t(info.triangle_count));
  std::vector<prt_material> mats(size_t(info.material_count));
  check(prt_scene_file_get(file, pos.data(), info.has_normals ? nrm.data() : nullptr, idx.data(), triMat.data(),
                           mats.data()));
  if (props.exists("no-emissive") || props.exists("no-emission"))  // MaterialFactory: emissiveEnabled
    for (prt_material& m : mats)
      m.emission[0] = m.emission[1] = m.emission[2] = 0.f;
  materialNames.assign(mats.size(), std::string());
  for (int32_t i = 0; i < info.material_count; ++i)
  {
    char name[256];
    check(prt_scene_file_material_name(file, i, name, sizeof(name)));
    materialNames[i] = name;
  }
  check(prt_scene_set_geometry(ctx, int32_t(info.vertex_count), pos.data(), info.has_normals ? nrm.data() : nullptr,
                               int32_t(info.triangle_count), idx.data(), triMat.data()));
  check(prt_scene_set_materials(ctx, int32_t(mats.size()), mats.data()));
  double buildMs = 0.0;
  check(prt_bvh_build(ctx, &buildMs));  // the Embree build of the intersector
}

Box3f DeviceHip::getSceneBounds()
{
  float lo[3], hi[3];
  check(prt_scene_bounds(ctx, lo, hi));
  return Box3f(vec3(lo), vec3(hi));
}

void DeviceHip::initLights(const Props& props)
{
  // Scene::initLights (render/scene.cpp:365-450): distant, ambient, then image and sun/sky lights
  const Box3f bounds = getSceneBounds();
  const float sceneRadius = length(bounds.high - bounds.low) * 0.5f;
  std::vector<prt_light> lights;
  const Vec3f distant = props.get("distantLight", Vec3f(zero));
  if (reduceMax(distant) > 0.f)
  {
    // DistantLight's default direction; the library normalizes it as DistantLight's ctor does
    const Vec3f dir = props.get("distantLightDir", Vec3f(0.3f, 1.0f, 0.1f));
    const float halfAngle = props.get("distantLightAngle", 9.35e-3f) * 0.5f;  // angular diameter -> half angle
    lights.push_back(prt_light{PRT_LIGHT_DISTANT, distant.x, distant.y, distant.z, dir.x, dir.y, dir.z, halfAngle});
  }
  const Vec3f ambient = props.get("ambientLight", Vec3f(zero));
  if (reduceMax(ambient) > 0.f)
    lights.push_back(prt_light{PRT_LIGHT_AMBIENT, ambient.x, ambient.y, ambient.z, 0.f, 0.f, 0.f, 0.f});
  HipEnvironment env;
  const bool wantsEnvironment = !props.get("imageLight", "").empty() || props.get("sunSkyLightDir", Vec3f(zero)) != Vec3f(zero);
  if (wantsEnvironment)
  {
    if (!environment || !environment(props, sceneRadius, env))
      throw std::invalid_argument("image and sun/sky lights need an EnvironmentProvider (setEnvironmentProvider)");
    lights.insert(lights.end(), env.lights.begin(), env.lights.end());
  }
  const int lutSize = env.lutSize > 0 ? env.lutSize : props.get("imageLightLutSize", 0);
  // with no light at all the library adds the default ambient light, as the reference does
  check(prt_scene_set_lights(ctx, int32_t(lights.size()), lights.empty() ? nullptr : lights.data(), env.width, env.height,
                             env.rgb.empty() ? nullptr : env.rgb.data(), lutSize));
}

// ---- renderer -------------------------------------------------------------------------

void DeviceHip::initRenderer(const Props& props, Props& stats)
{
  // DeviceCpu::initRenderer -> RendererFactoryStream::make("pathStream", ...)
  // (render/renderer_factory_stream.cpp:17-76, render/renderer_factory.cpp:9-17)
  const std::string type = props.get("type");
  if (type != "pathStream") throw std::invalid_argument("DeviceHip renders the pathStream renderer only: " + type);
  std::string sampler = props.get("sampler", "default");
  if (sampler == "default") sampler = "sobolBurley";
  if (sampler != "sobolBurley") throw std::invalid_argument("DeviceHip supports the sobolBurley sampler only: " + sampler);
  samplerType = sampler;
  prt_renderer_params p{props.get("maxDepth", 100), props.get("rrDepth", 8), props.get("minWeight", 1e-4f),
                        props.get("maxRadiance", float(posMax))};
  check(prt_renderer_init(ctx, &p));
  (void)stats;
}

void DeviceHip::initSampler(unsigned int seed)
{
  check(prt_sampler_init(ctx, seed));
}

void DeviceHip::render(Props& stats)
{
  // FrameBuffer::begin, PathRendererStream::render, FrameBuffer::finish (render/device_cpu.cpp:53-59)
  prt_stats st{};
  check(prt_render(ctx, 1, &st));
  stats.set("mray", st.mray);  // path_renderer_stream.h:175-178
  stats.set("ray", int64_t(st.rays));
  stats.set("spp", 1);
}

Props DeviceHip::hitProps(const prt_pixel_hit& h, bool withDepth) const
{
  // RendererStream::queryRay (render/renderer_stream.h:16-62) and Renderer::queryPixel's depth
  Props p;
  p.set("mat", size_t(h.material) < materialNames.size() ? materialNames[h.material] : std::string());
  p.set("matId", h.material);
  p.set("matType", h.material_type);
  p.set("p", vec3(h.p));
  p.set("Ng", vec3(h.Ng));
  p.set("N", vec3(h.N));
  p.set("U", vec3(h.U));
  p.set("V", vec3(h.V));
  p.set("primId", h.prim);
  p.set("uv", Vec2f(h.uv[0], h.uv[1]));
  p.set("T", vec3(h.T));
  p.set("eps", h.eps);
  p.set("dist", h.dist);
  if (withDepth) p.set("depth", h.depth);
  return p;
}

Props DeviceHip::queryPixel(const Vec2f& point)
{
  prt_pixel_hit hit;
  int32_t n = 0;
  check(prt_query_pixel(ctx, point.x, point.y, 1, &hit, &n));
  return n > 0 ? hitProps(hit, true) : Props();
}

void DeviceHip::queryPixel(const Vec2f& point, Array<Props>& hits)
{
  prt_pixel_hit h[32];
  int32_t n = 0;
  check(prt_query_pixel(ctx, point.x, point.y, 32, h, &n));  // renderer.h:83-122: up to 32 hits
  hits.clear();
  for (int32_t i = 0; i < n; ++i) hits.pushBack(hitProps(h[i], true));
}

Props DeviceHip::queryPixel(int x, int y)
{
  return queryPixel((Vec2f(float(x), float(y)) + 0.5f) / toFloat(frameSize));  // renderer.h:77-81
}

void DeviceHip::queryPixel(int x, int y, Array<Props>& hits)
{
  queryPixel((Vec2f(float(x), float(y)) + 0.5f) / toFloat(frameSize), hits);
}

Props DeviceHip::queryRay(const Ray& ray)
{
  const float q[8] = {ray.org.x, ray.org.y, ray.org.z, ray.dir.x, ray.dir.y, ray.dir.z, ray.near, ray.far};
  prt_pixel_hit hit;
  int32_t n = 0;
  check(prt_query_ray_hits(ctx, q, 1, &hit, &n));
  return n > 0 ? hitProps(hit, false) : Props();
}

bool DeviceHip::queryOcclusionRay(const Ray& ray)
{
  const float q[8] = {ray.org.x, ray.org.y, ray.org.z, ray.dir.x, ray.dir.y, ray.dir.z, ray.near, ray.far};
  int32_t occluded = 0;
  check(prt_query_occlusion(ctx, 1, q, &occluded));
  return occluded != 0;
}

// ---- camera ---------------------------------------------------------------------------

void DeviceHip::initCamera(const Props& props)
{
  // DeviceCpu::makeCamera (render/device_cpu.cpp:110-128): a pinhole camera unless the type is
  // "thinlens" with a non-zero lens radius
  const std::string type = props.get("type");
  const float lensRadius = props.get("lensRadius", 0.f);
  if (type != "pinhole" && type != "thinlens" && lensRadius != 0.f) throw std::invalid_argument("invalid camera type");
  const Vec3f position = props.get<Vec3f>("position");
  const Basis3f basis = props.get<Basis3f>("basis");
  prt_camera c{};
  const Vec3f* v[4] = {&position, &basis.U, &basis.V, &basis.N};
  float* dst[4] = {c.pos, c.U, c.V, c.N};
  for (int k = 0; k < 4; ++k) { dst[k][0] = v[k]->x; dst[k][1] = v[k]->y; dst[k][2] = v[k]->z; }
  c.fov = props.get<float>("fov");
  c.nearClip = props.get<float>("nearClip");
  c.lensRadius = (type == "pinhole") ? 0.f : lensRadius;
  c.focalDistance = props.get("focalDistance", 1.f);
  check(prt_set_camera(ctx, &c));
  camera = c;
  cameraRaster = Vec2i(props.get<int>("width"), props.get<int>("height"));
  hasCamera = true;
}

void DeviceHip::initCamera(const Props& props, const Props& prevProps, const Props& nextProps)
{
  // the previous and next cameras only feed the motion-vector AOVs, which this device rejects
  (void)prevProps;
  (void)nextProps;
  initCamera(props);
}

void DeviceHip::matrices(float* worldToView, float* viewToClip, float* worldToRaster)
{
  if (!hasCamera) throw std::logic_error("camera not initialized");
  check(prt_camera_matrices(&camera, cameraRaster.x, cameraRaster.y, worldToView, viewToClip, worldToRaster));
}

Mat4f DeviceHip::getWorldToViewD3D()
{
  float m[16];
  matrices(m, nullptr, nullptr);
  return mat4(m);
}

Mat4f DeviceHip::getViewToClipD3D()
{
  float m[16];
  matrices(nullptr, m, nullptr);
  return mat4(m);
}

Mat4f DeviceHip::getWorldToRaster()
{
  float m[16];
  matrices(nullptr, nullptr, m);
  return mat4(m);
}

// ---- jitter ---------------------------------------------------------------------------

bool DeviceHip::hasJitter()
{
  int32_t has = 0;
  check(prt_get_jitter(ctx, &has, nullptr));
  return has != 0;
}

Vec2f DeviceHip::getJitter()
{
  int32_t has = 0;
  float j[2] = {0.f, 0.f};
  check(prt_get_jitter(ctx, &has, j));
  if (!has) throw std::logic_error("no jitter set");
  return Vec2f(j[0], j[1]);
}

void DeviceHip::setJitter(const Vec2f& jitter)
{
  check(prt_set_jitter(ctx, 1, jitter.x, jitter.y));
}

void DeviceHip::disableJitter()
{
  check(prt_set_jitter(ctx, 0, 0.f, 0.f));
}

// ---- frame ----------------------------------------------------------------------------

void DeviceHip::initFrame(const Vec2i& size, const Props& props)
{
  // DeviceCpu::initFrame (render/device_cpu.cpp:178-232); "aovHdrHalf" enables the colour's half
  // buffer (AccumFlag::HalfBuffer, which DeviceCpu leaves off)
  for (const char* name : kUnsupportedAovs)
    if (props.exists(name)) throw std::invalid_argument(std::string("AOV not supported by DeviceHip: ") + name);
  uint32_t aov = 0;
  if (props.exists("aovHdrVar")) aov |= PRT_AOV_VAR;
  if (props.exists("aovHdrHalf")) aov |= PRT_AOV_HALF;
  if (props.exists("aovAlbedo")) aov |= PRT_AOV_ALBEDO;
  if (props.exists("aovAlbedo1")) aov |= PRT_AOV_ALBEDO1;
  if (props.exists("aovNormal")) aov |= PRT_AOV_NORMAL;
  if (props.exists("aovNormal1")) aov |= PRT_AOV_NORMAL1;
  if (props.exists("aovDepth")) aov |= PRT_AOV_DEPTH;
  check(prt_frame_init(ctx, size.x, size.y, aov));
  frameSize = size;
  halfBuffer = (aov & PRT_AOV_HALF) != 0;
  toneMapper.reset();  // a new FrameBuffer has no tone mapper, jitter or filter
  filterType = "box";
}

void DeviceHip::initFilter(const Props& props)
{
  // DeviceCpu::initFilter (render/device_cpu.cpp:235-282)
  if (frameSize.x == 0) return;
  const std::string type = props.get("filter");
  int32_t kind = PRT_FILTER_BOX;
  float width = 0.f, sigma = 0.f;
  int32_t bias = 1;
  if (type == "box")
    kind = PRT_FILTER_BOX;
  else if (type == "tent" || type == "triangle")
  {
    kind = PRT_FILTER_TRIANGLE;
    width = props.get("filterWidth", 2.f);
  }
  else if (type == "gauss" || type == "gaussian")
  {
    kind = PRT_FILTER_GAUSSIAN;
    width = props.get("filterWidth", 3.f);
    sigma = props.get("filterSigma", width / 6.f);
    bias = props.get("filterBias", 1);
  }
  else if (type == "blackmanHarris" || type == "blackmanharris")
  {
    kind = PRT_FILTER_BLACKMAN_HARRIS;
    width = props.get("filterWidth", 3.f);
  }
  else
    throw std::invalid_argument("invalid pixel filter type");
  const std::string mode = props.get("filterMode", "sample");
  if (mode == "splat") throw std::invalid_argument("the splat pixel filter mode is not supported");  // pixel_filter.h:65-70
  if (mode != "sample") throw std::invalid_argument("invalid pixel filter mode");
  check(prt_set_pixel_filter(ctx, kind, width, sigma, bias));
  filterType = type;
}

void DeviceHip::initToneMapper(const Props& props)
{
  // DeviceCpu::initToneMapper (render/device_cpu.cpp:284-305), with the reference's own tone mappers
  if (frameSize.x == 0) return;
  const std::string type = props.get("type");
  if (type == "linear") toneMapper = std::make_shared<LinearToneMapper>(props);
  else if (type == "reinhard") toneMapper = std::make_shared<ReinhardToneMapper>(props);
  else if (type == "aces") toneMapper = std::make_shared<AcesToneMapper>(props);
  else if (type == "filmic") toneMapper = std::make_shared<FilmicToneMapper>(props);
  else if (type == "log") toneMapper = std::make_shared<LogToneMapper>(props);
  else if (type == "none") toneMapper.reset();
  else throw std::invalid_argument("invalid tone mapper type");
}

void DeviceHip::clearFrame()
{
  check(prt_frame_clear(ctx));
}

void DeviceHip::blitFrame(Surface& dest)
{
  // AccumBuffer3f::blitLdr (image/accum_buffer.h:650-676): mean, tone mapper, BGR8
  std::vector<Vec3f> hdr(size_t(frameSize.x) * frameSize.y);
  check(prt_blit(ctx, PRT_BUF_COLOR, &hdr[0].x, nullptr));
  for (int y = 0; y < frameSize.y; ++y)
  {
    int* row = dest.getRow(y);
    for (int x = 0; x < frameSize.x; ++x)
    {
      Vec3f c = hdr[size_t(y) * frameSize.x + x];
      if (toneMapper) c = toneMapper->get(c);
      row[x] = encodeBgr8(c);
    }
  }
}

void DeviceHip::blitFrame(FrameAovs& d)
{
  // DeviceCpu::blitFrame(FrameAovs&) (render/device_cpu.cpp:321-376)
  const void* unsupported[] = {d.hdrUnclamp, d.hdrUnclampVar, d.hdrUnclampHalf, d.albedoVar, d.albedoHalf, d.diffuseAlbedo,
                               d.diffuseAlbedo1, d.specularAlbedo, d.specularAlbedo1, d.normalVar, d.normalHalf, d.depthVar,
                               d.depthHalf, d.hwDepth, d.hwDepthVar, d.hwDepthHalf, d.roughness, d.roughness1, d.alpha,
                               d.sh[0], d.motionBack, d.motionFore};
  for (const void* p : unsupported)
    if (p) throw std::invalid_argument("FrameAovs names a buffer DeviceHip does not produce");
  if (d.hdr) check(prt_blit(ctx, PRT_BUF_COLOR, &d.hdr->x, d.hdrVar ? &d.hdrVar->x : nullptr));
  if (d.hdrHalf)
  {
    if (!halfBuffer) throw std::logic_error("half buffer not enabled for the buffer");  // accum_buffer.h:611-612
    check(prt_blit_half(ctx, &d.hdrHalf->x));
  }
  if (d.ldr)  // AccumBuffer3f::blitLdr(Vec3f*): tone-mapped means (accum_buffer.h:678-706)
  {
    check(prt_blit(ctx, PRT_BUF_COLOR, &d.ldr->x, nullptr));
    if (toneMapper)
      for (size_t i = 0; i < size_t(frameSize.x) * frameSize.y; ++i) d.ldr[i] = toneMapper->get(d.ldr[i]);
  }
  if (d.albedo) check(prt_blit(ctx, PRT_BUF_ALBEDO, &d.albedo->x, nullptr));
  if (d.albedo1) check(prt_blit(ctx, PRT_BUF_ALBEDO1, &d.albedo1->x, nullptr));
  if (d.normal) check(prt_blit(ctx, PRT_BUF_NORMAL, &d.normal->x, nullptr));
  if (d.normal1) check(prt_blit(ctx, PRT_BUF_NORMAL1, &d.normal1->x, nullptr));
  if (d.depth) check(prt_blit(ctx, PRT_BUF_DEPTH, d.depth, nullptr));
}

void DeviceHip::blitFrameLdr(Surface& dest, const float* src, float offset, float scale)
{
  // render/device_cpu.cpp:378-392
  for (int y = 0; y < dest.height; ++y)
  {
    int* row = dest.getRow(y);
    for (int x = 0; x < dest.width; ++x) row[x] = encodeBgr8(Vec3f((src[size_t(y) * dest.width + x] + offset) * scale));
  }
}

void DeviceHip::blitFrameLdr(Surface& dest, const Vec3f* src, float offset, float scale)
{
  // render/device_cpu.cpp:394-408
  for (int y = 0; y < dest.height; ++y)
  {
    int* row = dest.getRow(y);
    for (int x = 0; x < dest.width; ++x) row[x] = encodeBgr8((src[size_t(y) * dest.width + x] + offset) * scale);
  }
}

void DeviceHip::blitFrameHdr(Surface& dest, const Vec3f* src)
{
  // render/device_cpu.cpp:410-430
  for (int y = 0; y < dest.height; ++y)
  {
    int* row = dest.getRow(y);
    for (int x = 0; x < dest.width; ++x)
    {
      Vec3f c = src[size_t(y) * dest.width + x];
      if (toneMapper) c = toneMapper->get(c);
      row[x] = encodeBgr8(c);
    }
  }
}

float DeviceHip::getFrameError()
{
  // AccumBuffer3f::getError(toneMapper) against the checkpoint (accum_buffer.h:752-805)
  float e = 0.f;
  check(prt_frame_error_exposure(ctx, toneMapper ? toneMapper->getExposure() : 1.f, &e));
  return e;
}

float DeviceHip::getFilterWidth() const
{
  float w = 0.f;
  check(prt_get_filter_width(ctx, &w));
  return w;
}

} // namespace prt
