// Host driver of the MI355X stream path tracer: the C ABI declared in include/prt_hip.h.
//
// It plays the roles of DeviceCpu (render/device_cpu.cpp), Scene (render/scene.cpp) and
// PathRendererStream::render (render/path_renderer_stream.h:150-187) for one HIP device:
// scene arrays are uploaded to HBM once, the BVH is built on the device, and each render pass
// launches the wavefront kernels of render.hip on the context's stream.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/prt_hip.h"
#include "internal.h"

using namespace prt;

namespace {

thread_local std::string gLastError;

int fail(int code, const std::string& msg)
{
  gLastError = msg;
  return code;
}

#define HIP_TRY(x)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return fail(e_ == hipErrorOutOfMemory ? PRT_ERR_NOMEM : PRT_ERR_HIP, \
                                      std::string(#x ": ") + hipGetErrorString(e_));          \
  } while (0)

template <class T>
struct DevBuf
{
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count)
  {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(&p, sizeof(T) * count);
  }
  void release()
  {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// ----------------------------------------------------------------------------------------
// Host math shared by the LUT precompute and the per-material constants (IEEE FP32)
// ----------------------------------------------------------------------------------------
struct H3 { float x, y, z; };
inline H3 hmk(float a, float b, float c) { return H3{a, b, c}; }
inline float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline H3 hcross(H3 a, H3 b) { return hmk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline H3 hscale(H3 a, float s) { return hmk(a.x * s, a.y * s, a.z * s); }
inline H3 hadd(H3 a, H3 b) { return hmk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline H3 hsub(H3 a, H3 b) { return hmk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline H3 hnorm(H3 a) { return hscale(a, 1.0f / std::sqrt(hdot(a, a))); }
inline float hsqr(float a) { return a * a; }
inline float hclamp(float x, float lo, float hi) { return std::min(std::max(x, lo), hi); }
inline float hlerp(float a, float b, float s) { return a + s * (b - a); }

// GGX with isotropic alpha (bsdf/ggx_distribution.h), host copy for the LUT quadrature
struct HostGGX
{
  float a;
  float D(H3 w) const
  {
    float c2 = hsqr(w.z);
    float e = (hsqr(w.x / a) + hsqr(w.y / a)) / c2;
    return 1.0f / (3.14159265358979323846f * a * a * hsqr(c2 * (1.f + e)));
  }
  float lambda(H3 w) const
  {
    float c2 = hsqr(w.z);
    return 0.5f * (-1.f + std::sqrt(1.f + (hsqr(w.x * a) + hsqr(w.y * a)) / c2));
  }
  float G1(H3 w) const { return 1.0f / (1.f + lambda(w)); }
  float G2(H3 wo, H3 wi) const { return 1.0f / (1.f + lambda(wo) + lambda(wi)); }
  H3 sampleVisible(H3 wo, float sx, float sy) const
  {
    const float pi = 3.14159265358979323846f;
    H3 V = hnorm(hmk(a * wo.x, a * wo.y, wo.z));
    H3 T1 = V.z < 0.9999f ? hnorm(hcross(V, hmk(0, 0, 1))) : hmk(1, 0, 0);
    H3 T2 = hcross(T1, V);
    float aa = 1.f / (1.f + V.z);
    float r = std::sqrt(sx);
    float phi = (sy < aa) ? sy / aa * pi : pi + (sy - aa) / (1.f - aa) * pi;
    float P1 = r * std::cos(phi), P2 = r * std::sin(phi) * ((sy < aa) ? 1.f : V.z);
    H3 wh = hadd(hadd(hscale(T1, P1), hscale(T2, P2)), hscale(V, std::sqrt(std::max(0.f, 1.f - P1 * P1 - P2 * P2))));
    return hnorm(hmk(a * wh.x, a * wh.y, std::max(0.f, wh.z)));
  }
};

inline float hostFresnelEx(float cosI, float& cosT, float eta)  // bsdf/optics.h:122-139
{
  cosT = 0.f;
  cosI = std::fabs(cosI);
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  if (!(k >= 0.f)) return 1.f;
  cosT = std::sqrt(k);
  float Rper = (eta * cosI - cosT) / (eta * cosI + cosT), Rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
  return 0.5f * (Rpar * Rpar + Rper * Rper);
}

// bsdf/albedo.h:16-120 single-scattering albedo estimators and albedo_tables.cpp:23-107 layout
// mode 0: GGX reflection (albedo.h:18-40); 1: dielectric reflection + refraction (:63-103);
// 2: dielectric reflection only (:124-151)
float lutIntegrate(float cosO, const HostGGX& m, float eta, int mode)
{
  const int n = 32;
  float sum = 0.f;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
    {
      float sx = std::min((float(i) + 0.5f) / float(n), 1.f - 1e-6f), sy = std::min((float(j) + 0.5f) / float(n), 1.f - 1e-6f);
      float c = std::max(cosO, 1e-6f);
      H3 wo = hmk(std::sqrt(std::max(1.f - c * c, 0.f)), 0.f, c);
      H3 wh = m.sampleVisible(wo, sx, sy);
      float cosOH = hdot(wo, wh);
      H3 wi = hsub(hscale(wh, 2.f * cosOH), wo);
      if (mode == 0)
      {
        if (wi.z > 0.f) sum += m.G2(wo, wi) * (1.0f / m.G1(wo));
        continue;
      }
      if (mode == 2)
      {
        float cosTH;
        float F = hostFresnelEx(std::max(cosOH, 0.f), cosTH, eta);  // optics.h:84-99 clamps cosI at 0
        if (wi.z > 0.f) sum += F * (m.G2(wo, wi) * (1.0f / m.G1(wo)));
        continue;
      }
      float cosTH;
      float F = hostFresnelEx(cosOH, cosTH, eta);
      float w = 0.f;
      if (wi.z > 0.f) w += F * (m.G2(wo, wi) * (1.0f / m.G1(wo)));
      H3 wt = hsub(hscale(hsub(hscale(wh, cosOH), wo), eta), hscale(wh, cosTH));
      if (wt.z < 0.f) w += (1.f - F) * (m.G2(wo, wt) * (1.0f / m.G1(wo)));
      sum += w;
    }
  return std::min(sum / float(n * n), 1.f);
}
float lutInterp1(float x, const float* f, int size)
{
  float xc = hclamp(x, 0.f, float(size - 1));
  float s = xc - std::floor(xc);
  int x0 = std::min(int(xc), size - 1), x1 = std::min(x0 + 1, size - 1);
  return hlerp(f[x0], f[x1], s);
}
float lutAvg(const float* f, int size)
{
  const int n = 1025;
  float sum = 0.f;
  for (int i = 0; i < n; ++i)
  {
    float c = float(i) / float(n - 1);
    sum += lutInterp1(c * float(size - 1), f, size) * c * ((i == 0 || i == n - 1) ? 0.5f : 1.f);
  }
  return std::min(2.f * (sum / float(n - 1)), 1.f);
}
// Sheen microfiber distribution (bsdf/sheen_distribution.h:10-78, Conty and Kulla 2017): host
// copy for the sheen albedo table
struct HostSheen
{
  float r;  // alpha, in (0, 1]
  float D(float c) const
  {
    float inv = 1.0f / r;
    return (2.f + inv) * std::pow(std::sqrt(std::max(1.f - c * c, 0.f)), inv) * (1.f / (2.f * 3.14159265358979323846f));
  }
  float L(float x) const
  {
    float w0 = hsqr(1.f - r), w1 = 1.f - w0;
    float a = w0 * 25.3245f + w1 * 21.5473f, b = w0 * 3.32435f + w1 * 3.82987f, c = w0 * 0.16801f + w1 * 0.19823f;
    float d = w0 * -1.27393f + w1 * -1.97760f, e = w0 * -4.85967f + w1 * -4.32054f;
    return a / (1.f + b * std::pow(x, c)) + d * x + e;
  }
  float lambda(float c) const { return c < 0.5f ? std::exp(L(c)) : std::exp(2.f * L(0.5f) - L(1.f - c)); }
  float lambdaI(float c) const { return std::pow(lambda(c), 1.f + 2.f * hsqr(hsqr(hsqr(1.f - c)))); }
  float G2(float co, float ci) const { return 1.0f / (1.f + lambda(co) + lambdaI(ci)); }
};
// bsdf/albedo.h:287-330: uniform hemisphere samples on a 32 x 32 midpoint grid
float sheenIntegrate(float cosO, const HostSheen& m)
{
  const int n = 32;
  const float pi = 3.14159265358979323846f;
  float c = std::max(cosO, 1e-6f);
  H3 wo = hmk(std::sqrt(std::max(1.f - c * c, 0.f)), 0.f, c);
  float sum = 0.f;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
    {
      float sx = std::min((float(i) + 0.5f) / float(n), 1.f - 1e-6f), sy = std::min((float(j) + 0.5f) / float(n), 1.f - 1e-6f);
      float phi = 2.f * pi * sx, st = std::sqrt(std::max(1.f - sy * sy, 0.f));
      H3 wi = hmk(std::cos(phi) * st, std::sin(phi) * st, sy);
      H3 wh = hnorm(hadd(wo, wi));
      sum += m.D(wh.z) * m.G2(c, wi.z) * (1.0f / (4.f * c * (1.f / (2.f * pi))));
    }
  return std::min(sum / float(n * n), 1.f);
}

void computeLuts(std::vector<float>& L)
{
  L.assign(PRT_LUT_FLOATS, 0.f);
  for (int j = 0; j < 32; ++j)
  {
    HostGGX m{std::max(hsqr(float(j) / 31.f), 0.001f)};
    float* d = L.data() + PRT_LUT_GGX_DIR + j * 32;
    for (int i = 0; i < 32; ++i) d[i] = lutIntegrate(float(i) / 31.f, m, 1.f, 0);
    L[PRT_LUT_GGX_AVG + j] = lutAvg(d, 32);
  }
  // MicrofacetDielectricAlbedoTable (eta in [1/3, 1] and [1, 3]) and
  // MicrofacetDielectricReflectionAlbedoTable (eta in [1/3, 1]), albedo_tables.cpp:94-116
  // MicrofacetSheenAlbedoTable (albedo_tables.cpp:155-170)
  for (int j = 0; j < 16; ++j)
  {
    HostSheen m{std::max(hsqr(float(j) / 15.f), 0.001f)};
    for (int i = 0; i < 16; ++i) L[PRT_LUT_SHEEN_DIR + j * 16 + i] = sheenIntegrate(float(i) / 15.f, m);
  }
  const int dirOff[3] = {PRT_LUT_DIEL_ETA_DIR, PRT_LUT_DIEL_RCPETA_DIR, PRT_LUT_DIEL_REFL_DIR};
  const int avgOff[3] = {PRT_LUT_DIEL_ETA_AVG, PRT_LUT_DIEL_RCPETA_AVG, PRT_LUT_DIEL_REFL_AVG};
  for (int table = 0; table < 3; ++table)
  {
    float minEta = table == 1 ? 1.f : 1.0f / 3.0f, maxEta = table == 1 ? 3.f : 1.f;
    float* dir = L.data() + dirOff[table];
    float* avg = L.data() + avgOff[table];
    for (int k = 0; k < 16; ++k)
    {
      HostGGX m{std::max(hsqr(float(k) / 15.f), 0.001f)};
      for (int j = 0; j < 16; ++j)
      {
        float eta = hlerp(minEta, maxEta, float(j) / 15.f);
        float* d = dir + (k * 16 + j) * 16;
        for (int i = 0; i < 16; ++i) d[i] = lutIntegrate(float(i) / 15.f, m, eta, table == 2 ? 2 : 1);
        avg[k * 16 + j] = lutAvg(d, 16);
      }
    }
  }
}
float lutInterp2(float px, float py, const float* f, int n)
{
  float xc = hclamp(px, 0.f, float(n - 1)), yc = hclamp(py, 0.f, float(n - 1));
  float sx = xc - std::floor(xc), sy = yc - std::floor(yc);
  int x0 = std::min(int(xc), n - 1), x1 = std::min(x0 + 1, n - 1), y0 = std::min(int(yc), n - 1), y1 = std::min(y0 + 1, n - 1);
  return hlerp(hlerp(f[x0 + y0 * n], f[x1 + y0 * n], sx), hlerp(f[x0 + y1 * n], f[x1 + y1 * n], sx), sy);
}
float dielAvg(const std::vector<float>& L, float eta, float roughness)  // albedo_tables.h:129-154
{
  if (eta <= 1.f)
  {
    float e = (eta - 1.0f / 3.0f) / (1.0f - 1.0f / 3.0f);
    return lutInterp2(e * 15.f, roughness * 15.f, L.data() + PRT_LUT_DIEL_ETA_AVG, 16);
  }
  float e = (eta - 1.f) / 2.f;
  return lutInterp2(e * 15.f, roughness * 15.f, L.data() + PRT_LUT_DIEL_RCPETA_AVG, 16);
}
float dielReflAvg(const std::vector<float>& L, float eta, float roughness)  // albedo_tables.h:187-194
{
  float e = (eta - 1.0f / 3.0f) / (1.0f - 1.0f / 3.0f);
  return lutInterp2(e * 15.f, roughness * 15.f, L.data() + PRT_LUT_DIEL_REFL_AVG, 16);
}
float fresnelAvg(float eta)  // bsdf/optics.h:144-175
{
  float re = 1.0f / eta;
  if (re >= 1.f) return (re - 1.f) / (4.08567f + 1.00071f * re);
  float re2 = re * re, re3 = re2 * re;
  return 0.997118f + 0.1014f * re - 0.965241f * re2 - 0.130607f * re3;
}

void hostMakeFrame(H3 N, H3& U, H3& V)  // math/basis3.h:173-180
{
  H3 U0 = hmk(0.f, N.z, -N.y), U1 = hmk(-N.z, 0.f, N.x);
  U = hnorm(hdot(U0, U0) > hdot(U1, U1) ? U0 : U1);
  V = hnorm(hcross(N, U));
}

bool isPowerOfTwo(int x) { return (x & (x - 1)) == 0; }

}  // namespace

// ---- the static scratch of prt_query_rays_device_noalloc -----------------------------------
// Traversal-stack overflow for a query that allocates nothing: kScratchBlocks persistent blocks of
// kTraceBlock lanes, kScratchOverflow (node, distance) entries per lane beyond the LDS stack:
// 128 x 256 x 2 x 64 ints = 16 MiB, held by every GPU the library runs on until the process ends.
// In host code `scratch` names a host-side stand-in; the device address, per device, comes from
// hipGetSymbolAddress at context creation (setupScratch) and only that address reaches kernels.
constexpr int kScratchBlocks = 128;
constexpr int kScratchOverflow = 64;
constexpr size_t kScratchInts = size_t(kScratchBlocks) * kTraceBlock * 2 * kScratchOverflow;
static __device__ __align__(256) int scratch[kScratchInts];
constexpr int kMaxDevices = 64;
static std::mutex gScratchMutex;
static int* gScratchPtr[kMaxDevices] = {};

// Records the device address of `scratch` on `device` (the current device) once.
static int setupScratch(int device)
{
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (device < 0 || device >= kMaxDevices) return fail(PRT_ERR_INVALID, "device ordinal beyond the scratch table");
  if (gScratchPtr[device]) return PRT_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return fail(PRT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  gScratchPtr[device] = static_cast<int*>(p);
  return PRT_OK;
}

static int* scratchOf(int device)
{
  std::lock_guard<std::mutex> lock(gScratchMutex);
  return (device >= 0 && device < kMaxDevices) ? gScratchPtr[device] : nullptr;
}

struct prt_context
{
  float boundsLo[3] = {1.f, 1.f, 1.f}, boundsHi[3] = {-1.f, -1.f, -1.f};  // scene box (empty: lo > hi)
  int anyTransparent = 0;  // some material lets shadow rays through
  int extMaterials = 0;    // some material is a thin dielectric, has opacity, an Oren-Nayar base,
                           // sheen or a coloured interior medium
  int stdMaterials = 0;    // some material is of kind 5
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t props{};
  int cus = 0;
  // scene
  int nverts = 0, ntris = 0, maxMaterialId = 0;
  DevBuf<float> pos, nrm;
  DevBuf<int> idx, triMat;
  std::vector<prt_material> materials;
  std::vector<prt_material_constants> consts;  // host copy of mats, before texture bindings
  DevBuf<prt_material_constants> mats;
  // image textures: per-vertex UVs, texture records, texels (decoded, rows bottom to top) and
  // the baseColor texture of each material (-1 none; reset by prt_scene_set_materials)
  DevBuf<float> uv;
  std::vector<prt_texture_dev> texHost;
  DevBuf<prt_texture_dev> tex;
  DevBuf<float4> texels;
  std::vector<int> baseTex;
  bool textured = false;  // some material has a texture bound
  DevBuf<float4> media;  // interior media: (r, g, b, ior) per medium ID, ID 0 clear
  std::vector<float> luts;
  DevBuf<float> lut;
  std::vector<prt_light_dev> lights;
  std::vector<float> lightCdf;
  DevBuf<float4> areaTris;
  DevBuf<float> areaCdf;
  DevBuf<int> areaMap;
  int areaCount = 0;
  float areaPower = 0.f;
  DevBuf<float4> envTex;
  DevBuf<float> envX, envY;
  DevBuf<int> envXG, envYG;  // guide tables of the sampling CDFs
  int envW = 0, envH = 0, envLW = 0, envLH = 0;
  float sceneRadius = 0.f;
  bool lightsReady = false;
  BvhData bvh;
  bool bvhReady = false;
  // two-level scene (geometry/instance.h): meshes are triangle ranges [meshBegin[m],
  // meshBegin[m + 1]) of the geometry, each with an object-space tree; the instances place them
  std::vector<int> meshBegin;
  std::vector<prt_instance> instances;
  std::vector<BvhData> blas;
  DevBuf<prt_bvh4_node> tlas;
  int tlasRoot = 0;
  DevBuf<prt_instance_dev> instDev;
  DevBuf<MeshView> meshDev;
  DevBuf<int> hitInst;
  bool instanced() const { return !instances.empty(); }
  // renderer
  prt_renderer_params params{100, 8, 1e-4f, 3.402823466e38f};
  bool rendererReady = false;
  uint32_t seed = 0, sampleIndex = 0;
  prt_camera cam{};
  bool camReady = false;
  // frame
  int W = 0, H = 0, npix = 0;
  uint32_t aovMask = 0;
  int count = 0;
  DevBuf<float> color, weight, color2, colorC, weightC, colorB, weightB;
  // pixel sample: jitter (FrameBuffer::setJitter) and pixel filter (DeviceCpu::initFilter); both
  // belong to the frame and are reset by prt_frame_init, as a new FrameBuffer resets them
  bool jitter = false;
  float jitterXY[2] = {0.f, 0.f};
  DevBuf<float> filterCdf;
  int filterSize = 0;
  float filterWidth = 0.f;
  DevBuf<float> albedo, albedoW, albedo1, albedo1W, normal, normalW, normal1, normal1W, depth, depthW;
  DevBuf<float> dbgL;
  DevBuf<uint32_t> dbgHash;
  DevBuf<int> dbgFirst;
  bool debug = false;
  bool countSteps = false;
  // bumped by every call that changes the scene, lights, BVH or renderer parameters (the device
  // pointers and values a batch frame copied into its kernel arguments)
  uint64_t generation = 0;  // traversal kernels built with the SIMD-utilisation counters
  // path state
  DevBuf<float4> rayO, rayD, hit, TP, Lr, shO, shD, dirL;
  DevBuf<int> flags, q0, q1, shadowQ, counters, spill;
  DevBuf<uint32_t> hash;
  DevBuf<unsigned long long> stats;
  DevBuf<RenderArgs> argsDev;
  RenderArgs* argsHost = nullptr;  // pinned staging copy
  LaunchCfg launch{};
  int spillStride = 0;
  // timing
  hipEvent_t ev[8] = {};
  std::vector<hipEvent_t> evPass[2];  // per-pass kernel timing events, double-buffered
  double kernelMs[4] = {0, 0, 0, 0};
  double lastStats[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  double launches[2] = {0, 0};
};

#define CTX_CHECK(c)                                                 \
  do {                                                               \
    if (!(c)) return fail(PRT_ERR_INVALID, "null context");          \
    hipError_t e_ = hipSetDevice((c)->device);                       \
    if (e_ != hipSuccess) return fail(PRT_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

namespace prt {
int setLastError(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace prt

extern "C" {

const char* prt_last_error(void) { return gLastError.c_str(); }

int prt_context_create(int device_ordinal, prt_context** out)
{
  if (!out) return fail(PRT_ERR_INVALID, "out is null");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device_ordinal < 0 || device_ordinal >= n) return fail(PRT_ERR_INVALID, "invalid device ordinal");
  HIP_TRY(hipSetDevice(device_ordinal));
  const int rcScratch = setupScratch(device_ordinal);
  if (rcScratch != PRT_OK) return rcScratch;
  prt_context* c = new prt_context();
  c->device = device_ordinal;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(PRT_ERR_HIP, hipGetErrorString(e)); }
  hipGetDeviceProperties(&c->props, device_ordinal);
  c->cus = c->props.multiProcessorCount;
  c->launch.gridTrace = c->cus * PRT_TRACE_BLOCKS_PER_CU;
  // one resident round of shading blocks: 4 blocks of 256 lanes per CU are the 4 waves per SIMD
  // the kernel is compiled for (measured on the 1080p workload: 4.17 ms of shading per pass
  // against 4.26 at 3, 4.57 at 5 and 4.65 at 8 blocks per CU)
  c->launch.gridShade = c->cus * 4;
  c->launch.gridOccluded = c->cus * PRT_OCCLUDED_BLOCKS_PER_CU;
  c->spillStride = std::max(c->launch.gridTrace, c->launch.gridOccluded) * kTraceBlock;
  for (auto& ev : c->ev) hipEventCreate(&ev);
  // albedo LUT precompute, as DeviceCpu's constructor does (render/device_cpu.cpp:20-23)
  computeLuts(c->luts);
  if (c->lut.alloc(PRT_LUT_FLOATS) != hipSuccess ||
      hipMemcpy(c->lut.p, c->luts.data(), sizeof(float) * PRT_LUT_FLOATS, hipMemcpyHostToDevice) != hipSuccess ||
      c->spill.alloc(2 * size_t(PRT_MAX_STACK - PRT_LDS_STACK) * c->spillStride) != hipSuccess ||
      c->stats.alloc(kStatCount) != hipSuccess || c->argsDev.alloc(1) != hipSuccess ||
      hipHostMalloc((void**)&c->argsHost, sizeof(RenderArgs), hipHostMallocDefault) != hipSuccess)
  {
    prt_context_destroy(c);
    return fail(PRT_ERR_NOMEM, "device allocation failed");
  }
  *out = c;
  return PRT_OK;
}

int prt_context_destroy(prt_context* c)
{
  if (!c) return PRT_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  freeBvh(c->bvh);
  for (auto& b : c->blas) freeBvh(b);
  if (c->argsHost) hipHostFree(c->argsHost);
  for (auto& ev : c->ev) if (ev) hipEventDestroy(ev);
  for (auto& set : c->evPass)
    for (auto& ev : set) if (ev) hipEventDestroy(ev);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PRT_OK;
}

int prt_debug_copy(prt_context* c, int64_t bytes, double* ms)
{
  CTX_CHECK(c);
  if (bytes <= 0 || bytes > (int64_t(1) << 36)) return fail(PRT_ERR_INVALID, "invalid copy size");
  DevBuf<char> a, b;
  HIP_TRY(a.alloc(size_t(bytes)));
  HIP_TRY(b.alloc(size_t(bytes)));
  HIP_TRY(hipMemsetAsync(a.p, 1, size_t(bytes), c->stream));
  HIP_TRY(hipEventRecord(c->ev[6], c->stream));
  HIP_TRY(hipMemcpyAsync(b.p, a.p, size_t(bytes), hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipEventRecord(c->ev[7], c->stream));
  HIP_TRY(hipEventSynchronize(c->ev[7]));
  float m = 0.f;
  HIP_TRY(hipEventElapsedTime(&m, c->ev[6], c->ev[7]));
  if (ms) *ms = m;
  return PRT_OK;
}

int prt_device_synchronize(prt_context* c)
{
  CTX_CHECK(c);
  HIP_TRY(hipDeviceSynchronize());
  return PRT_OK;
}

int prt_get_info(prt_context* c, char* buf, int len)
{
  CTX_CHECK(c);
  std::string s = std::string(c->props.name) + " (" + c->props.gcnArchName + "), " + std::to_string(c->cus) + " CUs, " +
                  std::to_string(c->props.totalGlobalMem >> 20) + " MiB";
  if (buf && len > 0)
  {
    std::strncpy(buf, s.c_str(), size_t(len) - 1);
    buf[len - 1] = 0;
  }
  return PRT_OK;
}

int prt_scene_set_geometry(prt_context* c, int32_t nverts, const float* positions, const float* normals, int32_t ntris,
                           const int32_t* indices, const int32_t* tri_material)
{
  CTX_CHECK(c);
  ++c->generation;
  if (nverts < 0 || ntris < 0 || (nverts > 0 && !positions) || (ntris > 0 && (!indices || !tri_material)))
    return fail(PRT_ERR_INVALID, "invalid geometry arrays");
  if (ntris >= (1 << 28)) return fail(PRT_ERR_INVALID, "too many triangles (limit 2^28)");
  for (int64_t i = 0; i < int64_t(ntris) * 3; ++i)
    if (indices[i] < 0 || indices[i] >= nverts) return fail(PRT_ERR_INVALID, "triangle index out of range");
  int maxMat = 0;
  for (int i = 0; i < ntris; ++i)
  {
    if (tri_material[i] < 1) return fail(PRT_ERR_INVALID, "material ID 0 is reserved for the environment");
    maxMat = std::max(maxMat, tri_material[i]);
  }
  c->maxMaterialId = maxMat;
  c->nverts = nverts;
  c->ntris = ntris;
  HIP_TRY(c->pos.alloc(size_t(nverts) * 3));
  HIP_TRY(c->idx.alloc(size_t(ntris) * 3));
  HIP_TRY(c->triMat.alloc(size_t(ntris)));
  if (nverts) HIP_TRY(hipMemcpy(c->pos.p, positions, sizeof(float) * 3 * nverts, hipMemcpyHostToDevice));
  if (ntris)
  {
    HIP_TRY(hipMemcpy(c->idx.p, indices, sizeof(int) * 3 * size_t(ntris), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->triMat.p, tri_material, sizeof(int) * size_t(ntris), hipMemcpyHostToDevice));
  }
  if (normals)
  {
    HIP_TRY(c->nrm.alloc(size_t(nverts) * 3));
    HIP_TRY(hipMemcpy(c->nrm.p, normals, sizeof(float) * 3 * nverts, hipMemcpyHostToDevice));
  }
  else c->nrm.release();
  c->uv.release();  // UVs belong to the vertices: set them again after new geometry
  c->meshBegin.clear();  // and so do the meshes and their instances
  c->instances.clear();
  // scene bounds over the referenced vertices (render/scene.cpp:81, :367)
  float lo[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f}, hi[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
  for (int64_t i = 0; i < int64_t(ntris) * 3; ++i)
    for (int k = 0; k < 3; ++k)
    {
      float v = positions[3 * int64_t(indices[i]) + k];
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  for (int k = 0; k < 3; ++k) { c->boundsLo[k] = lo[k]; c->boundsHi[k] = hi[k]; }
  H3 ext = ntris ? hmk(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) : hmk(0, 0, 0);
  c->sceneRadius = std::sqrt(hdot(ext, ext)) * 0.5f;
  c->bvhReady = false;
  c->lightsReady = false;
  return PRT_OK;
}

// The interior medium of a transmissive surface that is not thin-walled: colour
// transmission_color ^ transmission_depth (material/standard.cpp:48-49, :169-173) and IOR.  Media
// are deduplicated by (colour, IOR), ID 0 being the clear ambient medium
// (material/material_factory.cpp:20-21, :159-175); media holds (r, g, b, ior) per ID.
static int mediumOf(std::vector<float4>& media, float ior, const prt_material& s)
{
  const float d = std::max(s.transmission_depth, 0.f);
  float col[3];
  for (int j = 0; j < 3; ++j) col[j] = std::pow(hclamp(s.transmission_color[j], 0.f, 1.f), d);
  for (size_t t = 0; t < media.size(); ++t)
    if (media[t].x == col[0] && media[t].y == col[1] && media[t].z == col[2] && media[t].w == ior) return int(t);
  media.push_back(make_float4(col[0], col[1], col[2], ior));
  return int(media.size()) - 1;
}

// One lobe's constants (the per-hit constructor work of its reference class for the given,
// already clamped parameters).  kind: PRT_KIND_DIFFUSE .. PRT_KIND_THIN_DIELECTRIC or
// PRT_KIND_TRANSPARENT; etaArg[o]: the eta the dielectric, coating or thin dielectric is built
// with seen from outside (o = 0) and inside (o = 1), clamped to [1/3, 3]; coat*: the coating of
// the coated diffuse kind.
struct LobeParams
{
  int kind = PRT_KIND_DIFFUSE;
  float rgb[3] = {0.f, 0.f, 0.f}, edge[3] = {0.f, 0.f, 0.f};
  float roughness = 0.f, diffRough = 0.f;
  float etaArg[2] = {1.f, 1.f};
  float coatColor[3] = {1.f, 1.f, 1.f}, thickness = 0.f, scale = 0.f;
};

static void lobeConstants(const std::vector<float>& luts, const LobeParams& p, prt_material_constants& q)
{
  q.kind = p.kind;
  q.roughness = p.roughness;
  q.alpha = std::max(hsqr(q.roughness), 0.001f);  // bsdf/microfacet.h:10-24
  q.importance = 1.f;
  q.diffuseRoughness = p.diffRough;
  q.smooth = 1;
  const float* rgb = p.rgb;
  if (p.kind == PRT_KIND_DIFFUSE)
  {
    for (int j = 0; j < 3; ++j) q.c0[j] = rgb[j];
    q.importance = std::max(std::max(rgb[0], rgb[1]), rgb[2]);  // bsdf/diffuse.h:72-75
  }
  else if (p.kind == PRT_KIND_CONDUCTOR)
  {
    q.Eavg = lutInterp1(q.roughness * 31.f, luts.data() + PRT_LUT_GGX_AVG, 32);
    for (int j = 0; j < 3; ++j)
    {
      // bsdf/optics.h:214-240 (Gulbrandsen + Kulla-Conty average)
      float r = std::min(rgb[j], 0.99f), g = p.edge[j];
      float nMin = (1.f - r) / (1.f + r), nMax = (1.f + std::sqrt(r)) / (1.f - std::sqrt(r));
      float nn = g * nMin + (1.f - g) * nMax;
      float k2 = (hsqr(nn + 1.f) * r - hsqr(nn - 1.f)) / (1.f - r);
      q.c0[j] = nn;
      q.c1[j] = std::sqrt(std::max(k2, 0.f));
      float rr = rgb[j], r2 = rr * rr, r3 = r2 * rr, g2 = g * g, g3 = g2 * g;
      float Favg = 0.087237f + 0.0230685f * g - 0.0864902f * g2 + 0.0774594f * g3 + 0.782654f * rr - 0.136432f * r2 +
                   0.278708f * r3 + 0.19744f * g * rr + 0.0360605f * g2 * rr - 0.2586f * g * r2;
      q.c2[j] = hsqr(Favg) * q.Eavg / (1.f - Favg * (1.f - q.Eavg));  // rough_conductor.h:127-129
      q.rr[j] = rgb[j];
      q.edge[j] = g;
      q.Favg[j] = Favg;
    }
  }
  else if (p.kind == PRT_KIND_COATED)
  {
    // RoughCoatingBsdf over a diffuse base (bsdf/rough_coating.h:287-295): the member eta is
    // folded to <= 1, Ravg uses the argument as given
    float baseImportance = std::max(std::max(rgb[0], rgb[1]), rgb[2]);  // bsdf/diffuse.h:70-73
    for (int j = 0; j < 3; ++j)
    {
      q.c0[j] = rgb[j];
      q.c1[j] = p.coatColor[j];
      // getAlbedo (rough_coating.h:567-571): lerp(base, base * coat^thickness, scale)
      q.rr[j] = rgb[j] + p.scale * (rgb[j] * std::pow(p.coatColor[j], p.thickness) - rgb[j]);
    }
    q.c2[0] = p.thickness;
    q.c2[1] = p.scale;
    q.c2[2] = baseImportance;
    for (int o = 0; o < 2; ++o) q.eta[o] = p.etaArg[o] <= 1.f ? p.etaArg[o] : 1.0f / p.etaArg[o];
    q.Eavg = dielReflAvg(luts, p.etaArg[0], q.roughness) * p.scale;  // Ravg (one orientation per record)
    q.importance = std::max(1.f, baseImportance);                     // rough_coating.h:562-565
  }
  else if (p.kind == PRT_KIND_THIN_DIELECTRIC)
  {
    // ThinRoughDielectricBsdf (bsdf/thin_rough_dielectric.h:180-192)
    const float eta = p.etaArg[0];
    q.Eavg = lutInterp1(q.roughness * 31.f, luts.data() + PRT_LUT_GGX_AVG, 32);
    float Favg = fresnelAvg(eta);
    q.c1[0] = Favg * (1.f - q.Eavg) / (1.f - Favg * q.Eavg);  // fmsScale
    q.c2[0] = p.thickness;
    q.eta[0] = q.eta[1] = eta;
    for (int j = 0; j < 3; ++j)
    {
      q.c0[j] = rgb[j];
      q.rr[j] = std::pow(rgb[j], p.thickness);  // getAlbedo (:327-330)
    }
  }
  else if (p.kind == PRT_KIND_DIELECTRIC)
  {
    for (int o = 0; o < 2; ++o)
    {
      const float eta = p.etaArg[o];
      float Ee = dielAvg(luts, eta, q.roughness), Er = dielAvg(luts, 1.0f / eta, q.roughness);
      float Fe = fresnelAvg(eta), Fr = fresnelAvg(1.0f / eta);
      float a = (1.f - Fe) / (1.f - Er), b = (1.f - Fr) / (1.f - Ee);
      float x = b / (a + b);
      q.eta[o] = eta;
      q.EavgEta[o] = Ee;
      q.EavgRcpEta[o] = Er;
      q.fmsRatio[o] = 1.f - x * (1.f - Fe);  // rough_dielectric.h:283-291
      q.FavgEta[o] = Fe;
      q.FavgRcpEta[o] = Fr;
    }
  }
  else  // the transparent BTDF (bsdf/transparent.h:32-71): importance 1, no smooth lobe
  {
    q.smooth = 0;
  }
}

static float foldIor(float ior) { return ior < 1.f ? 1.0f / ior : ior; }

// The material records with their texture bindings applied, to the device.
static int uploadMaterialConstants(prt_context* c)
{
  std::vector<prt_material_constants> k = c->consts;
  c->textured = false;
  for (size_t i = 0; i < c->baseTex.size(); ++i)
  {
    k[i].texBase = c->baseTex[i] + 1;
    for (int j = 0; j < 3; ++j) k[i].texScale[j] = (&c->materials[i].r)[j];
    c->textured = c->textured || c->baseTex[i] >= 0;
  }
  HIP_TRY(c->mats.alloc(k.size()));
  HIP_TRY(hipMemcpy(c->mats.p, k.data(), sizeof(prt_material_constants) * k.size(), hipMemcpyHostToDevice));
  return PRT_OK;
}

int prt_scene_set_materials(prt_context* c, int32_t n, const prt_material* m)
{
  CTX_CHECK(c);
  ++c->generation;
  if (n < 1 || !m) return fail(PRT_ERR_INVALID, "need at least the reserved material 0");
  c->materials.assign(m, m + n);
  // the n materials' records, then the base records of kind-5 materials
  std::vector<prt_material_constants> k(n), extra;
  std::vector<float4> media(1, make_float4(1.f, 1.f, 1.f, 1.f));  // ID 0: the clear ambient medium
  const float eps = 1e-5f;  // material/standard.h:79
  for (int i = 0; i < n; ++i)
  {
    prt_material_constants& q = k[i];
    std::memset(&q, 0, sizeof(q));
    const prt_material& s = m[i];
    if (s.kind < PRT_KIND_DIFFUSE || s.kind > PRT_KIND_STANDARD) return fail(PRT_ERR_INVALID, "unknown material kind");
    if (!(s.transparency >= 0.f && s.transparency <= 1.f)) return fail(PRT_ERR_INVALID, "transparency must be in [0, 1]");
    for (int j = 0; j < 3; ++j)
      if (!(s.emission[j] >= 0.f)) return fail(PRT_ERR_INVALID, "emission must be non-negative");
    LobeParams p;
    p.kind = s.kind;
    for (int j = 0; j < 3; ++j)
    {
      p.rgb[j] = hclamp((&s.r)[j], 0.f, 1.f);
      p.edge[j] = hclamp((&s.er)[j], 0.f, 1.f);
    }
    p.roughness = hclamp(s.roughness, 0.f, 1.f);
    p.diffRough = hclamp(s.diffuse_roughness, 0.f, 1.f);
    const float ior = foldIor(s.ior);
    bool mediaInterface = false;
    if (s.kind == PRT_KIND_COATED)
    {
      // material/standard.cpp:512-518, 580-597; opaque, so never a media interface: eta = 1 / ior
      if (!(ior > 1.f)) return fail(PRT_ERR_INVALID, "coat IOR must differ from the ambient IOR 1");
      p.etaArg[0] = p.etaArg[1] = hclamp(1.0f / ior, 1.f / 3.f, 3.f);
      p.thickness = std::max(s.thickness, 0.f);
      p.scale = hclamp(s.weight, 0.f, 1.f);
      for (int j = 0; j < 3; ++j) p.coatColor[j] = p.edge[j];
    }
    else if (s.kind == PRT_KIND_THIN_DIELECTRIC)
    {
      // thin-walled, so always seen from outside: eta = ambientIor / specularIor (standard.cpp:429)
      if (!(ior > 1.f)) return fail(PRT_ERR_INVALID, "dielectric IOR must differ from the ambient IOR 1");
      p.etaArg[0] = p.etaArg[1] = hclamp(1.0f / ior, 1.f / 3.f, 3.f);
      p.thickness = std::max(s.thickness, 0.f);
    }
    else if (s.kind == PRT_KIND_DIELECTRIC)
    {
      // material/standard.cpp:436-437: outside eta = 1/ior, inside eta = ior, clamped to [1/3, 3]
      if (!(ior > 1.f)) return fail(PRT_ERR_INVALID, "dielectric IOR must differ from the ambient IOR 1");
      p.etaArg[0] = hclamp(1.0f / ior, 1.f / 3.f, 3.f);
      p.etaArg[1] = hclamp(ior, 1.f / 3.f, 3.f);
      mediaInterface = true;
    }
    if (s.kind != PRT_KIND_STANDARD) lobeConstants(c->luts, p, q);
    else
    {
      // StandardMaterial's bases (material/standard.cpp:409-566, ambient IOR 1), one set of base
      // records per orientation (fromOutside = thinWalled || !backfacing)
      if (!(s.ior > 0.f) || !(s.coat_ior > 0.f)) return fail(PRT_ERR_INVALID, "IORs must be positive");
      const bool thin = s.thin_walled != 0;
      const float base = hclamp(s.base_weight, 0.f, 1.f), metalness = hclamp(s.metalness, 0.f, 1.f);
      const float specular = hclamp(s.specular_weight, 0.f, 1.f);
      const float transmission = hclamp(s.transmission_weight, 0.f, 1.f);
      mediaInterface = s.transmission_weight != 0.f && !thin;  // MaterialType::MediaInterface (:159-175)
      q.kind = PRT_KIND_STANDARD;
      q.importance = 0.f;
      int flagsDiffuseOnly = 1;
      for (int o = 0; o < 2; ++o)
      {
        const bool fromOutside = thin || o == 0;
        const float eta = hclamp(fromOutside ? 1.0f / ior : ior, 1.f / 3.f, 3.f);
        std::vector<std::pair<LobeParams, float>> kids;
        const float dielectric = 1.f - metalness;
        if (dielectric > eps)
        {
          const float opaque = dielectric * (1.f - transmission);
          if (opaque > eps)  // (:428-505): diffuse, under a specular coating where the IOR differs
          {
            LobeParams c0 = p;
            for (int j = 0; j < 3; ++j) c0.rgb[j] = p.rgb[j] * base;
            if (std::fabs(ior - 1.f) > eps && specular > eps)
            {
              c0.kind = PRT_KIND_COATED;
              c0.etaArg[0] = c0.etaArg[1] = eta;
              c0.thickness = 1.f;
              c0.scale = specular;
              c0.coatColor[0] = c0.coatColor[1] = c0.coatColor[2] = 1.f;
            }
            else c0.kind = PRT_KIND_DIFFUSE;
            kids.push_back({c0, opaque});
          }
          const float translucent = dielectric * transmission * specular;
          if (translucent > eps)  // (:507-530)
          {
            LobeParams c1 = p;
            if (std::fabs(ior - 1.f) > eps)
            {
              if (thin)
              {
                c1.kind = PRT_KIND_THIN_DIELECTRIC;
                c1.etaArg[0] = c1.etaArg[1] = eta;
                for (int j = 0; j < 3; ++j) c1.rgb[j] = hclamp(s.transmission_color[j], 0.f, 1.f);
                c1.thickness = std::max(s.thickness, 0.f) / std::max(s.transmission_depth, 0.f);
              }
              else
              {
                c1.kind = PRT_KIND_DIELECTRIC;
                c1.etaArg[0] = c1.etaArg[1] = eta;
              }
            }
            else c1.kind = PRT_KIND_TRANSPARENT;
            kids.push_back({c1, translucent});
          }
        }
        const float metal = metalness * specular;
        if (metal > eps)  // (:533-546)
        {
          LobeParams c2 = p;
          c2.kind = PRT_KIND_CONDUCTOR;
          for (int j = 0; j < 3; ++j) c2.rgb[j] = p.rgb[j] * base;
          kids.push_back({c2, metal});
        }
        q.nKids = int(kids.size());
        const size_t first = extra.size();
        q.kidBase[o] = int(n + first) - i;  // relative to this record
        float impSum = 0.f;
        int anySmooth = 0;
        for (size_t t = 0; t < kids.size(); ++t)
        {
          prt_material_constants kq;
          std::memset(&kq, 0, sizeof(kq));
          lobeConstants(c->luts, kids[t].first, kq);
          kq.opacity = 1.f;
          q.kidScale[t] = kids[t].second;
          q.kidImp[t] = kq.importance * kids[t].second;
          impSum += q.kidImp[t];
          anySmooth |= kq.smooth;
          if (kids[t].first.kind != PRT_KIND_DIFFUSE) flagsDiffuseOnly = 0;
          extra.push_back(kq);
        }
        q.impSum = impSum;
        q.smooth = anySmooth;
        // the coat over the mix (:551-566)
        const float coat = hclamp(s.coat_weight, 0.f, 1.f), coatIor = foldIor(s.coat_ior);
        q.coat = (coat > eps && std::fabs(coatIor - 1.f) > eps) ? 1 : 0;
        if (q.coat)
        {
          const float etaArg = hclamp(fromOutside ? 1.0f / coatIor : coatIor, 1.f / 3.f, 3.f);
          q.coatRoughness = hclamp(s.coat_roughness, 0.f, 1.f);
          q.coatAlpha = std::max(hsqr(q.coatRoughness), 0.001f);
          q.coatEta[o] = etaArg <= 1.f ? etaArg : 1.0f / etaArg;
          q.coatRavg[o] = dielReflAvg(c->luts, etaArg, q.coatRoughness) * coat;
          q.coatScale = coat;
          q.coatThickness = std::max(s.coat_thickness, 0.f);
          for (int j = 0; j < 3; ++j) q.coatColor[j] = hclamp(s.coat_color[j], 0.f, 1.f);
          q.smooth = 1;
        }
        q.importance = q.coat ? std::max(1.f, impSum) : impSum;  // rough_coating.h:562-565
        // the albedo AOV of the mix (multi_bsdf.h:314-337), coated (rough_coating.h:567-571)
        float alb[3] = {0.f, 0.f, 0.f}, scaleSum = 0.f, total = 0.f;
        for (size_t t = 0; t < kids.size(); ++t)
        {
          const prt_material_constants& kq = extra[first + t];
          if (!(q.kidImp[t] > 0.f)) continue;
          total += q.kidScale[t];
          if (kq.kind == PRT_KIND_TRANSPARENT) continue;
          for (int j = 0; j < 3; ++j) alb[j] += (kq.kind == PRT_KIND_DIFFUSE ? kq.c0[j] : (kq.kind == PRT_KIND_DIELECTRIC ? 1.f : kq.rr[j])) * q.kidScale[t];
          scaleSum += q.kidScale[t];
        }
        for (int j = 0; j < 3; ++j)
        {
          float a = alb[j] * (total * (scaleSum == 0.f ? 3.402823466e38f : 1.0f / scaleSum));
          if (q.coat) a = a + q.coatScale * (a * std::pow(q.coatColor[j], q.coatThickness) - a);
          if (o == 0) q.rr[j] = a;  // the same in both orientations
        }
      }
      q.mixDiffuseOnly = flagsDiffuseOnly;
      q.roughness = q.coatRoughness;
      q.alpha = q.coat ? q.coatAlpha : 0.001f;
      q.diffuseRoughness = p.diffRough;
    }
    // opacity (material/standard.cpp:412-420, :613-627, eps = 1e-5, standard.h:79)
    q.opacity = hclamp(1.f - s.transparency, 0.f, 1.f);
    q.transp = 1.f - q.opacity;
    q.mix = q.transp > 1e-5f ? (q.opacity > 1e-5f ? 1 : 2) : 0;
    // MaterialType::Transparent (:152-166): opacity below 1, or thin-walled transmission
    q.transparent = (s.transparency != 0.f || s.kind == PRT_KIND_THIN_DIELECTRIC ||
                     (s.kind == PRT_KIND_STANDARD && s.thin_walled && s.transmission_weight != 0.f)) ? 1 : 0;
    // sheen (standard.cpp:599-611): colour and roughness clamped, alpha = roughnessToAlpha
    q.sheenScale = hclamp(s.sheen_weight, 0.f, 1.f);
    q.sheen = q.sheenScale > 1e-5f ? 1 : 0;
    for (int j = 0; j < 3; ++j) q.sheenColor[j] = hclamp(s.sheen_color[j], 0.f, 1.f);
    q.sheenRoughness = hclamp(s.sheen_roughness, 0.f, 1.f);
    q.sheenAlpha = std::max(hsqr(q.sheenRoughness), 0.001f);
    if (q.sheen) q.smooth = 1;
    for (int j = 0; j < 3; ++j) q.emission[j] = s.emission[j];
    q.emissive = (s.emission[0] > 0.f || s.emission[1] > 0.f || s.emission[2] > 0.f) ? 1 : 0;
    q.mediaInterface = mediaInterface ? 1 : 0;
    q.transmissive = (s.kind == PRT_KIND_DIELECTRIC || s.kind == PRT_KIND_THIN_DIELECTRIC ||
                      (s.kind == PRT_KIND_STANDARD && s.transmission_weight != 0.f)) ? 1 : 0;
    if (mediaInterface) q.mediumInt = mediumOf(media, ior, s);
  }
  c->anyTransparent = c->extMaterials = c->stdMaterials = 0;
  for (int i = 0; i < n; ++i)
  {
    c->anyTransparent |= k[i].transparent;
    c->stdMaterials |= k[i].kind == PRT_KIND_STANDARD ? 1 : 0;
    c->extMaterials |= (k[i].kind == PRT_KIND_THIN_DIELECTRIC || k[i].kind == PRT_KIND_STANDARD || k[i].mix != 0 ||
                        k[i].diffuseRoughness > 0.f || k[i].sheen) ? 1 : 0;
  }
  for (size_t t = 1; t < media.size(); ++t)  // a coloured medium absorbs (the base kernel skips that)
    if (media[t].x != 1.f || media[t].y != 1.f || media[t].z != 1.f) c->extMaterials = 1;
  k.insert(k.end(), extra.begin(), extra.end());
  c->consts = k;
  c->baseTex.assign(size_t(n), -1);
  int rc = uploadMaterialConstants(c);
  if (rc != PRT_OK) return rc;
  HIP_TRY(c->media.alloc(media.size()));
  HIP_TRY(hipMemcpy(c->media.p, media.data(), sizeof(float4) * media.size(), hipMemcpyHostToDevice));
  c->lightsReady = false;  // the area light depends on the materials: set the lights again
  return PRT_OK;
}

int prt_scene_set_uvs(prt_context* c, const float* uv, int32_t nverts)
{
  CTX_CHECK(c);
  ++c->generation;
  if (!uv)
  {
    c->uv.release();
    return PRT_OK;
  }
  if (nverts != c->nverts) return fail(PRT_ERR_INVALID, "UVs need one pair per vertex of the geometry");
  for (int64_t i = 0; i < int64_t(nverts) * 2; ++i)
    if (!std::isfinite(uv[i])) return fail(PRT_ERR_INVALID, "UVs must be finite");
  HIP_TRY(c->uv.alloc(size_t(nverts) * 2));
  if (nverts) HIP_TRY(hipMemcpy(c->uv.p, uv, sizeof(float) * 2 * size_t(nverts), hipMemcpyHostToDevice));
  return PRT_OK;
}

int prt_scene_set_textures(prt_context* c, int32_t n, const prt_texture* t)
{
  CTX_CHECK(c);
  ++c->generation;
  if (n < 0 || (n > 0 && !t)) return fail(PRT_ERR_INVALID, "invalid texture array");
  for (int b : c->baseTex)
    if (b >= n) return fail(PRT_ERR_INVALID, "a material is bound to a texture beyond the new set");
  // 8-bit sRGB decode: the reference's 2.2-gamma table (image/pixel.cpp:21-86, generated there
  // as pow(i / 255, 2.2) in float)
  float srgb[256];
  for (int i = 0; i < 256; ++i) srgb[i] = std::pow(float(i) / 255.0f, 2.2f);
  std::vector<prt_texture_dev> recs(static_cast<size_t>(n));
  int64_t total = 0;
  for (int i = 0; i < n; ++i)
  {
    const prt_texture& x = t[i];
    if (x.width < 1 || x.height < 1 || !x.texels) return fail(PRT_ERR_INVALID, "texture needs a size and texels");
    if (x.format < PRT_TEX_RGBA8_SRGB || x.format > PRT_TEX_RGB32F) return fail(PRT_ERR_INVALID, "unknown texture format");
    if ((x.address_u != PRT_TEX_REPEAT && x.address_u != PRT_TEX_CLAMP) ||
        (x.address_v != PRT_TEX_REPEAT && x.address_v != PRT_TEX_CLAMP))
      return fail(PRT_ERR_INVALID, "unknown texture address mode");
    prt_texture_dev& r = recs[size_t(i)];
    r.offset = int32_t(total);
    r.w = x.width;
    r.h = x.height;
    r.addrU = x.address_u == PRT_TEX_CLAMP ? PRT_TEX_ADDR_CLAMP : PRT_TEX_ADDR_REPEAT;
    r.addrV = x.address_v == PRT_TEX_CLAMP ? PRT_TEX_ADDR_CLAMP : PRT_TEX_ADDR_REPEAT;
    r.scaleU = x.uv_scale[0];
    r.scaleV = x.uv_scale[1];
    r.offU = x.uv_offset[0];
    r.offV = x.uv_offset[1];
    r.halfU = 0.5f / float(x.width);  // halfTexel = 0.5 / size (image_texture.h:163)
    r.halfV = 0.5f / float(x.height);
    r.pad = 0;
    total += int64_t(x.width) * x.height;
    if (total >= (int64_t(1) << 31)) return fail(PRT_ERR_INVALID, "textures exceed 2^31 texels");
  }
  std::vector<float4> texels(size_t(std::max<int64_t>(total, 1)));
  for (int i = 0; i < n; ++i)
  {
    const prt_texture& x = t[i];
    const prt_texture_dev& r = recs[size_t(i)];
    for (int y = 0; y < x.height; ++y)
      for (int xx = 0; xx < x.width; ++xx)
      {
        // texel row y counts from the bottom: image row height - 1 - y (image_texture.h:352-353)
        const size_t src = size_t(x.height - 1 - y) * x.width + xx;
        float4 v;
        if (x.format == PRT_TEX_RGB32F)
        {
          const float* f = static_cast<const float*>(x.texels) + 3 * src;
          v = make_float4(f[0], f[1], f[2], 1.f);
        }
        else
        {
          const uint8_t* b = static_cast<const uint8_t*>(x.texels) + 4 * src;
          if (x.format == PRT_TEX_RGBA8_SRGB)  // decodeBgra8 (image/pixel.h:80-91)
            v = make_float4(srgb[b[0]], srgb[b[1]], srgb[b[2]], float(b[3]) * (1.f / 255.f));
          else  // decodeBgra8Linear (:137-145)
            v = make_float4(float(b[0]) * (1.f / 255.f), float(b[1]) * (1.f / 255.f), float(b[2]) * (1.f / 255.f),
                            float(b[3]) * (1.f / 255.f));
        }
        texels[size_t(r.offset) + size_t(y) * x.width + xx] = v;
      }
  }
  c->texHost = recs;
  HIP_TRY(c->tex.alloc(std::max<size_t>(recs.size(), 1)));
  if (n) HIP_TRY(hipMemcpy(c->tex.p, recs.data(), sizeof(prt_texture_dev) * recs.size(), hipMemcpyHostToDevice));
  HIP_TRY(c->texels.alloc(texels.size()));
  HIP_TRY(hipMemcpy(c->texels.p, texels.data(), sizeof(float4) * texels.size(), hipMemcpyHostToDevice));
  return PRT_OK;
}

int prt_material_set_texture(prt_context* c, int32_t material, int32_t parameter, int32_t texture)
{
  CTX_CHECK(c);
  ++c->generation;
  if (material < 0 || material >= int32_t(c->materials.size())) return fail(PRT_ERR_INVALID, "invalid material ID");
  if (parameter != PRT_PARAM_BASE_COLOR) return fail(PRT_ERR_INVALID, "only the base colour can be textured");
  if (texture < -1 || texture >= int32_t(c->texHost.size())) return fail(PRT_ERR_INVALID, "invalid texture index");
  const int kind = c->materials[size_t(material)].kind;
  if (texture >= 0 && kind != PRT_KIND_DIFFUSE && kind != PRT_KIND_COATED)
    return fail(PRT_ERR_INVALID, "base colour textures apply to diffuse and coated materials");
  c->baseTex[size_t(material)] = texture;
  return uploadMaterialConstants(c);
}

// light/triangle_light.h:36-64 + render/scene.cpp:83-116: one TriangleLight over the emissive
// triangles of the mesh (the uploaded geometry is one TriangleMesh) in primitive order, picked
// by power |cross(e1, e2)| * luminance(emission).  The geometry is read back from the device
// only when some material emits.
static int buildAreaLight(prt_context* c)
{
  c->areaCount = 0;
  c->areaPower = 0.f;
  c->areaTris.release(); c->areaCdf.release(); c->areaMap.release();
  bool anyEmissive = false;
  for (const prt_material& m : c->materials)
    anyEmissive |= (m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f);
  if (!anyEmissive || c->ntris == 0) return PRT_OK;
  std::vector<int> triMat(c->ntris), idx(size_t(c->ntris) * 3);
  std::vector<float> pos(size_t(c->nverts) * 3);
  HIP_TRY(hipMemcpy(triMat.data(), c->triMat.p, sizeof(int) * triMat.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(idx.data(), c->idx.p, sizeof(int) * idx.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pos.data(), c->pos.p, sizeof(float) * pos.size(), hipMemcpyDeviceToHost));
  std::vector<int> map(c->ntris, -1);
  std::vector<float4> tris;
  std::vector<float> w;
  // a two-level scene: emissive triangles must lie in world space, in meshes placed exactly once
  // by the identity (the reference flattens emissive objects, render/scene.cpp:343-363)
  std::vector<char> worldMesh;
  if (c->instanced())
  {
    const int nMesh = int(c->meshBegin.size()) - 1;
    std::vector<int> uses(size_t(nMesh), 0);
    std::vector<char> ident(size_t(nMesh), 1);
    for (const prt_instance& I : c->instances)
    {
      ++uses[size_t(I.mesh)];
      const float id[12] = {1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f};
      for (int k = 0; k < 12; ++k)
        if (I.xfm[k] != id[k]) ident[size_t(I.mesh)] = 0;
    }
    worldMesh.resize(size_t(c->ntris));
    for (int m = 0; m < nMesh; ++m)
      for (int t = c->meshBegin[size_t(m)]; t < c->meshBegin[size_t(m) + 1]; ++t)
        worldMesh[size_t(t)] = uses[size_t(m)] == 1 && ident[size_t(m)];
  }
  for (int t = 0; t < c->ntris; ++t)
  {
    if (triMat[t] >= int(c->materials.size())) return fail(PRT_ERR_INVALID, "triangle material ID beyond the material table");
    const prt_material& m = c->materials[triMat[t]];
    if (!(m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f)) continue;
    if (!worldMesh.empty() && !worldMesh[size_t(t)])
      return fail(PRT_ERR_INVALID, "emissive triangles of a two-level scene must be in a mesh placed once by the identity");
    const float* p0 = &pos[3 * size_t(idx[3 * t])];
    const float* p1 = &pos[3 * size_t(idx[3 * t + 1])];
    const float* p2 = &pos[3 * size_t(idx[3 * t + 2])];
    H3 v0 = hmk(p0[0], p0[1], p0[2]), v1 = hmk(p1[0], p1[1], p1[2]), v2 = hmk(p2[0], p2[1], p2[2]);
    H3 nrm = hcross(hsub(v1, v0), hsub(v2, v0));
    map[t] = int(tris.size() / 4);
    tris.push_back(make_float4(v0.x, v0.y, v0.z, m.emission[0]));
    tris.push_back(make_float4(v1.x, v1.y, v1.z, m.emission[1]));
    tris.push_back(make_float4(v2.x, v2.y, v2.z, m.emission[2]));
    tris.push_back(make_float4(nrm.x, nrm.y, nrm.z, 0.f));
    float lum = 0.212671f * m.emission[0] + 0.715160f * m.emission[1] + 0.072169f * m.emission[2];
    w.push_back(std::sqrt(hdot(nrm, nrm)) * lum);
  }
  const int n = int(w.size());
  if (n == 0) return PRT_OK;
  std::vector<float> cdf(n + 1);
  cdf[0] = 0.f;  // sampling/distribution1d.cpp:18-32
  for (int i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + w[i - 1];
  float norm = cdf[n] == 0.f ? 0.f : 1.0f / cdf[n];
  for (int i = 1; i <= n; ++i) cdf[i] *= norm;
  cdf[n] = 1.f;
  float sumL = 0.f;  // getPower (triangle_light.h:152-162)
  for (int i = 0; i < n; ++i) sumL += w[i];
  c->areaPower = 3.14159265358979323846f * sumL;
  HIP_TRY(c->areaTris.alloc(tris.size()));
  HIP_TRY(c->areaCdf.alloc(cdf.size()));
  HIP_TRY(c->areaMap.alloc(map.size()));
  HIP_TRY(hipMemcpy(c->areaTris.p, tris.data(), sizeof(float4) * tris.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->areaCdf.p, cdf.data(), sizeof(float) * cdf.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->areaMap.p, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
  c->areaCount = n;
  return PRT_OK;
}

int prt_scene_set_lights(prt_context* c, int32_t n, const prt_light* lights, int32_t env_w, int32_t env_h, const float* env_rgb,
                         int32_t lut_max)
{
  CTX_CHECK(c);
  ++c->generation;
  if (n < 0 || n > PRT_MAX_LIGHTS || (n > 0 && !lights)) return fail(PRT_ERR_INVALID, "0..8 environment lights");
  if (c->materials.empty()) return fail(PRT_ERR_STATE, "materials must be set before the lights");
  {
    int rc = buildAreaLight(c);
    if (rc != PRT_OK) return rc;
  }
  prt_light defaultAmbient{};
  if (n == 0 && c->areaCount == 0)
  {
    // render/scene.cpp:432-438: no lights at all -> a default ambient light (Ke = 1)
    defaultAmbient.kind = PRT_LIGHT_AMBIENT;
    defaultAmbient.r = defaultAmbient.g = defaultAmbient.b = 1.f;
    lights = &defaultAmbient;
    n = 1;
  }
  bool needImage = false;
  for (int i = 0; i < n; ++i) needImage |= lights[i].kind == PRT_LIGHT_IMAGE;
  if (needImage && (!env_rgb || env_w <= 0 || env_h <= 0)) return fail(PRT_ERR_INVALID, "image light without an image");
  const float pi = 3.14159265358979323846f;
  std::vector<float> sumL;
  if (env_rgb && env_w > 0 && env_h > 0)
  {
    // light/image_light.h:20-54: importance LUT with 2x bucketing down to lut_max
    std::vector<float4> tex(size_t(env_w) * env_h);
    for (size_t i = 0; i < tex.size(); ++i) tex[i] = make_float4(env_rgb[3 * i], env_rgb[3 * i + 1], env_rgb[3 * i + 2], 0.f);
    int lw = env_w, lh = env_h, bucket = 1;
    if (lut_max > 0)
      while (std::max(lw, lh) > lut_max && lw % 2 == 0 && lh % 2 == 0) { lw /= 2; lh /= 2; bucket *= 2; }
    auto lum = [&](int x, int y) {
      const float* p = env_rgb + (size_t(y) * env_w + x) * 3;
      return 0.212671f * p[0] + 0.715160f * p[1] + 0.072169f * p[2];
    };
    std::vector<float> w(size_t(lw) * lh);
    for (int y = 0; y < lh; ++y)
      for (int x = 0; x < lw; ++x)
      {
        float sum = 0.f;
        for (int y2 = 0; y2 < bucket; ++y2)
          for (int x2 = 0; x2 < bucket; ++x2) sum += lum(x * bucket + x2, y * bucket + y2);
        w[size_t(y) * lw + x] = sum * std::sin((float(y) + 0.5f) / float(lh) * pi);
      }
    // sampling/distribution2d.cpp:16-34 and distribution1d.cpp:18-32
    auto init1d = [](float* cdf, const float* f, int size) {
      cdf[0] = 0.f;
      for (int i = 1; i < size + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1];
      float norm = (cdf[size] == 0.f) ? 0.f : 1.0f / cdf[size];
      for (int i = 1; i < size + 1; ++i) cdf[i] *= norm;
      cdf[size] = 1.f;
    };
    std::vector<float> xd(size_t(lh) * (lw + 1)), yd(lh + 1), fy(lh);
    for (int y = 0; y < lh; ++y)
    {
      fy[y] = 0.f;
      for (int x = 0; x < lw; ++x) fy[y] += w[size_t(y) * lw + x];
      init1d(xd.data() + size_t(y) * (lw + 1), w.data() + size_t(y) * lw, lw);
    }
    init1d(yd.data(), fy.data(), lh);
    // guide tables: the index the device's full search (distribution1d.cpp:61-80) returns at
    // each bin start g / PRT_ENV_GUIDE, and at 1 for the end
    auto lookup = [](const float* cdf, int size, float v) {
      int first = 0, len = size;
      while (len > 0)
      {
        int half = len >> 1, mid = first + half;
        if (v < cdf[mid]) len = half;
        else { first = mid + 1; len = len - half - 1; }
      }
      return first - 1 > 0 ? first - 1 : 0;
    };
    std::vector<int> xg(size_t(lh) * (PRT_ENV_GUIDE + 1)), yg(PRT_ENV_GUIDE + 1);
    for (int g = 0; g <= PRT_ENV_GUIDE; ++g)
    {
      const float v = float(g) / float(PRT_ENV_GUIDE);
      yg[g] = lookup(yd.data(), lh, v);
      for (int y = 0; y < lh; ++y) xg[size_t(y) * (PRT_ENV_GUIDE + 1) + g] = lookup(xd.data() + size_t(y) * (lw + 1), lw, v);
    }
    HIP_TRY(c->envXG.alloc(xg.size()));
    HIP_TRY(c->envYG.alloc(yg.size()));
    HIP_TRY(hipMemcpy(c->envXG.p, xg.data(), sizeof(int) * xg.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->envYG.p, yg.data(), sizeof(int) * yg.size(), hipMemcpyHostToDevice));
    HIP_TRY(c->envTex.alloc(tex.size()));
    HIP_TRY(c->envX.alloc(xd.size()));
    HIP_TRY(c->envY.alloc(yd.size()));
    HIP_TRY(hipMemcpy(c->envTex.p, tex.data(), sizeof(float4) * tex.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->envX.p, xd.data(), sizeof(float) * xd.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->envY.p, yd.data(), sizeof(float) * yd.size(), hipMemcpyHostToDevice));
    c->envW = env_w; c->envH = env_h; c->envLW = lw; c->envLH = lh;
  }
  c->lights.assign(n, prt_light_dev{});
  std::vector<float> power(n);
  const float R = c->sceneRadius;
  for (int i = 0; i < n; ++i)
  {
    const prt_light& l = lights[i];
    prt_light_dev& d = c->lights[i];
    std::memset(&d, 0, sizeof(d));
    d.kind = l.kind;
    d.r = l.r; d.g = l.g; d.b = l.b;
    float lumC = 0.212671f * l.r + 0.715160f * l.g + 0.072169f * l.b;
    if (l.kind == PRT_LIGHT_AMBIENT)
      power[i] = 4.f * hsqr(pi) * hsqr(R) * lumC;  // ambient_light.h:53-57
    else if (l.kind == PRT_LIGHT_DISTANT)
    {
      H3 dir = hnorm(hmk(l.dx, l.dy, l.dz));  // distant_light.h:31-36
      d.dx = dir.x; d.dy = dir.y; d.dz = dir.z;
      d.cosAngle = std::cos(l.angle);
      d.conePdf = 1.0f / (4.f * pi * hsqr(std::sin(0.5f * l.angle)));
      H3 U, V;
      hostMakeFrame(dir, U, V);
      d.fu[0] = U.x; d.fu[1] = U.y; d.fu[2] = U.z;
      d.fv[0] = V.x; d.fv[1] = V.y; d.fv[2] = V.z;
      power[i] = 4.f * hsqr(pi) * hsqr(std::sin(0.5f * l.angle)) * hsqr(R) * lumC;  // distant_light.h:73-77
    }
    else if (l.kind == PRT_LIGHT_IMAGE)
    {
      // light/image_light.h:137-148 over the image this light emits (the texels times its scale)
      float sumL = 0.f;
      for (size_t t = 0; t < size_t(env_w) * env_h; ++t)
      {
        const float* p = env_rgb + 3 * t;
        sumL += 0.212671f * (p[0] * l.r) + 0.715160f * (p[1] * l.g) + 0.072169f * (p[2] * l.b);
      }
      power[i] = 4.f * hsqr(pi) * hsqr(R) * (sumL / float(env_w * env_h));
    }
    else
      return fail(PRT_ERR_INVALID, "unknown light kind");
  }
  // render/scene.cpp:440-445: the light list is the area light (if any), then the
  // environment lights; one power distribution over all of them
  if (c->areaCount > 0) power.insert(power.begin(), c->areaPower);
  const int total = int(power.size());
  c->lightCdf.assign(total + 1, 0.f);
  for (int i = 1; i <= total; ++i) c->lightCdf[i] = c->lightCdf[i - 1] + power[i - 1];
  float norm = c->lightCdf[total] == 0.f ? 0.f : 1.0f / c->lightCdf[total];
  for (int i = 1; i <= total; ++i) c->lightCdf[i] *= norm;
  c->lightCdf[total] = 1.f;
  c->lightsReady = true;
  return PRT_OK;
}

// Top-level tree of a two-level scene, built on the host over the instances' world boxes: a
// 4-wide tree of nodes in the kernels' layout whose leaves each hold one instance (leaf code
// instance << 3), split top-down at the median of the longest centroid extent into four.
struct TlasBox { float lo[3], hi[3]; int inst; };
static int buildTlasNode(std::vector<prt_bvh4_node>& nodes, std::vector<TlasBox>& boxes, int begin, int end, int depth,
                         int& maxDepth)
{
  maxDepth = std::max(maxDepth, depth + 1);
  const int me = int(nodes.size());
  nodes.push_back(prt_bvh4_node{});
  // up to four groups: halve the range twice along the longest centroid axis
  int cuts[5] = {begin, begin, begin, begin, end};
  auto split = [&](int b, int e) {
    if (e - b < 2) return b + (e - b) / 2;
    float cl[3] = {3.4e38f, 3.4e38f, 3.4e38f}, ch[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int i = b; i < e; ++i)
      for (int k = 0; k < 3; ++k)
      {
        const float cc = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
        cl[k] = std::min(cl[k], cc);
        ch[k] = std::max(ch[k], cc);
      }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (ch[k] - cl[k] > ch[ax] - cl[ax]) ax = k;
    const int mid = b + (e - b) / 2;
    std::nth_element(boxes.begin() + b, boxes.begin() + mid, boxes.begin() + e, [&](const TlasBox& x, const TlasBox& y) {
      return x.lo[ax] + x.hi[ax] < y.lo[ax] + y.hi[ax];
    });
    return mid;
  };
  if (end - begin <= 4)
    for (int q = 0; q < 4; ++q) cuts[q] = std::min(begin + q, end);
  else
  {
    cuts[2] = split(begin, end);
    cuts[1] = split(begin, cuts[2]);
    cuts[3] = split(cuts[2], end);
  }
  cuts[4] = end;
  int child[4];
  float lo[4][3], hi[4][3];
  for (int q = 0; q < 4; ++q)
  {
    const int b = cuts[q], e = q == 3 ? end : cuts[q + 1];
    for (int k = 0; k < 3; ++k) { lo[q][k] = 3.4028235e38f; hi[q][k] = -3.4028235e38f; }
    if (e <= b) { child[q] = -1; continue; }
    for (int i = b; i < e; ++i)
      for (int k = 0; k < 3; ++k)
      {
        lo[q][k] = std::min(lo[q][k], boxes[i].lo[k]);
        hi[q][k] = std::max(hi[q][k], boxes[i].hi[k]);
      }
    child[q] = e - b == 1 ? int(0x80000000u | (uint32_t(boxes[b].inst) << 3)) : buildTlasNode(nodes, boxes, b, e, depth + 1, maxDepth);
  }
  prt_bvh4_node& n = nodes[size_t(me)];
  for (int q = 0; q < 4; ++q)
  {
    const bool empty = child[q] == -1;
    n.lox[q] = empty ? __builtin_inff() : lo[q][0];
    n.hix[q] = empty ? -__builtin_inff() : hi[q][0];
    n.loy[q] = empty ? __builtin_inff() : lo[q][1];
    n.hiy[q] = empty ? -__builtin_inff() : hi[q][1];
    n.loz[q] = empty ? __builtin_inff() : lo[q][2];
    n.hiz[q] = empty ? -__builtin_inff() : hi[q][2];
    n.child[q] = child[q];
    n.pad[q] = 0;
  }
  return me;
}

// Mesh trees, the top-level tree and the instance records of a two-level scene.
static int buildTwoLevel(prt_context* c)
{
  const int nMesh = int(c->meshBegin.size()) - 1;
  for (auto& b : c->blas) freeBvh(b);
  c->blas.assign(size_t(nMesh), BvhData{});
  double ms = 0.0;
  int blasDepth = 0;
  std::string err;
  for (int m = 0; m < nMesh; ++m)
  {
    const int b = c->meshBegin[size_t(m)], n = c->meshBegin[size_t(m) + 1] - b;
    hipError_t e = buildBvh(c->pos.p, c->idx.p + 3 * size_t(b), n, c->blas[size_t(m)], c->stream, err);
    if (e != hipSuccess) return fail(PRT_ERR_HIP, err);
    ms += c->blas[size_t(m)].buildMs;
    blasDepth = std::max(blasDepth, c->blas[size_t(m)].maxDepth);
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  // object-space boxes of the meshes from the uploaded vertices
  std::vector<float> pos(size_t(c->nverts) * 3);
  std::vector<int> idx(size_t(c->ntris) * 3);
  if (c->nverts) HIP_TRY(hipMemcpy(pos.data(), c->pos.p, sizeof(float) * pos.size(), hipMemcpyDeviceToHost));
  if (c->ntris) HIP_TRY(hipMemcpy(idx.data(), c->idx.p, sizeof(int) * idx.size(), hipMemcpyDeviceToHost));
  std::vector<std::array<double, 6>> mbox(static_cast<size_t>(nMesh));
  for (int m = 0; m < nMesh; ++m)
  {
    auto& bx = mbox[size_t(m)];
    bx = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int t = c->meshBegin[size_t(m)]; t < c->meshBegin[size_t(m) + 1]; ++t)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
        {
          const double v = pos[3 * size_t(idx[3 * size_t(t) + j]) + k];
          bx[k] = std::min(bx[k], v);
          bx[3 + k] = std::max(bx[3 + k], v);
        }
  }
  // world boxes: the eight corners through x' = L x + t (double), padded outward by a relative
  // 2^-20, as the builder pads triangle boxes, so rounding never culls an instance
  std::vector<TlasBox> boxes;
  double sceneLo[3] = {1e300, 1e300, 1e300}, sceneHi[3] = {-1e300, -1e300, -1e300};
  std::vector<prt_instance_dev> recs(c->instances.size());
  for (size_t i = 0; i < c->instances.size(); ++i)
  {
    const prt_instance& I = c->instances[i];
    const auto& bx = mbox[size_t(I.mesh)];
    prt_instance_dev& r = recs[i];
    std::memset(&r, 0, sizeof(r));
    r.mesh = I.mesh;
    const float* x = I.xfm;  // rows of [L | t]
    double L[9] = {x[0], x[1], x[2], x[4], x[5], x[6], x[8], x[9], x[10]}, T[3] = {x[3], x[7], x[11]};
    const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) + L[2] * (L[3] * L[7] - L[4] * L[6]);
    double A[9] = {(L[4] * L[8] - L[5] * L[7]) / det, (L[2] * L[7] - L[1] * L[8]) / det, (L[1] * L[5] - L[2] * L[4]) / det,
                   (L[5] * L[6] - L[3] * L[8]) / det, (L[0] * L[8] - L[2] * L[6]) / det, (L[2] * L[3] - L[0] * L[5]) / det,
                   (L[3] * L[7] - L[4] * L[6]) / det, (L[1] * L[6] - L[0] * L[7]) / det, (L[0] * L[4] - L[1] * L[3]) / det};
    for (int row = 0; row < 3; ++row)
    {
      for (int k = 0; k < 3; ++k)
      {
        r.inv[4 * row + k] = float(A[3 * row + k]);
        r.lin[3 * row + k] = float(L[3 * row + k]);
        r.nrm[3 * row + k] = float(A[3 * k + row]);  // (L^-1)^T
      }
      r.inv[4 * row + 3] = float(-(A[3 * row] * T[0] + A[3 * row + 1] * T[1] + A[3 * row + 2] * T[2]));
    }
    if (c->meshBegin[size_t(I.mesh) + 1] == c->meshBegin[size_t(I.mesh)]) continue;  // an empty mesh is never hit
    TlasBox wb;
    wb.inst = int(i);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int corner = 0; corner < 8; ++corner)
    {
      const double p[3] = {bx[(corner & 1) ? 3 : 0], bx[(corner & 2) ? 4 : 1], bx[(corner & 4) ? 5 : 2]};
      for (int k = 0; k < 3; ++k)
      {
        const double w = L[3 * k] * p[0] + L[3 * k + 1] * p[1] + L[3 * k + 2] * p[2] + T[k];
        lo[k] = std::min(lo[k], w);
        hi[k] = std::max(hi[k], w);
      }
    }
    for (int k = 0; k < 3; ++k)
    {
      sceneLo[k] = std::min(sceneLo[k], lo[k]);
      sceneHi[k] = std::max(sceneHi[k], hi[k]);
      const double pad = std::max(std::fabs(lo[k]), std::fabs(hi[k])) * 9.5367431640625e-7 + 1e-30;
      wb.lo[k] = std::nextafter(float(lo[k] - pad), -3.4028235e38f);
      wb.hi[k] = std::nextafter(float(hi[k] + pad), 3.4028235e38f);
    }
    boxes.push_back(wb);
  }
  std::vector<prt_bvh4_node> nodes;
  int tlasDepth = 0;
  if (boxes.size() == 1) c->tlasRoot = int(0x80000000u | (uint32_t(boxes[0].inst) << 3));
  else if (boxes.empty())
  {
    // nothing to hit: a root whose four slots are unused
    prt_bvh4_node n{};
    for (int q = 0; q < 4; ++q)
    {
      n.lox[q] = n.loy[q] = n.loz[q] = __builtin_inff();
      n.hix[q] = n.hiy[q] = n.hiz[q] = -__builtin_inff();
      n.child[q] = -1;
    }
    nodes.push_back(n);
    c->tlasRoot = 0;
    tlasDepth = 1;
  }
  else c->tlasRoot = buildTlasNode(nodes, boxes, 0, int(boxes.size()), 0, tlasDepth);
  // a wide node pushes at most three children, plus one marker per level change
  if (3 * (tlasDepth + blasDepth) + 1 >= PRT_MAX_STACK)
    return fail(PRT_ERR_INVALID, "two-level tree too deep for the traversal stack");
  HIP_TRY(c->tlas.alloc(std::max<size_t>(nodes.size(), 1)));
  if (!nodes.empty()) HIP_TRY(hipMemcpy(c->tlas.p, nodes.data(), sizeof(prt_bvh4_node) * nodes.size(), hipMemcpyHostToDevice));
  HIP_TRY(c->instDev.alloc(recs.size()));
  HIP_TRY(hipMemcpy(c->instDev.p, recs.data(), sizeof(prt_instance_dev) * recs.size(), hipMemcpyHostToDevice));
  std::vector<MeshView> mv(static_cast<size_t>(nMesh));
  for (int m = 0; m < nMesh; ++m)
    mv[size_t(m)] = MeshView{c->blas[size_t(m)].nodes, c->blas[size_t(m)].tris, c->blas[size_t(m)].root, c->meshBegin[size_t(m)]};
  HIP_TRY(c->meshDev.alloc(mv.size()));
  HIP_TRY(hipMemcpy(c->meshDev.p, mv.data(), sizeof(MeshView) * mv.size(), hipMemcpyHostToDevice));
  // scene bounds and radius (render/scene.cpp:367): the instances' world boxes, the mesh boxes'
  // corners through their transforms (Instance::getBounds, geometry/instance.h:52-55), in double
  // rounded to float; the light powers depend on the radius, so set the lights after this build
  for (int k = 0; k < 3; ++k) { c->boundsLo[k] = 1.f; c->boundsHi[k] = -1.f; }
  c->sceneRadius = 0.f;
  if (!boxes.empty())
  {
    for (int k = 0; k < 3; ++k)
    {
      c->boundsLo[k] = float(sceneLo[k]);
      c->boundsHi[k] = float(sceneHi[k]);
    }
    H3 ext = hmk(c->boundsHi[0] - c->boundsLo[0], c->boundsHi[1] - c->boundsLo[1], c->boundsHi[2] - c->boundsLo[2]);
    c->sceneRadius = std::sqrt(hdot(ext, ext)) * 0.5f;
  }
  c->lightsReady = false;
  c->bvh.maxDepth = tlasDepth + blasDepth;
  c->bvh.buildMs = ms + std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
  return PRT_OK;
}

int prt_bvh_build(prt_context* c, double* build_ms)
{
  CTX_CHECK(c);
  ++c->generation;
  if (!c->pos.p && c->ntris > 0) return fail(PRT_ERR_STATE, "no geometry");
  c->bvhReady = false;
  if (c->instanced())
  {
    int rc = buildTwoLevel(c);
    if (rc != PRT_OK) return rc;
  }
  else
  {
    std::string err;
    hipError_t e = buildBvh(c->pos.p, c->idx.p, c->ntris, c->bvh, c->stream, err);
    if (e != hipSuccess) return fail(PRT_ERR_HIP, err);
  }
  c->bvhReady = true;
  if (build_ms) *build_ms = c->bvh.buildMs;
  return PRT_OK;
}

int prt_scene_set_instances(prt_context* c, int32_t n_meshes, const int32_t* mesh_begin, int32_t n, const prt_instance* inst)
{
  CTX_CHECK(c);
  ++c->generation;
  c->bvhReady = false;
  if (n == 0)
  {
    c->meshBegin.clear();
    c->instances.clear();
    return PRT_OK;
  }
  if (n < 0 || n_meshes < 1 || !mesh_begin || !inst) return fail(PRT_ERR_INVALID, "invalid instance arrays");
  if (n >= (1 << 28)) return fail(PRT_ERR_INVALID, "too many instances (limit 2^28)");
  if (mesh_begin[0] != 0 || mesh_begin[n_meshes] != c->ntris) return fail(PRT_ERR_INVALID, "meshes must cover the triangles in order");
  for (int m = 0; m < n_meshes; ++m)
    if (mesh_begin[m + 1] < mesh_begin[m]) return fail(PRT_ERR_INVALID, "mesh ranges must not decrease");
  for (int i = 0; i < n; ++i)
  {
    if (inst[i].mesh < 0 || inst[i].mesh >= n_meshes) return fail(PRT_ERR_INVALID, "instance of an unknown mesh");
    for (int k = 0; k < 12; ++k)
      if (!std::isfinite(inst[i].xfm[k])) return fail(PRT_ERR_INVALID, "instance transforms must be finite");
    const float* x = inst[i].xfm;
    const double det = double(x[0]) * (double(x[5]) * x[10] - double(x[6]) * x[9]) - double(x[1]) * (double(x[4]) * x[10] - double(x[6]) * x[8]) +
                       double(x[2]) * (double(x[4]) * x[9] - double(x[5]) * x[8]);
    if (!(std::fabs(det) > 1e-30)) return fail(PRT_ERR_INVALID, "instance transforms must be invertible");
  }
  c->meshBegin.assign(mesh_begin, mesh_begin + n_meshes + 1);
  c->instances.assign(inst, inst + n);
  return PRT_OK;
}

int prt_renderer_init(prt_context* c, const prt_renderer_params* p)
{
  CTX_CHECK(c);
  ++c->generation;
  if (p)
  {
    if (p->maxDepth < 0 || p->maxDepth > 1000 || p->rrDepth < 0) return fail(PRT_ERR_INVALID, "invalid depth parameters");
    c->params = *p;
  }
  HIP_TRY(c->counters.alloc(size_t(2) * (c->params.maxDepth + 3)));
  c->rendererReady = true;
  c->sampleIndex = 0;  // PathRendererStream ctor calls initSampler(0) (:118-119)
  c->seed = 0;
  return PRT_OK;
}

int prt_sampler_init(prt_context* c, uint32_t seed)
{
  CTX_CHECK(c);
  c->seed = seed;
  c->sampleIndex = 0;
  return PRT_OK;
}

int prt_set_camera(prt_context* c, const prt_camera* cam)
{
  CTX_CHECK(c);
  if (!cam) return fail(PRT_ERR_INVALID, "camera is null");
  if (!(cam->fov > 0.f && cam->fov < 3.14159265f)) return fail(PRT_ERR_INVALID, "invalid fov");
  c->cam = *cam;
  c->camReady = true;
  return PRT_OK;
}

static hipError_t allocFrame(prt_context* c)
{
  const size_t n = size_t(c->npix);
  hipError_t e;
  auto A = [&](DevBuf<float>& b, size_t cnt, bool on) -> hipError_t {
    if (!on) { b.release(); return hipSuccess; }
    return b.alloc(cnt);
  };
  if ((e = A(c->color, 3 * n, true)) || (e = A(c->weight, n, true)) || (e = A(c->colorC, 3 * n, true)) ||
      (e = A(c->weightC, n, true)) || (e = A(c->color2, 3 * n, c->aovMask & PRT_AOV_VAR)) ||
      (e = A(c->colorB, 3 * n, c->aovMask & PRT_AOV_HALF)) || (e = A(c->weightB, n, c->aovMask & PRT_AOV_HALF)) ||
      (e = A(c->albedo, 3 * n, c->aovMask & PRT_AOV_ALBEDO)) || (e = A(c->albedoW, n, c->aovMask & PRT_AOV_ALBEDO)) ||
      (e = A(c->albedo1, 3 * n, c->aovMask & PRT_AOV_ALBEDO1)) || (e = A(c->albedo1W, n, c->aovMask & PRT_AOV_ALBEDO1)) ||
      (e = A(c->normal, 3 * n, c->aovMask & PRT_AOV_NORMAL)) || (e = A(c->normalW, n, c->aovMask & PRT_AOV_NORMAL)) ||
      (e = A(c->normal1, 3 * n, c->aovMask & PRT_AOV_NORMAL1)) || (e = A(c->normal1W, n, c->aovMask & PRT_AOV_NORMAL1)) ||
      (e = A(c->depth, n, c->aovMask & PRT_AOV_DEPTH)) || (e = A(c->depthW, n, c->aovMask & PRT_AOV_DEPTH)))
    return e;
  if ((e = c->rayO.alloc(n)) || (e = c->rayD.alloc(n)) || (e = c->hit.alloc(n)) || (e = c->TP.alloc(n)) || (e = c->Lr.alloc(n)) ||
      (e = c->shO.alloc(n)) || (e = c->shD.alloc(n)) || (e = c->dirL.alloc(n)) || (e = c->flags.alloc(n)) ||
      (e = c->q0.alloc(n)) || (e = c->q1.alloc(n)) || (e = c->shadowQ.alloc(n)))
    return e;
  return hipSuccess;
}

int prt_frame_clear(prt_context* c)
{
  CTX_CHECK(c);
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  const size_t n = size_t(c->npix);
  DevBuf<float>* bufs[] = {&c->color, &c->weight, &c->color2, &c->colorC, &c->weightC, &c->colorB, &c->weightB, &c->albedo, &c->albedoW, &c->albedo1,
                           &c->albedo1W, &c->normal, &c->normalW, &c->normal1, &c->normal1W, &c->depth, &c->depthW};
  for (auto* b : bufs)
    if (b->p) HIP_TRY(hipMemsetAsync(b->p, 0, sizeof(float) * b->n, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->count = 0;
  (void)n;
  return PRT_OK;
}

int prt_frame_init(prt_context* c, int32_t width, int32_t height, uint32_t aov_mask)
{
  CTX_CHECK(c);
  // render/path_renderer_stream.h:109-112
  if (width <= 0 || height <= 0 || width % 32 != 0 || height % 8 != 0)
    return fail(PRT_ERR_INVALID, "image size is not divisible by the tile size");
  if (width > 4096 || height > 4096) return fail(PRT_ERR_INVALID, "image size is too large");
  if (aov_mask & ~uint32_t(PRT_AOV_ALL)) return fail(PRT_ERR_INVALID, "unknown AOV bits");
  c->W = width;
  c->H = height;
  c->npix = width * height;
  c->aovMask = aov_mask;
  // a new frame has no jitter and the box filter (FrameBuffer ctor, image/frame_buffer.cpp:9-14)
  c->jitter = false;
  c->filterCdf.release();
  c->filterSize = 0;
  c->filterWidth = 0.f;
  hipError_t e = allocFrame(c);
  if (e != hipSuccess) return fail(PRT_ERR_NOMEM, hipGetErrorString(e));
  if (c->debug)
  {
    HIP_TRY(c->dbgL.alloc(size_t(3) * c->npix));
    HIP_TRY(c->dbgHash.alloc(c->npix));
    HIP_TRY(c->dbgFirst.alloc(c->npix));
    HIP_TRY(c->hash.alloc(c->npix));
  }
  return prt_frame_clear(c);
}

int prt_debug_enable(prt_context* c, int32_t enable)
{
  CTX_CHECK(c);
  c->debug = enable != 0;
  if (c->debug && c->npix)
  {
    HIP_TRY(c->dbgL.alloc(size_t(3) * c->npix));
    HIP_TRY(c->dbgHash.alloc(c->npix));
    HIP_TRY(c->dbgFirst.alloc(c->npix));
    HIP_TRY(c->hash.alloc(c->npix));
  }
  if (!c->debug) { c->dbgL.release(); c->dbgHash.release(); c->dbgFirst.release(); c->hash.release(); }
  return PRT_OK;
}

// Camera (camera/pinhole_camera.h:72-90 and thin_lens_camera.h:155-161) and frame size of the
// kernel arguments.
static void setCameraArgs(RenderArgs& a, const prt_camera& k, int W, int H)
{
  const float aspect = float(W) / float(H);

  const float hh = std::tan(0.5f * k.fov), hw = hh * aspect;
  H3 U = hmk(k.U[0], k.U[1], k.U[2]), V = hmk(k.V[0], k.V[1], k.V[2]), N = hmk(k.N[0], k.N[1], k.N[2]);
  auto basisMul = [&](float x, float y, float z) { return hadd(hadd(hscale(U, x), hscale(V, y)), hscale(N, z)); };
  H3 iO = basisMul(-hw, hh, -1.f), iDx = basisMul(2.f * hw, 0.f, 0.f), iDy = basisMul(0.f, -2.f * hh, 0.f);
  float* dst[3] = {a.imageO, a.imageDx, a.imageDy};
  H3 src[3] = {iO, iDx, iDy};
  for (int j = 0; j < 3; ++j) { dst[j][0] = src[j].x; dst[j][1] = src[j].y; dst[j][2] = src[j].z; }
  for (int j = 0; j < 3; ++j) { a.camPos[j] = k.pos[j]; a.camU[j] = k.U[j]; a.camV[j] = k.V[j]; a.camN[j] = k.N[j]; }
  a.lensRadius = k.lensRadius;
  a.focal = k.focalDistance;
  a.nearClip = k.nearClip;
  a.nearClipScale = k.nearClip / k.focalDistance;
  a.pinhole = k.lensRadius == 0.f;  // render/device_cpu.cpp:117-128
  a.W = W; a.H = H; a.npix = W * H;
  a.invW = 1.0f / float(W);
  a.invH = 1.0f / float(H);
}

// Camera matrices (camera/pinhole_camera.h:41-68) in the reference's operation order; Mat4 is
// stored by columns and multiplied column by column (math/mat4.h:64-74).
int prt_camera_matrices(const prt_camera* k, int32_t width, int32_t height, float* wv, float* vc, float* wr)
{
  if (!k || width <= 0 || height <= 0) return fail(PRT_ERR_INVALID, "invalid camera or image size");
  const float W = float(width), H = float(height), aspect = W / H;
  const float hh = std::tan(0.5f * k->fov), hw = hh * aspect;
  const H3 P = hmk(k->pos[0], k->pos[1], k->pos[2]);
  const H3 U = hmk(k->U[0], k->U[1], k->U[2]), V = hmk(k->V[0], k->V[1], k->V[2]), N = hmk(k->N[0], k->N[1], k->N[2]);
  // row-major values as written in the constructors
  const float worldToView[16] = {U.x, U.y, U.z, -hdot(U, P), V.x, V.y, V.z, -hdot(V, P), N.x, N.y, N.z, -hdot(N, P), 0, 0, 0, 1};
  const float viewToRaster[16] = {W / (2.f * hw), 0, -W / 2.f, 0, 0, -H / (2.f * hh), -H / 2.f, 0, 0, 0, 0, k->nearClip, 0, 0, -1, 0};
  if (wr)
  {
    // (a * b) column j = a[0] b[j].x + a[1] b[j].y + a[2] b[j].z + a[3] b[j].w, left to right
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 4; ++r)
      {
        float acc = viewToRaster[r * 4 + 0] * worldToView[0 * 4 + j];
        acc = acc + viewToRaster[r * 4 + 1] * worldToView[1 * 4 + j];
        acc = acc + viewToRaster[r * 4 + 2] * worldToView[2 * 4 + j];
        acc = acc + viewToRaster[r * 4 + 3] * worldToView[3 * 4 + j];
        wr[r * 4 + j] = acc;
      }
  }
  const H3 F = hmk(-N.x, -N.y, -N.z);
  if (wv)
  {
    const float m[16] = {U.x, V.x, F.x, 0, U.y, V.y, F.y, 0, U.z, V.z, F.z, 0, -hdot(U, P), -hdot(V, P), -hdot(F, P), 1};
    std::memcpy(wv, m, sizeof(m));
  }
  if (vc)
  {
    const float m[16] = {1.f / hw, 0, 0, 0, 0, 1.f / hh, 0, 0, 0, 0, 0, 1, 0, 0, k->nearClip, 0};
    std::memcpy(vc, m, sizeof(m));
  }
  return PRT_OK;
}

static RenderArgs makeArgs(prt_context* c)
{
  RenderArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pos = c->pos.p; a.nrm = c->nrm.p; a.idx = c->idx.p; a.triMat = c->triMat.p; a.mats = c->mats.p; a.media = c->media.p; a.lut = c->lut.p;
  a.uv = c->uv.p; a.tex = c->tex.p; a.texels = c->texels.p;
  a.anyTransparent = c->anyTransparent;
  a.bvh.nodes = c->bvh.nodes; a.bvh.tris = c->bvh.tris; a.bvh.root = c->bvh.root;
  if (c->instanced())
  {
    a.inst.nodes = c->tlas.p; a.inst.root = c->tlasRoot; a.inst.inst = c->instDev.p; a.inst.meshes = c->meshDev.p;
    a.hitInst = c->hitInst.p;
  }
  a.nLights = int(c->lights.size());
  a.lightOffset = c->areaCount > 0 ? 1 : 0;
  for (int i = 0; i < a.nLights; ++i) a.lights[i] = c->lights[i];
  for (int i = 0; i <= a.nLights + a.lightOffset; ++i) a.lightCdf[i] = c->lightCdf[i];
  a.areaTris = c->areaTris.p; a.areaCdf = c->areaCdf.p; a.areaMap = c->areaMap.p; a.areaCount = c->areaCount;
  a.env.texels = c->envTex.p; a.env.xDists = c->envX.p; a.env.yDist = c->envY.p;
  a.env.xGuide = c->envXG.p; a.env.yGuide = c->envYG.p;
  a.env.W = c->envW; a.env.H = c->envH; a.env.LW = c->envLW; a.env.LH = c->envLH;
  setCameraArgs(a, c->cam, c->W, c->H);
  a.jitter = c->jitter ? 1 : 0;
  a.jitterX = c->jitterXY[0];
  a.jitterY = c->jitterXY[1];
  a.filterCdf = c->filterSize ? c->filterCdf.p : nullptr;
  a.filterSize = c->filterSize;
  a.filterWidth = c->filterWidth;
  prt_frame_dev& f = a.fb;
  f.color = c->color.p; f.weight = c->weight.p; f.color2 = c->color2.p;
  f.colorB = c->colorB.p; f.weightB = c->weightB.p;
  f.albedo = c->albedo.p; f.albedoW = c->albedoW.p; f.albedo1 = c->albedo1.p; f.albedo1W = c->albedo1W.p;
  f.normal = c->normal.p; f.normalW = c->normalW.p; f.normal1 = c->normal1.p; f.normal1W = c->normal1W.p;
  f.depth = c->depth.p; f.depthW = c->depthW.p;
  f.dbgL = c->debug ? c->dbgL.p : nullptr;
  f.dbgHash = c->debug ? c->dbgHash.p : nullptr;
  f.dbgFirst = c->debug ? c->dbgFirst.p : nullptr;
  a.seed = c->seed;
  a.maxDepth = c->params.maxDepth;
  a.rrDepth = c->params.rrDepth;
  a.minWeight = c->params.minWeight;
  a.maxRadiance = c->params.maxRadiance;
  a.rayO = c->rayO.p; a.rayD = c->rayD.p; a.hit = c->hit.p; a.TP = c->TP.p; a.Lr = c->Lr.p; a.flags = c->flags.p;
  a.hash = c->debug ? c->hash.p : nullptr;
  a.shO = c->shO.p; a.shD = c->shD.p; a.dirL = c->dirL.p;
  a.queue[0] = c->q0.p; a.queue[1] = c->q1.p; a.shadowQ = c->shadowQ.p; a.counters = c->counters.p;
  a.spill = c->spill.p; a.spillStride = c->spillStride;
  a.stats = c->stats.p;
  return a;
}

int prt_render(prt_context* c, int32_t n_spp, prt_stats* stats)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (!c->lightsReady) return fail(PRT_ERR_STATE, "lights not set");
  if (!c->mats.p) return fail(PRT_ERR_STATE, "materials not set");
  if (c->maxMaterialId >= int(c->materials.size())) return fail(PRT_ERR_INVALID, "triangle material ID beyond the material table");
  if (!c->rendererReady) return fail(PRT_ERR_STATE, "renderer not initialized");
  if (!c->camReady) return fail(PRT_ERR_STATE, "camera not set");
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  if (n_spp < 0) return fail(PRT_ERR_INVALID, "negative sample count");
  if (c->instanced())
  {
    if (c->hitInst.n < size_t(c->npix)) HIP_TRY(c->hitInst.alloc(size_t(c->npix)));
  }
  RenderArgs a = makeArgs(c);
  *c->argsHost = a;
  HIP_TRY(hipMemcpyAsync(c->argsDev.p, c->argsHost, sizeof(RenderArgs), hipMemcpyHostToDevice, c->stream));
  const RenderArgs* ad = c->argsDev.p;
  HIP_TRY(hipMemsetAsync(c->stats.p, 0, sizeof(unsigned long long) * kStatCount, c->stream));
  for (int k = 0; k < 4; ++k) c->kernelMs[k] = 0.0;
  c->launches[0] = c->launches[1] = 0;
  // The first kFreeDepths depths of a pass are enqueued without waiting (a depth with no queued
  // paths launches kernels that find an empty queue and return); deeper ones check on the host
  // whether any path survives, so a large maxDepth does not launch hundreds of empty kernels.
  // The kernel times of pass s are read after pass s + 1 is enqueued (two event sets).
  constexpr int kFreeDepths = 9;
  const size_t nEv = 2 + 3 * size_t(a.maxDepth + 1);
  for (auto& set : c->evPass)
    while (set.size() < nEv)
    {
      hipEvent_t e = nullptr;
      HIP_TRY(hipEventCreate(&e));
      set.push_back(e);
    }
  int depths[2] = {0, 0};  // depths enqueued in each event set's pass
  auto readTimes = [&](int set) -> hipError_t {
    const std::vector<hipEvent_t>& ev = c->evPass[set];
    const int nd = depths[set];
    hipError_t e = hipEventSynchronize(ev[1 + 3 * nd]);
    if (e != hipSuccess) return e;
    float m = 0.f;
    hipEventElapsedTime(&m, ev[0], ev[1]);
    c->kernelMs[0] += m;
    for (int d = 0; d < nd; ++d)
      for (int k = 0; k < 3; ++k)
      {
        hipEventElapsedTime(&m, ev[1 + 3 * d + k], ev[2 + 3 * d + k]);
        c->kernelMs[1 + k] += m;
      }
    c->launches[0] += nd;
    c->launches[1] += nd;
    return hipSuccess;
  };
  auto t0 = std::chrono::high_resolution_clock::now();
  for (int s = 0; s < n_spp; ++s)
  {
    const std::vector<hipEvent_t>& ev = c->evPass[s & 1];
    // FrameBuffer::begin: checkpoint copy at power-of-two sample counts (accum_buffer.h:252-260)
    if (isPowerOfTwo(c->count))
    {
      HIP_TRY(hipMemcpyAsync(c->colorC.p, c->color.p, sizeof(float) * 3 * c->npix, hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->weightC.p, c->weight.p, sizeof(float) * c->npix, hipMemcpyDeviceToDevice, c->stream));
    }
    // AccumBuffer::finish sets classB for the next sample from the sample count (accum_buffer.h:36-47)
    const bool classB = (__builtin_popcount(uint32_t(c->count) & 0xaaaaaaaau) & 1) != 0;
    HIP_TRY(hipEventRecord(ev[0], c->stream));
    launchGenerate(ad, a.npix, c->sampleIndex, c->stream);
    HIP_TRY(hipEventRecord(ev[1], c->stream));
    int depth = 0;
    for (; depth <= a.maxDepth; ++depth)
    {
      launchIntersect(ad, depth, c->launch, c->countSteps, c->instanced(), c->stream);
      HIP_TRY(hipEventRecord(ev[2 + 3 * depth], c->stream));
      launchShade(ad, depth, c->sampleIndex, classB, c->launch,
                  c->stdMaterials ? 2 : (c->extMaterials || c->textured || c->instanced() ? 1 : 0), c->stream);
      HIP_TRY(hipEventRecord(ev[3 + 3 * depth], c->stream));
      launchOccluded(ad, depth, classB, c->launch, c->anyTransparent ? (c->stdMaterials ? 2 : 1) : 0, c->countSteps,
                     c->instanced(), c->stream);
      HIP_TRY(hipEventRecord(ev[4 + 3 * depth], c->stream));
      if (depth + 1 >= kFreeDepths && depth < a.maxDepth)
      {
        int alive = 0;
        HIP_TRY(hipMemcpyAsync(&alive, c->counters.p + 2 * (depth + 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (alive == 0) { ++depth; break; }
      }
    }
    depths[s & 1] = depth;
    HIP_TRY(hipGetLastError());
    if (s > 0) HIP_TRY(readTimes((s - 1) & 1));
    c->sampleIndex++;  // (:185-186)
    c->count++;        // FrameBuffer::finish (accum_buffer.h:263-267)
  }
  if (n_spp > 0) HIP_TRY(readTimes((n_spp - 1) & 1));
  HIP_TRY(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::high_resolution_clock::now();
  unsigned long long st[kStatCount] = {};
  HIP_TRY(hipMemcpy(st, c->stats.p, sizeof(st), hipMemcpyDeviceToHost));
  for (int k = 0; k < 10; ++k) c->lastStats[k] = double(st[k]);
  if (stats)
  {
    double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double rays = double(st[0]) + double(st[3]);
    stats->rays = rays;
    stats->ms = ms;
    stats->mray = ms > 0 ? rays / 1000.0 / ms : 0.0;
    stats->spp = n_spp;
  }
  return PRT_OK;
}

int prt_blit(prt_context* c, int32_t buffer, float* dst, float* var)
{
  CTX_CHECK(c);
  if (!dst) return fail(PRT_ERR_INVALID, "dst is null");
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  const size_t n = size_t(c->npix);
  DevBuf<float>* s = nullptr;
  DevBuf<float>* w = nullptr;
  int ch = 3;
  switch (buffer)
  {
    case PRT_BUF_COLOR: s = &c->color; w = &c->weight; break;
    case PRT_BUF_ALBEDO: s = &c->albedo; w = &c->albedoW; break;
    case PRT_BUF_ALBEDO1: s = &c->albedo1; w = &c->albedo1W; break;
    case PRT_BUF_NORMAL: s = &c->normal; w = &c->normalW; break;
    case PRT_BUF_NORMAL1: s = &c->normal1; w = &c->normal1W; break;
    case PRT_BUF_DEPTH: s = &c->depth; w = &c->depthW; ch = 1; break;
    default: return fail(PRT_ERR_INVALID, "unknown buffer");
  }
  if (!s->p) return fail(PRT_ERR_INVALID, "buffer not enabled in prt_frame_init");
  if (var && (buffer != PRT_BUF_COLOR || !c->color2.p)) return fail(PRT_ERR_INVALID, "sample variance not enabled for the buffer");
  std::vector<float> hs(ch * n), hw(n), h2;
  HIP_TRY(hipMemcpy(hs.data(), s->p, sizeof(float) * ch * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hw.data(), w->p, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (var)
  {
    h2.resize(3 * n);
    HIP_TRY(hipMemcpy(h2.data(), c->color2.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  }
  for (size_t i = 0; i < n; ++i)
  {
    float r = hw[i] == 0.f ? 3.402823466e38f : 1.0f / hw[i];  // rcpSafe
    for (int k = 0; k < ch; ++k) dst[i * ch + k] = hs[k * n + i] * r;
    if (var)  // accum_buffer.h:629-631
    {
      float r1 = std::max(hw[i] - 1.f, 1.f);
      for (int k = 0; k < 3; ++k)
        var[i * 3 + k] = std::max(h2[k * n + i] - hsqr(hs[k * n + i]) * r, 0.f) * (1.0f / r1);
    }
  }
  return PRT_OK;
}

// Path-replay adjoint: see include/prt_hip.h.  The primal image is the frame after spp sample
// passes (indices 0 .. spp-1 of the current seed); the loss and its per-pixel adjoint are formed
// on the host in double precision; the replay kernels run on the context stream.
int prt_adjoint(prt_context* c, const float* target, int32_t spp, double* loss, float* grad, int32_t grad_capacity, double* ms)
{
  CTX_CHECK(c);
  if (!target || !grad) return fail(PRT_ERR_INVALID, "target and grad are required");
  if (spp < 1) return fail(PRT_ERR_INVALID, "spp must be >= 1");
  if (c->textured) return fail(PRT_ERR_INVALID, "the adjoint does not differentiate textured materials");
  const int nMat = int(c->materials.size());
  const int nParams = 4 * nMat;
  if (grad_capacity < nParams) return fail(PRT_ERR_INVALID, "grad needs 4 floats per material");
  int rc = prt_frame_clear(c);
  if (rc != PRT_OK) return rc;
  c->sampleIndex = 0;
  rc = prt_render(c, spp, nullptr);
  if (rc != PRT_OK) return rc;
  const size_t n = size_t(c->npix);
  std::vector<float> sum(3 * n), w(n), delta(3 * n);
  HIP_TRY(hipMemcpy(sum.data(), c->color.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(w.data(), c->weight.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  double L = 0.0;
  const double norm = 1.0 / (3.0 * double(n));
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
    {
      double I = double(sum[k * n + i]) / double(w[i]);
      double r = I - double(target[3 * i + k]);
      L += r * r;
      delta[k * n + i] = float(2.0 * r * norm);
    }
  if (loss) *loss = L * norm;
  const int blocks = adjointBlocks(c->npix, c->spillStride);
  DevBuf<float> dDelta, partial, dGrad;
  HIP_TRY(dDelta.alloc(3 * n));
  HIP_TRY(partial.alloc(size_t(blocks) * nParams));
  HIP_TRY(dGrad.alloc(nParams));
  HIP_TRY(hipMemcpy(dDelta.p, delta.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemsetAsync(partial.p, 0, sizeof(float) * partial.n, c->stream));
  HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  for (int s = 0; s < spp; ++s)
    launchAdjoint(c->argsDev.p, blocks, dDelta.p, 1.0f / float(spp), nParams, partial.p, uint32_t(s), nullptr, c->instanced(),
                  c->stream);
  launchAdjointReduce(partial.p, blocks, nParams, dGrad.p, c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  HIP_TRY(hipMemcpyAsync(grad, dGrad.p, sizeof(float) * nParams, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
  if (ms) *ms = t;
  return PRT_OK;
}

int prt_debug_adjoint_replay(prt_context* c, uint32_t sample_index, float* radiance)
{
  CTX_CHECK(c);
  if (!radiance) return fail(PRT_ERR_INVALID, "radiance is null");
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  // the kernel arguments of the last prt_render call are on the device
  if (c->count == 0) return fail(PRT_ERR_STATE, "render at least one pass first");
  const size_t n = size_t(c->npix);
  const int blocks = adjointBlocks(c->npix, c->spillStride);
  DevBuf<float> dDelta, partial, dL;
  HIP_TRY(dDelta.alloc(3 * n));
  HIP_TRY(partial.alloc(size_t(blocks) * 4));
  HIP_TRY(dL.alloc(3 * n));
  HIP_TRY(hipMemsetAsync(dDelta.p, 0, sizeof(float) * 3 * n, c->stream));
  HIP_TRY(hipMemsetAsync(partial.p, 0, sizeof(float) * partial.n, c->stream));
  launchAdjoint(c->argsDev.p, blocks, dDelta.p, 1.f, 0, partial.p, sample_index, dL.p, c->instanced(), c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(radiance, dL.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return PRT_OK;
}

int prt_get_accum(prt_context* c, float* sum, float* weight)
{
  CTX_CHECK(c);
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  if (sum) HIP_TRY(hipMemcpy(sum, c->color.p, sizeof(float) * 3 * c->npix, hipMemcpyDeviceToHost));
  if (weight) HIP_TRY(hipMemcpy(weight, c->weight.p, sizeof(float) * c->npix, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_frame_error(prt_context* c, float* error) { return prt_frame_error_exposure(c, 1.f, error); }

int prt_frame_error_exposure(prt_context* c, float exposure, float* error)
{
  CTX_CHECK(c);
  if (!error) return fail(PRT_ERR_INVALID, "error is null");
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  const size_t n = size_t(c->npix);
  std::vector<float> s(3 * n), w(n), sc(3 * n), wc(n);
  HIP_TRY(hipMemcpy(s.data(), c->color.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(w.data(), c->weight.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(sc.data(), c->colorC.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(wc.data(), c->weightC.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  double err = 0.0;  // accum_buffer.h:752-805
  for (size_t i = 0; i < n; ++i)
  {
    float r = w[i] == 0.f ? 3.402823466e38f : 1.0f / w[i], rc = wc[i] == 0.f ? 3.402823466e38f : 1.0f / wc[i];
    float I[3], B[3];
    for (int k = 0; k < 3; ++k) { I[k] = s[k * n + i] * r; B[k] = sc[k * n + i] * rc; }
    if (exposure != 1.f)  // the tone mapper's exposure (accum_buffer.h:787-792)
      for (int k = 0; k < 3; ++k) { I[k] *= exposure; B[k] *= exposure; }
    float norm = I[0] + I[1] + I[2];
    if (norm < 1.f) norm = std::sqrt(norm);
    err += (std::fabs(I[0] - B[0]) + std::fabs(I[1] - B[1]) + std::fabs(I[2] - B[2])) / (norm + 1e-4f);
  }
  *error = float(err / double(n));
  return PRT_OK;
}

int prt_blit_half(prt_context* c, float* dst)
{
  CTX_CHECK(c);
  if (!dst) return fail(PRT_ERR_INVALID, "dst is null");
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  if (!c->colorB.p) return fail(PRT_ERR_INVALID, "half buffer not enabled for the buffer");
  const size_t n = size_t(c->npix);
  std::vector<float> s(3 * n), w(n);
  HIP_TRY(hipMemcpy(s.data(), c->colorB.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(w.data(), c->weightB.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i)  // accum_buffer.h:633-643
  {
    const float r = w[i] == 0.f ? 3.402823466e38f : 1.0f / w[i];
    for (int k = 0; k < 3; ++k) dst[i * 3 + k] = s[k * n + i] * r;
  }
  return PRT_OK;
}

int prt_set_jitter(prt_context* c, int32_t enable, float jx, float jy)
{
  CTX_CHECK(c);
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  if (enable && !(jx >= 0.f && jx <= 1.f && jy >= 0.f && jy <= 1.f))
    return fail(PRT_ERR_INVALID, "jitter outside [0, 1]");
  c->jitter = enable != 0;
  if (c->jitter) { c->jitterXY[0] = jx; c->jitterXY[1] = jy; }
  return PRT_OK;
}

int prt_get_jitter(prt_context* c, int32_t* has, float* jxy)
{
  CTX_CHECK(c);
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  if (has) *has = c->jitter ? 1 : 0;
  if (jxy && c->jitter) { jxy[0] = c->jitterXY[0]; jxy[1] = c->jitterXY[1]; }
  return PRT_OK;
}

// The filter's lookup table and its Distribution1D (image/pixel_filter.h:30-44 with the
// TriangleFilter, GaussianFilter and BlackmanHarrisFilter of :77-139; distribution1d.cpp:20-35).
int prt_set_pixel_filter(prt_context* c, int32_t type, float width, float sigma, int32_t bias)
{
  CTX_CHECK(c);
  if (!c->npix) return PRT_OK;  // DeviceCpu::initFilter returns without a frame buffer
  if (type == PRT_FILTER_BOX)
  {
    c->filterCdf.release();
    c->filterSize = 0;
    c->filterWidth = 0.f;
    return PRT_OK;
  }
  if (type != PRT_FILTER_TRIANGLE && type != PRT_FILTER_GAUSSIAN && type != PRT_FILTER_BLACKMAN_HARRIS)
    return fail(PRT_ERR_INVALID, "invalid pixel filter type");
  if (!(width > 0.f)) width = type == PRT_FILTER_TRIANGLE ? 2.f : 3.f;  // device_cpu.cpp:247-262 defaults
  if (type == PRT_FILTER_GAUSSIAN && !(sigma > 0.f)) sigma = width / 6.f;
  const float radius = width / 2.f;
  const float alpha = type == PRT_FILTER_GAUSSIAN ? -1.f / (2.f * sigma * sigma) : 0.f;
  const float beta = (type == PRT_FILTER_GAUSSIAN && bias) ? std::exp(alpha * radius * radius) : 0.f;
  auto eval = [&](float x) -> float {
    if (type == PRT_FILTER_TRIANGLE) return std::max(radius - std::fabs(x), 0.f);
    if (type == PRT_FILTER_GAUSSIAN) return std::max(std::exp(alpha * x * x) - beta, 0.f);
    x /= radius;
    if (x < -1.f || x > 1.f) return 0.f;
    const float t = (x + 1.f) * 3.14159265358979323846f;
    return 0.35875f - 0.48829f * std::cos(t) + 0.14128f * std::cos(2.f * t) - 0.01168f * std::cos(3.f * t);
  };
  const int n = 2048, size = 2 * n + 1;
  std::vector<float> cdf(size + 1);
  cdf[0] = 0.f;
  for (int i = 1; i <= size; ++i) cdf[i] = cdf[i - 1] + eval((float(i - 1 - n) / float(n)) * radius);
  const float norm = cdf[size] == 0.f ? 0.f : 1.0f / cdf[size];
  for (int i = 1; i <= size; ++i) cdf[i] *= norm;
  cdf[size] = 1.f;
  HIP_TRY(c->filterCdf.alloc(cdf.size()));
  HIP_TRY(hipMemcpy(c->filterCdf.p, cdf.data(), sizeof(float) * cdf.size(), hipMemcpyHostToDevice));
  c->filterSize = size;
  c->filterWidth = width;
  return PRT_OK;
}

int prt_get_filter_width(prt_context* c, float* width)
{
  CTX_CHECK(c);
  if (!width) return fail(PRT_ERR_INVALID, "width is null");
  *width = c->filterSize ? c->filterWidth : 0.f;  // device_cpu.h:73-78
  return PRT_OK;
}

static InstView instView(prt_context* c) { return InstView{c->tlas.p, c->tlasRoot, c->instDev.p, c->meshDev.p}; }

int prt_query_rays_device(prt_context* c, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (n < 0 || (n > 0 && (!rays || !prim))) return fail(PRT_ERR_INVALID, "invalid query arrays");
  if (n == 0) return PRT_OK;
  BvhView b{c->bvh.nodes, c->bvh.tris, c->bvh.root};
  if (c->instanced())
    launchQueryInst(instView(c), n, rays, prim, t, u, v, nullptr, 0, c->spill.p, c->spillStride, PRT_MAX_STACK - PRT_LDS_STACK,
                    c->launch.gridTrace, c->stream);
  else
    launchQuery(b, n, rays, prim, t, u, v, 0, c->spill.p, c->spillStride, PRT_MAX_STACK - PRT_LDS_STACK, c->launch.gridTrace,
                c->stream);
  HIP_TRY(hipGetLastError());
  return PRT_OK;
}

int prt_query_rays_device_noalloc(prt_context* c, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (n < 0 || (n > 0 && (!rays || !prim))) return fail(PRT_ERR_INVALID, "invalid query arrays");
  // a wide node pushes at most three children: the stack of a tree of L levels holds < 3 L entries
  if (3 * c->bvh.maxDepth >= PRT_LDS_STACK + kScratchOverflow)
    return fail(PRT_ERR_INVALID, "BVH of " + std::to_string(c->bvh.maxDepth) + " levels exceeds the static traversal scratch (at most " +
                                     std::to_string((PRT_LDS_STACK + kScratchOverflow - 1) / 3) + " levels)");
  if (c->instanced()) return fail(PRT_ERR_INVALID, "the allocation-free query takes one-level scenes");
  int* sp = scratchOf(c->device);
  if (!sp) return fail(PRT_ERR_STATE, "static scratch not set up on this device");
  if (n == 0) return PRT_OK;
  BvhView b{c->bvh.nodes, c->bvh.tris, c->bvh.root};
  launchQuery(b, n, rays, prim, t, u, v, 0, sp, kScratchBlocks * kTraceBlock, kScratchOverflow, kScratchBlocks, c->stream);
  HIP_TRY(hipGetLastError());
  return PRT_OK;
}

static int queryHost(prt_context* c, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v, int anyHit,
                     int32_t* inst = nullptr)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (n < 0 || (n > 0 && (!rays || !prim))) return fail(PRT_ERR_INVALID, "invalid query arrays");
  if (n == 0) return PRT_OK;
  DevBuf<float> dr, dt, du, dv;
  DevBuf<int> dp;
  HIP_TRY(dr.alloc(size_t(8) * n));
  HIP_TRY(dp.alloc(n));
  HIP_TRY(dt.alloc(n));
  HIP_TRY(du.alloc(n));
  HIP_TRY(dv.alloc(n));
  HIP_TRY(hipMemcpy(dr.p, rays, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
  BvhView b{c->bvh.nodes, c->bvh.tris, c->bvh.root};
  DevBuf<int> di;
  if (c->instanced())
  {
    HIP_TRY(di.alloc(n));
    launchQueryInst(instView(c), n, dr.p, dp.p, dt.p, du.p, dv.p, di.p, anyHit, c->spill.p, c->spillStride,
                    PRT_MAX_STACK - PRT_LDS_STACK, c->launch.gridTrace, c->stream);
  }
  else
    launchQuery(b, n, dr.p, dp.p, dt.p, du.p, dv.p, anyHit, c->spill.p, c->spillStride, PRT_MAX_STACK - PRT_LDS_STACK,
                c->launch.gridTrace, c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(prim, dp.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (inst)
  {
    if (di.p) HIP_TRY(hipMemcpy(inst, di.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    else
      for (int32_t i = 0; i < n; ++i) inst[i] = -1;
  }
  if (t) HIP_TRY(hipMemcpy(t, dt.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (u) HIP_TRY(hipMemcpy(u, du.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (v) HIP_TRY(hipMemcpy(v, dv.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_query_rays(prt_context* c, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v)
{
  return queryHost(c, n, rays, prim, t, u, v, 0);
}

int prt_query_rays_inst(prt_context* c, int32_t n, const float* rays, int32_t* prim, float* t, float* u, float* v, int32_t* inst)
{
  return queryHost(c, n, rays, prim, t, u, v, 0, inst);
}

int prt_query_occlusion(prt_context* c, int32_t n, const float* rays, int32_t* occluded)
{
  CTX_CHECK(c);
  if (c->anyTransparent && c->bvhReady && n > 0 && rays && occluded)
  {
    // the transparent-shadow filter needs the materials: the renderer's argument block
    RenderArgs a = makeArgs(c);
    DevBuf<RenderArgs> da;
    DevBuf<float> dr;
    DevBuf<int> docc;
    HIP_TRY(da.alloc(1));
    HIP_TRY(dr.alloc(size_t(8) * n));
    HIP_TRY(docc.alloc(n));
    HIP_TRY(hipMemcpy(da.p, &a, sizeof(a), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr.p, rays, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
    launchQueryOcclusionFilter(da.p, n, dr.p, docc.p, c->launch.gridTrace, c->instanced(), c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(occluded, docc.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    return PRT_OK;
  }
  int rc = queryHost(c, n, rays, occluded, nullptr, nullptr, nullptr, 1);
  if (rc != PRT_OK) return rc;
  for (int i = 0; i < n; ++i) occluded[i] = occluded[i] >= 0 ? 1 : 0;
  return PRT_OK;
}

int prt_scene_bounds(prt_context* c, float lo[3], float hi[3])
{
  CTX_CHECK(c);
  if (!lo || !hi) return fail(PRT_ERR_INVALID, "null bounds arrays");
  for (int k = 0; k < 3; ++k) { lo[k] = c->boundsLo[k]; hi[k] = c->boundsHi[k]; }
  return PRT_OK;
}

static int queryHits(prt_context* c, float px, float py, const float* ray8, int32_t max_hits, prt_pixel_hit* hits, int32_t* count)
{
  RenderArgs a = makeArgs(c);
  DevBuf<RenderArgs> da;
  DevBuf<prt_pixel_hit> dh;
  DevBuf<int> dn;
  DevBuf<float> dr;
  HIP_TRY(da.alloc(1));
  HIP_TRY(dh.alloc(max_hits));
  HIP_TRY(dn.alloc(1));
  HIP_TRY(hipMemcpy(da.p, &a, sizeof(a), hipMemcpyHostToDevice));
  if (ray8)
  {
    HIP_TRY(dr.alloc(8));
    HIP_TRY(hipMemcpy(dr.p, ray8, sizeof(float) * 8, hipMemcpyHostToDevice));
  }
  launchQueryPixel(da.p, px, py, max_hits, dh.p, dn.p, dr.p, c->instanced(), c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(count, dn.p, sizeof(int), hipMemcpyDeviceToHost));
  if (*count > 0) HIP_TRY(hipMemcpy(hits, dh.p, sizeof(prt_pixel_hit) * size_t(*count), hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_query_pixel(prt_context* c, float px, float py, int32_t max_hits, prt_pixel_hit* hits, int32_t* count)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (c->W <= 0 || c->H <= 0) return fail(PRT_ERR_STATE, "no frame: the pixel query needs the image size");
  if (max_hits < 1 || max_hits > 32 || !hits || !count) return fail(PRT_ERR_INVALID, "invalid pixel query arguments");
  return queryHits(c, px, py, nullptr, max_hits, hits, count);
}

int prt_query_ray_hits(prt_context* c, const float* ray, int32_t max_hits, prt_pixel_hit* hits, int32_t* count)
{
  CTX_CHECK(c);
  if (!c->bvhReady) return fail(PRT_ERR_STATE, "BVH not built");
  if (!c->mats.p) return fail(PRT_ERR_STATE, "materials not set");
  if (!ray || max_hits < 1 || max_hits > 32 || !hits || !count) return fail(PRT_ERR_INVALID, "invalid ray query arguments");
  return queryHits(c, 0.f, 0.f, ray, max_hits, hits, count);
}

int prt_debug_get(prt_context* c, float* sample_l, uint32_t* path_hash, int32_t* first_prim)
{
  CTX_CHECK(c);
  if (!c->debug || !c->dbgL.p) return fail(PRT_ERR_STATE, "debug outputs not enabled");
  if (sample_l) HIP_TRY(hipMemcpy(sample_l, c->dbgL.p, sizeof(float) * 3 * c->npix, hipMemcpyDeviceToHost));
  if (path_hash) HIP_TRY(hipMemcpy(path_hash, c->dbgHash.p, sizeof(uint32_t) * c->npix, hipMemcpyDeviceToHost));
  if (first_prim) HIP_TRY(hipMemcpy(first_prim, c->dbgFirst.p, sizeof(int) * c->npix, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_debug_sampler(prt_context* c, int32_t n, const uint32_t* in, int32_t ndim, float* o1, float* o2, float* o3)
{
  CTX_CHECK(c);
  if (n <= 0 || ndim <= 0 || !in || !o1 || !o2 || !o3) return fail(PRT_ERR_INVALID, "invalid sampler arrays");
  DevBuf<uint32_t> di;
  DevBuf<float> d1, d2, d3;
  HIP_TRY(di.alloc(size_t(3) * n));
  HIP_TRY(d1.alloc(size_t(n) * ndim));
  HIP_TRY(d2.alloc(size_t(n) * ndim * 2));
  HIP_TRY(d3.alloc(size_t(n) * ndim * 3));
  HIP_TRY(hipMemcpy(di.p, in, sizeof(uint32_t) * 3 * n, hipMemcpyHostToDevice));
  launchDebugSampler(n, di.p, ndim, d1.p, d2.p, d3.p, c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(o1, d1.p, sizeof(float) * n * ndim, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(o2, d2.p, sizeof(float) * n * ndim * 2, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(o3, d3.p, sizeof(float) * n * ndim * 3, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_debug_luts(prt_context* c, float* out, int32_t capacity, int32_t* count)
{
  CTX_CHECK(c);
  if (count) *count = PRT_LUT_FLOATS;
  if (!out) return PRT_OK;
  if (capacity < PRT_LUT_FLOATS) return fail(PRT_ERR_INVALID, "LUT buffer too small");
  std::memcpy(out, c->luts.data(), sizeof(float) * PRT_LUT_FLOATS);
  return PRT_OK;
}

int prt_debug_count_steps(prt_context* c, int32_t enable)
{
  CTX_CHECK(c);
  c->countSteps = enable != 0;
  return PRT_OK;
}

int prt_debug_trace_steps(prt_context* c, double* steps4)
{
  CTX_CHECK(c);
  if (!steps4) return fail(PRT_ERR_INVALID, "steps4 is null");
  for (int k = 0; k < 4; ++k) steps4[k] = c->lastStats[6 + k];
  return PRT_OK;
}

int prt_debug_light_cdf(prt_context* c, float* out, int32_t capacity, int32_t* count)
{
  CTX_CHECK(c);
  if (!c->lightsReady) return fail(PRT_ERR_STATE, "lights not set");
  const int n = int(c->lightCdf.size());
  if (count) *count = n;
  if (!out) return PRT_OK;
  if (capacity < n) return fail(PRT_ERR_INVALID, "light CDF buffer too small");
  std::memcpy(out, c->lightCdf.data(), sizeof(float) * n);
  return PRT_OK;
}

int prt_debug_bvh_stats(prt_context* c, int64_t* s)
{
  CTX_CHECK(c);
  if (!s) return fail(PRT_ERR_INVALID, "stats is null");
  s[0] = c->bvh.nodeCount;
  s[1] = c->bvh.ntris;
  s[2] = c->bvh.maxDepth;
  s[3] = c->ntris;
  return PRT_OK;
}

int prt_debug_kernel_ms(prt_context* c, double* ms4, double* counters6, double* launches2)
{
  CTX_CHECK(c);
  if (ms4) for (int k = 0; k < 4; ++k) ms4[k] = c->kernelMs[k];
  if (counters6) for (int k = 0; k < 6; ++k) counters6[k] = c->lastStats[k];
  if (launches2) { launches2[0] = c->launches[0]; launches2[1] = c->launches[1]; }
  return PRT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Persistent batch renderer (dataset generation): see include/prt_hip.h and batch.hip.
// ---------------------------------------------------------------------------------------
struct prt_batch
{
  int device = 0, cus = 0;
  hipStream_t stream = nullptr;
  struct Frame
  {
    prt_context* scene;
    prt_camera cam;
    int W, H, spp, nSnap;
    uint32_t seed, aov;
    int snap[PRT_BATCH_MAX_SNAPSHOTS];
    uint64_t generation;  // the scene context's generation when the frame's arguments were made
  };
  std::vector<Frame> frames;
  std::vector<size_t> outOffset;  // per frame: first float of its outputs in out
  std::vector<BatchFrameDev> desc;
  DevBuf<float> out;
  DevBuf<RenderArgs> args;
  DevBuf<BatchFrameDev> fdev;
  DevBuf<int> counter, spill;
  DevBuf<unsigned long long> samples;
  int spillStride = 0, nTiles = 0;
  int tilesPerFetch = 1, blocksPerCu = 2;
  bool prepared = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {
constexpr int kBatchMaxBlocksPerCu = 4;

void dropGraph(prt_batch* b)
{
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Outputs, kernel arguments and frame descriptors of every frame on the device.
int sceneReady(const prt_context* scene)
{
  if (!scene->bvhReady || !scene->lightsReady || !scene->mats.p || !scene->rendererReady)
    return fail(PRT_ERR_STATE, "scene context needs geometry, BVH, materials, lights and renderer parameters");
  if (scene->maxMaterialId >= int(scene->materials.size())) return fail(PRT_ERR_INVALID, "triangle material ID beyond the material table");
  return PRT_OK;
}

int batchPrepare(prt_batch* b)
{
  if (b->frames.empty()) return fail(PRT_ERR_STATE, "the batch has no frames");
  for (const auto& f : b->frames)  // a changed scene context must be complete again
  {
    const int rc = sceneReady(f.scene);
    if (rc != PRT_OK) return rc;
  }
  size_t total = 0;
  int tiles = 0;
  b->outOffset.clear();
  for (const auto& f : b->frames)
  {
    const size_t n = size_t(f.W) * f.H;
    b->outOffset.push_back(total);
    total += 3 * n * (1 + f.nSnap + ((f.aov & (PRT_AOV_ALBEDO1 | PRT_AOV_NORMAL1)) ? 2 : 0)) +
             ((f.aov & PRT_AOV_DEPTH) ? n : 0);
    tiles += (f.W / 8) * (f.H / 8);
  }
  HIP_TRY(b->out.alloc(total));
  HIP_TRY(b->args.alloc(b->frames.size()));
  HIP_TRY(b->fdev.alloc(b->frames.size()));
  std::vector<RenderArgs> args(b->frames.size());
  b->desc.assign(b->frames.size(), BatchFrameDev{});
  int tileBegin = 0;
  for (size_t i = 0; i < b->frames.size(); ++i)
  {
    auto& f = b->frames[i];
    RenderArgs a = makeArgs(f.scene);
    f.generation = f.scene->generation;
    setCameraArgs(a, f.cam, f.W, f.H);
    a.seed = f.seed;
    std::memset(&a.fb, 0, sizeof(a.fb));
    // the jitter and pixel filter belong to the scene context's own frame, not to batch frames
    a.jitter = 0;
    a.filterCdf = nullptr;
    a.filterSize = 0;
    args[i] = a;
    BatchFrameDev& d = b->desc[i];
    const size_t n = size_t(f.W) * f.H;
    float* base = b->out.p + b->outOffset[i];
    d.args = b->args.p + i;
    d.color = base;
    d.snap = base + 3 * n;
    const bool aovs = (f.aov & (PRT_AOV_ALBEDO1 | PRT_AOV_NORMAL1)) != 0;
    d.albedo1 = aovs ? base + 3 * n * (1 + f.nSnap) : nullptr;
    d.normal1 = aovs ? base + 3 * n * (2 + f.nSnap) : nullptr;
    d.depth = (f.aov & PRT_AOV_DEPTH) ? base + 3 * n * (1 + f.nSnap + (aovs ? 2 : 0)) : nullptr;
    d.W = f.W; d.H = f.H; d.npix = int(n); d.spp = f.spp; d.nSnap = f.nSnap;
    for (int k = 0; k < PRT_BATCH_MAX_SNAPSHOTS; ++k) d.snapSpp[k] = k < f.nSnap ? f.snap[k] : 0;
    d.tileBegin = tileBegin;
    d.tilesX = f.W / 8;
    tileBegin += (f.W / 8) * (f.H / 8);
  }
  b->nTiles = tiles;
  HIP_TRY(hipMemcpy(b->args.p, args.data(), sizeof(RenderArgs) * args.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->fdev.p, b->desc.data(), sizeof(BatchFrameDev) * b->desc.size(), hipMemcpyHostToDevice));
  b->spillStride = kBatchMaxBlocksPerCu * b->cus * kBatchBlockSize;
  HIP_TRY(b->spill.alloc(2 * size_t(PRT_MAX_STACK - PRT_LDS_STACK) * b->spillStride));
  HIP_TRY(b->counter.alloc(1));
  HIP_TRY(b->samples.alloc(1));
  b->prepared = true;
  dropGraph(b);
  return PRT_OK;
}

// A frame's scene context changed since its kernel arguments were made (its device buffers may
// have been reallocated): the arguments and the captured graph are made again.
bool batchStale(const prt_batch* b)
{
  for (const auto& f : b->frames)
    if (f.generation != f.scene->generation) return true;
  return false;
}

void batchLaunch(prt_batch* b)
{
  bool std = false, inst = false;
  for (const auto& f : b->frames)
  {
    std = std || f.scene->stdMaterials != 0;
    inst = inst || f.scene->instanced();
  }
  launchBatch(b->fdev.p, int(b->frames.size()), b->nTiles, b->counter.p, b->tilesPerFetch, b->blocksPerCu * b->cus, b->spill.p,
              b->spillStride, b->samples.p, std, inst, b->stream);
}
}  // namespace

extern "C" {

int prt_batch_create(int32_t device_ordinal, prt_batch** out)
{
  if (!out) return fail(PRT_ERR_INVALID, "out is null");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device_ordinal < 0 || device_ordinal >= n) return fail(PRT_ERR_INVALID, "invalid device ordinal");
  HIP_TRY(hipSetDevice(device_ordinal));
  prt_batch* b = new prt_batch();
  b->device = device_ordinal;
  hipDeviceProp_t props{};
  hipError_t e = hipGetDeviceProperties(&props, device_ordinal);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&b->ev[1]);
  if (e != hipSuccess) { prt_batch_destroy(b); return fail(PRT_ERR_HIP, hipGetErrorString(e)); }
  b->cus = props.multiProcessorCount;
  *out = b;
  return PRT_OK;
}

int prt_batch_destroy(prt_batch* b)
{
  if (!b) return PRT_OK;
  hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  dropGraph(b);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PRT_OK;
}

int prt_batch_add_frame(prt_batch* b, prt_context* scene, const prt_camera* cam, int32_t width, int32_t height, uint32_t seed,
                        int32_t spp, const int32_t* snapshot_spp, int32_t n_snapshots, uint32_t aov_mask, int32_t* frame_id)
{
  if (!b || !scene || !cam) return fail(PRT_ERR_INVALID, "null argument");
  if (scene->device != b->device) return fail(PRT_ERR_INVALID, "scene context on another device");
  const int rcScene = sceneReady(scene);
  if (rcScene != PRT_OK) return rcScene;
  if (width <= 0 || height <= 0 || width % 8 != 0 || height % 8 != 0 || width > 4096 || height > 4096)
    return fail(PRT_ERR_INVALID, "batch frame size must be a multiple of 8 and at most 4096");
  if (!(cam->fov > 0.f && cam->fov < 3.14159265f)) return fail(PRT_ERR_INVALID, "invalid fov");
  if (spp < 1) return fail(PRT_ERR_INVALID, "spp must be >= 1");
  if (n_snapshots < 0 || n_snapshots > PRT_BATCH_MAX_SNAPSHOTS || (n_snapshots > 0 && !snapshot_spp))
    return fail(PRT_ERR_INVALID, "0..4 snapshots");
  prt_batch::Frame f{};
  f.scene = scene;
  f.cam = *cam;
  f.W = width; f.H = height; f.spp = spp; f.nSnap = n_snapshots;
  f.seed = seed;
  f.aov = aov_mask;
  for (int k = 0; k < n_snapshots; ++k)
  {
    if (snapshot_spp[k] < 1 || snapshot_spp[k] > spp || (k > 0 && snapshot_spp[k] <= snapshot_spp[k - 1]))
      return fail(PRT_ERR_INVALID, "snapshot sample counts must increase within 1..spp");
    f.snap[k] = snapshot_spp[k];
  }
  b->frames.push_back(f);
  b->prepared = false;
  if (frame_id) *frame_id = int(b->frames.size()) - 1;
  return PRT_OK;
}

int prt_batch_tune(prt_batch* b, int32_t calibration_spp, int32_t* tiles_per_fetch, int32_t* blocks_per_cu, double* best_ms)
{
  if (!b) return fail(PRT_ERR_INVALID, "batch is null");
  if (calibration_spp < 1) return fail(PRT_ERR_INVALID, "calibration_spp must be >= 1");
  HIP_TRY(hipSetDevice(b->device));
  if (!b->prepared || batchStale(b))
  {
    int rc = batchPrepare(b);
    if (rc != PRT_OK) return rc;
  }
  // the frames' sample counts are lowered to calibration_spp while timing each candidate
  std::vector<BatchFrameDev> cal = b->desc;
  for (auto& d : cal) { d.spp = std::min(d.spp, int(calibration_spp)); d.nSnap = 0; }
  HIP_TRY(hipMemcpy(b->fdev.p, cal.data(), sizeof(BatchFrameDev) * cal.size(), hipMemcpyHostToDevice));
  const int fetches[4] = {1, 2, 4, 8};
  float best = 3.4e38f;
  int bestF = b->tilesPerFetch, bestB = b->blocksPerCu;
  for (int bp = 1; bp <= kBatchMaxBlocksPerCu; ++bp)
    for (int fi = 0; fi < 4; ++fi)
    {
      b->blocksPerCu = bp;
      b->tilesPerFetch = fetches[fi];
      batchLaunch(b);  // warm-up of this shape
      HIP_TRY(hipEventRecord(b->ev[0], b->stream));
      batchLaunch(b);
      HIP_TRY(hipEventRecord(b->ev[1], b->stream));
      HIP_TRY(hipEventSynchronize(b->ev[1]));
      HIP_TRY(hipGetLastError());
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
      if (ms < best) { best = ms; bestF = fetches[fi]; bestB = bp; }
    }
  b->tilesPerFetch = bestF;
  b->blocksPerCu = bestB;
  HIP_TRY(hipMemcpy(b->fdev.p, b->desc.data(), sizeof(BatchFrameDev) * b->desc.size(), hipMemcpyHostToDevice));
  dropGraph(b);
  if (tiles_per_fetch) *tiles_per_fetch = bestF;
  if (blocks_per_cu) *blocks_per_cu = bestB;
  if (best_ms) *best_ms = best;
  return PRT_OK;
}

int prt_batch_render(prt_batch* b, double* ms, double* samples)
{
  if (!b) return fail(PRT_ERR_INVALID, "batch is null");
  HIP_TRY(hipSetDevice(b->device));
  if (!b->prepared || batchStale(b))
  {
    int rc = batchPrepare(b);
    if (rc != PRT_OK) return rc;
  }
  if (!b->exec)
  {
    // capture the reset and the persistent launch once; later calls replay the graph
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    batchLaunch(b);
    hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
    if (e != hipSuccess) return fail(PRT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  unsigned long long n = 0;
  HIP_TRY(hipMemcpyAsync(&n, b->samples.p, sizeof(n), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
  if (ms) *ms = t;
  if (samples) *samples = double(n);
  return PRT_OK;
}

int prt_batch_get(prt_batch* b, int32_t frame, int32_t buffer, float* dst)
{
  if (!b || !dst) return fail(PRT_ERR_INVALID, "null argument");
  if (frame < 0 || frame >= int(b->frames.size())) return fail(PRT_ERR_INVALID, "invalid frame");
  if (!b->prepared) return fail(PRT_ERR_STATE, "batch not rendered");
  HIP_TRY(hipSetDevice(b->device));
  const BatchFrameDev& d = b->desc[frame];
  const float* src = nullptr;
  if (buffer == PRT_BATCH_COLOR) src = d.color;
  else if (buffer >= PRT_BATCH_SNAPSHOT && buffer < PRT_BATCH_SNAPSHOT + d.nSnap) src = d.snap + size_t(buffer - PRT_BATCH_SNAPSHOT) * 3 * d.npix;
  else if (buffer == PRT_BATCH_ALBEDO1) src = d.albedo1;
  else if (buffer == PRT_BATCH_NORMAL1) src = d.normal1;
  else if (buffer == PRT_BATCH_DEPTH)
  {
    if (!d.depth) return fail(PRT_ERR_INVALID, "buffer not produced for this frame");
    HIP_TRY(hipMemcpy(dst, d.depth, sizeof(float) * size_t(d.npix), hipMemcpyDeviceToHost));
    return PRT_OK;
  }
  if (!src) return fail(PRT_ERR_INVALID, "buffer not produced for this frame");
  const size_t n = size_t(d.npix);
  std::vector<float> h(3 * n);
  HIP_TRY(hipMemcpy(h.data(), src, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) dst[3 * i + k] = h[k * n + i];
  return PRT_OK;
}

}  // extern "C"

namespace prt {

int contextFrameImages(prt_context* c, FrameImages* out)
{
  CTX_CHECK(c);
  if (!c->npix) return fail(PRT_ERR_STATE, "frame not initialized");
  FrameImages f;
  f.device = c->device;
  f.W = c->W;
  f.H = c->H;
  f.color = c->color.p;
  f.colorW = c->weight.p;
  f.albedo = c->albedo1.p;
  f.albedoW = c->albedo1W.p;
  f.normal = c->normal1.p;
  f.normalW = c->normal1W.p;
  f.stream = c->stream;
  *out = f;
  return PRT_OK;
}

int batchFrameImages(prt_batch* b, int frame, int buffer, FrameImages* out)
{
  if (!b || !out) return fail(PRT_ERR_INVALID, "null argument");
  if (frame < 0 || frame >= int(b->frames.size())) return fail(PRT_ERR_INVALID, "invalid frame");
  if (!b->prepared) return fail(PRT_ERR_STATE, "batch not rendered");
  const BatchFrameDev& d = b->desc[frame];
  FrameImages f;
  f.device = b->device;
  f.W = d.W;
  f.H = d.H;
  if (buffer == PRT_BATCH_COLOR) f.color = d.color;
  else if (buffer >= PRT_BATCH_SNAPSHOT && buffer < PRT_BATCH_SNAPSHOT + d.nSnap)
    f.color = d.snap + size_t(buffer - PRT_BATCH_SNAPSHOT) * 3 * d.npix;
  else return fail(PRT_ERR_INVALID, "buffer must be the frame's colour or one of its snapshots");
  f.albedo = d.albedo1;
  f.normal = d.normal1;
  f.stream = b->stream;
  *out = f;
  return PRT_OK;
}

}  // namespace prt
