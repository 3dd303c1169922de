// Golden-vector harness: TEST INFRASTRUCTURE ONLY.
//
// Compiled in the build container against the reference's own sources (include path
// /root/reference, see oracle/Makefile.ref) and run once by tests/golden/make_golden.py to
// write the fixtures under tests/golden/.  Nothing in the product links or loads it.
//
// Every block calls the reference's SIMD (stream-path) code on inputs generated here and
// dumps inputs + outputs as raw little-endian arrays plus a manifest.  The reference code
// exercised is, per block:
//   sampler   sampling/sobol_burley_sampler_simd.h:36-85, sampling/sobol_burley.h:150-297
//   morton    image/morton.h:32-39 (+ image/morton.cpp tables)
//   camera    camera/pinhole_camera.h:72-133, camera/thin_lens_camera.h:155-182
//   warps     sampling/shape_sampler.h (cosine hemisphere, sphere, cone, disk)
//   dist      sampling/distribution1d.{h,cpp}, sampling/distribution2d.{h,cpp}
//   luts      bsdf/albedo_tables.cpp:173-186
//   bsdf      bsdf/diffuse.h:45-89, bsdf/rough_conductor.h:106-228,
//             bsdf/rough_dielectric.h:265-548, bsdf/multi_bsdf.h:151-260, bsdf/optics.h
//   coating   bsdf/rough_coating.h:270-593 over bsdf/diffuse.h:45-89, in a MultiBsdfSimd
//   sheen     bsdf/sheen.h:199-451 + bsdf/sheen_distribution.h over each lobe kind
//   lobes     bsdf/rough_diffuse.h:55-106, bsdf/thin_rough_dielectric.h:170-345,
//             bsdf/transparent.h:26-71, bsdf/multi_bsdf.h:151-459, composed as
//             material/standard.cpp:409-627 composes them (opacity mix, coat over rough diffuse)
//   standard  material/standard.cpp:409-627 restated for uniform parameters over the classes
//             above: weighted opaque / translucent / metal bases, a coat over any base, sheen,
//             opacity (bsdf/rough_coating.h with a generic base, bsdf/multi_bsdf.h)
//   postisect geometry/triangle_mesh.cpp:152-257, core/ray_simd.h:63-68,
//             core/shading_context.h:144-191 (makeFrame)
//   texcoord  geometry/triangle_mesh.cpp:191-200, :234-237 (UV set 0 at a hit, with and
//             without texcoords)
//   pixel     image/pixel.h:80-91, :137-145 (8-bit texel decode: the sRGB table of
//             image/pixel.cpp:21 and the linear /255)

#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sampling/sobol_burley_sampler_simd.h"
#include "image/morton.h"
#include "camera/thin_lens_camera.h"
#include "sampling/shape_sampler.h"
#include "sampling/distribution1d.h"
#include "sampling/distribution2d.h"
#include "bsdf/albedo_tables.h"
#include "bsdf/diffuse.h"
#include "bsdf/rough_conductor.h"
#include "bsdf/rough_dielectric.h"
#include "bsdf/multi_bsdf.h"
#include "bsdf/rough_coating.h"
#include "bsdf/rough_diffuse.h"
#include "bsdf/thin_rough_dielectric.h"
#include "bsdf/transparent.h"
#include "bsdf/sheen.h"
#include "geometry/triangle_mesh.h"
#include "image/pixel.h"

using namespace prt;

static std::string gOutDir;
static FILE* gManifest = nullptr;
static bool gFirst = true;

template <class T>
static void dump(const char* name, const char* dtype, const std::vector<T>& v, std::vector<int> shape)
{
  std::string path = gOutDir + "/" + name + ".bin";
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(1); }
  fwrite(v.data(), sizeof(T), v.size(), f);
  fclose(f);
  fprintf(gManifest, "%s\n  \"%s\": {\"dtype\": \"%s\", \"shape\": [", gFirst ? "" : ",", name, dtype);
  for (size_t i = 0; i < shape.size(); ++i) fprintf(gManifest, "%s%d", i ? ", " : "", shape[i]);
  fprintf(gManifest, "]}");
  gFirst = false;
}

// Harness-local input generator (not reference code): 32-bit xorshift, floats in [0,1).
struct InputRng
{
  uint32_t s;
  explicit InputRng(uint32_t seed) : s(seed ? seed : 1u) {}
  uint32_t u() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; }
  float f() { return float(u() >> 8) * (1.0f / 16777216.0f); }
  Vec3f dir()
  {
    for (;;)
    {
      Vec3f d(2.f*f()-1.f, 2.f*f()-1.f, 2.f*f()-1.f);
      float l2 = dot(d, d);
      if (l2 > 1e-4f && l2 <= 1.f) return d * (1.0f / std::sqrt(l2));
    }
  }
};

static void goldenSampler()
{
  const uint32_t seeds[] = {0u, 1u, 0x9e3779b9u};
  const uint32_t pixelIds[] = {0u, 1u, 511u | (511u << 16), 1919u | (1079u << 16),
                               37u | (200u << 16), 4095u | (4095u << 16), 12u | (3u << 16), 800u | (9u << 16)};
  const uint32_t indices[] = {0u, 1u, 2u, 3u, 7u, 15u, 63u, 64u, 1000u, 65535u, 1u << 20, 0x7fffffffu};
  const int nDim = 64;
  std::vector<uint32_t> in;   // seed, pixelId, index
  std::vector<float> out1, out2, out3;
  for (uint32_t seed : seeds)
    for (uint32_t pid : pixelIds)
      for (uint32_t idx : indices)
      {
        in.push_back(seed); in.push_back(pid); in.push_back(idx);
        SobolBurleySamplerSimd<false> sampler;
        sampler.init(4 + 7 * 100, 0, seed);
        SobolBurleySamplerSimd<false>::State st;
        sampler.setSample(vbool(one), st, vuint(idx), vuint(pid));
        for (int d = 0; d < nDim; ++d)
        {
          vfloat a = sampler.get1D(vbool(one), st, d);
          Vec2vf b = sampler.get2D(vbool(one), st, d);
          Vec3vf c = sampler.get3D(vbool(one), st, d);
          out1.push_back(a[0]);
          out2.push_back(b.x[0]); out2.push_back(b.y[0]);
          out3.push_back(c.x[0]); out3.push_back(c.y[0]); out3.push_back(c.z[0]);
        }
      }
  int n = (int)(in.size() / 3);
  dump("sampler_in", "u4", in, {n, 3});
  dump("sampler_1d", "f4", out1, {n, nDim});
  dump("sampler_2d", "f4", out2, {n, nDim, 2});
  dump("sampler_3d", "f4", out3, {n, nDim, 3});
}

static void goldenMorton()
{
  std::vector<int32_t> out;
  for (int taskId = 0; taskId < 1024; taskId += simdSize)
  {
    Vec2vi p = getMorton8x8Step<32>(taskId);
    for (int k = 0; k < simdSize; ++k) { out.push_back(p.x[k]); out.push_back(p.y[k]); }
  }
  dump("morton32", "i4", out, {1024, 2});
}

static void goldenCamera()
{
  // in: [width, height, fov, nearClip, lensRadius, focalDistance, posx, posy, posz]
  struct Cfg { int w, h; float fov, nearClip, lensRadius, focal; Vec3f pos; };
  const Cfg cfgs[] = {
    {512, 512, 45.f * float(pi) / 180.f, 0.01f, 0.f, 1.f, Vec3f(0.f)},
    {1920, 1080, 45.f * float(pi) / 180.f, 0.01f, 0.05f, 15.f, Vec3f(0.f)},
    {640, 360, 30.f * float(pi) / 180.f, 0.05f, 0.2f, 4.f, Vec3f(1.f, 2.f, -3.f)},
  };
  const Basis3f basis(Vec3f(-1.f, 0.f, 0.f), Vec3f(0.f, 1.f, 0.f), Vec3f(0.f, 0.f, -1.f));
  InputRng rng(1234);
  const int nRay = 256;
  std::vector<float> cfgOut, sIn, rayOut;
  for (const Cfg& c : cfgs)
  {
    cfgOut.insert(cfgOut.end(), {float(c.w), float(c.h), c.fov, c.nearClip, c.lensRadius, c.focal, c.pos.x, c.pos.y, c.pos.z});
    Props props;
    props.set("position", c.pos);
    props.set("basis", basis);
    props.set("nearClip", c.nearClip);
    props.set("fov", c.fov);
    props.set("width", c.w);
    props.set("height", c.h);
    props.set("lensRadius", c.lensRadius);
    props.set("focalDistance", c.focal);
    PinholeCamera pin(props);
    ThinLensCamera thin(props);
    for (int i = 0; i < nRay; i += simdSize)
    {
      CameraSampleSimd cs;
      for (int k = 0; k < simdSize; ++k)
      {
        cs.image.x[k] = rng.f(); cs.image.y[k] = rng.f();
        cs.lens.x[k] = rng.f(); cs.lens.y[k] = rng.f();
        sIn.insert(sIn.end(), {cs.image.x[k], cs.image.y[k], cs.lens.x[k], cs.lens.y[k]});
      }
      RaySimd r0, r1;
      pin.getRay(r0, cs);
      thin.getRay(r1, cs);
      for (int k = 0; k < simdSize; ++k)
      {
        rayOut.insert(rayOut.end(), {r0.org.x[k], r0.org.y[k], r0.org.z[k], r0.dir.x[k], r0.dir.y[k], r0.dir.z[k], r0.near[k],
                                     r1.org.x[k], r1.org.y[k], r1.org.z[k], r1.dir.x[k], r1.dir.y[k], r1.dir.z[k], r1.near[k]});
      }
    }
  }
  int nc = 3;
  dump("camera_cfg", "f4", cfgOut, {nc, 9});
  dump("camera_s", "f4", sIn, {nc, nRay, 4});
  dump("camera_rays", "f4", rayOut, {nc, nRay, 2, 7});
}

static void goldenWarps()
{
  InputRng rng(777);
  const int n = 512;
  std::vector<float> sIn, outCos, outSph, outCone, outDisk;
  const float angles[2] = {0.1f, 0.00445f};
  for (int i = 0; i < n; i += simdSize)
  {
    Vec2vf s;
    for (int k = 0; k < simdSize; ++k) { s.x[k] = rng.f(); s.y[k] = rng.f(); sIn.push_back(s.x[k]); sIn.push_back(s.y[k]); }
    vfloat pc, ps;
    Vec3vf c = cosineSampleHemisphere(pc, s);
    Vec3vf sp = uniformSampleSphere(ps, s);
    Vec2vf d = uniformSampleDisk(s);
    for (int k = 0; k < simdSize; ++k)
    {
      outCos.insert(outCos.end(), {c.x[k], c.y[k], c.z[k], pc[k]});
      outSph.insert(outSph.end(), {sp.x[k], sp.y[k], sp.z[k], ps[k]});
      outDisk.insert(outDisk.end(), {d.x[k], d.y[k]});
    }
    for (float a : angles)
    {
      vfloat pa;
      Vec3vf cn = uniformSampleCone(pa, s, vfloat(a));
      for (int k = 0; k < simdSize; ++k) outCone.insert(outCone.end(), {cn.x[k], cn.y[k], cn.z[k], pa[k]});
    }
  }
  dump("warp_s", "f4", sIn, {n, 2});
  dump("warp_cosine", "f4", outCos, {n, 4});
  dump("warp_sphere", "f4", outSph, {n, 4});
  dump("warp_disk", "f4", outDisk, {n, 2});
  dump("warp_cone", "f4", outCone, {n / simdSize, 2, simdSize, 4});
}

static void goldenDistributions()
{
  // 1D: light-power style weights, including a zero bucket
  const float w1[5] = {0.2f, 5.0f, 0.0f, 1.3f, 0.01f};
  Distribution1D d1(w1, 5);
  InputRng rng(99);
  const int n = 256;
  std::vector<float> s1In, s1Out;
  std::vector<int32_t> i1Out;
  for (int i = 0; i < n; i += simdSize)
  {
    vfloat s;
    for (int k = 0; k < simdSize; ++k) { s[k] = rng.f(); s1In.push_back(s[k]); }
    vfloat pdf;
    vfloat sr = s;
    vint idx = d1.sampleReuse(vbool(one), pdf, sr);
    for (int k = 0; k < simdSize; ++k) { i1Out.push_back(idx[k]); s1Out.push_back(pdf[k]); s1Out.push_back(sr[k]); }
  }
  std::vector<float> p1;
  for (int i = -1; i <= 5; ++i) p1.push_back(d1.pdf(i));
  dump("dist1_w", "f4", std::vector<float>(w1, w1 + 5), {5});
  dump("dist1_s", "f4", s1In, {n});
  dump("dist1_idx", "i4", i1Out, {n});
  dump("dist1_pdf_sr", "f4", s1Out, {n, 2});
  dump("dist1_pdf_i", "f4", p1, {7});

  // 2D: 16x8 grid with a zero row and zero cells
  const int W = 16, H = 8;
  std::vector<float> w2(W * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      w2[y * W + x] = (y == 3) ? 0.f : ((x == 5) ? 0.f : (0.1f + rng.f() * 4.f * float(1 + (x * y) % 3)));
  Distribution2D d2(w2.data(), Vec2i(W, H));
  std::vector<float> s2In, s2Out;
  std::vector<int32_t> i2Out;
  for (int i = 0; i < n; i += simdSize)
  {
    Vec2vf s;
    for (int k = 0; k < simdSize; ++k) { s.x[k] = rng.f(); s.y[k] = rng.f(); s2In.push_back(s.x[k]); s2In.push_back(s.y[k]); }
    vfloat pdf;
    Vec2vf sr = s;
    Vec2vi idx = d2.sampleReuse(vbool(one), pdf, sr);
    for (int k = 0; k < simdSize; ++k)
    {
      i2Out.push_back(idx.x[k]); i2Out.push_back(idx.y[k]);
      s2Out.insert(s2Out.end(), {pdf[k], sr.x[k], sr.y[k]});
    }
  }
  std::vector<float> p2;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      p2.push_back(d2.pdf(Vec2i(x, y)));
  dump("dist2_w", "f4", w2, {H, W});
  dump("dist2_s", "f4", s2In, {n, 2});
  dump("dist2_idx", "i4", i2Out, {n, 2});
  dump("dist2_pdf_sr", "f4", s2Out, {n, 3});
  dump("dist2_pdf_cells", "f4", p2, {H, W});
}

static void goldenLuts()
{
  auto v = [](const float* p, int n) { return std::vector<float>(p, p + n); };
  dump("lut_ggx_dir", "f4", v(MicrofacetAlbedoTable::dir, 32 * 32), {32, 32});
  dump("lut_ggx_avg", "f4", v(MicrofacetAlbedoTable::avg, 32), {32});
  dump("lut_diel_eta_dir", "f4", v(MicrofacetDielectricAlbedoTable::etaDir, 16 * 16 * 16), {16, 16, 16});
  dump("lut_diel_eta_avg", "f4", v(MicrofacetDielectricAlbedoTable::etaAvg, 16 * 16), {16, 16});
  dump("lut_diel_rcpeta_dir", "f4", v(MicrofacetDielectricAlbedoTable::rcpEtaDir, 16 * 16 * 16), {16, 16, 16});
  dump("lut_diel_rcpeta_avg", "f4", v(MicrofacetDielectricAlbedoTable::rcpEtaAvg, 16 * 16), {16, 16});
  dump("lut_diel_refl_dir", "f4", v(MicrofacetDielectricReflectionAlbedoTable::etaDir, 16 * 16 * 16), {16, 16, 16});
  dump("lut_diel_refl_avg", "f4", v(MicrofacetDielectricReflectionAlbedoTable::etaAvg, 16 * 16), {16, 16});
  dump("lut_sheen_dir", "f4", v(MicrofacetSheenAlbedoTable::dir, 16 * 16), {16, 16});
}

// BSDF kinds: 0 diffuse, 1 rough conductor (artistic Fresnel), 2 rough dielectric.  Each is
// wrapped in a MultiBsdfSimd with weight 1, as StandardMaterial builds it (material/standard.cpp:428-573).
static void goldenBsdf()
{
  InputRng rng(4242);
  const int nPerKind = 768;
  // params: kind, r,g,b (albedo / reflectivity), er,eg,eb (edge tint), roughness, eta
  std::vector<float> params, geo, samp, outSample, outEval;
  std::vector<int32_t> outType;
  int n = 0;
  for (int kind = 0; kind < 3; ++kind)
  {
    for (int i = 0; i < nPerKind; i += simdSize)
    {
      // per-batch material (uniform across lanes, as in a material bucket)
      Vec3f col(0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f());
      Vec3f edge(rng.f(), rng.f(), rng.f());
      float roughness = (i % 64 == 0) ? 0.0f : std::sqrt(0.01f + 0.5f * rng.f());
      float eta = (kind == 2 && (i / simdSize) % 2) ? 1.5f : 1.f / 1.5f;

      Basis3vf frame;
      ShadingContextSimd* ctx = new ShadingContextSimd();
      Vec3vf wo, wi;
      Vec3vf s3;
      for (int k = 0; k < simdSize; ++k)
      {
        Vec3f N = rng.dir();
        Vec3f U, V;
        makeFrame(U, V, N);
        Vec3f Ng = N;
        if (rng.f() < 0.3f) { Ng = normalize(N + 0.2f * rng.dir()); if (dot(Ng, N) < 0.2f) Ng = N; }
        Vec3f o = rng.dir();
        if (rng.f() < 0.85f && dot(o, Ng) < 0.f) o = -o;
        Vec3f w = rng.dir();
        if (kind != 2 && rng.f() < 0.7f && dot(w, N) < 0.f) w = -w;
        frame.U.x[k] = U.x; frame.U.y[k] = U.y; frame.U.z[k] = U.z;
        frame.V.x[k] = V.x; frame.V.y[k] = V.y; frame.V.z[k] = V.z;
        frame.N.x[k] = N.x; frame.N.y[k] = N.y; frame.N.z[k] = N.z;
        ctx->Ng.x[k] = Ng.x; ctx->Ng.y[k] = Ng.y; ctx->Ng.z[k] = Ng.z;
        wo.x[k] = o.x; wo.y[k] = o.y; wo.z[k] = o.z;
        wi.x[k] = w.x; wi.y[k] = w.y; wi.z[k] = w.z;
        s3.x[k] = rng.f(); s3.y[k] = rng.f(); s3.z[k] = rng.f();
        params.insert(params.end(), {float(kind), col.x, col.y, col.z, edge.x, edge.y, edge.z, roughness, eta});
        geo.insert(geo.end(), {U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z, Ng.x, Ng.y, Ng.z, o.x, o.y, o.z, w.x, w.y, w.z});
        samp.insert(samp.end(), {s3.x[k], s3.y[k], s3.z[k]});
      }
      vbool m = one;
      BsdfSimd* child = nullptr;
      DiffuseBrdfSimd* diffuse = nullptr;
      RoughConductorBrdfSimd<FresnelConductorArtistic<vfloat>>* cond = nullptr;
      RoughDielectricBsdfSimd* diel = nullptr;
      if (kind == 0)
        child = diffuse = new DiffuseBrdfSimd(m, &frame, Vec3vf(col));
      else if (kind == 1)
        child = cond = new RoughConductorBrdfSimd<FresnelConductorArtistic<vfloat>>(m, &frame, FresnelConductorArtistic<vfloat>(Vec3vf(col), Vec3vf(edge)), vfloat(roughness), vfloat(0.f));
      else
        child = diel = new RoughDielectricBsdfSimd(m, &frame, vfloat(eta), vfloat(roughness), vfloat(0.f));
      MultiBsdfSimd multi;
      multi.add(m, child, vfloat(1.f));

      Vec3vf swi = zero;
      vfloat spdf = zero;
      vint stype = zero;
      Vec3vf sw = multi.sample(m, *ctx, wo, swi, spdf, stype, BsdfSampleSimd(s3));
      vfloat epdf;
      Vec3vf ev = multi.eval(m, *ctx, wo, wi, epdf);
      for (int k = 0; k < simdSize; ++k)
      {
        outSample.insert(outSample.end(), {sw.x[k], sw.y[k], sw.z[k], swi.x[k], swi.y[k], swi.z[k], spdf[k]});
        outType.push_back(stype[k]);
        outEval.insert(outEval.end(), {ev.x[k], ev.y[k], ev.z[k], epdf[k]});
      }
      delete child;
      delete ctx;
      n += simdSize;
    }
  }
  dump("bsdf_params", "f4", params, {n, 9});
  dump("bsdf_geo", "f4", geo, {n, 18});
  dump("bsdf_s", "f4", samp, {n, 3});
  dump("bsdf_sample", "f4", outSample, {n, 7});
  dump("bsdf_sample_type", "i4", outType, {n});
  dump("bsdf_eval", "f4", outEval, {n, 4});
}

// Coated diffuse: RoughCoatingBsdfSimd over a DiffuseBrdfSimd, wrapped in a MultiBsdfSimd with
// weight 1, as StandardMaterial builds its specular layer over the opaque base
// (material/standard.cpp:512-518) and its coat layer (:580-597).  The eta handed to the coating
// is ambientIor / coatIor (fromOutside), clamped to [1/3, 3]; color, thickness and weight vary.
// params: r,g,b (diffuse), cr,cg,cb (coat color), roughness, eta, thickness, weight
static void goldenCoating()
{
  InputRng rng(5151);
  const int nTotal = 1024;
  std::vector<float> params, geo, samp, outSample, outEval;
  std::vector<int32_t> outType;
  int n = 0;
  for (int i = 0; i < nTotal; i += simdSize)
  {
    const int batch = i / simdSize;
    Vec3f col(0.02f + 0.96f * rng.f(), 0.02f + 0.96f * rng.f(), 0.02f + 0.96f * rng.f());
    Vec3f coatCol = (batch % 3 == 0) ? Vec3f(1.f) : Vec3f(0.2f + 0.8f * rng.f(), 0.2f + 0.8f * rng.f(), 0.2f + 0.8f * rng.f());
    float thickness = (batch % 3 == 0) ? 1.f : 2.f * rng.f();
    float weight = (batch % 4 == 0) ? 1.f : 0.2f + 0.8f * rng.f();
    float roughness = std::sqrt(0.005f + 0.6f * rng.f());
    const float iors[4] = {1.5f, 1.3f, 2.0f, 2.8f};
    float eta = 1.f / iors[batch % 4];

    Basis3vf frame;
    ShadingContextSimd* ctx = new ShadingContextSimd();
    Vec3vf wo, wi, s3;
    for (int k = 0; k < simdSize; ++k)
    {
      Vec3f N = rng.dir();
      Vec3f U, V;
      makeFrame(U, V, N);
      Vec3f Ng = N;
      if (rng.f() < 0.3f) { Ng = normalize(N + 0.2f * rng.dir()); if (dot(Ng, N) < 0.2f) Ng = N; }
      Vec3f o = rng.dir();
      if (rng.f() < 0.9f && dot(o, N) < 0.f) o = -o;
      Vec3f w = rng.dir();
      if (rng.f() < 0.8f && dot(w, N) < 0.f) w = -w;
      frame.U.x[k] = U.x; frame.U.y[k] = U.y; frame.U.z[k] = U.z;
      frame.V.x[k] = V.x; frame.V.y[k] = V.y; frame.V.z[k] = V.z;
      frame.N.x[k] = N.x; frame.N.y[k] = N.y; frame.N.z[k] = N.z;
      ctx->Ng.x[k] = Ng.x; ctx->Ng.y[k] = Ng.y; ctx->Ng.z[k] = Ng.z;
      wo.x[k] = o.x; wo.y[k] = o.y; wo.z[k] = o.z;
      wi.x[k] = w.x; wi.y[k] = w.y; wi.z[k] = w.z;
      s3.x[k] = rng.f(); s3.y[k] = rng.f(); s3.z[k] = rng.f();
      params.insert(params.end(), {col.x, col.y, col.z, coatCol.x, coatCol.y, coatCol.z, roughness, eta, thickness, weight});
      geo.insert(geo.end(), {U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z, Ng.x, Ng.y, Ng.z, o.x, o.y, o.z, w.x, w.y, w.z});
      samp.insert(samp.end(), {s3.x[k], s3.y[k], s3.z[k]});
    }
    vbool m = one;
    DiffuseBrdfSimd* diffuse = new DiffuseBrdfSimd(m, &frame, Vec3vf(col));
    RoughCoatingBsdfSimd* coat = new RoughCoatingBsdfSimd(m, &frame, diffuse, vfloat(eta), Vec3vf(coatCol), vfloat(thickness),
                                                         vfloat(roughness), vfloat(0.f), vfloat(weight));
    MultiBsdfSimd multi;
    multi.add(m, coat, vfloat(1.f));

    Vec3vf swi = zero;
    vfloat spdf = zero;
    vint stype = zero;
    Vec3vf sw = multi.sample(m, *ctx, wo, swi, spdf, stype, BsdfSampleSimd(s3));
    vfloat epdf;
    Vec3vf ev = multi.eval(m, *ctx, wo, wi, epdf);
    Vec3vf alb = multi.getAlbedo(m);
    for (int k = 0; k < simdSize; ++k)
    {
      outSample.insert(outSample.end(), {sw.x[k], sw.y[k], sw.z[k], swi.x[k], swi.y[k], swi.z[k], spdf[k]});
      outType.push_back(stype[k]);
      outEval.insert(outEval.end(), {ev.x[k], ev.y[k], ev.z[k], epdf[k], alb.x[k], alb.y[k], alb.z[k]});
    }
    delete coat;
    delete diffuse;
    delete ctx;
    n += simdSize;
  }
  dump("coat_params", "f4", params, {n, 10});
  dump("coat_geo", "f4", geo, {n, 18});
  dump("coat_s", "f4", samp, {n, 3});
  dump("coat_sample", "f4", outSample, {n, 7});
  dump("coat_sample_type", "i4", outType, {n});
  dump("coat_eval", "f4", outEval, {n, 7});
}

static void goldenPostIntersect()
{
  InputRng rng(31337);
  const int nTri = 64;
  const int nRay = 512;
  for (int withNormals = 0; withNormals < 2; ++withNormals)
  {
    TriangleMesh mesh;
    mesh.alloc(nTri, nTri * 3, withNormals != 0, 0, false);
    std::vector<float> verts, norms;
    std::vector<int32_t> idx;
    for (int t = 0; t < nTri; ++t)
    {
      Vec3f c(20.f * rng.f() - 10.f, 8.f * rng.f(), 5.f + 25.f * rng.f());
      for (int j = 0; j < 3; ++j)
      {
        Vec3f p = c + 1.5f * rng.dir();
        mesh.setPosition(t * 3 + j, p);
        verts.insert(verts.end(), {p.x, p.y, p.z});
        if (withNormals)
        {
          Vec3f nn = rng.dir();
          mesh.setNormal(t * 3 + j, nn);
          norms.insert(norms.end(), {nn.x, nn.y, nn.z});
        }
      }
      mesh.indices[t] = Vec3i(t * 3, t * 3 + 1, t * 3 + 2);
      idx.insert(idx.end(), {t * 3, t * 3 + 1, t * 3 + 2});
    }
    mesh.materialIds[0] = 1;

    std::vector<float> rayIn, ctxOut, frameOut;
    std::vector<int32_t> primIn;
    for (int i = 0; i < nRay; i += simdSize)
    {
      RaySimd ray;
      HitSimd hit;
      for (int k = 0; k < simdSize; ++k)
      {
        int prim = int(rng.u() % nTri);
        float u = rng.f(), v = rng.f();
        if (u + v > 1.f) { u = 1.f - u; v = 1.f - v; }
        Vec3f o(20.f * rng.f() - 10.f, 8.f * rng.f() - 4.f, 30.f * rng.f() - 5.f);
        Vec3f d = rng.dir();
        float t = 0.5f + 40.f * rng.f();
        ray.org.x[k] = o.x; ray.org.y[k] = o.y; ray.org.z[k] = o.z;
        ray.dir.x[k] = d.x; ray.dir.y[k] = d.y; ray.dir.z[k] = d.z;
        ray.near[k] = 0.f; ray.far[k] = t;
        hit.primId[k] = prim; hit.geomId[k] = 0; hit.instId[k] = -1; hit.u[k] = u; hit.v[k] = v;
        rayIn.insert(rayIn.end(), {o.x, o.y, o.z, d.x, d.y, d.z, t, u, v});
        primIn.push_back(prim);
      }
      ShadingContextSimd* ctx = new ShadingContextSimd();
      vbool m = one;
      mesh.postIntersect(m, ray, hit, *ctx, 0);
      for (int k = 0; k < simdSize; ++k)
        ctxOut.insert(ctxOut.end(), {ctx->p.x[k], ctx->p.y[k], ctx->p.z[k], ctx->eps[k],
                                     ctx->Ng.x[k], ctx->Ng.y[k], ctx->Ng.z[k],
                                     ctx->f.U.x[k], ctx->f.U.y[k], ctx->f.U.z[k],
                                     ctx->f.V.x[k], ctx->f.V.y[k], ctx->f.V.z[k],
                                     ctx->f.N.x[k], ctx->f.N.y[k], ctx->f.N.z[k]});
      // Orientation flip toward the viewer, as render/scene.h:61-72 applies after postIntersect
      vbool backfacing = m & (dot(ctx->Ng, ray.dir) > 0.f);
      set(backfacing, ctx->Ng, -ctx->Ng);
      set(backfacing, ctx->f.N, -ctx->f.N);
      ctx->begin();
      Vec3vf wo = -ray.dir;
      const Basis3vf* fr = ctx->makeFrame(m, wo);
      for (int k = 0; k < simdSize; ++k)
        frameOut.insert(frameOut.end(), {fr->U.x[k], fr->U.y[k], fr->U.z[k], fr->V.x[k], fr->V.y[k], fr->V.z[k], fr->N.x[k], fr->N.y[k], fr->N.z[k]});
      delete ctx;
    }
    std::string sfx = withNormals ? "_n" : "";
    dump(("pi_verts" + sfx).c_str(), "f4", verts, {nTri * 3, 3});
    if (withNormals) dump("pi_normals_n", "f4", norms, {nTri * 3, 3});
    dump(("pi_indices" + sfx).c_str(), "i4", idx, {nTri, 3});
    dump(("pi_rays" + sfx).c_str(), "f4", rayIn, {nRay, 9});
    dump(("pi_prim" + sfx).c_str(), "i4", primIn, {nRay});
    dump(("pi_ctx" + sfx).c_str(), "f4", ctxOut, {nRay, 16});
    dump(("pi_frame" + sfx).c_str(), "f4", frameOut, {nRay, 9});
  }
}

// Camera matrices of PinholeCamera (camera/pinhole_camera.h:41-68; ThinLensCamera inherits them):
// worldToViewD3D, viewToClipD3D and worldToRaster, written row-major (m[r][c] = column c, row r).
// in: [width, height, fov, nearClip, pos xyz, U xyz, V xyz, N xyz]
// StandardMaterial's lobes for uniform parameters (material/standard.cpp:409-627, Default mode
// with an ambient IOR of 1): the opaque base (diffuse or rough diffuse, under a specular coating
// of weight specularWeight when the IOR differs from the ambient one), the translucent base (rough
// dielectric, thin rough dielectric when thin-walled, the transparent BTDF at the ambient IOR) and
// the metal base (rough conductor, artistic Fresnel) in one MultiBsdfSimd weighted
// (1 - metalness)(1 - transmission), (1 - metalness) transmission specular and metalness specular;
// a rough coating of weight coatWeight over that mix; the sheen layer; the opacity mix.  The
// orientation follows fromOutside = thinWalled || !backfacing.  wantOpaque = false builds the
// lobes getTransparency uses (:633-661).  Row parameters (std_params, 33 columns): baseWeight,
// baseColor rgb, metalness, diffuseRoughness, specularWeight, specularColor rgb,
// specularRoughness, specularIOR, transmissionWeight, transmissionColor rgb, transmissionDepth,
// thinWalled, thickness, coatWeight, coatColor rgb, coatRoughness, coatIOR, coatThickness,
// sheenWeight, sheenColor rgb, sheenRoughness, opacity, backfacing.
struct BsdfPool
{
  std::vector<BsdfSimd*> v;
  template <class T, class... A>
  T* make(A&&... a)
  {
    T* p = new T(std::forward<A>(a)...);
    v.push_back(p);
    return p;
  }
  ~BsdfPool()
  {
    for (BsdfSimd* b : v) delete b;
  }
};

static const BsdfSimd* standardLobes(vbool m, const Basis3vf* frame, const float* p, bool wantOpaque, BsdfPool& pool)
{
  const float eps = 1e-5f;  // material/standard.h:79
  const float ambientIor = 1.f;
  auto cl = [](float x) { return std::min(std::max(x, 0.f), 1.f); };
  auto rcpIor = [](float x) { x = std::max(x, 0.f); return x < 1.f ? 1.f / x : x; };
  const float opacity = cl(p[31]), transparency = 1.f - opacity;
  const BsdfSimd* bsdf = nullptr;
  if (opacity > eps)
  {
    const float base = cl(p[0]);
    const Vec3f baseColor(cl(p[1]), cl(p[2]), cl(p[3]));
    const float metalness = cl(p[4]), diffRough = cl(p[5]), specular = cl(p[6]);
    const Vec3f specColor(cl(p[7]), cl(p[8]), cl(p[9]));
    const float specRough = cl(p[10]), specIor = rcpIor(p[11]);
    const Vec3f transColor(cl(p[13]), cl(p[14]), cl(p[15]));
    const float transDepth = std::max(p[16], 0.f);
    const bool thin = p[17] != 0.f;
    const bool fromOutside = thin ? true : !(p[32] != 0.f);
    MultiBsdfSimd* baseBsdf = pool.make<MultiBsdfSimd>();
    const float dielectric = 1.f - metalness;
    if (dielectric > eps)
    {
      const float transmission = cl(p[12]);
      float eta = fromOutside ? ambientIor / specIor : specIor / ambientIor;
      eta = std::min(std::max(eta, 1.f / 3.f), 3.f);
      if (wantOpaque)
      {
        const float opaque = dielectric * (1.f - transmission);
        if (opaque > eps)
        {
          BsdfSimd* opaqueBsdf;
          const Vec3vf diffuseColor(baseColor * base);
          if (diffRough != 0.f) opaqueBsdf = pool.make<RoughDiffuseBrdfSimd>(m, frame, diffuseColor, vfloat(diffRough));
          else opaqueBsdf = pool.make<DiffuseBrdfSimd>(m, frame, diffuseColor);
          if (std::fabs(specIor - ambientIor) > eps && specular > eps)
            opaqueBsdf = pool.make<RoughCoatingBsdfSimd>(m, frame, opaqueBsdf, vfloat(eta), Vec3vf(1.f), vfloat(1.f),
                                                         vfloat(specRough), vfloat(0.f), vfloat(specular));
          baseBsdf->add(m, opaqueBsdf, vfloat(opaque));
        }
      }
      const float translucent = dielectric * transmission * specular;
      if (translucent > eps)
      {
        BsdfSimd* tb;
        if (std::fabs(specIor - ambientIor) > eps)
        {
          if (thin)
            tb = pool.make<ThinRoughDielectricBsdfSimd>(m, frame, vfloat(eta), vfloat(specRough), vfloat(0.f), Vec3vf(transColor),
                                                        vfloat(std::max(p[18], 0.f) / transDepth));
          else tb = pool.make<RoughDielectricBsdfSimd>(m, frame, vfloat(eta), vfloat(specRough), vfloat(0.f));
        }
        else tb = pool.make<TransparentBtdfSimd>();
        baseBsdf->add(m, tb, vfloat(translucent));
      }
    }
    if (wantOpaque)
    {
      const float metal = metalness * specular;
      if (metal > eps)
      {
        FresnelConductorArtistic<vfloat> fresnel(Vec3vf(baseColor * base), Vec3vf(specColor));
        BsdfSimd* metalBsdf = pool.make<RoughConductorBrdfSimd<FresnelConductorArtistic<vfloat>>>(m, frame, fresnel, vfloat(specRough), vfloat(0.f));
        baseBsdf->add(m, metalBsdf, vfloat(metal));
      }
    }
    bsdf = baseBsdf;
    const float coat = cl(p[19]), coatIor = rcpIor(p[24]);
    if (coat > eps && std::fabs(coatIor - ambientIor) > eps)
    {
      float coatEta = fromOutside ? ambientIor / coatIor : coatIor / ambientIor;
      coatEta = std::min(std::max(coatEta, 1.f / 3.f), 3.f);
      const Vec3f coatColor(cl(p[20]), cl(p[21]), cl(p[22]));
      bsdf = pool.make<RoughCoatingBsdfSimd>(m, frame, bsdf, vfloat(coatEta), Vec3vf(coatColor), vfloat(std::max(p[25], 0.f)),
                                             vfloat(cl(p[23])), vfloat(0.f), vfloat(coat));
    }
    const float sheen = cl(p[26]);
    if (sheen > eps)
      bsdf = pool.make<SheenBsdfSimd>(m, frame, bsdf, Vec3vf(Vec3f(cl(p[27]), cl(p[28]), cl(p[29]))), vfloat(cl(p[30])), vfloat(sheen));
  }
  if (transparency > eps)
  {
    BsdfSimd* tr = pool.make<TransparentBtdfSimd>();
    if (bsdf)
    {
      MultiBsdfSimd* blend = pool.make<MultiBsdfSimd>();
      blend->add(m, bsdf, vfloat(opacity));
      blend->add(m, tr, vfloat(transparency));
      bsdf = blend;
    }
    else bsdf = tr;
  }
  return bsdf;
}

// Variants by batch: 0 partial metalness; 1 partial transmission (rough dielectric beside the
// coated diffuse base), seen from outside or inside; 2 a coat over metal; 3 a coat over glass,
// from outside or inside; 4 every lobe at once with sheen, opacity and a rough diffuse base;
// 5 thin-walled partial transmission under a coat (at the ambient IOR in every third batch: the
// transparent BTDF); 6 base and specular weights below 1 with partial metalness; 7 a coat over a
// diffuse-only base (specular weight 0, the Kelemen energy term), with and without metal.
static void goldenStandard()
{
  InputRng rng(8484);
  const int nTotal = 1024;
  std::vector<float> params, geo, samp, outSample, outEval, outAlb, outNormal, outTransp;
  std::vector<int32_t> outType;
  int n = 0;
  for (int i = 0; i < nTotal; i += simdSize)
  {
    const int batch = i / simdSize, variant = batch % 8;
    float p[33];
    auto u = [&](float a, float b) { return a + (b - a) * rng.f(); };
    p[0] = 1.f;
    p[1] = u(0.05f, 0.95f); p[2] = u(0.05f, 0.95f); p[3] = u(0.05f, 0.95f);
    p[4] = 0.f;
    p[5] = 0.f;
    p[6] = 1.f;
    p[7] = u(0.3f, 1.f); p[8] = u(0.3f, 1.f); p[9] = u(0.3f, 1.f);
    p[10] = std::sqrt(0.01f + 0.5f * rng.f());
    const float iors[3] = {1.5f, 1.33f, 2.2f};
    p[11] = iors[batch % 3];
    p[12] = 0.f;
    p[13] = u(0.2f, 1.f); p[14] = u(0.2f, 1.f); p[15] = u(0.2f, 1.f);
    p[16] = u(0.5f, 2.f);
    p[17] = 0.f;
    p[18] = 1.f;
    p[19] = 0.f;
    p[20] = u(0.3f, 1.f); p[21] = u(0.3f, 1.f); p[22] = u(0.3f, 1.f);
    p[23] = std::sqrt(0.01f + 0.3f * rng.f());
    p[24] = (batch % 2) ? 1.6f : 1.45f;
    p[25] = 0.5f;
    p[26] = 0.f;
    p[27] = u(0.2f, 1.f); p[28] = u(0.2f, 1.f); p[29] = u(0.2f, 1.f);
    p[30] = u(0.1f, 0.9f);
    p[31] = 1.f;
    p[32] = 0.f;
    const bool inside = (batch / 8) % 2 == 1;
    switch (variant)
    {
      case 0: p[4] = u(0.2f, 0.8f); break;
      case 1: p[12] = u(0.2f, 0.8f); p[32] = inside ? 1.f : 0.f; break;
      case 2: p[4] = 1.f; p[19] = u(0.3f, 1.f); break;
      case 3: p[12] = 1.f; p[19] = u(0.3f, 1.f); p[32] = inside ? 1.f : 0.f; break;
      case 4:
        p[4] = u(0.2f, 0.5f); p[12] = u(0.2f, 0.6f); p[19] = u(0.3f, 0.9f); p[26] = u(0.2f, 0.9f); p[31] = u(0.5f, 0.95f);
        p[5] = u(0.1f, 0.9f);
        break;
      case 5:
        p[12] = u(0.3f, 0.9f); p[17] = 1.f; p[18] = u(0.2f, 1.5f); p[19] = u(0.3f, 0.9f);
        if ((batch / 8) % 3 == 0) p[11] = 1.f;
        break;
      case 6: p[0] = u(0.3f, 0.9f); p[6] = u(0.2f, 0.9f); p[4] = u(0.2f, 0.8f); break;
      default: p[6] = 0.f; p[19] = u(0.3f, 1.f); p[4] = ((batch / 8) % 2) ? u(0.2f, 0.6f) : 0.f; break;
    }

    Basis3vf frame;
    ShadingContextSimd* ctx = new ShadingContextSimd();
    Vec3vf wo, wi, s3;
    for (int k = 0; k < simdSize; ++k)
    {
      Vec3f N = rng.dir();
      Vec3f U, V;
      makeFrame(U, V, N);
      Vec3f Ng = N;
      if (rng.f() < 0.3f) { Ng = normalize(N + 0.2f * rng.dir()); if (dot(Ng, N) < 0.2f) Ng = N; }
      Vec3f o = rng.dir();
      if (rng.f() < 0.9f && dot(o, N) < 0.f) o = -o;
      Vec3f w = rng.dir();
      if (rng.f() < 0.6f && dot(w, N) < 0.f) w = -w;
      frame.U.x[k] = U.x; frame.U.y[k] = U.y; frame.U.z[k] = U.z;
      frame.V.x[k] = V.x; frame.V.y[k] = V.y; frame.V.z[k] = V.z;
      frame.N.x[k] = N.x; frame.N.y[k] = N.y; frame.N.z[k] = N.z;
      ctx->Ng.x[k] = Ng.x; ctx->Ng.y[k] = Ng.y; ctx->Ng.z[k] = Ng.z;
      wo.x[k] = o.x; wo.y[k] = o.y; wo.z[k] = o.z;
      wi.x[k] = w.x; wi.y[k] = w.y; wi.z[k] = w.z;
      s3.x[k] = rng.f(); s3.y[k] = rng.f(); s3.z[k] = rng.f();
      params.insert(params.end(), p, p + 33);
      geo.insert(geo.end(), {U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z, Ng.x, Ng.y, Ng.z, o.x, o.y, o.z, w.x, w.y, w.z});
      samp.insert(samp.end(), {s3.x[k], s3.y[k], s3.z[k]});
    }
    vbool m = one;
    BsdfPool pool, poolT;
    const BsdfSimd* bsdf = standardLobes(m, &frame, p, true, pool);
    Vec3vf swi = zero;
    vfloat spdf = zero;
    vint stype = zero;
    Vec3vf sw = bsdf->sample(m, *ctx, wo, swi, spdf, stype, BsdfSampleSimd(s3));
    vfloat epdf;
    Vec3vf ev = bsdf->eval(m, *ctx, wo, wi, epdf);
    Vec3vf alb = bsdf->getAlbedo(m);
    Vec3vf nrm = bsdf->getNormal(m);
    // StandardMaterial::getTransparency: a transmissive material evaluates the lobes built without
    // the opaque and metal bases, any other returns 1 - opacity
    Vec3vf tr;
    if (p[12] != 0.f) tr = standardLobes(m, &frame, p, false, poolT)->getTransparency(m, *ctx, wo);
    else tr = Vec3vf(vfloat(1.f - std::min(std::max(p[31], 0.f), 1.f)));
    for (int k = 0; k < simdSize; ++k)
    {
      outSample.insert(outSample.end(), {sw.x[k], sw.y[k], sw.z[k], swi.x[k], swi.y[k], swi.z[k], spdf[k]});
      outType.push_back(stype[k]);
      outEval.insert(outEval.end(), {ev.x[k], ev.y[k], ev.z[k], epdf[k]});
      outAlb.insert(outAlb.end(), {alb.x[k], alb.y[k], alb.z[k]});
      outNormal.insert(outNormal.end(), {nrm.x[k], nrm.y[k], nrm.z[k]});
      outTransp.insert(outTransp.end(), {tr.x[k], tr.y[k], tr.z[k]});
    }
    delete ctx;
    n += simdSize;
  }
  dump("std_params", "f4", params, {n, 33});
  dump("std_geo", "f4", geo, {n, 18});
  dump("std_s", "f4", samp, {n, 3});
  dump("std_sample", "f4", outSample, {n, 7});
  dump("std_sample_type", "i4", outType, {n});
  dump("std_eval", "f4", outEval, {n, 4});
  dump("std_albedo", "f4", outAlb, {n, 3});
  dump("std_normal", "f4", outNormal, {n, 3});
  dump("std_transparency", "f4", outTransp, {n, 3});
}

static void goldenCameraMatrices()
{
  InputRng rng(4321);
  std::vector<float> in, out;
  const int nCam = 6;
  for (int i = 0; i < nCam; ++i)
  {
    const int w = 32 * (8 + int(rng.f() * 52)), h = 8 * (8 + int(rng.f() * 120));
    const float fov = (20.f + 70.f * rng.f()) * float(pi) / 180.f, nearClip = 0.005f + 0.1f * rng.f();
    const Vec3f pos(20.f * rng.f() - 10.f, 20.f * rng.f() - 10.f, 20.f * rng.f() - 10.f);
    Vec3f N = i == 0 ? Vec3f(0.f, 0.f, -1.f) : rng.dir();
    Vec3f U, V;
    if (i == 0) { U = Vec3f(-1.f, 0.f, 0.f); V = Vec3f(0.f, 1.f, 0.f); }
    else makeFrame(U, V, N);
    const Basis3f basis(U, V, N);
    Props props;
    props.set("position", pos);
    props.set("basis", basis);
    props.set("nearClip", nearClip);
    props.set("fov", fov);
    props.set("width", w);
    props.set("height", h);
    PinholeCamera cam(props);
    in.insert(in.end(), {float(w), float(h), fov, nearClip, pos.x, pos.y, pos.z, U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z});
    for (const Mat4f* m : {&cam.worldToViewD3D, &cam.viewToClipD3D, &cam.worldToRaster})
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
        {
          const Vec4f& col = (*m)[c];
          out.push_back(r == 0 ? col.x : (r == 1 ? col.y : (r == 2 ? col.z : col.w)));
        }
  }
  dump("cam_mat_in", "f4", in, {nCam, 16});
  dump("cam_mat_out", "f4", out, {nCam, 3, 16});
}

// The remaining lobes, composed as StandardMaterial composes them (material/standard.cpp:
// 409-627): the lobe in a MultiBsdfSimd of weight 1; with opacity < 1 - 1e-5 that in turn in a
// MultiBsdfSimd with a TransparentBtdfSimd (scales opacity, 1 - opacity); opacity <= 1e-5 leaves
// the TransparentBtdfSimd alone.  Variants by batch: 0 rough diffuse, 1 coating over a rough
// diffuse base, 2 thin rough dielectric, 3 rough diffuse with transparency, 4 thin dielectric
// with transparency, 5 rough conductor with transparency (fully transparent every 4th batch).
// params: kind, r,g,b, er,eg,eb, roughness, ior, thickness, weight, transparency, diffuseRoughness
// (kinds as the HIP path numbers them: 0 diffuse, 1 conductor, 3 coated, 4 thin dielectric).
// transparency: what StandardMaterial::getTransparency returns (:633-661) for the same inputs.
static void goldenLobes()
{
  InputRng rng(6262);
  const int nTotal = 768;
  std::vector<float> params, geo, samp, outSample, outEval, outAlb, outNormal, outTransp;
  std::vector<int32_t> outType;
  int n = 0;
  for (int i = 0; i < nTotal; i += simdSize)
  {
    const int batch = i / simdSize, variant = batch % 6;
    const int kind = (variant == 0 || variant == 3) ? 0 : (variant == 1 ? 3 : (variant == 5 ? 1 : 4));
    Vec3f col(0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f());
    Vec3f edge(rng.f(), rng.f(), rng.f());
    Vec3f coatCol(0.3f + 0.7f * rng.f(), 0.3f + 0.7f * rng.f(), 0.3f + 0.7f * rng.f());
    float roughness = std::sqrt(0.01f + 0.5f * rng.f());
    float diffRough = (kind == 0 || kind == 3) ? 0.05f + 0.95f * rng.f() : 0.f;
    const float iors[3] = {1.5f, 1.33f, 2.2f};
    float ior = iors[batch % 3];
    float thickness = kind == 4 ? 0.2f + 1.8f * rng.f() : (kind == 3 ? 2.f * rng.f() : 0.f);
    float weight = kind == 3 ? 0.3f + 0.7f * rng.f() : 0.f;
    float transparency = variant >= 3 ? 0.1f + 0.8f * rng.f() : 0.f;
    if (variant == 5 && (batch / 6) % 4 == 0) transparency = 1.f;
    const float eta = std::min(std::max(1.f / ior, 1.f / 3.f), 3.f);

    Basis3vf frame;
    ShadingContextSimd* ctx = new ShadingContextSimd();
    Vec3vf wo, wi, s3;
    for (int k = 0; k < simdSize; ++k)
    {
      Vec3f N = rng.dir();
      Vec3f U, V;
      makeFrame(U, V, N);
      Vec3f Ng = N;
      if (rng.f() < 0.3f) { Ng = normalize(N + 0.2f * rng.dir()); if (dot(Ng, N) < 0.2f) Ng = N; }
      Vec3f o = rng.dir();
      if (rng.f() < 0.9f && dot(o, N) < 0.f) o = -o;
      Vec3f w = rng.dir();
      if (rng.f() < 0.7f && dot(w, N) < 0.f) w = -w;
      frame.U.x[k] = U.x; frame.U.y[k] = U.y; frame.U.z[k] = U.z;
      frame.V.x[k] = V.x; frame.V.y[k] = V.y; frame.V.z[k] = V.z;
      frame.N.x[k] = N.x; frame.N.y[k] = N.y; frame.N.z[k] = N.z;
      ctx->Ng.x[k] = Ng.x; ctx->Ng.y[k] = Ng.y; ctx->Ng.z[k] = Ng.z;
      wo.x[k] = o.x; wo.y[k] = o.y; wo.z[k] = o.z;
      wi.x[k] = w.x; wi.y[k] = w.y; wi.z[k] = w.z;
      s3.x[k] = rng.f(); s3.y[k] = rng.f(); s3.z[k] = rng.f();
      const Vec3f e3 = kind == 3 ? coatCol : edge;
      params.insert(params.end(), {float(kind), col.x, col.y, col.z, e3.x, e3.y, e3.z, roughness, ior, thickness, weight,
                                   transparency, diffRough});
      geo.insert(geo.end(), {U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z, Ng.x, Ng.y, Ng.z, o.x, o.y, o.z, w.x, w.y, w.z});
      samp.insert(samp.end(), {s3.x[k], s3.y[k], s3.z[k]});
    }
    vbool m = one;
    BsdfSimd* base = nullptr;
    BsdfSimd* child = nullptr;
    if (kind == 0)
      child = new RoughDiffuseBrdfSimd(m, &frame, Vec3vf(col), vfloat(diffRough));
    else if (kind == 3)
    {
      base = new RoughDiffuseBrdfSimd(m, &frame, Vec3vf(col), vfloat(diffRough));
      child = new RoughCoatingBsdfSimd(m, &frame, base, vfloat(eta), Vec3vf(coatCol), vfloat(thickness), vfloat(roughness),
                                       vfloat(0.f), vfloat(weight));
    }
    else if (kind == 1)
      child = new RoughConductorBrdfSimd<FresnelConductorArtistic<vfloat>>(m, &frame, FresnelConductorArtistic<vfloat>(Vec3vf(col), Vec3vf(edge)), vfloat(roughness), vfloat(0.f));
    else
      child = new ThinRoughDielectricBsdfSimd(m, &frame, vfloat(eta), vfloat(roughness), vfloat(0.f), Vec3vf(col), vfloat(thickness));
    const float opacity = std::min(std::max(1.f - transparency, 0.f), 1.f), transp = 1.f - opacity;
    const float epsM = 1e-5f;  // material/standard.h:79
    MultiBsdfSimd inner, outer;
    TransparentBtdfSimd transparent;
    const BsdfSimd* bsdf = nullptr;
    if (opacity > epsM)
    {
      inner.add(m, child, vfloat(1.f));
      bsdf = &inner;
    }
    if (transp > epsM)
    {
      if (bsdf)
      {
        outer.add(m, &inner, vfloat(opacity));
        outer.add(m, &transparent, vfloat(transp));
        bsdf = &outer;
      }
      else bsdf = &transparent;
    }

    Vec3vf swi = zero;
    vfloat spdf = zero;
    vint stype = zero;
    Vec3vf sw = bsdf->sample(m, *ctx, wo, swi, spdf, stype, BsdfSampleSimd(s3));
    vfloat epdf;
    Vec3vf ev = bsdf->eval(m, *ctx, wo, wi, epdf);
    Vec3vf alb = bsdf->getAlbedo(m);
    Vec3vf nrm = bsdf->getNormal(m);
    // StandardMaterial::getTransparency: a thin-walled transmissive material evaluates its
    // lobes (here the same composition), any other returns 1 - opacity
    Vec3vf tr = kind == 4 ? bsdf->getTransparency(m, *ctx, wo) : Vec3vf(vfloat(1.f - opacity));
    for (int k = 0; k < simdSize; ++k)
    {
      outSample.insert(outSample.end(), {sw.x[k], sw.y[k], sw.z[k], swi.x[k], swi.y[k], swi.z[k], spdf[k]});
      outType.push_back(stype[k]);
      outEval.insert(outEval.end(), {ev.x[k], ev.y[k], ev.z[k], epdf[k]});
      outAlb.insert(outAlb.end(), {alb.x[k], alb.y[k], alb.z[k]});
      outNormal.insert(outNormal.end(), {nrm.x[k], nrm.y[k], nrm.z[k]});
      outTransp.insert(outTransp.end(), {tr.x[k], tr.y[k], tr.z[k]});
    }
    delete child;
    delete base;
    delete ctx;
    n += simdSize;
  }
  dump("lobe_params", "f4", params, {n, 13});
  dump("lobe_geo", "f4", geo, {n, 18});
  dump("lobe_s", "f4", samp, {n, 3});
  dump("lobe_sample", "f4", outSample, {n, 7});
  dump("lobe_sample_type", "i4", outType, {n});
  dump("lobe_eval", "f4", outEval, {n, 4});
  dump("lobe_albedo", "f4", outAlb, {n, 3});
  dump("lobe_normal", "f4", outNormal, {n, 3});
  dump("lobe_transparency", "f4", outTransp, {n, 3});
}

// The sheen layer (material/standard.cpp:599-611): SheenBsdfSimd over the lobe's MultiBsdfSimd,
// under the opacity mix when there is one.  Variants by batch: 0 diffuse, 1 rough conductor,
// 2 coated diffuse, 3 thin rough dielectric, 4 rough diffuse with transparency, 5 rough
// dielectric (from outside).  params: the 13 columns of the lobe rows, then sheen weight, sheen
// color rgb, sheen roughness.
static void goldenSheen()
{
  InputRng rng(7373);
  const int nTotal = 768;
  std::vector<float> params, geo, samp, outSample, outEval, outAlb, outNormal, outTransp;
  std::vector<int32_t> outType;
  int n = 0;
  for (int i = 0; i < nTotal; i += simdSize)
  {
    const int batch = i / simdSize, variant = batch % 6;
    const int kinds[6] = {0, 1, 3, 4, 0, 2};
    const int kind = kinds[variant];
    Vec3f col(0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f(), 0.05f + 0.9f * rng.f());
    Vec3f edge(rng.f(), rng.f(), rng.f());
    Vec3f coatCol(1.f);
    float roughness = std::sqrt(0.01f + 0.5f * rng.f());
    float diffRough = variant == 4 ? 0.1f + 0.9f * rng.f() : 0.f;
    float ior = (batch % 2) ? 1.5f : 1.4f;
    float thickness = kind == 4 ? 0.2f + 1.8f * rng.f() : 1.f;
    float weight = kind == 3 ? 1.f : 0.f;
    float transparency = variant == 4 ? 0.1f + 0.8f * rng.f() : 0.f;
    float sheenW = 0.2f + 0.8f * rng.f();
    Vec3f sheenCol(0.2f + 0.8f * rng.f(), 0.2f + 0.8f * rng.f(), 0.2f + 0.8f * rng.f());
    float sheenR = 0.05f + 0.95f * rng.f();
    const float eta = std::min(std::max(1.f / ior, 1.f / 3.f), 3.f);

    Basis3vf frame;
    ShadingContextSimd* ctx = new ShadingContextSimd();
    Vec3vf wo, wi, s3;
    for (int k = 0; k < simdSize; ++k)
    {
      Vec3f N = rng.dir();
      Vec3f U, V;
      makeFrame(U, V, N);
      Vec3f Ng = N;
      if (rng.f() < 0.3f) { Ng = normalize(N + 0.2f * rng.dir()); if (dot(Ng, N) < 0.2f) Ng = N; }
      Vec3f o = rng.dir();
      if (rng.f() < 0.9f && dot(o, N) < 0.f) o = -o;
      Vec3f w = rng.dir();
      if (rng.f() < 0.7f && dot(w, N) < 0.f) w = -w;
      frame.U.x[k] = U.x; frame.U.y[k] = U.y; frame.U.z[k] = U.z;
      frame.V.x[k] = V.x; frame.V.y[k] = V.y; frame.V.z[k] = V.z;
      frame.N.x[k] = N.x; frame.N.y[k] = N.y; frame.N.z[k] = N.z;
      ctx->Ng.x[k] = Ng.x; ctx->Ng.y[k] = Ng.y; ctx->Ng.z[k] = Ng.z;
      wo.x[k] = o.x; wo.y[k] = o.y; wo.z[k] = o.z;
      wi.x[k] = w.x; wi.y[k] = w.y; wi.z[k] = w.z;
      s3.x[k] = rng.f(); s3.y[k] = rng.f(); s3.z[k] = rng.f();
      const Vec3f e3 = kind == 3 ? coatCol : edge;
      params.insert(params.end(), {float(kind), col.x, col.y, col.z, e3.x, e3.y, e3.z, roughness, ior, thickness, weight,
                                   transparency, diffRough, sheenW, sheenCol.x, sheenCol.y, sheenCol.z, sheenR});
      geo.insert(geo.end(), {U.x, U.y, U.z, V.x, V.y, V.z, N.x, N.y, N.z, Ng.x, Ng.y, Ng.z, o.x, o.y, o.z, w.x, w.y, w.z});
      samp.insert(samp.end(), {s3.x[k], s3.y[k], s3.z[k]});
    }
    vbool m = one;
    BsdfSimd* base = nullptr;
    BsdfSimd* child = nullptr;
    if (kind == 0)
    {
      if (diffRough > 0.f) child = new RoughDiffuseBrdfSimd(m, &frame, Vec3vf(col), vfloat(diffRough));
      else child = new DiffuseBrdfSimd(m, &frame, Vec3vf(col));
    }
    else if (kind == 1)
      child = new RoughConductorBrdfSimd<FresnelConductorArtistic<vfloat>>(m, &frame, FresnelConductorArtistic<vfloat>(Vec3vf(col), Vec3vf(edge)), vfloat(roughness), vfloat(0.f));
    else if (kind == 2)
      child = new RoughDielectricBsdfSimd(m, &frame, vfloat(eta), vfloat(roughness), vfloat(0.f));
    else if (kind == 3)
    {
      base = new DiffuseBrdfSimd(m, &frame, Vec3vf(col));
      child = new RoughCoatingBsdfSimd(m, &frame, base, vfloat(eta), Vec3vf(coatCol), vfloat(thickness), vfloat(roughness),
                                       vfloat(0.f), vfloat(weight));
    }
    else
      child = new ThinRoughDielectricBsdfSimd(m, &frame, vfloat(eta), vfloat(roughness), vfloat(0.f), Vec3vf(col), vfloat(thickness));
    MultiBsdfSimd inner, outer;
    inner.add(m, child, vfloat(1.f));
    SheenBsdfSimd sheen(m, &frame, &inner, Vec3vf(sheenCol), vfloat(sheenR), vfloat(sheenW));
    TransparentBtdfSimd transparent;
    const BsdfSimd* bsdf = &sheen;
    const float opacity = 1.f - transparency;
    if (transparency > 1e-5f)
    {
      outer.add(m, &sheen, vfloat(opacity));
      outer.add(m, &transparent, vfloat(transparency));
      bsdf = &outer;
    }
    Vec3vf swi = zero;
    vfloat spdf = zero;
    vint stype = zero;
    Vec3vf sw = bsdf->sample(m, *ctx, wo, swi, spdf, stype, BsdfSampleSimd(s3));
    vfloat epdf;
    Vec3vf ev = bsdf->eval(m, *ctx, wo, wi, epdf);
    Vec3vf alb = bsdf->getAlbedo(m);
    Vec3vf nrm = bsdf->getNormal(m);
    Vec3vf tr = kind == 4 ? bsdf->getTransparency(m, *ctx, wo) : Vec3vf(vfloat(1.f - opacity));
    for (int k = 0; k < simdSize; ++k)
    {
      outSample.insert(outSample.end(), {sw.x[k], sw.y[k], sw.z[k], swi.x[k], swi.y[k], swi.z[k], spdf[k]});
      outType.push_back(stype[k]);
      outEval.insert(outEval.end(), {ev.x[k], ev.y[k], ev.z[k], epdf[k]});
      outAlb.insert(outAlb.end(), {alb.x[k], alb.y[k], alb.z[k]});
      outNormal.insert(outNormal.end(), {nrm.x[k], nrm.y[k], nrm.z[k]});
      outTransp.insert(outTransp.end(), {tr.x[k], tr.y[k], tr.z[k]});
    }
    delete child;
    delete base;
    delete ctx;
    n += simdSize;
  }
  dump("sheen_params", "f4", params, {n, 18});
  dump("sheen_geo", "f4", geo, {n, 18});
  dump("sheen_s", "f4", samp, {n, 3});
  dump("sheen_sample", "f4", outSample, {n, 7});
  dump("sheen_sample_type", "i4", outType, {n});
  dump("sheen_eval", "f4", outEval, {n, 4});
  dump("sheen_albedo", "f4", outAlb, {n, 3});
  dump("sheen_normal", "f4", outNormal, {n, 3});
  dump("sheen_transparency", "f4", outTransp, {n, 3});
}

// UV set 0 at random hits of a mesh with texcoords (and the same hits on one without).
static void goldenTexcoords()
{
  InputRng rng(4242);
  const int nTri = 32, nHit = 256;
  std::vector<float> uvIn, hitIn, uvOut, uvOutNone;
  std::vector<int32_t> primIn;
  TriangleMesh mesh, bare;
  mesh.alloc(nTri, nTri * 3, false, 1, false);
  bare.alloc(nTri, nTri * 3, false, 0, false);
  for (int t = 0; t < nTri; ++t)
  {
    for (int j = 0; j < 3; ++j)
    {
      Vec3f p(10.f * rng.f(), 10.f * rng.f(), 10.f * rng.f());
      mesh.setPosition(t * 3 + j, p);
      bare.setPosition(t * 3 + j, p);
      Vec2f uv(4.f * rng.f() - 2.f, 4.f * rng.f() - 2.f);
      mesh.setTexcoord(0, t * 3 + j, uv);
      uvIn.insert(uvIn.end(), {uv.x, uv.y});
    }
    mesh.indices[t] = Vec3i(t * 3, t * 3 + 1, t * 3 + 2);
    bare.indices[t] = Vec3i(t * 3, t * 3 + 1, t * 3 + 2);
  }
  for (int i = 0; i < nHit; i += simdSize)
  {
    RaySimd ray;
    HitSimd hit;
    for (int k = 0; k < simdSize; ++k)
    {
      int prim = int(rng.u() % nTri);
      float u = rng.f(), v = rng.f();
      if (u + v > 1.f) { u = 1.f - u; v = 1.f - v; }
      ray.org.x[k] = 0.f; ray.org.y[k] = 0.f; ray.org.z[k] = -1.f;
      ray.dir.x[k] = 0.f; ray.dir.y[k] = 0.f; ray.dir.z[k] = 1.f;
      ray.near[k] = 0.f; ray.far[k] = 5.f;
      hit.primId[k] = prim; hit.geomId[k] = 0; hit.instId[k] = -1; hit.u[k] = u; hit.v[k] = v;
      hitIn.insert(hitIn.end(), {u, v});
      primIn.push_back(prim);
    }
    vbool m = one;
    ShadingContextSimd* ctx = new ShadingContextSimd();
    mesh.postIntersect(m, ray, hit, *ctx, 0);
    for (int k = 0; k < simdSize; ++k) uvOut.insert(uvOut.end(), {ctx->uv[0].x[k], ctx->uv[0].y[k]});
    delete ctx;
    ctx = new ShadingContextSimd();
    bare.postIntersect(m, ray, hit, *ctx, 0);
    for (int k = 0; k < simdSize; ++k) uvOutNone.insert(uvOutNone.end(), {ctx->uv[0].x[k], ctx->uv[0].y[k]});
    delete ctx;
  }
  dump("texcoord_vertex_uv", "<f4", uvIn, {nTri * 3, 2});
  dump("texcoord_hit_uv", "<f4", hitIn, {nHit, 2});
  dump("texcoord_prim", "<i4", primIn, {nHit});
  dump("texcoord_out", "<f4", uvOut, {nHit, 2});
  dump("texcoord_out_none", "<f4", uvOutNone, {nHit, 2});
}

// 8-bit texel decode of every byte value: sRGB (the 2.2-gamma table) and linear.
static void goldenPixelDecode()
{
  std::vector<float> srgb, lin;
  for (int i = 0; i < 256; ++i)
  {
    const int c = (i << 16) | (i << 8) | i | (i << 24);
    Vec4f a = decodeBgra8(c), b = decodeBgra8Linear(c);
    srgb.insert(srgb.end(), {a.x, a.y, a.z, a.w});
    lin.insert(lin.end(), {b.x, b.y, b.z, b.w});
  }
  dump("pixel_srgb8", "<f4", srgb, {256, 4});
  dump("pixel_linear8", "<f4", lin, {256, 4});
}

int main(int argc, char** argv)
{
  if (argc < 2) { fprintf(stderr, "usage: %s outdir\n", argv[0]); return 2; }
  gOutDir = argv[1];
  gManifest = fopen((gOutDir + "/manifest.json").c_str(), "w");
  fprintf(gManifest, "{");
  precomputeAlbedoTables();
  goldenSampler();
  goldenMorton();
  goldenCamera();
  goldenWarps();
  goldenDistributions();
  goldenLuts();
  goldenBsdf();
  goldenCoating();
  goldenPostIntersect();
  goldenLobes();
  goldenSheen();
  goldenStandard();
  goldenCameraMatrices();
  goldenTexcoords();
  goldenPixelDecode();
  fprintf(gManifest, "\n}\n");
  fclose(gManifest);
  return 0;
}
