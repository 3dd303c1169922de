// Per-path building blocks shared by the wavefront kernels (render.hip) and the path-replay
// adjoint (adjoint.hip): camera rays, post-intersection, light evaluation and sampling, the
// radiance clamp.  Both replay identical paths because they call the same code.
#pragma once

#include "internal.h"

namespace prt {

PRT_DEV uint32_t hashMix(uint32_t h, uint32_t k)  // math/hash.h:12-30 (murmurHash3Mix)
{
  k *= 0xcc9e2d51u; k = (k << 15) | (k >> 17); k *= 0x1b873593u;
  h ^= k; h = ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
  return h;
}

PRT_DEV f3 ld3(const float* p) { return f3(p[0], p[1], p[2]); }

// integrator.h:27-46
PRT_DEV f3 clampL(f3 L, float maxRadiance)
{
  f3 c(0.f < L.x ? L.x : 0.f, 0.f < L.y ? L.y : 0.f, 0.f < L.z ? L.z : 0.f);
  float y = luminance(c);
  if (y > maxRadiance) c = c * (maxRadiance * (1.0f / y));
  return c;
}

// render/scene.h:137-157
PRT_DEV f3 evalEnv(const RenderArgs& a, f3 wi, float& pdf, uint32_t& hitMask)
{
  f3 v(0.f);
  pdf = 0.f;
  hitMask = 0;
  for (int i = 0; i < a.nLights; ++i)
  {
    float cp;
    v = v + envEvalOne(a.lights[i], a.env, wi, cp);
    if (cp > 0.f) hitMask |= 1u << i;
    const int k = a.lightOffset + i;
    pdf += cp * (a.lightCdf[k + 1] - a.lightCdf[k]);
  }
  return v;
}

// light/triangle_light.h:84-100 + render/scene.h:159-195: the area light seen along wi at
// distance dist on primitive prim
PRT_DEV f3 evalArea(const RenderArgs& a, f3 wi, float dist, int prim, float& pdf)
{
  pdf = 0.f;
  int tri = a.areaMap[prim];
  float pickProb = dist1dPdf(a.areaCdf, a.areaCount, tri);
  float4 nn = a.areaTris[4 * tri + 3];
  float dirDotN = dot(wi, f3(nn.x, nn.y, nn.z));
  if (dirDotN >= 0.f) return f3(0.f);
  pdf = 2.0f * dist * dist * rcpSafe(fabsf(dirDotN)) * pickProb;
  pdf *= a.lightCdf[1] - a.lightCdf[0];
  return f3(a.areaTris[4 * tri].w, a.areaTris[4 * tri + 1].w, a.areaTris[4 * tri + 2].w);
}

// light/triangle_light.h:123-150: triangle by power, point uniform on it (shape_sampler.h:
// 135-152, Heitz 2019 map); wi is left unchanged when the sample faces away
PRT_DEV f3 sampleArea(const RenderArgs& a, f3 p, f3& wi, float& pdf, float& dist, float sx, float sy, float sc)
{
  pdf = 0.f;
  int tri = dist1dLookup(a.areaCdf, a.areaCount, sc);
  float pickProb = a.areaCdf[tri + 1] - a.areaCdf[tri];
  float4 A = a.areaTris[4 * tri], B = a.areaTris[4 * tri + 1], C = a.areaTris[4 * tri + 2], Nn = a.areaTris[4 * tri + 3];
  float u = sx, v = sy;
  if (v > u) { u *= 0.5f; v -= u; }
  else { v *= 0.5f; u -= v; }
  f3 v0(A.x, A.y, A.z), v1(B.x, B.y, B.z), v2(C.x, C.y, C.z);
  f3 dir = (v2 + u * (v0 - v2) + v * (v1 - v2)) - p;
  dist = length(dir);
  float dirDotN = dot(dir, f3(Nn.x, Nn.y, Nn.z));
  if (!(dirDotN < 0.f)) return f3(0.f);
  wi = dir * (1.0f / dist);
  pdf = 2.0f * dist * dist * dist * rcpSafe(fabsf(dirDotN)) * pickProb;
  return f3(A.w, B.w, C.w) * (1.0f / pdf);
}

// render/scene.h:215-250 (Distribution1D::sampleReuse for the pick, s.c remapped); the light
// list is [area light,] environment lights.  dist: distance to the sampled point (area light)
// or kPosMax.
PRT_DEV f3 sampleLight(const RenderArgs& a, f3 p, f3& wi, float& pdf, float& dist, float sx, float sy, float sc, int& picked)
{
  picked = 0;
  dist = kPosMax;
  const int total = a.lightOffset + a.nLights;
  if (total == 1)
    return a.lightOffset ? sampleArea(a, p, wi, pdf, dist, sx, sy, sc) : envSampleOne(a.lights[0], a.env, wi, pdf, sx, sy);
  float pick;
  int id = dist1dSampleReuse(a.lightCdf, total, pick, sc);
  picked = id;
  f3 w = (a.lightOffset && id == 0) ? sampleArea(a, p, wi, pdf, dist, sx, sy, sc)
                                     : envSampleOne(a.lights[id - a.lightOffset], a.env, wi, pdf, sx, sy);
  pdf *= pick;
  return w * (1.0f / pick);
}

// geometry/triangle_mesh.cpp:152-213 + render/scene.h:61-72
// Beer-Lambert transmittance of a path segment of length t through a medium of colour (r, g, b)
// (path_renderer_stream.h:652-657, pow(mediumColor, ray.far)).
PRT_DEV f3 mediumTransmittance(float4 medium, float t)
{
  return f3(powf(medium.x, t), powf(medium.y, t), powf(medium.z, t));
}

// clamp(x, lo, hi) as math/math.h:83-86 writes it: min(lo < x ? x : lo, hi)
PRT_DEV float clampRef(float x, float lo, float hi) { return fminf(lo < x ? x : lo, hi); }

// Bilinear lookup (image/image_texture.h:186-227): half-texel offset, repeat or clamp per axis,
// the four decoded texels interpolated as lerp(a, b, s) = a + s (b - a) (math/math.h:88-92),
// first along x.
PRT_DEV f3 textureLookup(const prt_texture_dev& t, const float4* texels, float u, float v)
{
  const float x = u - t.halfU, y = v - t.halfV;
  const float fw = float(t.w), fh = float(t.h);
  const float xr = t.addrU == PRT_TEX_ADDR_REPEAT ? (x - floorf(x)) * fw : clampRef(x * fw, 0.f, float(t.w - 1));
  const float yr = t.addrV == PRT_TEX_ADDR_REPEAT ? (y - floorf(y)) * fh : clampRef(y * fh, 0.f, float(t.h - 1));
  const float sx = xr - floorf(xr), sy = yr - floorf(yr);
  const int x0 = min(max(int(xr), 0), t.w - 1), y0 = min(max(int(yr), 0), t.h - 1);
  const int x1 = t.addrU == PRT_TEX_ADDR_REPEAT ? (x0 + 1 == t.w ? 0 : x0 + 1) : min(x0 + 1, t.w - 1);
  const int y1 = t.addrV == PRT_TEX_ADDR_REPEAT ? (y0 + 1 == t.h ? 0 : y0 + 1) : min(y0 + 1, t.h - 1);
  const float4* T = texels + t.offset;
  const float4 c00 = T[y0 * t.w + x0], c11 = T[y1 * t.w + x1], c01 = T[y1 * t.w + x0], c10 = T[y0 * t.w + x1];
  const f3 a00(c00.x, c00.y, c00.z), a11(c11.x, c11.y, c11.z), a01(c01.x, c01.y, c01.z), a10(c10.x, c10.y, c10.z);
  const f3 c0 = a00 + sx * (a10 - a00), c1 = a01 + sx * (a11 - a01);
  return c0 + sy * (c1 - c0);
}

// A textured baseColor (kinds 0 and 3): baseColor = clamp(texture(uv) * scale)
// (material/material_param.h:37-41, material/standard.cpp:424) replaces the record's colour and
// what depends on it: the diffuse colour and importance (bsdf/diffuse.h:70-75) and, under a
// coating, the base importance, the coating's importance and its albedo (bsdf/rough_coating.h:
// 562-571).  UV set 0 interpolates barycentrically (geometry/triangle_mesh.cpp:191-200).
PRT_DEV void applyBaseTexture(const RenderArgs& a, const prt_material_constants& m, int prim, float u, float v, Bsdf& b)
{
  const prt_texture_dev& t = a.tex[m.texBase - 1];
  float tu = u, tv = v;
  if (a.uv)
  {
    const int i0 = a.idx[3 * prim], i1 = a.idx[3 * prim + 1], i2 = a.idx[3 * prim + 2];
    // b0 uv0 + b1 uv1 + b2 uv2 in the reference build's evaluation: two fused multiply-adds
    const float b0 = 1.f - u - v;
    tu = fmaf(v, a.uv[2 * i2], fmaf(u, a.uv[2 * i1], b0 * a.uv[2 * i0]));
    tv = fmaf(v, a.uv[2 * i2 + 1], fmaf(u, a.uv[2 * i1 + 1], b0 * a.uv[2 * i0 + 1]));
  }
  const f3 c = textureLookup(t, a.texels, tu * t.scaleU + t.offU, tv * t.scaleV + t.offV);
  const f3 rgb(clampRef(c.x * m.texScale[0], 0.f, 1.f), clampRef(c.y * m.texScale[1], 0.f, 1.f),
               clampRef(c.z * m.texScale[2], 0.f, 1.f));
  b.c0 = rgb;
  const float imp = fmaxf(fmaxf(rgb.x, rgb.y), rgb.z);
  if (b.kind == PRT_KIND_COATED)
  {
    b.c2.z = imp;
    b.importance = fmaxf(1.f, imp);
    b.rr = f3(rgb.x + b.c2.y * (rgb.x * powf(b.c1.x, b.c2.x) - rgb.x), rgb.y + b.c2.y * (rgb.y * powf(b.c1.y, b.c2.x) - rgb.y),
              rgb.z + b.c2.y * (rgb.z * powf(b.c1.z, b.c2.x) - rgb.z));
  }
  else
    b.importance = imp;
}

// inst >= 0: the hit is on that instance of its mesh: the mesh's post-intersection runs in
// object space, then the frame's tangents go back to world space by the linear part and the
// normals by its inverse transpose, each renormalised (geometry/instance.h:17-26); the hit
// point is the world ray's.
PRT_DEV Ctx postIntersect(const RenderArgs& a, f3 o, f3 d, float t, int prim, float u, float v, int inst = -1)
{
  Ctx c;
  c.p = o + t * d;
  c.eps = fmaxf(t, fmaxf(fmaxf(fabsf(c.p.x), fabsf(c.p.y)), fabsf(c.p.z))) * 0x1.fp-18f;
  int i0 = a.idx[3 * prim], i1 = a.idx[3 * prim + 1], i2 = a.idx[3 * prim + 2];
  f3 p0 = ld3(a.pos + 3 * i0), p1 = ld3(a.pos + 3 * i1), p2 = ld3(a.pos + 3 * i2);
  f3 Ng = normalize(cross(p1 - p0, p2 - p0));
  f3 N = Ng;
  if (a.nrm)
  {
    // a triangle of a geometry without normals carries zero vertex normals in the shared array:
    // it keeps N = Ng, as a TriangleMesh without normals does (geometry/triangle_mesh.cpp:176-189)
    float b0 = 1.f - u - v;
    f3 Ns = b0 * ld3(a.nrm + 3 * i0) + u * ld3(a.nrm + 3 * i1) + v * ld3(a.nrm + 3 * i2);
    if (lengthSqr(Ns) > 0.f)
    {
      N = normalize(Ns);
      if (dot(N, Ng) < 0.f) N = -N;
    }
  }
  makeFrame(c.f.U, c.f.V, N);
  if (inst >= 0)
  {
    const prt_instance_dev& I = a.inst.inst[inst];
    c.f.U = normalize(xfmRows(I.lin, c.f.U));
    c.f.V = normalize(xfmRows(I.lin, c.f.V));
    N = normalize(xfmRows(I.nrm, N));
    Ng = normalize(xfmRows(I.nrm, Ng));
  }
  if (dot(Ng, d) > 0.f) { Ng = -Ng; N = -N; }
  c.Ng = Ng;
  c.f.N = N;
  return c;
}

// Transparent-shadow filter (render/transparent_filter.h:42-93): a hit on a transparent
// material multiplies the shadow ray's transmittance T by the material's transparency seen
// along -d and lets the ray go on until T falls below 1e-5; a hit on any other material
// occludes.  The product does not depend on the order in which the hits are found.  kStd: the
// scene has kind-5 materials.
template <bool kStd = true>
struct ShadowFilterT
{
  const RenderArgs* a;
  f3 o, d, T;
  uint32_t crossed;  // transparent surfaces passed (part of the path's identity in debug hashes)
  PRT_DEV bool operator()(int prim, float t, float u, float v, int inst = -1)
  {
    const prt_material_constants& m = a->mats[a->triMat[prim]];
    if (!m.transparent) { T = f3(0.f); return true; }
    ++crossed;
    Ctx c = postIntersect(*a, o, d, t, prim, u, v, inst);
    Bsdf b = makeBsdf(m, c, -d, false, true, kStd);
    T = T * bsdfTransparency(a->lut, b, -d);
    return reduceMax(T) < 1e-5f;
  }
};
using ShadowFilter = ShadowFilterT<true>;


// PixelFilter::sample for one coordinate (image/pixel_filter.h:57-63): a continuous sample of
// the filter's tabulated distribution (Distribution1D::sampleContinuous,
// sampling/distribution1d.h:104-108), centred and scaled to the filter width.
PRT_DEV float filterSample(const RenderArgs& a, float s)
{
  float pdf;
  const int i = dist1dSampleReuse(a.filterCdf, a.filterSize, pdf, s);
  const float x = (float(i) + s) * (1.0f / float(a.filterSize));
  return (x - 0.5f) * a.filterWidth;
}

// Camera ray of pixel (x, y) for sample sampleIndex (path_renderer_stream.h:209-268): the pixel
// sample from the jitter or from dimension set 0 (through the pixel filter), the lens sample
// from set 2.
PRT_DEV void cameraRay(const RenderArgs& a, int x, int y, uint32_t sampleIndex, f3& org, f3& dir, float& tnear)
{
  uint32_t pixelId = uint32_t(x) | (uint32_t(y) << 16);  // path_renderer_stream.h:219-220
  Sampler smp;
  smp.set(a.seed, sampleIndex, pixelId);
  float psx, psy, lsx, lsy;
  if (a.jitter)  // (:228-231)
  {
    psx = a.jitterX;
    psy = a.jitterY;
  }
  else
  {
    smp.get2D(0, psx, psy);  // sampleDimPixel (integrator.h:52)
    if (a.filterCdf)         // (:234-235)
    {
      psx = filterSample(a, psx) + 0.5f;
      psy = filterSample(a, psy) + 0.5f;
    }
  }
  smp.get2D(2, lsx, lsy);  // sampleDimLens (integrator.h:53)
  float imx = (float(x) + psx) * a.invW, imy = (float(y) + psy) * a.invH;
  f3 imageO = ld3(a.imageO), imageDx = ld3(a.imageDx), imageDy = ld3(a.imageDy);
  org = ld3(a.camPos);
  if (a.pinhole)  // camera/pinhole_camera.h:128-133
  {
    dir = imageO + imx * imageDx + imy * imageDy;
    float len = length(dir);
    dir = dir * (1.0f / len);
    tnear = a.nearClip * len;
  }
  else  // camera/thin_lens_camera.h:175-182, sampling/shape_sampler.h:169-179
  {
    float r = sqrtf(lsx), sp, cp;
    sincosf(2.f * kPi * lsy, &sp, &cp);
    float lx = r * cp * a.lensRadius, ly = r * sp * a.lensRadius;
    f3 off = lx * ld3(a.camU) + ly * ld3(a.camV) + 0.f * ld3(a.camN);
    dir = a.focal * (imageO + imx * imageDx + imy * imageDy) - off;
    float len = length(dir);
    org = org + off;
    dir = dir * (1.0f / len);
    tnear = a.nearClipScale * len;
  }
}

}  // namespace prt
