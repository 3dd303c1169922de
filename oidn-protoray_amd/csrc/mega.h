// One whole path of the stream path tracer traced by one lane (camera ray to termination), for
// the kernels that own a pixel for a whole sample: the path-replay adjoint and the persistent
// batch renderer.  It makes exactly the decisions kShade / kOccluded make for the same pixel and
// sample (same building blocks, same order of radiance accumulation), so its radiance equals
// the wavefront renderer's sample.
#pragma once

#include "path.h"

namespace prt {

// First-hit AOVs of a sample (path_renderer_stream.h:301-330, :475-519, :739-819): albedo,
// shading normal and view-axis depth of the first shaded vertex; zero (and an infinite depth) for
// a miss; 1 and -direction for an emitter.
struct FirstHit
{
  f3 albedo, normal;
  float depth;  // distance of the first hit along the view axis (+inf for a miss)
};

// Everything a vertex hook may read about a shaded vertex, after its next-event estimate and
// its BSDF sample (the shadow test has been done; L includes the unoccluded estimate).
struct PathVertex
{
  int depth, matId;
  const prt_material_constants* m;
  const Bsdf* b;
  const Ctx* ctx;
  f3 wo;
  bool nee;   // unoccluded next-event estimate made here
  f3 dirL;    // its contribution
  f3 wiL;     // its light direction
  bool ext;   // the path continues along wi
  f3 wi;
  int type;   // the sampled lobe's type (BsdfType)
  f3 L;       // radiance gathered so far, through this vertex's next-event estimate
};

struct NoVertexHook
{
  PRT_DEV void operator()(const PathVertex&) const {}
};

// Traces sample sampleIndex of pixel (x, y); returns the path's radiance (unclamped).  fh, when
// not null, receives the first-hit AOVs; hook(vertex) runs at every shaded vertex.  kStd: the
// scene may hold kind-5 materials (false keeps their code out of the kernel); kInst: a two-level
// scene (its rays go through the top-level tree and the instances' mesh trees, and the hit's
// frame through the instance's transform).
template <bool kStd, bool kInst = false, class Hook>
PRT_DEV f3 tracePathMega(const RenderArgs& a, TraceStack& st, int x, int y, uint32_t sampleIndex, FirstHit* fh, Hook&& hook,
                         uint32_t& nodes, uint32_t& tris)
{
  const uint32_t pixelId = uint32_t(x) | (uint32_t(y) << 16);
  f3 o, d;
  float tnear;
  cameraRay(a, x, y, sampleIndex, o, d, tnear);
  f3 T(1.f), L(0.f);
  float lastPdf = 1.f;
  int lastType = BsdfType::SpecularReflection, mediumId = 0;
  bool aov1 = false, aov2 = false;
  if (fh)
  {
    fh->albedo = fh->normal = f3(0.f);
    fh->depth = __int_as_float(0x7f800000);
  }
  for (int depth = 0;; ++depth)
  {
    const int sampleDim = 4 + 7 * depth;
    float th, u = 0.f, v = 0.f;
    int inst = -1;
    const int prim = kInst ? traceRayInst<false>(a.inst, o, d, tnear, kPosMax, th, u, v, inst, st, nodes, tris)
                           : traceRay<false>(a.bvh, o, d, tnear, kPosMax, th, u, v, st, nodes, tris);
    const int matId = prim >= 0 ? a.triMat[prim] : 0;
    if (matId == 0)  // miss (path_renderer_stream.h:285-415)
    {
      float lightPdf;
      uint32_t envMask;
      f3 Le = evalEnv(a, d, lightPdf, envMask);
      float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
      L = L + T * w * Le;
      break;
    }
    const prt_material_constants& m = a.mats[matId];
    if (m.emissive)  // area-light hit (:419-582)
    {
      float lightPdf;
      f3 Le = evalArea(a, d, th, prim, lightPdf);
      float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
      L = L + T * w * Le;
      if (fh && !aov1)
      {
        fh->albedo = f3(1.f);
        fh->normal = normalize(-d);
        fh->depth = dot((o + d * th) - ld3(a.camPos), -ld3(a.camN));
      }
      break;
    }
    if (depth >= a.maxDepth) break;
    Ctx ctx = postIntersect(a, o, d, th, prim, u, v, inst);
    const f3 wo = -d;
    const bool inside = m.mediaInterface && (mediumId == m.mediumInt);  // material/material.h:54-61
    Bsdf b = makeBsdf(m, ctx, wo, inside, true, kStd);
    if (m.texBase) applyBaseTexture(a, m, prim, u, v, b);
    Sampler smp;
    smp.set(a.seed, sampleIndex, pixelId);
    if (mediumId > 0) T = T * mediumTransmittance(a.media[mediumId], th);  // volume transmission (:652-657)
    const bool mAov1 = !aov1;
    bool mAov2 = !aov2;
    const bool mRR = !(mAov1 || mAov2);
    // next-event estimate (:659-718) with its shadow ray (:951-1002)
    f3 dirL(0.f), wiL(0.f);
    bool nee = false;
    if (b.mix != 2 && b.smooth)  // a smooth lobe (all but the transparent BTDF)
    {
      float s0, s1, s2;
      smp.get3D(sampleDim + 3, s0, s1, s2);
      float lightPdf, lightDist;
      int picked;
      f3 lw = sampleLight(a, ctx.p, wiL, lightPdf, lightDist, s0, s1, s2, picked);
      if (reduceMax(lw) > 0.f && lightPdf > 0.f)
      {
        float bsdfPdf;
        f3 f = bsdfEval(a.lut, b, ctx, wo, wiL, bsdfPdf);
        if (reduceMax(f) > 0.f)
        {
          f3 dirT = T * f;
          if (depth >= a.rrDepth)
          {
            float contProb = mRR ? fminf(luminance(dirT * (1.0f / bsdfPdf)), 0.95f) : 1.f;
            bsdfPdf *= contProb;
          }
          float w = misWeight(lightPdf, bsdfPdf);
          dirL = dirT * w * lw;
          float ts, us = 0.f, vs = 0.f;
          if (a.anyTransparent)
          {
            ShadowFilterT<kStd> flt{&a, ctx.p, wiL, f3(1.f), 0u};
            int si;
            nee = (kInst ? traceRayInst<true>(a.inst, ctx.p, wiL, ctx.eps, lightDist - ctx.eps, ts, us, vs, si, st, nodes, tris, flt)
                         : traceRay<true>(a.bvh, ctx.p, wiL, ctx.eps, lightDist - ctx.eps, ts, us, vs, st, nodes, tris, flt)) < 0;
            dirL = dirL * flt.T;
          }
          else
          {
            int si;
            nee = (kInst ? traceRayInst<true>(a.inst, ctx.p, wiL, ctx.eps, lightDist - ctx.eps, ts, us, vs, si, st, nodes, tris)
                         : traceRay<true>(a.bvh, ctx.p, wiL, ctx.eps, lightDist - ctx.eps, ts, us, vs, st, nodes, tris)) < 0;
          }
        }
      }
    }
    // BSDF sample (:720-730), AOV flags (:739-874), throughput, RR (:877-887)
    float s0, s1, s2;
    smp.get3D(sampleDim + 0, s0, s1, s2);
    f3 wi(0.f);
    float pdf = 0.f;
    int type = 0;
    f3 bw = bsdfSample(a.lut, b, ctx, wo, wi, pdf, type, s0, s1, s2);
    bool ext = reduceMax(bw) > 0.f && pdf > 0.f;
    const int extType = ext ? type : 0;
    if (mAov1)
    {
      aov1 = true;
      if (fh)
      {
        fh->albedo = bsdfAlbedo(b);
        fh->normal = bsdfNormal(b);
        fh->depth = dot(ctx.p - ld3(a.camPos), -ld3(a.camN));
      }
    }
    mAov2 = mAov2 && (extType & BsdfType::Delta) == 0;
    if (mAov2) aov2 = true;
    T = T * bw;
    ext = ext && reduceMax(T) >= a.minWeight;
    if (depth >= a.rrDepth)
    {
      float contProb = mRR ? fminf(luminance(T), 0.95f) : 1.f;
      ext = ext && smp.get1D(sampleDim + 6) <= contProb;
      T = T * (1.0f / contProb);
      pdf *= contProb;
    }
    // the shadow ray's radiance joins after the shading of this depth, as in the wavefront kernels
    if (nee) L = L + dirL;
    PathVertex pv{depth, matId, &m, &b, &ctx, wo, nee, dirL, wiL, ext, wi, type, L};
    hook(pv);
    if (!ext) break;
    o = ctx.p;
    tnear = ctx.eps;
    d = wi;
    if ((type & BsdfType::Continuation) == 0) { lastPdf = pdf; lastType = type; }
    if (m.mediaInterface && dot(wo, ctx.Ng) * dot(wi, ctx.Ng) < 0.f)
      mediumId = (mediumId == m.mediumInt) ? 0 : m.mediumInt;
  }
  return L;
}

}  // namespace prt
