// Wavefront kernels of the stream path tracer on gfx950.
//
// One sample pass (render/path_renderer_stream.h:150-187 with renderTile :195-1073) is:
//   kGenerate                  camera rays for every pixel            (:209-268)
//   for depth = 0 .. maxDepth:
//     kIntersect               closest hit of the queued paths         (:280)
//     kShade                   miss / hit shading, NEE shadow ray,     (:285-415, :586-949)
//                              BSDF sample, RR, compaction of survivors
//     kOccluded                any hit of the shadow rays, unoccluded  (:951-1065)
//                              direct light added, terminated paths accumulated
// The reference works on 1024-path streams per 32x32 tile with a counting sort by material;
// here all paths of the frame form one queue, compacted with wave-aggregated atomics, and
// per-path state lives in HBM at a fixed slot per pixel (8x8 tiles).  Each pixel gets exactly
// one sample per pass, so the FP32 accumulators are updated without atomics and in sample
// order, as AccumBuffer3f::addSerial does (image/accum_buffer.h:439-486).
#include "path.h"
#include "../../include/prt_hip.h"

namespace prt {

namespace {

PRT_DEV int laneId() { return __lane_id(); }

// Reserves one queue entry per lane with pred set; returns the entry index (or -1).
PRT_DEV int waveAppend(int* counter, bool pred)
{
  uint64_t mask = __ballot(pred);
  if (mask == 0) return -1;
  int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (laneId() == leader) base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader);
  uint64_t below = mask & ((1ull << laneId()) - 1ull);
  return pred ? base + __popcll(below) : -1;
}

PRT_DEV unsigned long long waveSum(unsigned long long v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  return v;
}

PRT_DEV uint32_t waveMax(uint32_t v)
{
  for (int off = 32; off > 0; off >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), off)));
  return v;
}


// Coarse direction cell (1/32 per axis) folded into the debug path hash.
PRT_DEV uint32_t dirCell(f3 d)
{
  uint32_t qx = uint32_t(int(floorf(d.x * 32.f)) + 64) & 0xffu;
  uint32_t qy = uint32_t(int(floorf(d.y * 32.f)) + 64) & 0xffu;
  uint32_t qz = uint32_t(int(floorf(d.z * 32.f)) + 64) & 0xffu;
  return qx | (qy << 8) | (qz << 16);
}

PRT_DEV void addAov3(float* buf, float* w, int npix, int pix, f3 c)
{
  if (!buf) return;
  buf[pix] += c.x; buf[npix + pix] += c.y; buf[2 * npix + pix] += c.z;
  w[pix] += 1.f;
}


// Adds the path's radiance as this pass's sample of its pixel (path_renderer_stream.h:323,1026);
// classB: the pass is a class-B sample and also goes to the half buffer (image/accum_buffer.h:
// 36-47, 455-470).
PRT_DEV void finalizePath(const RenderArgs& a, int slot, f3 L, uint32_t hash, bool classB)
{
  int x, y;
  slotToPixel(slot, a.W, x, y);
  int pix = y * a.W + x;
  f3 c = clampL(L, a.maxRadiance);
  const prt_frame_dev& fb = a.fb;
  fb.color[pix] += c.x;
  fb.color[a.npix + pix] += c.y;
  fb.color[2 * a.npix + pix] += c.z;
  fb.weight[pix] += 1.f;
  if (fb.colorB && classB)
  {
    fb.colorB[pix] += c.x;
    fb.colorB[a.npix + pix] += c.y;
    fb.colorB[2 * a.npix + pix] += c.z;
    fb.weightB[pix] += 1.f;
  }
  if (fb.color2)
  {
    fb.color2[pix] += c.x * c.x;
    fb.color2[a.npix + pix] += c.y * c.y;
    fb.color2[2 * a.npix + pix] += c.z * c.z;
  }
  if (fb.dbgL)
  {
    fb.dbgL[pix] = c.x; fb.dbgL[a.npix + pix] = c.y; fb.dbgL[2 * a.npix + pix] = c.z;
    fb.dbgHash[pix] = hash;
  }
}

__global__ void __launch_bounds__(256) kGenerate(const RenderArgs* __restrict__ ap, uint32_t sampleIndex)
{
  const RenderArgs& a = *ap;
  int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < 2 * (a.maxDepth + 2); k += blockDim.x) a.counters[k] = 0;
  if (slot >= a.npix) return;
  int x, y;
  slotToPixel(slot, a.W, x, y);
  f3 org, dir;
  float tnear;
  cameraRay(a, x, y, sampleIndex, org, dir, tnear);
  a.rayO[slot] = make_float4(org.x, org.y, org.z, tnear);
  a.rayD[slot] = make_float4(dir.x, dir.y, dir.z, 0.f);
  a.TP[slot] = make_float4(1.f, 1.f, 1.f, 1.f);  // throughput 1, lastBsdfPdf 1
  a.Lr[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
  a.flags[slot] = BsdfType::SpecularReflection;  // lastBsdfType, medium 0, no AOVs yet
  if (a.hash) a.hash[slot] = 0x9e3779b9u;
}

// Closest hits of the queued paths of depth `depth`.  kCount: also count the traversal loop
// iterations per lane and per wave (the SIMD utilisation), in a separate build that profiling
// passes use, so the timed kernel keeps its registers and occupancy.
template <bool kCount, bool kInst>
__global__ void __launch_bounds__(kTraceBlock) kIntersect(const RenderArgs* __restrict__ ap, int depth)
{
  const RenderArgs& a = *ap;
  __shared__ TraceLds<kTraceBlock> lds;
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  TraceStack st = makeStack(lds, a.spill, a.spillStride, gtid);
  const int count = depth == 0 ? a.npix : a.counters[2 * depth];
  const int* q = a.queue[depth & 1];
  uint32_t nodes = 0, tris = 0, rays = 0, waveSteps = 0;
  auto one = [&](int i) {
    int slot = depth == 0 ? i : ldg(q + i);
    float4 ro = ldg(a.rayO + slot), rd = ldg(a.rayD + slot);
    float t, u = 0.f, v = 0.f;
    int prim;
    if (kInst)
    {
      int inst = -1;
      prim = traceRayInst<false>(a.inst, f3(ro.x, ro.y, ro.z), f3(rd.x, rd.y, rd.z), ro.w, kPosMax, t, u, v, inst, st, nodes, tris);
      stg(a.hitInst + slot, inst);
    }
    else
      prim = traceRay<false>(a.bvh, f3(ro.x, ro.y, ro.z), f3(rd.x, rd.y, rd.z), ro.w, kPosMax, t, u, v, st, nodes, tris);
    stg(a.hit + slot, make_float4(t, u, v, __int_as_float(prim)));
    ++rays;
  };
  const int stride = gridDim.x * blockDim.x;
  if (!kCount)
    for (int i = gtid; i < count; i += stride) one(i);
  else
    // the lanes of a wave run the same number of iterations here, so the wave's steps for a batch
    // of rays (its longest lane's) are a cross-lane maximum
    for (int base = gtid - laneId(); base < count; base += stride)
    {
      const uint32_t before = st.steps;
      if (base + laneId() < count) one(base + laneId());
      waveSteps += waveMax(st.steps - before);
    }
  unsigned long long r = waveSum(rays), nv = waveSum(nodes), tt = waveSum(tris);
  if (laneId() == 0 && r)
  {
    atomicAdd(&a.stats[0], r);
    atomicAdd(&a.stats[1], nv);
    atomicAdd(&a.stats[2], tt);
  }
  if (kCount)
  {
    unsigned long long ls = waveSum(st.steps);
    if (laneId() == 0 && r)
    {
      atomicAdd(&a.stats[6], ls);
      atomicAdd(&a.stats[7], (unsigned long long)waveSteps);
    }
  }
}

// Waves per SIMD the shading kernel is compiled for (its register budget: 128 VGPRs at 4, with
// 44 bytes of spills).  Measured on the 1080p workload: 4.68 ms of shading per sample pass at 4,
// 4.86 at 3 (151 VGPRs, no spills), 5.17 at 5.
#ifndef PRT_SHADE_WAVES_PER_EU
#define PRT_SHADE_WAVES_PER_EU 4
#endif
// kLevel: 0 the four original lobes and clear media (this kernel keeps their register budget);
// 1 the scene also has thin dielectrics, opacity, Oren-Nayar bases, sheen or coloured media;
// 2 it also has kind-5 materials (StandardMaterial's weighted bases and coats).
template <int kLevel>
__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(PRT_SHADE_WAVES_PER_EU)))
kShade(const RenderArgs* __restrict__ ap, int depth, uint32_t sampleIndex, bool classB)
{
  const RenderArgs& a = *ap;
  const int count = depth == 0 ? a.npix : a.counters[2 * depth];
  const int* qIn = a.queue[depth & 1];
  int* qOut = a.queue[(depth + 1) & 1];
  int* cntOut = a.counters + 2 * (depth + 1);
  int* cntShadow = a.counters + 2 * depth + 1;
  const int sampleDim = 4 + 7 * depth;  // sampleDimBaseSize + sampleDimBounceSize * depth
  const int stride = gridDim.x * blockDim.x;
  // all lanes of a wave run the same number of iterations so the appends stay converged
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int waveBase = gtid - laneId();
  for (int base = waveBase; base < count; base += stride)
  {
    const int i = base + laneId();
    const bool active = i < count;
    bool wantExt = false, wantShadow = false;
    int slot = 0;
    if (active)
    {
      slot = depth == 0 ? i : qIn[i];
      int x, y;
      slotToPixel(slot, a.W, x, y);
      const int pix = y * a.W + x;
      const uint32_t pixelId = uint32_t(x) | (uint32_t(y) << 16);
      float4 h = a.hit[slot];
      int prim = __float_as_int(h.w);
      float4 ro = a.rayO[slot], rd = a.rayD[slot], tp = a.TP[slot], lr = a.Lr[slot];
      f3 o(ro.x, ro.y, ro.z), d(rd.x, rd.y, rd.z), T(tp.x, tp.y, tp.z), L(lr.x, lr.y, lr.z);
      float lastPdf = tp.w;
      int fl = a.flags[slot];
      int lastType = fl & PRT_F_LASTTYPE_MASK;
      uint32_t hsh = a.hash ? hashMix(a.hash[slot], uint32_t(prim + 1)) : 0u;
      // two-level scenes: the hit's instance (part of the path's identity too)
      const int inst = (kLevel >= 1 && a.hitInst && prim >= 0) ? a.hitInst[slot] : -1;
      if (a.hash && inst >= 0) hsh = hashMix(hsh, 0x40000000u | uint32_t(inst));
      if (depth == 0 && a.fb.dbgFirst) a.fb.dbgFirst[pix] = prim;
      int matId = prim >= 0 ? a.triMat[prim] : 0;
      fl &= ~(PRT_F_DIRECT | PRT_F_TERMINATE);
      if (matId == 0)
      {
        // miss (path_renderer_stream.h:285-415): environment light with MIS
        float lightPdf;
        uint32_t envMask;
        f3 Le = evalEnv(a, d, lightPdf, envMask);
        if (a.hash) hsh = hashMix(hsh, 0x100u | envMask);
        float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
        L = L + T * w * Le;
        if (!(fl & PRT_F_AOV1))
        {
          addAov3(a.fb.albedo1, a.fb.albedo1W, a.npix, pix, f3(0.f));
          addAov3(a.fb.normal1, a.fb.normal1W, a.npix, pix, f3(0.f));
          if (a.fb.depth) { a.fb.depth[pix] += __int_as_float(0x7f800000); a.fb.depthW[pix] += 1.f; }
        }
        if (!(fl & PRT_F_AOV2))
        {
          addAov3(a.fb.albedo, a.fb.albedoW, a.npix, pix, f3(0.f));
          addAov3(a.fb.normal, a.fb.normalW, a.npix, pix, f3(0.f));
        }
        finalizePath(a, slot, L, hsh, classB);
      }
      else if (a.mats[matId].emissive)
      {
        // area-light hit (:419-582), at any depth: emission with MIS, then the path ends
        float lightPdf;
        f3 Le = evalArea(a, d, h.x, prim, lightPdf);
        float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
        L = L + T * w * Le;
        f3 Nv = normalize(-d);
        if (!(fl & PRT_F_AOV1))
        {
          addAov3(a.fb.albedo1, a.fb.albedo1W, a.npix, pix, f3(1.f));
          addAov3(a.fb.normal1, a.fb.normal1W, a.npix, pix, Nv);
          if (a.fb.depth)
          {
            a.fb.depth[pix] += dot((o + d * h.x) - ld3(a.camPos), -ld3(a.camN));
            a.fb.depthW[pix] += 1.f;
          }
        }
        if (!(fl & PRT_F_AOV2))
        {
          addAov3(a.fb.albedo, a.fb.albedoW, a.npix, pix, f3(1.f));
          addAov3(a.fb.normal, a.fb.normalW, a.npix, pix, Nv);
        }
        finalizePath(a, slot, L, hsh, classB);
      }
      else if (depth >= a.maxDepth)
      {
        finalizePath(a, slot, L, hsh, classB);  // no shading beyond maxDepth (:615)
      }
      else
      {
        const prt_material_constants& m = a.mats[matId];
        Ctx ctx = postIntersect(a, o, d, h.x, prim, h.y, h.z, inst);
        f3 wo = -d;
        int mediumId = fl >> PRT_F_MEDIUM_SHIFT;
        bool inside = m.mediaInterface && (mediumId == m.mediumInt);  // material/material.h:54-61
        Bsdf b = makeBsdf(m, ctx, wo, inside, kLevel >= 1, kLevel >= 2);
        if (kLevel >= 1 && m.texBase) applyBaseTexture(a, m, prim, h.y, h.z, b);
        Sampler smp;
        smp.set(a.seed, sampleIndex, pixelId);  // resetSample (:641-642)
        if (kLevel >= 1 && mediumId > 0) T = T * mediumTransmittance(a.media[mediumId], h.x);  // volume transmission (:652-657)
        bool mAov1 = !(fl & PRT_F_AOV1), mAov2 = !(fl & PRT_F_AOV2);
        bool mRR = !(mAov1 || mAov2);
        // direct light (:659-718) when the BSDF has a smooth lobe: all but the transparent BTDF
        f3 dirL(0.f);
        if (b.mix != 2 && b.smooth)
        {
          float s0, s1, s2;
          smp.get3D(sampleDim + 3, s0, s1, s2);
          f3 wiL(0.f);
          float lightPdf, lightDist;
          int picked;
          f3 lw = sampleLight(a, ctx.p, wiL, lightPdf, lightDist, s0, s1, s2, picked);
          if (a.hash) hsh = hashMix(hashMix(hsh, 0x200u | uint32_t(picked)), dirCell(wiL));
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
              a.shO[slot] = make_float4(ctx.p.x, ctx.p.y, ctx.p.z, ctx.eps);
              a.shD[slot] = make_float4(wiL.x, wiL.y, wiL.z, lightDist - ctx.eps);  // [eps, dist - eps] (:686)
              a.dirL[slot] = make_float4(dirL.x, dirL.y, dirL.z, 0.f);
              wantShadow = true;
              fl |= PRT_F_DIRECT;
            }
          }
        }
        // BSDF sample (:720-730)
        float s0, s1, s2;
        smp.get3D(sampleDim + 0, s0, s1, s2);
        f3 wi(0.f);
        float pdf = 0.f;
        int type = 0;
        f3 bw = bsdfSample(a.lut, b, ctx, wo, wi, pdf, type, s0, s1, s2);
        bool ext = reduceMax(bw) > 0.f && pdf > 0.f;
        int extType = ext ? type : 0;
        if (a.hash)
        {
          hsh = hashMix(hsh, uint32_t(extType));
          if (ext) hsh = hashMix(hsh, dirCell(wi));
        }
        if (mAov1)  // first-hit AOVs (:739-819)
        {
          addAov3(a.fb.albedo1, a.fb.albedo1W, a.npix, pix, bsdfAlbedo(b));
          addAov3(a.fb.normal1, a.fb.normal1W, a.npix, pix, bsdfNormal(b));
          if (a.fb.depth)
          {
            a.fb.depth[pix] += dot(ctx.p - ld3(a.camPos), -ld3(a.camN));
            a.fb.depthW[pix] += 1.f;
          }
          fl |= PRT_F_AOV1;
        }
        mAov2 = mAov2 && (extType & BsdfType::Delta) == 0;  // (:828-832)
        if (mAov2)
        {
          addAov3(a.fb.albedo, a.fb.albedoW, a.npix, pix, bsdfAlbedo(b));
          addAov3(a.fb.normal, a.fb.normalW, a.npix, pix, bsdfNormal(b));
          fl |= PRT_F_AOV2;
        }
        T = T * bw;
        ext = ext && reduceMax(T) >= a.minWeight;
        if (depth >= a.rrDepth)  // Russian roulette (:880-887)
        {
          float contProb = mRR ? fminf(luminance(T), 0.95f) : 1.f;
          ext = ext && smp.get1D(sampleDim + 6) <= contProb;
          T = T * (1.0f / contProb);
          pdf *= contProb;
        }
        if (ext)
        {
          // next ray (:892-912)
          a.rayO[slot] = make_float4(ctx.p.x, ctx.p.y, ctx.p.z, ctx.eps);
          a.rayD[slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
          if ((type & BsdfType::Continuation) == 0) { lastPdf = pdf; lastType = type; }
          if (m.mediaInterface && dot(wo, ctx.Ng) * dot(wi, ctx.Ng) < 0.f)  // material/material.h:84-95
            mediumId = (mediumId == m.mediumInt) ? 0 : m.mediumInt;
          a.TP[slot] = make_float4(T.x, T.y, T.z, lastPdf);
          wantExt = true;
        }
        fl = (fl & ~(PRT_F_LASTTYPE_MASK | (0xffff << PRT_F_MEDIUM_SHIFT))) | (lastType & PRT_F_LASTTYPE_MASK) |
             (mediumId << PRT_F_MEDIUM_SHIFT);
        if (!ext)
        {
          if (wantShadow) fl |= PRT_F_TERMINATE;  // accumulate after the shadow test
          else finalizePath(a, slot, L, hsh, classB);
        }
      }
      a.flags[slot] = fl;
      if (a.hash) a.hash[slot] = hsh;
    }
    int e = waveAppend(cntOut, wantExt);
    if (wantExt) qOut[e] = slot;
    int s = waveAppend(cntShadow, wantShadow);
    if (wantShadow) a.shadowQ[s] = slot;
  }
}

// One shadow ray of the queue (path_renderer_stream.h:951-1065).  kFilter: 0 opaque scene, 1
// transparent materials, 2 transparent materials and kind-5 materials; kInst: a two-level scene.
template <int kFilter, bool kInst>
PRT_DEV void occludedRay(const RenderArgs& a, int i, ShadowStack& st, uint32_t& nodes, uint32_t& tris, uint32_t& rays, bool classB)
{
  {
    int slot = ldg(a.shadowQ + i);
    float4 so = ldg(a.shO + slot), sd = ldg(a.shD + slot);
    float t, u, v;
    const f3 o(so.x, so.y, so.z), d(sd.x, sd.y, sd.z);
    int prim;
    f3 Tsh(1.f);
    uint32_t crossed = 0;
    int inst;
    if (kFilter)
    {
      ShadowFilterT<(kFilter >= 2)> flt{&a, o, d, f3(1.f), 0u};
      prim = kInst ? traceRayInst<true>(a.inst, o, d, so.w, sd.w, t, u, v, inst, st, nodes, tris, flt)
                   : traceRay<true>(a.bvh, o, d, so.w, sd.w, t, u, v, st, nodes, tris, flt);
      Tsh = flt.T;
      crossed = flt.crossed;
    }
    else
      prim = kInst ? traceRayInst<true>(a.inst, o, d, so.w, sd.w, t, u, v, inst, st, nodes, tris)
                   : traceRay<true>(a.bvh, o, d, so.w, sd.w, t, u, v, st, nodes, tris);
    ++rays;
    float4 lr = ldg(a.Lr + slot);
    f3 L(lr.x, lr.y, lr.z);
    if (prim < 0)  // unoccluded: add the direct light times the shadow transmittance (:991-1002)
    {
      float4 dl = ldg(a.dirL + slot);
      L = L + (kFilter ? f3(dl.x, dl.y, dl.z) * Tsh : f3(dl.x, dl.y, dl.z));
      stg(a.Lr + slot, make_float4(L.x, L.y, L.z, 0.f));
    }
    uint32_t hsh = 0u;
    if (a.hash)  // the visibility event is part of the path's identity
    {
      hsh = hashMix(a.hash[slot], prim < 0 ? 1u + 4u * crossed : 2u);  // + the transparent surfaces crossed
      a.hash[slot] = hsh;
    }
    if (a.flags[slot] & PRT_F_TERMINATE) finalizePath(a, slot, L, hsh, classB);
  }
}

// Shadow rays of depth `depth`.  kFilter: the scene has transparent materials, and the shadow
// rays gather the transmittance of the transparent surfaces they cross (path_renderer_stream.h:
// 951-1002 with the TransparentShadowFilter of render/transparent_filter.h); 2: some of them are
// kind-5 materials.
template <int kFilter, bool kCount, bool kInst>
__global__ void __launch_bounds__(kTraceBlock) kOccluded(const RenderArgs* __restrict__ ap, int depth, bool classB)
{
  const RenderArgs& a = *ap;
  __shared__ ShadowLds<kTraceBlock> lds;
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  ShadowStack st = makeShadowStack(lds, a.spill, a.spillStride, gtid);
  const int count = a.counters[2 * depth + 1];
  uint32_t nodes = 0, tris = 0, rays = 0, waveSteps = 0;
  const int stride = gridDim.x * blockDim.x;
  if (!kCount)
    for (int i = gtid; i < count; i += stride) occludedRay<kFilter, kInst>(a, i, st, nodes, tris, rays, classB);
  else
    for (int base = gtid - laneId(); base < count; base += stride)
    {
      const uint32_t before = st.steps;
      if (base + laneId() < count) occludedRay<kFilter, kInst>(a, base + laneId(), st, nodes, tris, rays, classB);
      waveSteps += waveMax(st.steps - before);
    }
  unsigned long long r = waveSum(rays), nv = waveSum(nodes), tt = waveSum(tris);
  if (laneId() == 0 && r)
  {
    atomicAdd(&a.stats[3], r);
    atomicAdd(&a.stats[4], nv);
    atomicAdd(&a.stats[5], tt);
  }
  if (kCount)
  {
    unsigned long long ls = waveSum(st.steps);
    if (laneId() == 0 && r)
    {
      atomicAdd(&a.stats[8], ls);
      atomicAdd(&a.stats[9], (unsigned long long)waveSteps);
    }
  }
}

__global__ void __launch_bounds__(kTraceBlock) kQuery(BvhView bvh, int n, const float* __restrict__ rays, int* __restrict__ prim,
                                                      float* __restrict__ tOut, float* __restrict__ uOut, float* __restrict__ vOut,
                                                      int anyHit, int* spill, int spillStride, int spillCap)
{
  __shared__ TraceLds<kTraceBlock> lds;
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  TraceStack st = makeStack(lds, spill, spillStride, gtid, spillCap);
  uint32_t nodes = 0, tris = 0;
  for (int i = gtid; i < n; i += gridDim.x * blockDim.x)
  {
    const float* r = rays + 8 * i;
    f3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
    float t, u = 0.f, v = 0.f;
    int p = anyHit ? traceRay<true>(bvh, o, d, r[6], r[7], t, u, v, st, nodes, tris)
                   : traceRay<false>(bvh, o, d, r[6], r[7], t, u, v, st, nodes, tris);
    prim[i] = p;
    if (tOut) tOut[i] = p >= 0 ? t : r[7];
    if (uOut) uOut[i] = u;
    if (vOut) vOut[i] = v;
  }
}

// kQuery on a two-level scene; inst (may be null) receives the hit's instance.
__global__ void __launch_bounds__(kTraceBlock) kQueryInst(InstView iv, int n, const float* __restrict__ rays, int* __restrict__ prim,
                                                          float* __restrict__ tOut, float* __restrict__ uOut,
                                                          float* __restrict__ vOut, int* __restrict__ instOut, int anyHit,
                                                          int* spill, int spillStride, int spillCap)
{
  __shared__ TraceLds<kTraceBlock> lds;
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  TraceStack st = makeStack(lds, spill, spillStride, gtid, spillCap);
  uint32_t nodes = 0, tris = 0;
  for (int i = gtid; i < n; i += gridDim.x * blockDim.x)
  {
    const float* r = rays + 8 * i;
    f3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
    float t, u = 0.f, v = 0.f;
    int inst = -1;
    int p = anyHit ? traceRayInst<true>(iv, o, d, r[6], r[7], t, u, v, inst, st, nodes, tris)
                   : traceRayInst<false>(iv, o, d, r[6], r[7], t, u, v, inst, st, nodes, tris);
    prim[i] = p;
    if (tOut) tOut[i] = p >= 0 ? t : r[7];
    if (uOut) uOut[i] = u;
    if (vOut) vOut[i] = v;
    if (instOut) instOut[i] = p >= 0 ? inst : -1;
  }
}

// Shadow-style occlusion queries in a scene with transparent materials: the transparent-shadow
// filter decides (render/renderer_stream.h:64-97).
template <bool kInst>
__global__ void __launch_bounds__(kTraceBlock) kQueryOcclusionFilter(const RenderArgs* __restrict__ ap, int n,
                                                                     const float* __restrict__ rays, int* __restrict__ occ)
{
  const RenderArgs& a = *ap;
  __shared__ TraceLds<kTraceBlock> lds;
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  TraceStack st = makeStack(lds, a.spill, a.spillStride, gtid);
  uint32_t nodes = 0, tris = 0;
  for (int i = gtid; i < n; i += gridDim.x * blockDim.x)
  {
    const float* r = rays + 8 * i;
    f3 o(r[0], r[1], r[2]), d(r[3], r[4], r[5]);
    float t, u, v;
    ShadowFilter flt{&a, o, d, f3(1.f), 0u};
    int inst;
    occ[i] = (kInst ? traceRayInst<true>(a.inst, o, d, r[6], r[7], t, u, v, inst, st, nodes, tris, flt)
                    : traceRay<true>(a.bvh, o, d, r[6], r[7], t, u, v, st, nodes, tris, flt)) >= 0 ? 1 : 0;
  }
}

// Renderer::queryPixel (render/renderer.h:82-115): one lane walks the hits of the pixel's ray;
// with ray8 (origin, direction, near, far) the hits of that ray instead (RendererStream::queryRay,
// render/renderer_stream.h:16-62), their depth being the distance along the ray.
template <bool kInst>
__global__ void __launch_bounds__(64) kQueryPixel(const RenderArgs* __restrict__ ap, float px, float py, int maxHits,
                                                  prt_pixel_hit* __restrict__ out, int* __restrict__ count,
                                                  const float* __restrict__ ray8)
{
  const RenderArgs& a = *ap;
  __shared__ TraceLds<64> lds;
  TraceStack st = makeStack(lds, a.spill, a.spillStride, threadIdx.x);
  if (threadIdx.x != 0) return;
  // camera ray with the lens sample at the centre (pinhole_camera.h:128-133, thin_lens_camera.h:175-182)
  auto ray = [&](float x, float y, f3& d, float& tnear) {
    d = ld3(a.imageO) + x * ld3(a.imageDx) + y * ld3(a.imageDy);
    if (!a.pinhole) d = a.focal * d;
    float len = length(d);
    d = d * (1.0f / len);
    tnear = (a.pinhole ? a.nearClip : a.nearClipScale) * len;
  };
  f3 dc, d, o;
  float tc, tnear, tfar = kPosMax;
  if (ray8)
  {
    o = f3(ray8[0], ray8[1], ray8[2]);
    d = f3(ray8[3], ray8[4], ray8[5]);
    dc = d;
    tnear = ray8[6];
    tfar = ray8[7];
  }
  else
  {
    ray(0.5f, 0.5f, dc, tc);
    ray(fminf(fmaxf(px, 0.f), 1.f), fminf(fmaxf(py, 0.f), 1.f), d, tnear);
    o = ld3(a.camPos);
  }
  uint32_t nodes = 0, tris = 0;
  int n = 0;
  for (; n < maxHits; ++n)
  {
    float t, u = 0.f, v = 0.f;
    int inst = -1;
    const int prim = kInst ? traceRayInst<false>(a.inst, o, d, tnear, tfar, t, u, v, inst, st, nodes, tris)
                           : traceRay<false>(a.bvh, o, d, tnear, tfar, t, u, v, st, nodes, tris);
    if (prim < 0) break;
    const Ctx c = postIntersect(a, o, d, t, prim, u, v, inst);
    const int matId = a.triMat[prim];
    const prt_material_constants& m = a.mats[matId];
    f3 T(0.f);
    if (m.transparent) T = bsdfTransparency(a.lut, makeBsdf(m, c, -d, false), -d);
    prt_pixel_hit& h = out[n];
    h.prim = prim;
    h.material = matId;
    h.material_type = (m.transparent ? 1 : 0) | (m.transmissive ? 4 : 0) | (m.mediaInterface ? 8 : 0) | (m.emissive ? 16 : 0);
    const f3 vals[6] = {c.p, c.Ng, c.f.N, c.f.U, c.f.V, T};
    float* dst[6] = {h.p, h.Ng, h.N, h.U, h.V, h.T};
    for (int k = 0; k < 6; ++k) { dst[k][0] = vals[k].x; dst[k][1] = vals[k].y; dst[k][2] = vals[k].z; }
    h.uv[0] = u;
    h.uv[1] = v;
    h.eps = c.eps;
    h.dist = t;
    h.depth = t * dot(d, dc);
    if (!(reduceMax(T) >= 1e-5f)) { ++n; break; }
    tnear = t + c.eps * 16.f;  // the next hit (renderer.h:111-113)
  }
  *count = n;
}

__global__ void kDebugSampler(int n, const uint32_t* __restrict__ in, int ndim, float* o1, float* o2, float* o3)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Sampler s;
  s.set(in[3 * i], in[3 * i + 2], in[3 * i + 1]);
  for (int d = 0; d < ndim; ++d)
  {
    o1[i * ndim + d] = s.get1D(d);
    s.get2D(d, o2[(i * ndim + d) * 2], o2[(i * ndim + d) * 2 + 1]);
    s.get3D(d, o3[(i * ndim + d) * 3], o3[(i * ndim + d) * 3 + 1], o3[(i * ndim + d) * 3 + 2]);
  }
}

}  // namespace

void launchGenerate(const RenderArgs* a, int npix, uint32_t sampleIndex, hipStream_t s)
{
  kGenerate<<<(npix + 255) / 256, 256, 0, s>>>(a, sampleIndex);
}
void launchIntersect(const RenderArgs* a, int depth, const LaunchCfg& c, bool count, bool inst, hipStream_t s)
{
  if (inst)
  {
    if (count) kIntersect<true, true><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth);
    else kIntersect<false, true><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth);
  }
  else if (count) kIntersect<true, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth);
  else kIntersect<false, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth);
}
void launchShade(const RenderArgs* a, int depth, uint32_t sampleIndex, bool classB, const LaunchCfg& c, int level, hipStream_t s)
{
  if (level >= 2) kShade<2><<<c.gridShade, kShadeBlock, 0, s>>>(a, depth, sampleIndex, classB);
  else if (level == 1) kShade<1><<<c.gridShade, kShadeBlock, 0, s>>>(a, depth, sampleIndex, classB);
  else kShade<0><<<c.gridShade, kShadeBlock, 0, s>>>(a, depth, sampleIndex, classB);
}
void launchOccluded(const RenderArgs* a, int depth, bool classB, const LaunchCfg& c, int filter, bool count, bool inst,
                    hipStream_t s)
{
  if (inst)  // two-level scenes: no counting build
  {
    if (filter >= 2) kOccluded<2, false, true><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
    else if (filter == 1) kOccluded<1, false, true><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
    else kOccluded<0, false, true><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
    return;
  }
  if (count)
  {
    if (filter >= 2) kOccluded<2, true, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
    else if (filter == 1) kOccluded<1, true, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
    else kOccluded<0, true, false><<<c.gridOccluded, kTraceBlock, 0, s>>>(a, depth, classB);
  }
  else if (filter >= 2) kOccluded<2, false, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
  else if (filter == 1) kOccluded<1, false, false><<<c.gridTrace, kTraceBlock, 0, s>>>(a, depth, classB);
  else kOccluded<0, false, false><<<c.gridOccluded, kTraceBlock, 0, s>>>(a, depth, classB);
}
void launchQuery(const BvhView& bvh, int n, const float* rays, int* prim, float* t, float* u, float* v, int anyHit, int* spill,
                 int spillStride, int spillCap, int grid, hipStream_t s)
{
  kQuery<<<grid, kTraceBlock, 0, s>>>(bvh, n, rays, prim, t, u, v, anyHit, spill, spillStride, spillCap);
}
void launchQueryInst(const InstView& iv, int n, const float* rays, int* prim, float* t, float* u, float* v, int* inst,
                     int anyHit, int* spill, int spillStride, int spillCap, int grid, hipStream_t s)
{
  kQueryInst<<<grid, kTraceBlock, 0, s>>>(iv, n, rays, prim, t, u, v, inst, anyHit, spill, spillStride, spillCap);
}
void launchQueryOcclusionFilter(const RenderArgs* a, int n, const float* rays, int* occ, int grid, bool inst, hipStream_t s)
{
  if (inst) kQueryOcclusionFilter<true><<<grid, kTraceBlock, 0, s>>>(a, n, rays, occ);
  else kQueryOcclusionFilter<false><<<grid, kTraceBlock, 0, s>>>(a, n, rays, occ);
}
void launchQueryPixel(const RenderArgs* a, float px, float py, int maxHits, prt_pixel_hit* out, int* count, const float* ray8,
                      bool inst, hipStream_t s)
{
  if (inst) kQueryPixel<true><<<1, 64, 0, s>>>(a, px, py, maxHits, out, count, ray8);
  else kQueryPixel<false><<<1, 64, 0, s>>>(a, px, py, maxHits, out, count, ray8);
}
void launchDebugSampler(int n, const uint32_t* in, int ndim, float* o1, float* o2, float* o3, hipStream_t s)
{
  kDebugSampler<<<(n + 63) / 64, 64, 0, s>>>(n, in, ndim, o1, o2, o3);
}

}  // namespace prt
