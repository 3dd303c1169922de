// Path-replay adjoint of the stream path tracer (differentiable rendering; no counterpart in
// the reference).
//
// Loss: mean over pixels and channels of (I - target)^2, I = the spp-sample image.  With the
// per-pixel adjoint delta = dLoss/dI, the gradient of a material parameter theta is
//   sum over samples (1/spp) sum over path vertices v with that material, over channels c, of
//     delta_c * [ C_nee,v,c * df_c(wo, wl)/dtheta / f_c(wo, wl) + S_v,c * df_c(wo, wi)/dtheta / f_c(wo, wi) ]
// where C_nee,v is the unoccluded next-event contribution made at v and S_v the radiance that
// arrives through the BSDF-sampled continuation of v (everything the path gathers after v's
// next-event estimate).  Sampling is detached (directions, MIS weights, Russian roulette and
// light picks are those of the primal pass), so this is the derivative of the primal estimator
// with its random numbers fixed.  Nothing is stored per path: each sample is replayed from the
// Sobol-Burley seed twice, first to recover its radiance, then to accumulate the gradient, with
// the same building blocks (path.h) the wavefront kernels use, so the replay follows the primal
// paths exactly.
//
// Reduction: each lane records (material, 4 partials) per vertex in LDS; the block sums them per
// parameter in a fixed order and adds the result to its own slice of a partial buffer; a final
// kernel sums the slices in block order.  The result is deterministic.
#include "path.h"
#include "grad.h"
#include <algorithm>

namespace prt {

namespace {

constexpr int kAdjBlock = 256;

__global__ void __launch_bounds__(kAdjBlock) kAdjoint(const RenderArgs* __restrict__ ap, const float* __restrict__ delta,
                                                      float invSpp, int nParams, float* __restrict__ partial,
                                                      uint32_t sampleIndex, float* __restrict__ replayL)
{
  const RenderArgs& a = *ap;
  __shared__ TraceLds<kAdjBlock> lds;
  __shared__ int recMat[PRT_ADJ_MAX_VERTICES][kAdjBlock];
  __shared__ float4 recG[PRT_ADJ_MAX_VERTICES][kAdjBlock];
  // the grid covers at most spillStride lanes (the traversal overflow area), so it strides
  TraceStack st = makeStack(lds, a.spill, a.spillStride, blockIdx.x * kAdjBlock + threadIdx.x);
  float* out = partial + size_t(blockIdx.x) * nParams;
  uint32_t nodes = 0, tris = 0;
  for (int base = blockIdx.x * kAdjBlock; base < a.npix; base += gridDim.x * kAdjBlock)
  {
  const int slot = base + threadIdx.x;
  for (int v = 0; v < PRT_ADJ_MAX_VERTICES; ++v) recMat[v][threadIdx.x] = -1;
  if (slot < a.npix)
  {
    int x, y;
    slotToPixel(slot, a.W, x, y);
    const int pix = y * a.W + x;
    const uint32_t pixelId = uint32_t(x) | (uint32_t(y) << 16);
    const f3 dl(delta[pix] * invSpp, delta[a.npix + pix] * invSpp, delta[2 * a.npix + pix] * invSpp);
    f3 Ltotal(0.f);
    for (int pass = 0; pass < 2; ++pass)
    {
      const bool grad = pass == 1;
      f3 o, d;
      float tnear;
      cameraRay(a, x, y, sampleIndex, o, d, tnear);
      f3 T(1.f), L(0.f);
      float lastPdf = 1.f;
      int lastType = BsdfType::SpecularReflection, mediumId = 0;
      bool aov1 = false, aov2 = false;
      for (int depth = 0;; ++depth)
      {
        const int sampleDim = 4 + 7 * depth;
        float th, u = 0.f, v = 0.f;
        const int prim = traceRay<false>(a.bvh, o, d, tnear, kPosMax, th, u, v, st, nodes, tris);
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
          break;
        }
        if (depth >= a.maxDepth) break;
        Ctx ctx = postIntersect(a, o, d, th, prim, u, v);
        const f3 wo = -d;
        const bool inside = (m.kind == PRT_KIND_DIELECTRIC) && (mediumId == m.mediumInt);
        Bsdf b = makeBsdf(m, ctx, wo, inside);
        Sampler smp;
        smp.set(a.seed, sampleIndex, pixelId);
        const bool mAov1 = !aov1;
        bool mAov2 = !aov2;
        const bool mRR = !(mAov1 || mAov2);
        // next-event estimate (:659-718) with its shadow ray (:951-1002)
        f3 dirL(0.f), wiL(0.f);
        bool nee = false;
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
              nee = traceRay<true>(a.bvh, ctx.p, wiL, ctx.eps, lightDist - ctx.eps, ts, us, vs, st, nodes, tris) < 0;
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
        if (mAov1) aov1 = true;
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
        // the shadow ray's radiance joins after the shading of this depth, as in the primal
        if (nee) L = L + dirL;
        if (grad && depth < PRT_ADJ_MAX_VERTICES && (m.kind == PRT_KIND_DIFFUSE || m.kind == PRT_KIND_COATED))
        {
          float4 gv = make_float4(0.f, 0.f, 0.f, 0.f);
          if (nee)
          {
            BsdfGrad g = bsdfEvalGrad(a.lut, b, ctx, wo, wiL);
            f3 wgt = dl * dirL;
            if (g.value.x > 0.f) { gv.x += wgt.x * g.dAlbedo.x / g.value.x; gv.w += wgt.x * g.dAlpha.x / g.value.x; }
            if (g.value.y > 0.f) { gv.y += wgt.y * g.dAlbedo.y / g.value.y; gv.w += wgt.y * g.dAlpha.y / g.value.y; }
            if (g.value.z > 0.f) { gv.z += wgt.z * g.dAlbedo.z / g.value.z; gv.w += wgt.z * g.dAlpha.z / g.value.z; }
          }
          if (ext)
          {
            BsdfGrad g = bsdfEvalGrad(a.lut, b, ctx, wo, wi);
            f3 wgt = dl * (Ltotal - L);
            if (g.value.x > 0.f) { gv.x += wgt.x * g.dAlbedo.x / g.value.x; gv.w += wgt.x * g.dAlpha.x / g.value.x; }
            if (g.value.y > 0.f) { gv.y += wgt.y * g.dAlbedo.y / g.value.y; gv.w += wgt.y * g.dAlpha.y / g.value.y; }
            if (g.value.z > 0.f) { gv.z += wgt.z * g.dAlbedo.z / g.value.z; gv.w += wgt.z * g.dAlpha.z / g.value.z; }
          }
          recMat[depth][threadIdx.x] = matId;
          recG[depth][threadIdx.x] = gv;
        }
        if (!ext) break;
        o = ctx.p;
        tnear = ctx.eps;
        d = wi;
        if ((type & BsdfType::Continuation) == 0) { lastPdf = pdf; lastType = type; }
        if (m.kind == PRT_KIND_DIELECTRIC && dot(wo, ctx.Ng) * dot(wi, ctx.Ng) < 0.f)
          mediumId = (mediumId == m.mediumInt) ? 0 : m.mediumInt;
      }
      Ltotal = L;
    }
    if (replayL)
    {
      replayL[pix] = Ltotal.x; replayL[a.npix + pix] = Ltotal.y; replayL[2 * a.npix + pix] = Ltotal.z;
    }
  }
  __syncthreads();
  // block reduction in a fixed order: parameter p = 4 * material + (r, g, b, alpha)
  for (int p = threadIdx.x; p < nParams; p += kAdjBlock)
  {
    const int mat = p >> 2, comp = p & 3;
    float sum = 0.f;
    for (int v = 0; v < PRT_ADJ_MAX_VERTICES; ++v)
      for (int t = 0; t < kAdjBlock; ++t)
        if (recMat[v][t] == mat)
        {
          float4 g = recG[v][t];
          sum += comp == 0 ? g.x : comp == 1 ? g.y : comp == 2 ? g.z : g.w;
        }
    out[p] += sum;
  }
  __syncthreads();
  }
}

__global__ void kAdjointReduce(const float* __restrict__ partial, int nBlocks, int nParams, float* __restrict__ grad)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nParams) return;
  float s = 0.f;
  for (int b = 0; b < nBlocks; ++b) s += partial[size_t(b) * nParams + p];
  grad[p] = s;
}

}  // namespace

// Blocks of the adjoint kernel: one per 256 pixels, at most one lane per column of the
// traversal overflow area (spillStride lanes).
int adjointBlocks(int npix, int spillStride) { return std::max(1, std::min((npix + kAdjBlock - 1) / kAdjBlock, spillStride / kAdjBlock)); }

void launchAdjoint(const RenderArgs* a, int blocks, const float* delta, float invSpp, int nParams, float* partial,
                   uint32_t sampleIndex, float* replayL, hipStream_t s)
{
  kAdjoint<<<blocks, kAdjBlock, 0, s>>>(a, delta, invSpp, nParams, partial, sampleIndex, replayL);
}

void launchAdjointReduce(const float* partial, int nBlocks, int nParams, float* grad, hipStream_t s)
{
  kAdjointReduce<<<(nParams + 255) / 256, 256, 0, s>>>(partial, nBlocks, nParams, grad);
}

}  // namespace prt
