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
// Sobol-Burley seed twice (tracePathMega, which makes the wavefront kernels' decisions), first to
// recover its radiance, then to accumulate the gradient at every vertex.
//
// Reduction: each lane records (material, 4 partials) per vertex in LDS; the block sums them per
// parameter in a fixed order and adds the result to its own slice of a partial buffer; a final
// kernel sums the slices in block order.  The result is deterministic.
#include "mega.h"
#include "grad.h"
#include <algorithm>

namespace prt {

namespace {

constexpr int kAdjBlock = 256;

// kInst: a two-level scene (the replays trace through its top-level and mesh trees).
template <bool kInst>
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
  // path vertices are recorded PRT_ADJ_MAX_VERTICES depths at a time: a deeper path is replayed
  // once more per further group of depths (shading happens at depths 0 .. maxDepth - 1)
  const int groups = max(1, (a.maxDepth + PRT_ADJ_MAX_VERTICES - 1) / PRT_ADJ_MAX_VERTICES);
  for (int base = blockIdx.x * kAdjBlock; base < a.npix; base += gridDim.x * kAdjBlock)
  {
  const int slot = base + threadIdx.x;
  int x = 0, y = 0, pix = 0;
  f3 dl(0.f), Ltotal(0.f);
  if (slot < a.npix)
  {
    slotToPixel(slot, a.W, x, y);
    pix = y * a.W + x;
    dl = f3(delta[pix] * invSpp, delta[a.npix + pix] * invSpp, delta[2 * a.npix + pix] * invSpp);
    // the first replay recovers the sample's radiance; the next ones replay the same path with
    // the gradient hook
    Ltotal = tracePathMega<true, kInst>(a, st, x, y, sampleIndex, nullptr, NoVertexHook(), nodes, tris);
    if (replayL)
    {
      replayL[pix] = Ltotal.x; replayL[a.npix + pix] = Ltotal.y; replayL[2 * a.npix + pix] = Ltotal.z;
    }
  }
  for (int group = 0; group < groups && nParams > 0; ++group)
  {
    const int v0 = group * PRT_ADJ_MAX_VERTICES;
    for (int v = 0; v < PRT_ADJ_MAX_VERTICES; ++v) recMat[v][threadIdx.x] = -1;
    if (slot < a.npix)
    {
      auto hook = [&](const PathVertex& pv) {
        if (pv.depth < v0 || pv.depth >= v0 + PRT_ADJ_MAX_VERTICES) return;
        float4 gv = make_float4(0.f, 0.f, 0.f, 0.f);
        auto add = [&](f3 dir, f3 radiance) {
          BsdfGrad g = bsdfEvalGrad(a.lut, *pv.b, *pv.m, *pv.ctx, pv.wo, dir);
          f3 wgt = dl * radiance;
          if (g.value.x > 0.f) { gv.x += wgt.x * g.dAlbedo.x / g.value.x; gv.w += wgt.x * g.dAlpha.x / g.value.x; }
          if (g.value.y > 0.f) { gv.y += wgt.y * g.dAlbedo.y / g.value.y; gv.w += wgt.y * g.dAlpha.y / g.value.y; }
          if (g.value.z > 0.f) { gv.z += wgt.z * g.dAlbedo.z / g.value.z; gv.w += wgt.z * g.dAlpha.z / g.value.z; }
        };
        if (pv.nee) add(pv.wiL, pv.dirL);           // the next-event estimate made here
        // everything gathered after it, through a smooth lobe (a delta or continuation lobe's
        // weight holds no differentiated parameter)
        if (pv.ext && !(pv.type & BsdfType::Delta)) add(pv.wi, Ltotal - pv.L);
        recMat[pv.depth - v0][threadIdx.x] = pv.matId;
        recG[pv.depth - v0][threadIdx.x] = gv;
      };
      tracePathMega<true, kInst>(a, st, x, y, sampleIndex, nullptr, hook, nodes, tris);
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
                   uint32_t sampleIndex, float* replayL, bool inst, hipStream_t s)
{
  if (inst) kAdjoint<true><<<blocks, kAdjBlock, 0, s>>>(a, delta, invSpp, nParams, partial, sampleIndex, replayL);
  else kAdjoint<false><<<blocks, kAdjBlock, 0, s>>>(a, delta, invSpp, nParams, partial, sampleIndex, replayL);
}

void launchAdjointReduce(const float* partial, int nBlocks, int nParams, float* grad, hipStream_t s)
{
  kAdjointReduce<<<(nParams + 255) / 256, 256, 0, s>>>(partial, nBlocks, nParams, grad);
}

}  // namespace prt
