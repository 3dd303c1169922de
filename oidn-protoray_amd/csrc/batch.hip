// Persistent batch renderer: many frames of uneven resolution, from several scenes, rendered in
// one launch (dataset generation; no counterpart in the reference, whose batch mode renders the
// frames one after another, main/render/render_window.cpp:630-698).
//
// The frames' 8x8 pixel tiles form one list (frame after frame, tiles row-major).  Each wave of
// the persistent grid takes tilesPerFetch consecutive tiles with one atomic on the batch counter
// and renders, per lane (one pixel of the tile), every sample of its frame in order with
// tracePathMega: the same decisions as the wavefront renderer, so a frame's pixels equal
// prt_render's for the same scene, camera, seed and sample count.  A tile belongs to one wave for
// all its samples, so the per-pixel sums live in registers and are written once, without atomics
// and in sample order; the snapshot means (e.g. after 1 and 4 samples) are written on the way.
#include "mega.h"

namespace prt {

namespace {

constexpr int kBatchBlock = kBatchBlockSize;

// kStd: some frame's scene holds kind-5 materials; kInst: some frame's scene is two-level (each
// frame then takes the traversal of its own scene).
template <bool kStd, bool kInst>
__global__ void __launch_bounds__(kBatchBlock) kBatch(const BatchFrameDev* __restrict__ frames, int nFrames, int nTiles,
                                                      int* __restrict__ counter, int tilesPerFetch, int* __restrict__ spill,
                                                      int spillStride, unsigned long long* __restrict__ samples)
{
  __shared__ TraceLds<kBatchBlock> lds;
  TraceStack st = makeStack(lds, spill, spillStride, blockIdx.x * kBatchBlock + threadIdx.x);
  const int lane = __lane_id();
  uint32_t nodes = 0, tris = 0;
  unsigned long long done = 0;
  for (;;)
  {
    int first = 0;
    if (lane == 0) first = atomicAdd(counter, tilesPerFetch);
    first = __shfl(first, 0);
    if (first >= nTiles) break;
    const int last = min(first + tilesPerFetch, nTiles);
    for (int tile = first; tile < last; ++tile)
    {
      // the frame holding this tile: the last frame whose first tile is <= tile
      int lo = 0, hi = nFrames - 1;
      while (lo < hi)
      {
        int mid = (lo + hi + 1) >> 1;
        if (frames[mid].tileBegin <= tile) lo = mid; else hi = mid - 1;
      }
      const BatchFrameDev& f = frames[lo];
      const RenderArgs& a = *f.args;
      const int lt = tile - f.tileBegin;
      const int x = (lt % f.tilesX) * 8 + (lane & 7), y = (lt / f.tilesX) * 8 + (lane >> 3);
      const int pix = y * f.W + x;
      f3 sum(0.f), alb(0.f), nrm(0.f);
      float dep = 0.f;
      int k = 0;
      FirstHit fh;
      const bool wantFirst = f.albedo1 || f.depth;
      for (int s = 0; s < f.spp; ++s)
      {
        f3 L = kInst && a.inst.nodes
                   ? tracePathMega<kStd, true>(a, st, x, y, uint32_t(s), wantFirst ? &fh : nullptr, NoVertexHook(), nodes, tris)
                   : tracePathMega<kStd, false>(a, st, x, y, uint32_t(s), wantFirst ? &fh : nullptr, NoVertexHook(), nodes, tris);
        sum = sum + clampL(L, a.maxRadiance);
        if (f.albedo1) { alb = alb + fh.albedo; nrm = nrm + fh.normal; }
        if (f.depth) dep += fh.depth;
        if (k < f.nSnap && s + 1 == f.snapSpp[k])
        {
          const float r = 1.0f / float(s + 1);
          float* o = f.snap + size_t(k) * 3 * f.npix;
          o[pix] = sum.x * r; o[f.npix + pix] = sum.y * r; o[2 * f.npix + pix] = sum.z * r;
          ++k;
        }
      }
      const float r = f.spp > 0 ? 1.0f / float(f.spp) : 0.f;
      f.color[pix] = sum.x * r; f.color[f.npix + pix] = sum.y * r; f.color[2 * f.npix + pix] = sum.z * r;
      if (f.albedo1)
      {
        f.albedo1[pix] = alb.x * r; f.albedo1[f.npix + pix] = alb.y * r; f.albedo1[2 * f.npix + pix] = alb.z * r;
        f.normal1[pix] = nrm.x * r; f.normal1[f.npix + pix] = nrm.y * r; f.normal1[2 * f.npix + pix] = nrm.z * r;
      }
      if (f.depth) f.depth[pix] = dep * r;
      done += f.spp;
    }
  }
  for (int off = 32; off > 0; off >>= 1) done += __shfl_down(done, off);
  if (lane == 0 && done) atomicAdd(samples, done);
}

// Clears the batch's tile counter and sample count (a kernel, not a memset, so the batch can be
// captured in a graph and replayed).
__global__ void kBatchReset(int* counter, unsigned long long* samples)
{
  *counter = 0;
  *samples = 0;
}

}  // namespace

void launchBatch(const BatchFrameDev* frames, int nFrames, int nTiles, int* counter, int tilesPerFetch, int grid, int* spill,
                 int spillStride, unsigned long long* samples, bool std, bool inst, hipStream_t s)
{
  kBatchReset<<<1, 1, 0, s>>>(counter, samples);
  if (std && inst)
    kBatch<true, true><<<grid, kBatchBlock, 0, s>>>(frames, nFrames, nTiles, counter, tilesPerFetch, spill, spillStride, samples);
  else if (std)
    kBatch<true, false><<<grid, kBatchBlock, 0, s>>>(frames, nFrames, nTiles, counter, tilesPerFetch, spill, spillStride, samples);
  else if (inst)
    kBatch<false, true><<<grid, kBatchBlock, 0, s>>>(frames, nFrames, nTiles, counter, tilesPerFetch, spill, spillStride, samples);
  else
    kBatch<false, false><<<grid, kBatchBlock, 0, s>>>(frames, nFrames, nTiles, counter, tilesPerFetch, spill, spillStride, samples);
}

}  // namespace prt
