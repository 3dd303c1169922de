// Internal interfaces between the host driver (context.hip), the BVH builder (bvh.hip) and
// the wavefront kernels (render.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <string>
#include "prt_types.h"
#include "dev_math.h"
#include "trace.h"

struct prt_pixel_hit;  // include/prt_hip.h
struct prt_context;
struct prt_batch;

namespace prt {

struct BvhData
{
  prt_bvh4_node* nodes = nullptr;
  float4* tris = nullptr;
  int root = 0;
  int ntris = 0;
  long long nodeCount = 0;  // 4-wide nodes
  int maxDepth = 0;         // levels of the 4-wide tree
  double buildMs = 0.0;
};

// Records the message returned by prt_last_error() and returns code (context.hip).
int setLastError(int code, const std::string& msg);

hipError_t buildBvh(const float* dPos, const int* dIdx, int ntris, BvhData& out, hipStream_t stream, std::string& err);
void freeBvh(BvhData& b);

// Everything a sample pass reads, passed by value to each kernel.
struct RenderArgs
{
  // scene
  const float* pos;
  const float* nrm;  // may be null
  const int* idx;
  const int* triMat;
  const prt_material_constants* mats;
  const float4* media;  // interior media (r, g, b, ior) by medium ID; ID 0 is clear
  const float* lut;
  // image textures: per-vertex UVs (set 0, 2 floats per vertex, null: the hit's barycentric
  // u, v stand in, geometry/triangle_mesh.cpp:131), texture records and their texels
  const float* uv;
  const prt_texture_dev* tex;
  const float4* texels;
  BvhView bvh;
  // two-level scenes: the top-level tree, instances and meshes (inst.nodes null: one level)
  InstView inst;
  int* hitInst;  // per path slot: the instance of the closest hit (two-level scenes only)
  int anyTransparent;  // some material lets shadow rays through (render/scene.cpp:452-462)
  prt_light_dev lights[PRT_MAX_LIGHTS];  // environment lights
  float lightCdf[PRT_MAX_LIGHTS + 2];    // power distribution over [area light,] environment lights
  int nLights;                           // environment lights
  int lightOffset;                       // 1 if the area light heads the light list
  // area light (light/triangle_light.h): emissive triangles as 4 float4 (v0 | r, v1 | g, v2 | b,
  // n = cross(v1 - v0, v2 - v0)), their power CDF, and primitive -> emissive index
  const float4* areaTris;
  const float* areaCdf;
  const int* areaMap;
  int areaCount;
  EnvView env;
  // camera (precomputed as PinholeCamera / ThinLensCamera constructors do)
  float camPos[3], imageO[3], imageDx[3], imageDy[3], camU[3], camV[3], camN[3];
  float lensRadius, focal, nearClip, nearClipScale;
  int pinhole;
  // pixel sample (path_renderer_stream.h:228-236): a fixed sub-pixel jitter for every pixel
  // (FrameBuffer::setJitter), else the sampler's dimension-set 0 value, warped by the pixel
  // filter's tabulated CDF when one is set (image/pixel_filter.h:30-66, PixelFilterMode::Sample)
  int jitter;
  float jitterX, jitterY;
  const float* filterCdf;  // filterSize + 1 entries, or null (box filter)
  int filterSize;
  float filterWidth;
  // frame
  int W, H, npix;
  float invW, invH;
  prt_frame_dev fb;
  // renderer
  uint32_t seed;
  int maxDepth, rrDepth;
  float minWeight, maxRadiance;
  // path state, indexed by path slot (8x8 pixel tiles, 64 consecutive slots per tile)
  float4* rayO;   // origin, tnear
  float4* rayD;   // direction, unused
  float4* hit;    // t, u, v, prim bits
  float4* TP;     // throughput, lastBsdfPdf
  float4* Lr;     // radiance, unused
  int* flags;     // lastBsdfType | AOV flags | direct/terminate flags | medium << 16
  uint32_t* hash; // per-path event hash (debug), may be null
  float4* shO;    // shadow ray origin, tnear
  float4* shD;    // shadow ray direction, tfar
  float4* dirL;   // unshadowed direct light
  int* queue[2];
  int* shadowQ;
  int* counters;  // [2*d] paths queued for depth d (d >= 1), [2*d+1] shadow rays of depth d
  int* spill;     // traversal stack overflow, 2 x (PRT_MAX_STACK - PRT_LDS_STACK) x spillStride (nodes, distances)
  int spillStride;
  // closest hit: [0] rays, [1] inner-node visits, [2] triangle tests; shadow: [3..5]; traversal
  // loop iterations summed over lanes and over waves (the longest lane's per call): closest hit
  // [6], [7], shadow [8], [9]
  unsigned long long* stats;
};

// One frame of a persistent batch (batch.hip): its scene and camera as kernel arguments, its
// outputs (planar [3][npix] means), its sample count and snapshot counts, and its place in the
// batch's tile list (8x8 tiles, row-major, tilesX per row).
#define PRT_BATCH_MAX_SNAPSHOTS 4
struct BatchFrameDev
{
  const RenderArgs* args;
  float* color;    // mean of all spp samples
  float* snap;     // [nSnap][3][npix]: means after snapSpp[k] samples
  float* albedo1;  // first-hit albedo / normal means (both null when not requested)
  float* normal1;
  float* depth;    // first-hit view-axis depth mean (1 x npix, null when not requested)
  int W, H, npix, spp, nSnap;
  int snapSpp[PRT_BATCH_MAX_SNAPSHOTS];
  int tileBegin, tilesX;
};
constexpr int kBatchBlockSize = 256;
void launchBatch(const BatchFrameDev* frames, int nFrames, int nTiles, int* counter, int tilesPerFetch, int grid, int* spill,
                 int spillStride, unsigned long long* samples, bool std, bool inst, hipStream_t s);

// Kernel launchers (render.hip)
struct LaunchCfg
{
  int gridTrace;  // blocks of the persistent traversal kernels
  int gridShade;  // blocks of the persistent shading kernel
  int gridOccluded;  // blocks of the opaque-scene shadow-ray kernel
};
// The kernels read RenderArgs from device memory (one copy per render call); passing the
// ~1 KB record by value would make every lane copy it to scratch.
void launchGenerate(const RenderArgs* a, int npix, uint32_t sampleIndex, hipStream_t s);
void launchIntersect(const RenderArgs* a, int depth, const LaunchCfg& c, bool count, bool inst, hipStream_t s);
// level: 0 the original lobes, 1 the extended lobes and coloured media, 2 kind-5 materials too;
// filter: 0 opaque scene, 1 transparent materials, 2 transparent materials and kind-5 materials
void launchShade(const RenderArgs* a, int depth, uint32_t sampleIndex, bool classB, const LaunchCfg& c, int level, hipStream_t s);
void launchOccluded(const RenderArgs* a, int depth, bool classB, const LaunchCfg& c, int filter, bool count, bool inst,
                    hipStream_t s);
void launchQueryOcclusionFilter(const RenderArgs* a, int n, const float* rays, int* occ, int grid, bool inst, hipStream_t s);
void launchQueryPixel(const RenderArgs* a, float px, float py, int maxHits, prt_pixel_hit* out, int* count, const float* ray8,
                      bool inst, hipStream_t s);
void launchQuery(const BvhView& bvh, int n, const float* rays, int* prim, float* t, float* u, float* v, int anyHit,
                 int* spill, int spillStride, int spillCap, int grid, hipStream_t s);
// The same on a two-level scene; inst (may be null) receives the hit's instance.
void launchQueryInst(const InstView& iv, int n, const float* rays, int* prim, float* t, float* u, float* v, int* inst,
                     int anyHit, int* spill, int spillStride, int spillCap, int grid, hipStream_t s);
// Path-replay adjoint (adjoint.hip): partial is [blocks][nParams] (zeroed by the caller before the
// first sample pass), grad [nParams]; delta is the per-pixel dLoss/dI (planar rgb).
int adjointBlocks(int npix, int spillStride);
void launchAdjoint(const RenderArgs* a, int blocks, const float* delta, float invSpp, int nParams, float* partial,
                   uint32_t sampleIndex, float* replayL, bool inst, hipStream_t s);
void launchAdjointReduce(const float* partial, int nBlocks, int nParams, float* grad, hipStream_t s);
void launchDebugSampler(int n, const uint32_t* in, int ndim, float* o1, float* o2, float* o3, hipStream_t s);

constexpr int kTraceBlock = 256;
constexpr int kStatCount = 12;
constexpr int kShadeBlock = 256;

// Planar FP32 images of a frame that rendering left on the device, as the denoiser reads them:
// [3][W*H] sums with [W*H] sample weights (the frame buffer of a context) or [3][W*H] means
// (weights null: a batch frame's outputs).  An auxiliary image is null when the frame has none.
struct FrameImages
{
  int device = 0, W = 0, H = 0;
  const float *color = nullptr, *colorW = nullptr;
  const float *albedo = nullptr, *albedoW = nullptr;
  const float *normal = nullptr, *normalW = nullptr;
  hipStream_t stream = nullptr;  // the stream the producing kernels were enqueued on
};
// A context's frame: colour and the first-hit albedo / normal AOVs (context.hip).
int contextFrameImages(prt_context* c, FrameImages* out);
// A batch frame: colour (buffer PRT_BATCH_COLOR) or a snapshot, with the frame's first-hit
// albedo / normal means (context.hip).
int batchFrameImages(prt_batch* b, int frame, int buffer, FrameImages* out);

// Slot <-> pixel mapping: slot = (tile * 64) + (ly * 8 + lx), tiles row-major over W/8 columns.
PRT_DEV void slotToPixel(int slot, int W, int& x, int& y)
{
  int tile = slot >> 6, within = slot & 63;
  int tilesX = W >> 3;
  x = (tile % tilesX) * 8 + (within & 7);
  y = (tile / tilesX) * 8 + (within >> 3);
}

}  // namespace prt
