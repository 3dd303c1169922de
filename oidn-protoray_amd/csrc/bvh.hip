// On-device BVH build for gfx950 (replaces the Embree build, accel/embree_intersector.cpp:11-90).
//
// Linear BVH (Karras 2012): 63-bit Morton codes of triangle centroids, a radix sort
// (hipcub/rocPRIM onesweep), one thread per inner node to find its key range and split,
// bottom-up bounding boxes with one atomic arrival counter per node, then a top-down collapse
// into 4-wide 128-byte nodes (greedy surface-area opening).  Subtrees covering <= PRT_LEAF_MAX
// consecutive (Morton-ordered) triangles become leaves, and the triangles are stored in that
// order as (v0, e1, e2) records for the traversal kernels.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <chrono>
#include "internal.h"

namespace prt {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void atomicMinF(float* a, float v)
{
  if (v >= 0.f) atomicMin(reinterpret_cast<int*>(a), __float_as_int(v));
  else atomicMax(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}
__device__ __forceinline__ void atomicMaxF(float* a, float v)
{
  if (v >= 0.f) atomicMax(reinterpret_cast<int*>(a), __float_as_int(v));
  else atomicMin(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}

__device__ __forceinline__ float3 ldPos(const float* pos, int i) { return make_float3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); }

// Triangle centroids and the centroid bounds (block reduction, then one atomic per block).
__global__ void __launch_bounds__(kBlock) kCentroids(const float* __restrict__ pos, const int* __restrict__ idx, int n,
                                                     float4* __restrict__ cent, float* __restrict__ cbox)
{
  __shared__ float red[6][kBlock];
  int t = blockIdx.x * kBlock + threadIdx.x;
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  if (t < n)
  {
    float3 a = ldPos(pos, idx[3 * t]), b = ldPos(pos, idx[3 * t + 1]), c = ldPos(pos, idx[3 * t + 2]);
    float cx = 0.5f * (fminf(fminf(a.x, b.x), c.x) + fmaxf(fmaxf(a.x, b.x), c.x));
    float cy = 0.5f * (fminf(fminf(a.y, b.y), c.y) + fmaxf(fmaxf(a.y, b.y), c.y));
    float cz = 0.5f * (fminf(fminf(a.z, b.z), c.z) + fmaxf(fmaxf(a.z, b.z), c.z));
    cent[t] = make_float4(cx, cy, cz, 0.f);
    lo[0] = hi[0] = cx; lo[1] = hi[1] = cy; lo[2] = hi[2] = cz;
  }
  for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = lo[k]; red[3 + k][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1)
  {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k)
      {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) { atomicMinF(&cbox[k], red[k][0]); atomicMaxF(&cbox[3 + k], red[3 + k][0]); }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ void __launch_bounds__(kBlock) kMorton(const float4* __restrict__ cent, int n, const float* __restrict__ cbox,
                                                  uint64_t* __restrict__ keys, int* __restrict__ vals)
{
  int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  float4 c = cent[t];
  float q[3] = {c.x, c.y, c.z};
  uint64_t m[3];
  for (int k = 0; k < 3; ++k)
  {
    float ext = cbox[3 + k] - cbox[k];
    float u = ext > 0.f ? (q[k] - cbox[k]) / ext : 0.f;
    float f = fminf(fmaxf(u * 2097152.0f, 0.f), 2097151.0f);
    m[k] = (uint64_t)f;
  }
  keys[t] = (spread21(m[0]) << 2) | (spread21(m[1]) << 1) | spread21(m[2]);
  vals[t] = t;
}

__device__ __forceinline__ int delta(const uint64_t* keys, int n, int i, int j)
{
  if (j < 0 || j >= n) return -1;
  uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz(uint32_t(i ^ j));
  return __clzll(a ^ b);
}

// Karras 2012, Fig. 4: inner node i covers [first, last] and splits at gamma.
// Child codes: >= 0 inner node, < 0 leaf (~sorted index).
__global__ void __launch_bounds__(kBlock) kHierarchy(const uint64_t* __restrict__ keys, int n, int* __restrict__ childL,
                                                     int* __restrict__ childR, int* __restrict__ parentInner,
                                                     int* __restrict__ parentLeaf, int2* __restrict__ range)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  do
  {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? d : 0);
  int lo = min(i, j), hi = max(i, j);
  int cl = (lo == gamma) ? ~gamma : gamma;
  int cr = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  childL[i] = cl;
  childR[i] = cr;
  range[i] = make_int2(lo, hi);
  if (cl < 0) parentLeaf[~cl] = i; else parentInner[cl] = i;
  if (cr < 0) parentLeaf[~cr] = i; else parentInner[cr] = i;
  if (i == 0) parentInner[0] = -1;
}

struct Box { float lo[3], hi[3]; };

__device__ __forceinline__ void storeBox(float* b, const Box& x)
{
  for (int k = 0; k < 3; ++k) { b[k] = x.lo[k]; b[3 + k] = x.hi[k]; }
}
// Loads that bypass the non-coherent per-CU cache: the box was written by another CU.
__device__ __forceinline__ Box loadBoxCoherent(const float* b)
{
  Box x;
  for (int k = 0; k < 3; ++k)
  {
    x.lo[k] = __int_as_float(__hip_atomic_load(reinterpret_cast<const int*>(b + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    x.hi[k] = __int_as_float(__hip_atomic_load(reinterpret_cast<const int*>(b + 3 + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  return x;
}

// Leaf boxes, then bottom-up unions: the second thread to arrive at a node computes it.
__global__ void __launch_bounds__(kBlock) kBoxes(const float* __restrict__ pos, const int* __restrict__ idx,
                                                 const int* __restrict__ vals, int n, const int* __restrict__ childL,
                                                 const int* __restrict__ childR, const int* __restrict__ parentInner,
                                                 const int* __restrict__ parentLeaf, int* __restrict__ arrive,
                                                 float* __restrict__ leafBox, float* __restrict__ innerBox)
{
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  int prim = vals[k];
  float3 a = ldPos(pos, idx[3 * prim]), b = ldPos(pos, idx[3 * prim + 1]), c = ldPos(pos, idx[3 * prim + 2]);
  Box x;
  x.lo[0] = fminf(fminf(a.x, b.x), c.x); x.hi[0] = fmaxf(fmaxf(a.x, b.x), c.x);
  x.lo[1] = fminf(fminf(a.y, b.y), c.y); x.hi[1] = fmaxf(fmaxf(a.y, b.y), c.y);
  x.lo[2] = fminf(fminf(a.z, b.z), c.z); x.hi[2] = fmaxf(fmaxf(a.z, b.z), c.z);
  storeBox(leafBox + 6 * k, x);
  if (n == 1) return;
  int p = parentLeaf[k];
  while (p >= 0)
  {
    __threadfence();
    if (atomicAdd(&arrive[p], 1) == 0) return;  // the sibling is not done yet
    __threadfence();
    int cl = childL[p], cr = childR[p];
    Box bl = loadBoxCoherent(cl < 0 ? leafBox + 6 * (~cl) : innerBox + 6 * cl);
    Box br = loadBoxCoherent(cr < 0 ? leafBox + 6 * (~cr) : innerBox + 6 * cr);
    for (int q = 0; q < 3; ++q) { x.lo[q] = fminf(bl.lo[q], br.lo[q]); x.hi[q] = fmaxf(bl.hi[q], br.hi[q]); }
    storeBox(innerBox + 6 * p, x);
    p = parentInner[p];
  }
}

__device__ __forceinline__ int leafCode(int first, int count) { return int(0x80000000u | (uint32_t(first) << 3) | uint32_t(count - 1)); }

// Pads a box by 2^-20 of its magnitude per axis so that rounding in the slab test can never
// cull a hit the triangle test reports.
__device__ __forceinline__ void padBox(Box& b)
{
  for (int k = 0; k < 3; ++k)
  {
    float m = fmaxf(fabsf(b.lo[k]), fabsf(b.hi[k])) * 9.5367431640625e-7f + 1e-30f;
    b.lo[k] -= m;
    b.hi[k] += m;
  }
}

__device__ __forceinline__ float halfArea(const float* b)
{
  float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

// Top-down collapse of the binary tree into 4-wide nodes, one level of the wide tree per
// launch.  Each wide node starts from the two children of its binary node and repeatedly
// opens the child with the largest surface area until it has four children or none can be
// opened; children covering <= PRT_LEAF_MAX triangles are leaves.
__global__ void __launch_bounds__(kBlock) kCollapse(const int2* __restrict__ frontier, int count, int2* __restrict__ next,
                                                    int* __restrict__ nextCount, int* __restrict__ nodeCounter,
                                                    const int* __restrict__ childL, const int* __restrict__ childR,
                                                    const int2* __restrict__ range, const float* __restrict__ leafBox,
                                                    const float* __restrict__ innerBox, prt_bvh4_node* __restrict__ nodes)
{
  int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= count) return;
  int2 e = frontier[f];
  int b = e.x, k = e.y;
  int codes[4];
  int n = 2;
  codes[0] = childL[b];
  codes[1] = childR[b];
  auto openable = [&](int c) { return c >= 0 && (range[c].y - range[c].x + 1) > PRT_LEAF_MAX; };
  while (n < 4)
  {
    int best = -1;
    float bestA = -1.f;
    for (int q = 0; q < n; ++q)
      if (openable(codes[q]))
      {
        float a = halfArea(innerBox + 6 * codes[q]);
        if (a > bestA) { bestA = a; best = q; }
      }
    if (best < 0) break;
    int c = codes[best];
    codes[best] = childL[c];
    codes[n++] = childR[c];
  }
  prt_bvh4_node nd;
  for (int q = 0; q < 4; ++q)
  {
    nd.pad[q] = 0;
    if (q >= n)
    {
      nd.lox[q] = nd.loy[q] = nd.loz[q] = 3.4e38f;
      nd.hix[q] = nd.hiy[q] = nd.hiz[q] = -3.4e38f;
      nd.child[q] = -1;  // empty slot: never visited (traversal masks child == -1)
      continue;
    }
    int c = codes[q];
    Box bb;
    const float* src = c < 0 ? leafBox + 6 * (~c) : innerBox + 6 * c;
    for (int a = 0; a < 3; ++a) { bb.lo[a] = src[a]; bb.hi[a] = src[3 + a]; }
    padBox(bb);
    nd.lox[q] = bb.lo[0]; nd.hix[q] = bb.hi[0];
    nd.loy[q] = bb.lo[1]; nd.hiy[q] = bb.hi[1];
    nd.loz[q] = bb.lo[2]; nd.hiz[q] = bb.hi[2];
    if (c < 0)
      nd.child[q] = leafCode(~c, 1);
    else if (!openable(c))
    {
      int2 r = range[c];
      nd.child[q] = leafCode(r.x, r.y - r.x + 1);
    }
    else
    {
      int j = atomicAdd(nodeCounter, 1);
      next[atomicAdd(nextCount, 1)] = make_int2(c, j);
      nd.child[q] = j;
    }
  }
  nodes[k] = nd;
}

// Triangle records in leaf order: (v0, prim), (e1 = v1 - v0), (e2 = v2 - v0).
__global__ void __launch_bounds__(kBlock) kTris(const float* __restrict__ pos, const int* __restrict__ idx,
                                                const int* __restrict__ vals, int n, float4* __restrict__ tris)
{
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  int prim = vals[k];
  float3 a = ldPos(pos, idx[3 * prim]), b = ldPos(pos, idx[3 * prim + 1]), c = ldPos(pos, idx[3 * prim + 2]);
  tris[3 * k + 0] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
  tris[3 * k + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
  tris[3 * k + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.f);
}

inline unsigned blocks(long long n) { return unsigned((n + kBlock - 1) / kBlock); }

}  // namespace

#define BVH_CHECK(x)                                                  \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) { err = std::string(#x ": ") + hipGetErrorString(e_); return e_; } \
  } while (0)

hipError_t buildBvh(const float* dPos, const int* dIdx, int n, BvhData& out, hipStream_t stream, std::string& err)
{
  freeBvh(out);
  auto t0 = std::chrono::high_resolution_clock::now();
  out.ntris = n;
  if (n <= 0)
  {
    // empty scene: one degenerate triangle (det == 0, never hit) under a single leaf
    BVH_CHECK(hipMalloc(&out.tris, 3 * sizeof(float4)));
    BVH_CHECK(hipMemsetAsync(out.tris, 0, 3 * sizeof(float4), stream));
    BVH_CHECK(hipMalloc(&out.nodes, sizeof(prt_bvh4_node)));
    out.root = int(0x80000000u);
    out.nodeCount = 0;
    out.maxDepth = 1;
    BVH_CHECK(hipStreamSynchronize(stream));
    return hipSuccess;
  }
  float4* cent = nullptr;
  float* cbox = nullptr;
  uint64_t *keys = nullptr, *keys2 = nullptr;
  int *vals = nullptr, *vals2 = nullptr, *childL = nullptr, *childR = nullptr, *parentInner = nullptr, *parentLeaf = nullptr;
  int *arrive = nullptr, *counts = nullptr;
  int2 *range = nullptr, *front0 = nullptr, *front1 = nullptr;
  float *leafBox = nullptr, *innerBox = nullptr;
  void* temp = nullptr;
  size_t tempBytes = 0;
  int inner = n > 1 ? n - 1 : 1;
  BVH_CHECK(hipMalloc(&cent, sizeof(float4) * n));
  BVH_CHECK(hipMalloc(&cbox, sizeof(float) * 8));
  BVH_CHECK(hipMalloc(&keys, sizeof(uint64_t) * n));
  BVH_CHECK(hipMalloc(&keys2, sizeof(uint64_t) * n));
  BVH_CHECK(hipMalloc(&vals, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&vals2, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&childL, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&childR, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&parentInner, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&parentLeaf, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&arrive, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&range, sizeof(int2) * inner));
  BVH_CHECK(hipMalloc(&front0, sizeof(int2) * inner));
  BVH_CHECK(hipMalloc(&front1, sizeof(int2) * inner));
  BVH_CHECK(hipMalloc(&counts, sizeof(int) * 2));
  BVH_CHECK(hipMalloc(&leafBox, sizeof(float) * 6 * size_t(n)));
  BVH_CHECK(hipMalloc(&innerBox, sizeof(float) * 6 * size_t(inner)));
  BVH_CHECK(hipMalloc(&out.nodes, sizeof(prt_bvh4_node) * inner));
  BVH_CHECK(hipMalloc(&out.tris, sizeof(float4) * 3 * size_t(n)));

  const float initBox[8] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f, 0.f, 0.f};
  BVH_CHECK(hipMemcpyAsync(cbox, initBox, sizeof(initBox), hipMemcpyHostToDevice, stream));
  BVH_CHECK(hipMemsetAsync(arrive, 0, sizeof(int) * inner, stream));
  kCentroids<<<blocks(n), kBlock, 0, stream>>>(dPos, dIdx, n, cent, cbox);
  kMorton<<<blocks(n), kBlock, 0, stream>>>(cent, n, cbox, keys, vals);
  BVH_CHECK(hipGetLastError());
  BVH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, keys, keys2, vals, vals2, n, 0, 63, stream));
  BVH_CHECK(hipMalloc(&temp, tempBytes));
  BVH_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, keys, keys2, vals, vals2, n, 0, 63, stream));
  if (n > 1)
    kHierarchy<<<blocks(n - 1), kBlock, 0, stream>>>(keys2, n, childL, childR, parentInner, parentLeaf, range);
  kBoxes<<<blocks(n), kBlock, 0, stream>>>(dPos, dIdx, vals2, n, childL, childR, parentInner, parentLeaf, arrive, leafBox, innerBox);
  kTris<<<blocks(n), kBlock, 0, stream>>>(dPos, dIdx, vals2, n, out.tris);
  BVH_CHECK(hipGetLastError());
  int levels = 1, wideNodes = 0;
  if (n > 1)
  {
    // level-by-level collapse into 4-wide nodes; wide node 0 = binary root 0
    int2 root = make_int2(0, 0);
    int one = 1;
    BVH_CHECK(hipMemcpyAsync(front0, &root, sizeof(int2), hipMemcpyHostToDevice, stream));
    BVH_CHECK(hipMemcpyAsync(counts, &one, sizeof(int), hipMemcpyHostToDevice, stream));
    int count = 1;
    levels = 0;
    while (count > 0)
    {
      BVH_CHECK(hipMemsetAsync(counts + 1, 0, sizeof(int), stream));
      kCollapse<<<blocks(count), kBlock, 0, stream>>>(front0, count, front1, counts + 1, counts, childL, childR, range, leafBox,
                                                      innerBox, out.nodes);
      BVH_CHECK(hipGetLastError());
      int h[2] = {0, 0};
      BVH_CHECK(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, stream));
      BVH_CHECK(hipStreamSynchronize(stream));
      wideNodes = h[0];
      count = h[1];
      std::swap(front0, front1);
      ++levels;
    }
  }
  BVH_CHECK(hipStreamSynchronize(stream));
  out.root = n > 1 ? 0 : int(0x80000000u);
  out.nodeCount = n > 1 ? wideNodes : 0;
  out.maxDepth = levels;
  hipFree(cent); hipFree(cbox); hipFree(keys); hipFree(keys2); hipFree(vals); hipFree(vals2);
  hipFree(childL); hipFree(childR); hipFree(parentInner); hipFree(parentLeaf); hipFree(arrive); hipFree(range);
  hipFree(front0); hipFree(front1); hipFree(counts); hipFree(leafBox); hipFree(innerBox); hipFree(temp);
  auto t1 = std::chrono::high_resolution_clock::now();
  out.buildMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (3 * out.maxDepth >= PRT_MAX_STACK)  // a wide node pushes at most three children
  {
    err = "BVH deeper than the traversal stack (" + std::to_string(out.maxDepth) + " levels)";
    return hipErrorInvalidValue;
  }
  return hipSuccess;
}

void freeBvh(BvhData& b)
{
  if (b.nodes) hipFree(b.nodes);
  if (b.tris) hipFree(b.tris);
  b.nodes = nullptr;
  b.tris = nullptr;
}

}  // namespace prt
