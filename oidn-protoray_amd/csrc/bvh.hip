// On-device BVH build for gfx950 (replaces the Embree build, accel/embree_intersector.cpp:11-90).
//
// Locally-ordered clustering (PLOC, Meister and Bittner 2018): triangles sorted by 63-bit
// Morton codes of their centroids (hipcub/rocPRIM onesweep radix sort) start as clusters; each
// iteration finds every cluster's nearest neighbour within PRT_PLOC_RADIUS positions of the
// sorted cluster list (smallest surface area of the merged box, ties to the smaller index),
// merges mutual nearest neighbours into a new binary node and compacts the list with a prefix
// sum, until one cluster is left.  The binary tree is then collapsed top-down into 4-wide
// 128-byte nodes (greedy surface-area opening; the inner children of a node get consecutive
// node indices); subtrees of <= PRT_LEAF_MAX triangles become
// leaves whose triangles are written contiguously as (v0, e1, e2) records.
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

struct Box { float lo[3], hi[3]; };

// Neighbour window of the clustering.  Measured on the 1080p workload (4096 icospheres) with
// the greedy collapse: radius 2 / 3 / 4 / 6 / 8 / 16 / 32 -> 110.3 / 112.8 / 112.3 / 111.2 /
// 109.5 / 107.0 / 97.9 Msamples/s (22.4 / 21.3 / 21.5 / 21.9 / 22.7 / 24.4 / 26.9 node fetches
// per ray); with the cost-driven collapse, radius 3 / 4 / 6 -> closest-hit traversal 6.31 /
// 6.36 / 6.39 ms per pass (21.1 / 20.7 / 21.0 node fetches per ray).
#ifndef PRT_PLOC_RADIUS
#define PRT_PLOC_RADIUS 3
#endif
constexpr int kRadius = PRT_PLOC_RADIUS;

// Initial clusters: one per sorted triangle (leaf code ~k), with its box and count 1.
__global__ void __launch_bounds__(kBlock) kLeafInit(const float* __restrict__ pos, const int* __restrict__ idx,
                                                    const int* __restrict__ vals, int n, float* __restrict__ leafBox,
                                                    int* __restrict__ clus, float4* __restrict__ clo, float4* __restrict__ chi,
                                                    int* __restrict__ ccnt)
{
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  int prim = vals[k];
  float3 a = ldPos(pos, idx[3 * prim]), b = ldPos(pos, idx[3 * prim + 1]), c = ldPos(pos, idx[3 * prim + 2]);
  float4 lo = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z), 0.f);
  float4 hi = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z), 0.f);
  float* lb = leafBox + 6 * size_t(k);
  lb[0] = lo.x; lb[1] = lo.y; lb[2] = lo.z; lb[3] = hi.x; lb[4] = hi.y; lb[5] = hi.z;
  clus[k] = ~k;
  clo[k] = lo;
  chi[k] = hi;
  ccnt[k] = 1;
}

// Half the surface area of a box given by its corners (the same expression wherever a cost is
// computed and later recomputed, so the collapse's comparisons reproduce the clustering's).
__device__ __forceinline__ float boxHalfArea(float lx, float ly, float lz, float hx, float hy, float hz)
{
  float dx = hx - lx, dy = hy - ly, dz = hz - lz;
  return dx * dy + dy * dz + dz * dx;
}

// Cost model of the wide tree (surface-area heuristic): entering a wide node costs `node` per
// unit of its area, a leaf `tri` per triangle per unit of its area; leaves hold at most
// `leafMax` triangles.  C(x, j) is the least cost of the subtree of binary node x cut into at
// most j subtrees (each a leaf or a wide node); the collapse reads C to place the cuts.
struct CollapseCost
{
  float node, tri;
  int leafMax;
};
static_assert(PRT_LEAF_MAX >= 1 && PRT_LEAF_MAX <= PRT_LEAF_CAP, "leaf size beyond the leaf reference's count field");
inline CollapseCost collapseCost() { return CollapseCost{1.0f, 0.5f, PRT_LEAF_MAX}; }

// Sorted triangle k's leaf cost, the same for every j (a triangle is never cut).
__device__ __forceinline__ float4 subtreeCost(int c, const float4* cost, const float* leafBox, float tri)
{
  if (c >= 0) return cost[c];
  const float* b = leafBox + 6 * size_t(~c);
  const float a = tri * boxHalfArea(b[0], b[1], b[2], b[3], b[4], b[5]);
  return make_float4(a, a, a, a);
}

__device__ __forceinline__ float costAt(float4 c, int j) { return j == 1 ? c.x : j == 2 ? c.y : j == 3 ? c.z : c.w; }

// Least cost of the two children cut into at most j >= 2 subtrees together, and the left
// child's share k of the cuts (the smallest k on ties).
__device__ __forceinline__ float splitCost(float4 l, float4 r, int j, int& kBest)
{
  kBest = 1;
  float best = costAt(l, 1) + costAt(r, j - 1);
  for (int k = 2; k < j; ++k)
  {
    const float v = costAt(l, k) + costAt(r, j - k);
    if (v < best) { best = v; kBest = k; }
  }
  return best;
}

// Cost of a new binary node over children l, r with `cnt` triangles and half area `area`.
__device__ __forceinline__ float4 nodeCost(float4 l, float4 r, int cnt, float area, const CollapseCost& cc)
{
  int k;
  const float d2 = splitCost(l, r, 2, k), d3 = splitCost(l, r, 3, k), d4 = splitCost(l, r, 4, k);
  const float inner = cc.node * area + d4;
  const float leaf = cnt <= cc.leafMax ? cc.tri * float(cnt) * area : __int_as_float(0x7f800000);
  const float c1 = fminf(leaf, inner);
  return make_float4(c1, fminf(c1, d2), fminf(c1, d3), fminf(c1, d4));
}

__device__ __forceinline__ float mergedArea(float4 alo, float4 ahi, float4 blo, float4 bhi)
{
  float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return dx * dy + dy * dz + dz * dx;
}

// Nearest neighbour of every cluster among the kRadius clusters on either side; the block's
// window of boxes (plus halos) is staged in LDS.  Candidates are scanned in increasing index
// and only a strictly smaller area replaces the best, so ties go to the smaller index; the
// merged area is symmetric, so the globally smallest (area, pair) is always a mutual pair and
// every iteration merges at least one pair.
__global__ void __launch_bounds__(kBlock) kNearest(const float4* __restrict__ clo, const float4* __restrict__ chi, int m,
                                                   int* __restrict__ nn)
{
  __shared__ float4 sLo[kBlock + 2 * kRadius], sHi[kBlock + 2 * kRadius];
  const int base = blockIdx.x * kBlock - kRadius;
  for (int t = threadIdx.x; t < kBlock + 2 * kRadius; t += kBlock)
  {
    int g = base + t;
    if (g >= 0 && g < m) { sLo[t] = clo[g]; sHi[t] = chi[g]; }
  }
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const int li = threadIdx.x + kRadius;
  const float4 alo = sLo[li], ahi = sHi[li];
  float best = __int_as_float(0x7f800000);
  int bj = -1;
  const int dlo = max(-kRadius, -i), dhi = min(kRadius, m - 1 - i);
  for (int d = dlo; d <= dhi; ++d)
  {
    if (d == 0) continue;
    float a = mergedArea(alo, ahi, sLo[li + d], sHi[li + d]);
    if (a < best) { best = a; bj = i + d; }
  }
  if (bj < 0) bj = i + (i > 0 ? -1 : 1);  // non-finite boxes: fall back to a neighbour
  nn[i] = bj;
}

// Per cluster: bit 0 = survives the iteration (unmerged, or the lower index of a merged pair,
// which holds the new node), bit 32 = creates a new node.  An exclusive prefix sum of these
// words gives each survivor its compacted position (low half) and each new node its ID offset
// (high half); neither half can carry (both sums are below 2^32).
__global__ void __launch_bounds__(kBlock) kMergeFlags(const int* __restrict__ nn, int m, unsigned long long* __restrict__ flags)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  int j = nn[i];
  bool mutual = nn[j] == i;
  unsigned long long f = (!mutual || i < j) ? 1ull : 0ull;
  if (mutual && i < j) f |= 1ull << 32;
  flags[i] = f;
}

__global__ void kTotals(const unsigned long long* __restrict__ flags, const unsigned long long* __restrict__ scan, int m,
                        int* __restrict__ out)
{
  unsigned long long t = scan[m - 1] + flags[m - 1];
  out[0] = int(t & 0xffffffffull);  // surviving clusters
  out[1] = int(t >> 32);            // new nodes
}

__global__ void __launch_bounds__(kBlock) kCompact(const int* __restrict__ clus, const float4* __restrict__ clo,
                                                   const float4* __restrict__ chi, const int* __restrict__ ccnt,
                                                   const int* __restrict__ nn, const unsigned long long* __restrict__ flags,
                                                   const unsigned long long* __restrict__ scan, int m, int nodeBase,
                                                   int* __restrict__ clus2, float4* __restrict__ clo2, float4* __restrict__ chi2,
                                                   int* __restrict__ ccnt2, int* __restrict__ childL, int* __restrict__ childR,
                                                   int* __restrict__ nodeCnt, float* __restrict__ innerBox,
                                                   const float* __restrict__ leafBox, float4* __restrict__ cost, CollapseCost cc)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  unsigned long long f = flags[i];
  if (!(f & 1ull)) return;
  unsigned long long sc = scan[i];
  int p = int(sc & 0xffffffffull);
  if (f >> 32)
  {
    int j = nn[i];
    int id = nodeBase + int(sc >> 32);
    float4 alo = clo[i], ahi = chi[i], blo = clo[j], bhi = chi[j];
    float4 lo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.f);
    float4 hi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.f);
    childL[id] = clus[i];
    childR[id] = clus[j];
    int cnt = ccnt[i] + ccnt[j];
    nodeCnt[id] = cnt;
    float* ib = innerBox + 6 * size_t(id);
    ib[0] = lo.x; ib[1] = lo.y; ib[2] = lo.z; ib[3] = hi.x; ib[4] = hi.y; ib[5] = hi.z;
    // the children were made in earlier iterations, so their costs are final
    cost[id] = nodeCost(subtreeCost(clus[i], cost, leafBox, cc.tri), subtreeCost(clus[j], cost, leafBox, cc.tri), cnt,
                        boxHalfArea(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z), cc);
    clus2[p] = id;
    clo2[p] = lo;
    chi2[p] = hi;
    ccnt2[p] = cnt;
  }
  else
  {
    clus2[p] = clus[i];
    clo2[p] = clo[i];
    chi2[p] = chi[i];
    ccnt2[p] = ccnt[i];
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

// Triangle record (v0, prim), (e1 = v1 - v0), (e2 = v2 - v0) of sorted triangle k at slot o.
__device__ __forceinline__ void writeTri(const float* pos, const int* idx, const int* vals, int k, float4* tris, int o)
{
  int prim = vals[k];
  float3 a = ldPos(pos, idx[3 * prim]), b = ldPos(pos, idx[3 * prim + 1]), c = ldPos(pos, idx[3 * prim + 2]);
  tris[3 * size_t(o) + 0] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
  tris[3 * size_t(o) + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
  tris[3 * size_t(o) + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.f);
}

// Writes a 4-wide node: the n child boxes (padded) and references; empty slots get an empty box.
__device__ void writeNode(prt_bvh4_node* dst, const Box* bb, const int* ref, int n)
{
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
    nd.lox[q] = bb[q].lo[0]; nd.hix[q] = bb[q].hi[0];
    nd.loy[q] = bb[q].lo[1]; nd.hiy[q] = bb[q].hi[1];
    nd.loz[q] = bb[q].lo[2]; nd.hiz[q] = bb[q].hi[2];
    nd.child[q] = ref[q];
  }
  *dst = nd;
}

// Top-down collapse of the binary tree into 4-wide nodes, one level of the wide tree per
// launch.  A frontier entry is (binary node, wide node, first triangle slot of its subtree).
// The wide node of binary node b takes the cut of b's subtree into at most four subtrees that
// the cost C chose (C(b, 4) below b's own cost): a subtree x given j cuts stays whole when
// C(x, j) = C(x, 1), else its cuts are split between its children as C's minimum splits them.
// A whole subtree is a leaf when it is one triangle or its leaf cost is C(x, 1), else a wide
// node; inner children get consecutive wide-node indices, and leaf triangles are written at
// consecutive slots in child order (subtree slots follow the children's order, so they never
// overlap).
__global__ void __launch_bounds__(kBlock) kCollapse(const int4* __restrict__ frontier, int count, int4* __restrict__ next,
                                                    int* __restrict__ nextCount, int* __restrict__ nodeCounter,
                                                    const int* __restrict__ childL, const int* __restrict__ childR,
                                                    const int* __restrict__ nodeCnt, const float* __restrict__ leafBox,
                                                    const float* __restrict__ innerBox, const float4* __restrict__ cost,
                                                    CollapseCost cc, prt_bvh4_node* __restrict__ nodes,
                                                    const float* __restrict__ pos, const int* __restrict__ idx,
                                                    const int* __restrict__ vals, float4* __restrict__ tris)
{
  int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= count) return;
  int4 e = frontier[f];
  int b = e.x, k = e.y, slot = e.z;
  // cut b's subtree: (subtree, cuts) pairs on a stack, left child on top, so the subtrees come
  // out in left-to-right order; the cuts on the stack and in codes[] add up to at most four
  int codes[4];
  int n = 0;
  int2 stk[4];
  int sp = 0;
  {
    int kl;
    splitCost(subtreeCost(childL[b], cost, leafBox, cc.tri), subtreeCost(childR[b], cost, leafBox, cc.tri), 4, kl);
    stk[sp++] = make_int2(childR[b], 4 - kl);
    stk[sp++] = make_int2(childL[b], kl);
  }
  while (sp > 0)
  {
    const int2 x = stk[--sp];
    if (x.x < 0 || x.y == 1) { codes[n++] = x.x; continue; }
    const float4 cx = cost[x.x];
    if (costAt(cx, x.y) == cx.x) { codes[n++] = x.x; continue; }
    int kl;
    splitCost(subtreeCost(childL[x.x], cost, leafBox, cc.tri), subtreeCost(childR[x.x], cost, leafBox, cc.tri), x.y, kl);
    stk[sp++] = make_int2(childR[x.x], x.y - kl);
    stk[sp++] = make_int2(childL[x.x], kl);
  }
  // a whole inner subtree is a leaf when its leaf cost is its cost
  auto isLeaf = [&](int c) {
    if (c < 0) return true;
    const int cnt = nodeCnt[c];
    if (cnt > cc.leafMax) return false;
    const float* ib = innerBox + 6 * size_t(c);
    return cc.tri * float(cnt) * boxHalfArea(ib[0], ib[1], ib[2], ib[3], ib[4], ib[5]) == cost[c].x;
  };
  // child boxes (padded) and references; inner children get consecutive wide-node indices
  Box bb[4];
  int ref[4];
  bool leaf[4];
  int nInner = 0;
  for (int q = 0; q < n; ++q)
  {
    const int c = codes[q];
    const float* src = c < 0 ? leafBox + 6 * size_t(~c) : innerBox + 6 * size_t(c);
    for (int a = 0; a < 3; ++a) { bb[q].lo[a] = src[a]; bb[q].hi[a] = src[3 + a]; }
    padBox(bb[q]);
    leaf[q] = isLeaf(c);
    if (!leaf[q]) ++nInner;
  }
  const int base = nInner ? atomicAdd(nodeCounter, nInner) : 0;
  const int fbase = nInner ? atomicAdd(nextCount, nInner) : 0;
  int inner = 0;
  for (int q = 0; q < n; ++q)
  {
    int c = codes[q];
    const int cnt = c < 0 ? 1 : nodeCnt[c];
    if (leaf[q])
    {
      ref[q] = leafCode(slot, cnt);
      // gather the subtree's triangles (at most PRT_LEAF_CAP leaves, 2 PRT_LEAF_CAP - 1 nodes)
      int tstk[2 * PRT_LEAF_CAP];
      int tsp = 0, o = slot;
      tstk[tsp++] = c;
      while (tsp > 0)
      {
        int x = tstk[--tsp];
        if (x < 0) writeTri(pos, idx, vals, ~x, tris, o++);
        else { tstk[tsp++] = childR[x]; tstk[tsp++] = childL[x]; }
      }
    }
    else
    {
      const int j = base + inner;
      next[fbase + inner] = make_int4(c, j, slot, 0);
      ++inner;
      ref[q] = j;
    }
    slot += cnt;
  }
  writeNode(nodes + k, bb, ref, n);
}

// One-triangle scene: the root is a leaf over triangle slot 0.
__global__ void kCollapseSingle(const float* pos, const int* idx, const int* vals, float4* tris) { writeTri(pos, idx, vals, 0, tris, 0); }

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
  int *vals = nullptr, *vals2 = nullptr, *childL = nullptr, *childR = nullptr, *nodeCnt = nullptr, *nn = nullptr;
  int *clus[2] = {nullptr, nullptr}, *ccnt[2] = {nullptr, nullptr}, *counts = nullptr;
  float4 *clo[2] = {nullptr, nullptr}, *chi[2] = {nullptr, nullptr};
  unsigned long long *flags = nullptr, *scan = nullptr;
  int4 *front0 = nullptr, *front1 = nullptr;
  float *leafBox = nullptr, *innerBox = nullptr;
  float4* cost = nullptr;
  const CollapseCost cc = collapseCost();
  void *temp = nullptr, *scanTemp = nullptr;
  size_t tempBytes = 0, scanBytes = 0;
  const int inner = n > 1 ? n - 1 : 1;
  BVH_CHECK(hipMalloc(&cent, sizeof(float4) * n));
  BVH_CHECK(hipMalloc(&cbox, sizeof(float) * 8));
  BVH_CHECK(hipMalloc(&keys, sizeof(uint64_t) * n));
  BVH_CHECK(hipMalloc(&keys2, sizeof(uint64_t) * n));
  BVH_CHECK(hipMalloc(&vals, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&vals2, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&childL, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&childR, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&nodeCnt, sizeof(int) * inner));
  BVH_CHECK(hipMalloc(&nn, sizeof(int) * n));
  BVH_CHECK(hipMalloc(&flags, sizeof(unsigned long long) * n));
  BVH_CHECK(hipMalloc(&scan, sizeof(unsigned long long) * n));
  for (int q = 0; q < 2; ++q)
  {
    BVH_CHECK(hipMalloc(&clus[q], sizeof(int) * n));
    BVH_CHECK(hipMalloc(&ccnt[q], sizeof(int) * n));
    BVH_CHECK(hipMalloc(&clo[q], sizeof(float4) * n));
    BVH_CHECK(hipMalloc(&chi[q], sizeof(float4) * n));
  }
  BVH_CHECK(hipMalloc(&front0, sizeof(int4) * inner));
  BVH_CHECK(hipMalloc(&front1, sizeof(int4) * inner));
  BVH_CHECK(hipMalloc(&counts, sizeof(int) * 4));
  BVH_CHECK(hipMalloc(&leafBox, sizeof(float) * 6 * size_t(n)));
  BVH_CHECK(hipMalloc(&innerBox, sizeof(float) * 6 * size_t(inner)));
  BVH_CHECK(hipMalloc(&cost, sizeof(float4) * size_t(inner)));
  BVH_CHECK(hipMalloc(&out.nodes, sizeof(prt_bvh4_node) * inner));
  BVH_CHECK(hipMalloc(&out.tris, sizeof(float4) * 3 * size_t(n)));

  const float initBox[8] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f, 0.f, 0.f};
  BVH_CHECK(hipMemcpyAsync(cbox, initBox, sizeof(initBox), hipMemcpyHostToDevice, stream));
  kCentroids<<<blocks(n), kBlock, 0, stream>>>(dPos, dIdx, n, cent, cbox);
  kMorton<<<blocks(n), kBlock, 0, stream>>>(cent, n, cbox, keys, vals);
  BVH_CHECK(hipGetLastError());
  BVH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, keys, keys2, vals, vals2, n, 0, 63, stream));
  BVH_CHECK(hipMalloc(&temp, tempBytes));
  BVH_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, keys, keys2, vals, vals2, n, 0, 63, stream));
  kLeafInit<<<blocks(n), kBlock, 0, stream>>>(dPos, dIdx, vals2, n, leafBox, clus[0], clo[0], chi[0], ccnt[0]);
  BVH_CHECK(hipGetLastError());
  BVH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, flags, scan, n, stream));
  BVH_CHECK(hipMalloc(&scanTemp, scanBytes));
  // clustering: every iteration merges at least one pair, so n - 1 iterations bound the loop
  int m = n, nodeBase = 0, cur = 0, iters = 0;
  while (m > 1)
  {
    if (++iters > n) { err = "BVH clustering did not converge"; return hipErrorUnknown; }
    kNearest<<<blocks(m), kBlock, 0, stream>>>(clo[cur], chi[cur], m, nn);
    kMergeFlags<<<blocks(m), kBlock, 0, stream>>>(nn, m, flags);
    BVH_CHECK(hipGetLastError());
    BVH_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTemp, scanBytes, flags, scan, m, stream));
    kTotals<<<1, 1, 0, stream>>>(flags, scan, m, counts);
    kCompact<<<blocks(m), kBlock, 0, stream>>>(clus[cur], clo[cur], chi[cur], ccnt[cur], nn, flags, scan, m, nodeBase,
                                               clus[cur ^ 1], clo[cur ^ 1], chi[cur ^ 1], ccnt[cur ^ 1], childL, childR, nodeCnt,
                                               innerBox, leafBox, cost, cc);
    BVH_CHECK(hipGetLastError());
    int h[2] = {0, 0};
    BVH_CHECK(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, stream));
    BVH_CHECK(hipStreamSynchronize(stream));
    if (h[1] < 1 || h[0] != m - h[1]) { err = "BVH clustering made no progress"; return hipErrorUnknown; }
    m = h[0];
    nodeBase += h[1];
    cur ^= 1;
  }
  int levels = 1, wideNodes = 0;
  if (n == 1)
  {
    kCollapseSingle<<<1, 1, 0, stream>>>(dPos, dIdx, vals2, out.tris);
    BVH_CHECK(hipGetLastError());
  }
  else
  {
    // the one cluster left is the root; level-by-level collapse into 4-wide nodes, wide node 0
    // = the root, its triangles from slot 0
    int rootNode = -1;
    BVH_CHECK(hipMemcpyAsync(&rootNode, clus[cur], sizeof(int), hipMemcpyDeviceToHost, stream));
    BVH_CHECK(hipStreamSynchronize(stream));
    if (rootNode < 0 || rootNode >= inner) { err = "BVH clustering left no root node"; return hipErrorUnknown; }
    int4 root = make_int4(rootNode, 0, 0, 0);
    int one = 1;
    BVH_CHECK(hipMemcpyAsync(front0, &root, sizeof(int4), hipMemcpyHostToDevice, stream));
    BVH_CHECK(hipMemcpyAsync(counts, &one, sizeof(int), hipMemcpyHostToDevice, stream));
    int count = 1;
    levels = 0;
    while (count > 0)
    {
      BVH_CHECK(hipMemsetAsync(counts + 1, 0, sizeof(int), stream));
      kCollapse<<<blocks(count), kBlock, 0, stream>>>(front0, count, front1, counts + 1, counts, childL, childR, nodeCnt, leafBox,
                                                      innerBox, cost, cc, out.nodes, dPos, dIdx, vals2, out.tris);
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
  hipFree(childL); hipFree(childR); hipFree(nodeCnt); hipFree(nn); hipFree(flags); hipFree(scan);
  for (int q = 0; q < 2; ++q) { hipFree(clus[q]); hipFree(ccnt[q]); hipFree(clo[q]); hipFree(chi[q]); }
  hipFree(front0); hipFree(front1); hipFree(counts); hipFree(leafBox); hipFree(innerBox); hipFree(cost); hipFree(temp); hipFree(scanTemp);
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
