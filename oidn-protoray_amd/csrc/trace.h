// BVH traversal on gfx950: closest hit and any hit, one ray per lane, with a traversal stack
// of (node, entry distance) in LDS (PRT_LDS_STACK entries per lane, lane-interleaved so a
// wave's pushes hit 64 distinct banks) and a per-lane overflow area in HBM for deeper trees;
// closest-hit traversal drops popped entries that start beyond the current hit.  Replaces
// Embree's rtcIntersect8/16 and rtcOccluded8/16 (accel/embree_intersector_stream.h:66-131).
#pragma once

#include "dev_math.h"

namespace prt {

struct BvhView
{
  const prt_bvh4_node* nodes;
  const float4* tris;  // 3 float4 per triangle in leaf order: (v0, prim bits), (e1, 0), (e2, 0)
  int root;            // root reference (inner node 0, or a leaf when the scene is tiny)
};

// Moller-Trumbore, two-sided, accepting near <= t <= far.  Contraction is disabled so the
// result is bit-identical to the CPU oracle evaluating the same expressions.
PRT_DEV bool intersectTriangle(float4 a, float4 b, float4 c, f3 o, f3 d, float tnear, float tfar, float& t, float& u, float& v)
{
#pragma clang fp contract(off)
  f3 v0(a.x, a.y, a.z), e1(b.x, b.y, b.z), e2(c.x, c.y, c.z);
  f3 pvec(d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x);
  float det = e1.x * pvec.x + e1.y * pvec.y + e1.z * pvec.z;
  if (det == 0.f) return false;
  float inv = 1.0f / det;
  f3 tv(o.x - v0.x, o.y - v0.y, o.z - v0.z);
  u = (tv.x * pvec.x + tv.y * pvec.y + tv.z * pvec.z) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  f3 q(tv.y * e1.z - tv.z * e1.y, tv.z * e1.x - tv.x * e1.z, tv.x * e1.y - tv.y * e1.x);
  v = (d.x * q.x + d.y * q.y + d.z * q.z) * inv;
  if (!(v >= 0.f && u + v <= 1.f)) return false;
  t = (e2.x * q.x + e2.y * q.y + e2.z * q.z) * inv;
  return t >= tnear && t <= tfar;
}

// Slab tests of the four children of a wide node.  The far distance is widened by
// 1 + 2*gamma(3) and the boxes are padded at build time, so rounding never culls a hit the
// triangle test reports.  Returns the entry distance per child (+inf where missed or empty).
// kTwoPlus: the node has at least two children (every node of the device builder's trees; a
// host-built top-level node over one instance has one), so slots 0 and 1 are never empty.
template <bool kTwoPlus = false>
PRT_DEV void boxQuad(const prt_bvh4_node* nodes, int n, f3 o, f3 id, float tnear, float tfar, float t[4], int c[4])
{
  const float4* p = reinterpret_cast<const float4*>(nodes + n);
  float4 lx = ldg(p), hx = ldg(p + 1), ly = ldg(p + 2), hy = ldg(p + 3), lz = ldg(p + 4), hz = ldg(p + 5);
  int4 ch = ldg(reinterpret_cast<const int4*>(p) + 6);
  const float g = 1.0000004f;
  const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
  const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
  const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
  c[0] = ch.x; c[1] = ch.y; c[2] = ch.z; c[3] = ch.w;
#pragma unroll
  for (int q = 0; q < 4; ++q)
  {
    float x0 = (lxs[q] - o.x) * id.x, x1 = (hxs[q] - o.x) * id.x;
    float y0 = (lys[q] - o.y) * id.y, y1 = (hys[q] - o.y) * id.y;
    float z0 = (lzs[q] - o.z) * id.z, z1 = (hzs[q] - o.z) * id.z;
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tnear));
    float tf = fminf(fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * g, tfar);
    t[q] = (tn <= tf && ((kTwoPlus && q < 2) || c[q] != -1)) ? tn : __int_as_float(0x7f800000);
  }
}

PRT_DEV void cswap(float& ta, int& ca, float& tb, int& cb)
{
  if (tb < ta) { float t = ta; ta = tb; tb = t; int c = ca; ca = cb; cb = c; }
}

// Per-lane traversal stack of (node, entry distance): the first kEntries entries in LDS (two
// lane-interleaved columns, so a wave's accesses hit 64 distinct banks), deeper entries in the
// lane's overflow column in HBM (nodes in rows [0, cap), the distances in the rows after them;
// cap = PRT_MAX_STACK - kEntries except for the static-scratch query).  kDist = false: a stack
// for any-hit traversal, which never reads the entry distances, so only nodes are kept and the
// same LDS holds more entries.
template <int kEntries, bool kDist>
struct TraceStackT
{
  static constexpr bool kHasDist = kDist;
  int* lds;       // this lane's node column: lds[k * 64]
  float* ldsT;    // this lane's distance column: ldsT[k * 64] (kDist)
  int* spill;     // this lane's overflow column: spill[k * stride]
  int stride;
  int cap;        // overflow entries per lane: nodes in rows [0, cap), distances in [cap, 2 cap)
  int sp;
  uint32_t steps;  // traversal loop iterations of this lane (inner-node and leaf visits)
  // The build rejects trees deeper than PRT_MAX_STACK (and the static-scratch query trees deeper
  // than its capacity), so the last check never drops an entry; it only keeps every store inside
  // the overflow area.
  PRT_DEV void push(int v, float t)
  {
    if (sp < kEntries)
    {
      lds[sp * 64] = v;
      if (kDist) ldsT[sp * 64] = t;
    }
    else if (sp < kEntries + cap)
    {
      stg(spill + (sp - kEntries) * stride, v);
      if (kDist) stg(spill + (sp - kEntries + cap) * stride, __float_as_int(t));
    }
    else return;
    ++sp;
  }
  PRT_DEV int pop(float& t)
  {
    --sp;
    if (sp < kEntries)
    {
      t = kDist ? ldsT[sp * 64] : 0.f;
      return lds[sp * 64];
    }
    t = kDist ? __int_as_float(ldg(spill + (sp - kEntries + cap) * stride)) : 0.f;
    return ldg(spill + (sp - kEntries) * stride);
  }
};
using TraceStack = TraceStackT<PRT_LDS_STACK, true>;
using ShadowStack = TraceStackT<PRT_LDS_STACK_ANYHIT, false>;

// LDS a block of kThreads lanes needs for its stacks (nodes then distances).
template <int kThreads>
struct TraceLds
{
  int node[PRT_LDS_STACK * kThreads];
  float dist[PRT_LDS_STACK * kThreads];
};
template <int kThreads>
struct ShadowLds
{
  int node[PRT_LDS_STACK_ANYHIT * kThreads];
};
// This lane's stack over the block's LDS and the overflow area (gtid-th column of spill).
template <int kThreads>
PRT_DEV TraceStack makeStack(TraceLds<kThreads>& l, int* spill, int stride, int gtid,
                             int cap = PRT_MAX_STACK - PRT_LDS_STACK)
{
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  TraceStack st;
  st.lds = l.node + wave * (PRT_LDS_STACK * 64) + lane;
  st.ldsT = l.dist + wave * (PRT_LDS_STACK * 64) + lane;
  st.spill = spill + gtid;
  st.stride = stride;
  st.cap = cap;
  st.sp = 0;
  st.steps = 0;
  return st;
}
template <int kThreads>
PRT_DEV ShadowStack makeShadowStack(ShadowLds<kThreads>& l, int* spill, int stride, int gtid)
{
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  ShadowStack st;
  st.lds = l.node + wave * (PRT_LDS_STACK_ANYHIT * 64) + lane;
  st.ldsT = nullptr;
  st.spill = spill + gtid;
  st.stride = stride;
  st.cap = PRT_MAX_STACK - PRT_LDS_STACK_ANYHIT;
  st.sp = 0;
  st.steps = 0;
  return st;
}

// Any-hit filter that accepts every hit (a plain occlusion test).
struct AcceptHit
{
  PRT_DEV bool operator()(int, float, float, float, int = -1) const { return true; }
};

// A two-level scene (geometry/instance.h, accel/embree_intersector.cpp:100-129): a top-level
// tree over the instances' world boxes whose leaves each hold one instance, and per mesh its
// object-space tree.  A mesh is a contiguous range of the scene's triangles: its trees report
// triangle triBegin + local index.
struct MeshView
{
  const prt_bvh4_node* nodes;
  const float4* tris;
  int root;
  int triBegin;
};
struct InstView
{
  const prt_bvh4_node* nodes;  // top level
  int root;
  const prt_instance_dev* inst;
  const MeshView* meshes;
};

// The world -> object map of an instance, written out in one order (no contraction in the
// path kernels) so the CPU oracle transforms rays bit for bit the same.
PRT_DEV f3 xfmPointInv(const prt_instance_dev& I, f3 p)
{
  return f3(I.inv[0] * p.x + I.inv[1] * p.y + I.inv[2] * p.z + I.inv[3],
            I.inv[4] * p.x + I.inv[5] * p.y + I.inv[6] * p.z + I.inv[7],
            I.inv[8] * p.x + I.inv[9] * p.y + I.inv[10] * p.z + I.inv[11]);
}
PRT_DEV f3 xfmVectorInv(const prt_instance_dev& I, f3 v)
{
  return f3(I.inv[0] * v.x + I.inv[1] * v.y + I.inv[2] * v.z, I.inv[4] * v.x + I.inv[5] * v.y + I.inv[6] * v.z,
            I.inv[8] * v.x + I.inv[9] * v.y + I.inv[10] * v.z);
}
PRT_DEV f3 xfmRows(const float* m, f3 v)
{
  return f3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// One traversal step per loop iteration: a lane at an inner node visits it, a lane at a leaf
// tests its triangles, and a lane that has nothing to descend into pops, all in the same
// iteration.  Written with `continue` after a descent, the loop has two back edges, and the
// compiler turns it into a descent loop nested in the leaf loop: lanes that reach a leaf then
// idle until every lane of the wave has reached one (measured: closest-hit traversal 12 %
// slower, shadow rays 15 %).  The empty asm hides the flag's value from jump threading, which
// would otherwise restore the second back edge.
PRT_DEV void keepOneLoop(int& popNext) { __asm__ volatile("" : "+v"(popNext)); }

// Closest hit (kAnyHit = false) or any hit.  Ties in t go to the smaller primitive ID, so the
// result does not depend on the traversal order.  Returns the primitive ID or -1; nodeVisits
// counts wide-node fetches.  An any-hit traversal offers each hit in [tnear, tfar] to
// filter(prim, t, u, v), which accepts it (the ray is occluded) or lets the ray go on, as
// Embree's filter callbacks do (accel/embree_intersector.cpp:154-215).
template <bool kAnyHit, class Filter = AcceptHit, class Stack = TraceStack>
PRT_DEV int traceRay(const BvhView& bvh, f3 o, f3 d, float tnear, float tfar, float& tHit, float& uHit, float& vHit,
                     Stack& st, uint32_t& nodeVisits, uint32_t& triTests, Filter&& filter = Filter())
{
  static_assert(kAnyHit || Stack::kHasDist, "closest-hit traversal needs the entry distances");
  f3 id(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int best = -1;
  st.sp = 0;
  int node = bvh.root;
  const float inf = __int_as_float(0x7f800000);
  if constexpr (kAnyHit)
  {
    // any hit: a leaf reached by a descent or a pop is tested in the same iteration (the ray may
    // end there), then the next entry is popped; measured on the 1080p workload's shadow rays:
    // 4.44 against 4.60 ms per pass with the closest-hit order below, which in turn is faster for
    // closest hit (6.36 against 7.02)
    for (;;)
    {
      ++st.steps;
      int popNext = 0;
      if (node >= 0)
      {
        float t[4];
        int c[4];
        ++nodeVisits;
        boxQuad<true>(bvh.nodes, node, o, id, tnear, tfar, t, c);
        cswap(t[0], c[0], t[1], c[1]);
        cswap(t[2], c[2], t[3], c[3]);
        cswap(t[0], c[0], t[2], c[2]);
        cswap(t[1], c[1], t[3], c[3]);
        cswap(t[1], c[1], t[2], c[2]);
        popNext = 1;
        if (t[0] != inf)
        {
          if (t[3] != inf) st.push(c[3], t[3]);
          if (t[2] != inf) st.push(c[2], t[2]);
          if (t[1] != inf) st.push(c[1], t[1]);
          node = c[0];
          popNext = 0;
        }
      }
      keepOneLoop(popNext);
      if (popNext)
      {
        if (st.sp == 0) { tHit = tfar; return -1; }
        float tEntry;
        node = st.pop(tEntry);
      }
      if (node < 0)
      {
        const int code = node & 0x7fffffff;
        const int first = code >> 3, count = (code & 7) + 1;
        for (int k = 0; k < count; ++k)
        {
          const float4* tp = bvh.tris + 3 * (first + k);
          float4 a = ldg(tp), b = ldg(tp + 1), cc = ldg(tp + 2);
          float th, u, v;
          ++triTests;
          if (intersectTriangle(a, b, cc, o, d, tnear, tfar, th, u, v))
          {
            const int prim = __float_as_int(a.w);
            if (filter(prim, th, u, v)) { tHit = th; return prim; }
          }
        }
        if (st.sp == 0) { tHit = tfar; return -1; }
        float tEntry;
        node = st.pop(tEntry);
      }
    }
  }
  for (;;)
  {
    ++st.steps;
    int popNext = 1;
    if (node >= 0)
    {
      float t[4];
      int c[4];
      ++nodeVisits;
      boxQuad<true>(bvh.nodes, node, o, id, tnear, tfar, t, c);
      // closest first: sorting network on (t, child), misses sort to the end as +inf
      cswap(t[0], c[0], t[1], c[1]);
      cswap(t[2], c[2], t[3], c[3]);
      cswap(t[0], c[0], t[2], c[2]);
      cswap(t[1], c[1], t[3], c[3]);
      cswap(t[1], c[1], t[2], c[2]);
      if (t[0] != inf)
      {
        if (t[3] != inf) st.push(c[3], t[3]);
        if (t[2] != inf) st.push(c[2], t[2]);
        if (t[1] != inf) st.push(c[1], t[1]);
        node = c[0];
        popNext = 0;
      }
    }
    else
    {
      int code = node & 0x7fffffff;
      int first = code >> 3, count = (code & 7) + 1;
      for (int k = 0; k < count; ++k)
      {
        const float4* tp = bvh.tris + 3 * (first + k);
        float4 a = ldg(tp), b = ldg(tp + 1), cc = ldg(tp + 2);
        float th, u, v;
        ++triTests;
        if (intersectTriangle(a, b, cc, o, d, tnear, tfar, th, u, v))
        {
          int prim = __float_as_int(a.w);
          if (kAnyHit)
          {
            if (filter(prim, th, u, v)) { tHit = th; return prim; }
            continue;
          }
          if (th < tfar || best < 0 || prim < best)
          {
            tfar = th; best = prim; uHit = u; vHit = v;
          }
        }
      }
    }
    // next node: skip entries that start beyond the closest hit found so far (an entry at
    // exactly tfar may still hold a tie with a smaller primitive ID)
    keepOneLoop(popNext);
    if (popNext)
    {
      float tEntry;
      do
      {
        if (st.sp == 0) { tHit = tfar; return best; }
        node = st.pop(tEntry);
      } while (!kAnyHit && tEntry > tfar);
    }
  }
}

// Stack marker of the two-level traversal: leave the current instance (restore the world ray).
// Inner-node indices stay below it; it is pushed with entry distance -inf, so the closest-hit
// skip never drops it.
constexpr int kLeaveInstance = 0x7ffffffe;

// Closest or any hit in a two-level scene.  At a top-level leaf the ray moves into the
// instance's object space (the parameter t keeps its meaning, the direction is not
// renormalised, as in Embree's instance traversal) and the mesh's tree is traversed with the
// same stack above a marker; popping the marker restores the world ray.  Ties in t go to the
// smaller (instance, triangle) pair.  Returns the triangle (scene index) or -1, and its instance.
template <bool kAnyHit, class Filter = AcceptHit, class Stack = TraceStack>
PRT_DEV int traceRayInst(const InstView& iv, f3 wo, f3 wd, float tnear, float tfar, float& tHit, float& uHit, float& vHit,
                         int& instHit, Stack& st, uint32_t& nodeVisits, uint32_t& triTests, Filter&& filter = Filter())
{
  static_assert(kAnyHit || Stack::kHasDist, "closest-hit traversal needs the entry distances");
  const f3 wid(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
  f3 o = wo, d = wd, id = wid;
  int best = -1, bestInst = -1;
  st.sp = 0;
  int node = iv.root;
  const prt_bvh4_node* nodes = iv.nodes;
  const float4* tris = nullptr;
  int inst = -1, triBegin = 0;
  const float inf = __int_as_float(0x7f800000);
  for (;;)
  {
    ++st.steps;
    int popNext = 1;
    if (node == kLeaveInstance)
    {
      o = wo; d = wd; id = wid;
      nodes = iv.nodes;
      inst = -1;
    }
    else if (node >= 0)
    {
      float t[4];
      int c[4];
      ++nodeVisits;
      boxQuad(nodes, node, o, id, tnear, tfar, t, c);
      cswap(t[0], c[0], t[1], c[1]);
      cswap(t[2], c[2], t[3], c[3]);
      cswap(t[0], c[0], t[2], c[2]);
      cswap(t[1], c[1], t[3], c[3]);
      cswap(t[1], c[1], t[2], c[2]);
      if (t[0] != inf)
      {
        if (t[3] != inf) st.push(c[3], t[3]);
        if (t[2] != inf) st.push(c[2], t[2]);
        if (t[1] != inf) st.push(c[1], t[1]);
        node = c[0];
        popNext = 0;
      }
    }
    else if (inst < 0)  // a top-level leaf: enter its instance
    {
      inst = (node & 0x7fffffff) >> 3;
      const prt_instance_dev& I = iv.inst[inst];
      const MeshView& M = iv.meshes[I.mesh];
      st.push(kLeaveInstance, -inf);
      o = xfmPointInv(I, wo);
      d = xfmVectorInv(I, wd);
      id = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
      nodes = M.nodes;
      tris = M.tris;
      triBegin = M.triBegin;
      node = M.root;
      popNext = 0;
    }
    else
    {
      int code = node & 0x7fffffff;
      int first = code >> 3, count = (code & 7) + 1;
      for (int k = 0; k < count; ++k)
      {
        const float4* tp = tris + 3 * (first + k);
        float4 a = ldg(tp), b = ldg(tp + 1), cc = ldg(tp + 2);
        float th, u, v;
        ++triTests;
        if (intersectTriangle(a, b, cc, o, d, tnear, tfar, th, u, v))
        {
          int prim = triBegin + __float_as_int(a.w);
          if (kAnyHit)
          {
            if (filter(prim, th, u, v, inst)) { tHit = th; instHit = inst; return prim; }
            continue;
          }
          if (th < tfar || best < 0 || inst < bestInst || (inst == bestInst && prim < best))
          {
            tfar = th; best = prim; bestInst = inst; uHit = u; vHit = v;
          }
        }
      }
    }
    keepOneLoop(popNext);
    if (popNext)
    {
      float tEntry;
      do
      {
        if (st.sp == 0) { tHit = tfar; instHit = bestInst; return best; }
        node = st.pop(tEntry);
      } while (!kAnyHit && tEntry > tfar);
    }
  }
}

}  // namespace prt
