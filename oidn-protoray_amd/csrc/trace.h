// BVH traversal on gfx950: closest hit and any hit, one ray per lane, with the traversal
// stack in LDS (PRT_LDS_STACK entries per lane, lane-interleaved so a wave's pushes hit 64
// distinct banks) and a per-lane overflow area in HBM for deeper trees.  Replaces Embree's
// rtcIntersect8/16 and rtcOccluded8/16 (accel/embree_intersector_stream.h:66-131).
#pragma once

#include "dev_math.h"

namespace prt {

struct BvhView
{
  const prt_bvh_node* nodes;
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

// Slab test of both children of a node.  The far distance is widened by 1 + 2*gamma(3) and
// the boxes are padded at build time, so rounding never culls a hit the triangle test reports.
PRT_DEV void boxPair(const prt_bvh_node* nodes, int n, f3 o, f3 id, float tnear, float tfar, bool& h0, bool& h1, float& t0, float& t1,
                     int& c0, int& c1)
{
  const float4* p = reinterpret_cast<const float4*>(nodes + n);
  float4 bx = p[0], by = p[1], bz = p[2];
  int4 ch = reinterpret_cast<const int4*>(p)[3];
  // child 0: (bx.x, bx.z) (by.x, by.z) (bz.x, bz.z); child 1: (bx.y, bx.w) ...
  float ax0 = (bx.x - o.x) * id.x, ax1 = (bx.z - o.x) * id.x;
  float ay0 = (by.x - o.y) * id.y, ay1 = (by.z - o.y) * id.y;
  float az0 = (bz.x - o.z) * id.z, az1 = (bz.z - o.z) * id.z;
  float bx0 = (bx.y - o.x) * id.x, bx1 = (bx.w - o.x) * id.x;
  float by0 = (by.y - o.y) * id.y, by1 = (by.w - o.y) * id.y;
  float bz0 = (bz.y - o.z) * id.z, bz1 = (bz.w - o.z) * id.z;
  const float g = 1.0000004f;
  float an = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), tnear));
  float af = fminf(fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fmaxf(az0, az1)) * g, tfar);
  float bn = fmaxf(fmaxf(fminf(bx0, bx1), fminf(by0, by1)), fmaxf(fminf(bz0, bz1), tnear));
  float bf = fminf(fminf(fminf(fmaxf(bx0, bx1), fmaxf(by0, by1)), fmaxf(bz0, bz1)) * g, tfar);
  h0 = an <= af;
  h1 = bn <= bf;
  t0 = an;
  t1 = bn;
  c0 = ch.x;
  c1 = ch.y;
}

struct TraceStack
{
  int* lds;       // this lane's column: lds[k * 64]
  int* spill;     // this lane's overflow column: spill[k * stride]
  int stride;
  int sp;
  // The build rejects trees deeper than PRT_MAX_STACK, so the last check never drops an entry;
  // it only keeps every store inside the allocated overflow area.
  PRT_DEV void push(int v)
  {
    if (sp < PRT_LDS_STACK) lds[sp * 64] = v;
    else if (sp < PRT_MAX_STACK) spill[(sp - PRT_LDS_STACK) * stride] = v;
    else return;
    ++sp;
  }
  PRT_DEV int pop()
  {
    --sp;
    return sp < PRT_LDS_STACK ? lds[sp * 64] : spill[(sp - PRT_LDS_STACK) * stride];
  }
};

// Closest hit.  Ties in t go to the smaller primitive ID, so the result does not depend on
// the traversal order.  Returns the primitive ID or -1; nodeVisits counts inner-node fetches.
template <bool kAnyHit>
PRT_DEV int traceRay(const BvhView& bvh, f3 o, f3 d, float tnear, float tfar, float& tHit, float& uHit, float& vHit,
                     TraceStack& st, uint32_t& nodeVisits, uint32_t& triTests)
{
  f3 id(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int best = -1;
  st.sp = 0;
  int node = bvh.root;
  for (;;)
  {
    if (node >= 0)
    {
      bool h0, h1;
      float t0, t1;
      int c0, c1;
      ++nodeVisits;
      boxPair(bvh.nodes, node, o, id, tnear, tfar, h0, h1, t0, t1, c0, c1);
      if (h0 && h1)
      {
        if (t1 < t0) { int tmp = c0; c0 = c1; c1 = tmp; }
        st.push(c1);
        node = c0;
        continue;
      }
      if (h0) { node = c0; continue; }
      if (h1) { node = c1; continue; }
    }
    else
    {
      int code = node & 0x7fffffff;
      int first = code >> 3, count = (code & 7) + 1;
      for (int k = 0; k < count; ++k)
      {
        const float4* tp = bvh.tris + 3 * (first + k);
        float4 a = tp[0], b = tp[1], c = tp[2];
        float t, u, v;
        ++triTests;
        if (intersectTriangle(a, b, c, o, d, tnear, tfar, t, u, v))
        {
          int prim = __float_as_int(a.w);
          if (kAnyHit) { tHit = t; return prim; }
          if (t < tfar || best < 0 || prim < best)
          {
            tfar = t; best = prim; uHit = u; vHit = v;
          }
        }
      }
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  tHit = tfar;
  return best;
}

}  // namespace prt
