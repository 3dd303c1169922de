// Oracle: scene, CPU BVH, per-path integrator and C API.  TEST INFRASTRUCTURE ONLY.
//
// This is the clean-room CPU restatement of ProtoRay's stream path tracer
// (render/path_renderer_stream.h:195-1073) used as the parity checker for the HIP path and
// as the timed CPU baseline.  The reference processes 1024-path SoA streams sorted by
// material; paths never interact, so this restatement runs each path to completion with the
// same per-path state machine, which yields the same per-path results.  Triangle
// intersection is a Moller-Trumbore test evaluated with FP contraction off (see Makefile),
// so hit distances and barycentrics are bit-identical to the HIP kernels, which evaluate
// the same expression order with contraction disabled.
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>
#include "orc_core.h"

namespace orc {

// ---------------------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------------------
struct Tri { V3 v0, e1, e2; int prim; };
struct Node { float bmin[3], bmax[3]; int first; int count; };  // count == 0: inner, first = left child

struct Scene
{
  std::vector<V3> pos, nrm;
  std::vector<int> idx, triMat;
  std::vector<Material> mats;
  std::vector<Light> lights;
  std::vector<float> envRgb, envX, envY;
  EnvImage env;
  LightSet ls;
  AreaLight area;
  float sceneRadius = 0.f;
  V3 bmin, bmax;
  std::vector<Node> nodes;
  std::vector<Tri> tris;
  Luts luts;
  // Interior medium ID per material and the media's colours.  Media are deduplicated by
  // (colour, IOR) with ID 0 the clear ambient medium (material/material_factory.cpp:20-21,
  // 159-175); new media get IDs 1, 2, ... in material order.
  std::vector<int> matMedium;
  std::vector<V3> mediumColor;
  bool anyTransparent = false;  // render/scene.cpp:452-462: shadow rays use the transparent filter
  // pixel sample (path_renderer_stream.h:228-236): a fixed jitter, or the sampler's value warped
  // by the pixel filter's tabulated CDF (image/pixel_filter.h:30-66) when filterCdf is not empty
  bool jitter = false;
  float jitterXY[2] = {0.f, 0.f};
  std::vector<float> filterCdf;
  float filterWidth = 0.f;
  // image textures: UV set 0 (2 floats per vertex, empty: the hit's barycentrics), textures,
  // and the baseColor texture of each material (-1 none)
  std::vector<float> uv;
  std::vector<Texture> textures;
  std::vector<int> baseTex;
  // two-level scene (geometry/instance.h): meshes are triangle ranges [meshBegin[m],
  // meshBegin[m + 1]) with object-space trees; instances place them (empty: one level)
  struct Instance { int mesh; float inv[12], lin[9], nrm[9]; };
  struct MeshTree { std::vector<Node> nodes; std::vector<Tri> tris; };
  std::vector<int> meshBegin;
  std::vector<Instance> inst;
  std::vector<MeshTree> meshTrees;
};

// A textured baseColor at a hit: clamp(texture(uv) * (r, g, b)) (material/material_param.h:
// 37-41, material/standard.cpp:424); UV set 0 interpolated barycentrically
// (geometry/triangle_mesh.cpp:50-53, :83-90), or the hit's (u, v) without texcoords (:131).
static void hitUV(const Scene& s, int prim, float u, float v, float& tu, float& tv)
{
  tu = u;
  tv = v;
  if (!s.uv.empty())
  {
    const int i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    // b0 uv0 + b1 uv1 + b2 uv2 as the reference's build evaluates it: contracted into two fused
    // multiply-adds (GCC's default -ffp-contract=fast on the AVX2 + FMA build)
    const float b1 = u, b2 = v, b0 = 1.f - b1 - b2;
    tu = std::fma(b2, s.uv[2 * i2], std::fma(b1, s.uv[2 * i1], b0 * s.uv[2 * i0]));
    tv = std::fma(b2, s.uv[2 * i2 + 1], std::fma(b1, s.uv[2 * i1 + 1], b0 * s.uv[2 * i0 + 1]));
  }
}

static V3 texturedBase(const Scene& s, const Material& m, int prim, float u, float v, int tex)
{
  float tu, tv;
  hitUV(s, prim, u, v, tu, tv);
  const V3 c = textureGet(s.textures[size_t(tex)], tu, tv);
  return V3(clampSel(c.x * m.r, 0.f, 1.f), clampSel(c.y * m.g, 0.f, 1.f), clampSel(c.z * m.b, 0.f, 1.f));
}

// The colour of a material's interior medium: pow(clamp(transmissionColor), transmissionDepth)
// (material/standard.cpp:48-49, :169-173).
static V3 interiorColor(const Material& m)
{
  const float d = std::max(m.transmissionDepth, 0.f);
  auto ch = [&](int k) { return std::pow(std::min(std::max(m.transmissionColor[k], 0.f), 1.f), d); };
  return V3(ch(0), ch(1), ch(2));
}

static void assignMedia(Scene& s)
{
  std::vector<float> iors;
  s.matMedium.assign(s.mats.size(), 0);
  s.mediumColor.assign(1, V3(1.f));
  s.anyTransparent = false;
  for (size_t m = 0; m < s.mats.size(); ++m) s.anyTransparent = s.anyTransparent || isTransparent(s.mats[m]);
  for (size_t m = 0; m < s.mats.size(); ++m)
  {
    if (!isMediaInterface(s.mats[m])) continue;
    const float ior = interfaceIor(s.mats[m]);
    const V3 col = interiorColor(s.mats[m]);
    // a clear medium at the ambient IOR is the ambient medium itself
    int id = (ior == 1.f && col.x == 1.f && col.y == 1.f && col.z == 1.f) ? 0 : -1;
    for (size_t k = 0; k < iors.size(); ++k)
      if (iors[k] == ior && s.mediumColor[k + 1].x == col.x && s.mediumColor[k + 1].y == col.y &&
          s.mediumColor[k + 1].z == col.z)
        id = int(k) + 1;
    if (id < 0)
    {
      iors.push_back(ior);
      s.mediumColor.push_back(col);
      id = int(iors.size());
    }
    s.matMedium[m] = id;
  }
}

// light/triangle_light.h:36-64 + render/scene.cpp:83-116: the emissive triangles of the mesh
// (one TriangleMesh: the flattened geometry) in primitive order, weights |n| * luminance
static void buildAreaLight(Scene& s)
{
  AreaLight& a = s.area;
  a = AreaLight();
  a.triMap.assign(s.triMat.size(), -1);
  std::vector<float> w;
  for (size_t t = 0; t < s.triMat.size(); ++t)
  {
    const Material& m = s.mats[s.triMat[t]];
    if (s.triMat[t] == 0 || !isEmissive(m)) continue;
    V3 p0 = s.pos[s.idx[3 * t]], p1 = s.pos[s.idx[3 * t + 1]], p2 = s.pos[s.idx[3 * t + 2]];
    V3 n = cross(p1 - p0, p2 - p0);
    V3 c(m.emission[0], m.emission[1], m.emission[2]);
    a.triMap[t] = int(a.v0.size());
    a.v0.push_back(p0); a.v1.push_back(p1); a.v2.push_back(p2); a.n.push_back(n); a.color.push_back(c);
    w.push_back(length(n) * luminance(c));
  }
  if (a.v0.empty()) return;
  a.cdf.resize(a.v0.size() + 1);
  dist1dInit(a.cdf.data(), w.data(), int(w.size()));
  float sumL = 0.f;
  for (size_t i = 0; i < a.v0.size(); ++i) sumL += length(a.n[i]) * luminance(a.color[i]);
  a.power = kPi * sumL;
}

// geometry/triangle_mesh.cpp / render/scene.cpp: bounds over all vertices referenced
static void computeBounds(Scene& s)
{
  s.bmin = V3(kPosMax);
  s.bmax = V3(-kPosMax);
  for (size_t t = 0; t < s.idx.size(); ++t)
  {
    s.bmin = vmin(s.bmin, s.pos[s.idx[t]]);
    s.bmax = vmax(s.bmax, s.pos[s.idx[t]]);
  }
  s.sceneRadius = length(s.bmax - s.bmin) * 0.5f;  // render/scene.cpp:367
}

// ---------------------------------------------------------------------------------------
// Moller-Trumbore, two-sided, hit accepted for near <= t <= far.  Same expression order as
// the HIP kernels; compiled with -ffp-contract=off.
// ---------------------------------------------------------------------------------------
static inline bool intersectTri(const Tri& tr, V3 o, V3 d, float tnear, float tfar, float& t, float& u, float& v)
{
  V3 pvec = cross(d, tr.e2);
  float det = dot(tr.e1, pvec);
  if (det == 0.f) return false;
  float inv = 1.0f / det;
  V3 tvec = o - tr.v0;
  u = dot(tvec, pvec) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  V3 qvec = cross(tvec, tr.e1);
  v = dot(d, qvec) * inv;
  if (!(v >= 0.f && u + v <= 1.f)) return false;
  t = dot(tr.e2, qvec) * inv;
  return t >= tnear && t <= tfar;
}

// ---------------------------------------------------------------------------------------
// Binned-SAH BVH2 (CPU acceleration for the oracle; structure is irrelevant to results)
// ---------------------------------------------------------------------------------------
struct BuildRef { float bmin[3], bmax[3], c[3]; int tri; };

static void buildRecursive(std::vector<Node>& nodes, std::vector<BuildRef>& refs, int nodeIdx, int begin, int end, int depth)
{
  Node& n = nodes[nodeIdx];
  float bmin[3] = {kPosMax, kPosMax, kPosMax}, bmax[3] = {-kPosMax, -kPosMax, -kPosMax};
  float cmin[3] = {kPosMax, kPosMax, kPosMax}, cmax[3] = {-kPosMax, -kPosMax, -kPosMax};
  for (int i = begin; i < end; ++i)
    for (int a = 0; a < 3; ++a)
    {
      bmin[a] = std::min(bmin[a], refs[i].bmin[a]); bmax[a] = std::max(bmax[a], refs[i].bmax[a]);
      cmin[a] = std::min(cmin[a], refs[i].c[a]); cmax[a] = std::max(cmax[a], refs[i].c[a]);
    }
  for (int a = 0; a < 3; ++a) { nodes[nodeIdx].bmin[a] = bmin[a]; nodes[nodeIdx].bmax[a] = bmax[a]; }
  int count = end - begin;
  if (count <= 4 || depth > 60)
  {
    nodes[nodeIdx].first = begin;
    nodes[nodeIdx].count = count;
    return;
  }
  const int NB = 16;
  int bestAxis = -1, bestSplit = -1;
  float bestCost = kPosMax;
  for (int a = 0; a < 3; ++a)
  {
    float ext = cmax[a] - cmin[a];
    if (!(ext > 0.f)) continue;
    int cnt[NB] = {0};
    float lo[NB][3], hi[NB][3];
    for (int b = 0; b < NB; ++b) for (int k = 0; k < 3; ++k) { lo[b][k] = kPosMax; hi[b][k] = -kPosMax; }
    float scale = NB / ext;
    for (int i = begin; i < end; ++i)
    {
      int b = std::min(NB - 1, int((refs[i].c[a] - cmin[a]) * scale));
      cnt[b]++;
      for (int k = 0; k < 3; ++k) { lo[b][k] = std::min(lo[b][k], refs[i].bmin[k]); hi[b][k] = std::max(hi[b][k], refs[i].bmax[k]); }
    }
    float rArea[NB];
    int rCnt[NB];
    float l[3] = {kPosMax, kPosMax, kPosMax}, h[3] = {-kPosMax, -kPosMax, -kPosMax};
    int c = 0;
    for (int b = NB - 1; b > 0; --b)
    {
      c += cnt[b];
      for (int k = 0; k < 3; ++k) { l[k] = std::min(l[k], lo[b][k]); h[k] = std::max(h[k], hi[b][k]); }
      float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
      rArea[b] = c ? (dx * dy + dy * dz + dz * dx) : 0.f;
      rCnt[b] = c;
    }
    for (int k = 0; k < 3; ++k) { l[k] = kPosMax; h[k] = -kPosMax; }
    c = 0;
    for (int b = 0; b < NB - 1; ++b)
    {
      c += cnt[b];
      for (int k = 0; k < 3; ++k) { l[k] = std::min(l[k], lo[b][k]); h[k] = std::max(h[k], hi[b][k]); }
      float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
      float lArea = c ? (dx * dy + dy * dz + dz * dx) : 0.f;
      if (c == 0 || rCnt[b + 1] == 0) continue;
      float cost = lArea * c + rArea[b + 1] * rCnt[b + 1];
      if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
    }
  }
  int mid;
  if (bestAxis < 0)
  {
    mid = begin + count / 2;  // all centroids equal: split in the middle
  }
  else
  {
    float scale = NB / (cmax[bestAxis] - cmin[bestAxis]);
    auto it = std::partition(refs.begin() + begin, refs.begin() + end, [&](const BuildRef& r) {
      return std::min(NB - 1, int((r.c[bestAxis] - cmin[bestAxis]) * scale)) <= bestSplit;
    });
    mid = int(it - refs.begin());
    if (mid == begin || mid == end) mid = begin + count / 2;
  }
  int left = (int)nodes.size();
  nodes.push_back(Node());
  nodes.push_back(Node());
  nodes[nodeIdx].first = left;
  nodes[nodeIdx].count = 0;
  (void)n;
  buildRecursive(nodes, refs, left, begin, mid, depth + 1);
  buildRecursive(nodes, refs, left + 1, mid, end, depth + 1);
}

// A tree over the scene triangles [begin, end) (their Tri records carry the scene index).
static void buildTree(const Scene& s, int begin, int end, std::vector<Node>& nodes, std::vector<Tri>& tris)
{
  int nt = end - begin;
  std::vector<BuildRef> refs(nt);
  for (int k = 0; k < nt; ++k)
  {
    const int t = begin + k;
    V3 p0 = s.pos[s.idx[3 * t]], p1 = s.pos[s.idx[3 * t + 1]], p2 = s.pos[s.idx[3 * t + 2]];
    V3 lo = vmin(p0, vmin(p1, p2)), hi = vmax(p0, vmax(p1, p2));
    BuildRef& r = refs[k];
    for (int a = 0; a < 3; ++a)
    {
      // pad by a relative 2^-20 so a box never culls a hit the triangle test reports
      float pad = std::max(std::fabs(lo[a]), std::fabs(hi[a])) * 9.5367431640625e-7f + 1e-30f;
      r.bmin[a] = lo[a] - pad; r.bmax[a] = hi[a] + pad;
      r.c[a] = 0.5f * (lo[a] + hi[a]);
    }
    r.tri = t;
  }
  nodes.clear();
  nodes.reserve(size_t(nt) * 2 / 2 + 16);
  nodes.push_back(Node());
  if (nt > 0) buildRecursive(nodes, refs, 0, 0, nt, 0);
  else
  {
    // an empty tree: a leaf with no triangles and an empty box (never entered)
    nodes[0].count = 0; nodes[0].first = 0;
    for (int a = 0; a < 3; ++a) { nodes[0].bmin[a] = 1; nodes[0].bmax[a] = -1; }
  }
  tris.resize(nt);
  for (int i = 0; i < nt; ++i)
  {
    int t = refs[i].tri;
    V3 p0 = s.pos[s.idx[3 * t]], p1 = s.pos[s.idx[3 * t + 1]], p2 = s.pos[s.idx[3 * t + 2]];
    tris[i].v0 = p0; tris[i].e1 = p1 - p0; tris[i].e2 = p2 - p0; tris[i].prim = t;
  }
}

static void buildBvh(Scene& s) { buildTree(s, 0, (int)s.triMat.size(), s.nodes, s.tris); }

struct RayHit { V3 o, d; float tnear, tfar; int prim; float u, v; int inst = -1; };

// The world -> object map of an instance (rows of [A | b]), in the kernels' expression order.
static V3 xfmPointInv(const Scene::Instance& I, V3 p)
{
  return V3(I.inv[0] * p.x + I.inv[1] * p.y + I.inv[2] * p.z + I.inv[3], I.inv[4] * p.x + I.inv[5] * p.y + I.inv[6] * p.z + I.inv[7],
            I.inv[8] * p.x + I.inv[9] * p.y + I.inv[10] * p.z + I.inv[11]);
}
static V3 xfmVectorInv(const Scene::Instance& I, V3 v)
{
  return V3(I.inv[0] * v.x + I.inv[1] * v.y + I.inv[2] * v.z, I.inv[4] * v.x + I.inv[5] * v.y + I.inv[6] * v.z,
            I.inv[8] * v.x + I.inv[9] * v.y + I.inv[10] * v.z);
}
static V3 xfmRows(const float* m, V3 v)
{
  return V3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

static inline bool boxHit(const Node& n, V3 o, V3 inv, float tnear, float tfar, float& tEntry)
{
  float t0 = tnear, t1 = tfar;
  for (int a = 0; a < 3; ++a)
  {
    float oa = a == 0 ? o.x : (a == 1 ? o.y : o.z);
    float ia = a == 0 ? inv.x : (a == 1 ? inv.y : inv.z);
    float ta = (n.bmin[a] - oa) * ia, tb = (n.bmax[a] - oa) * ia;
    if (ta > tb) std::swap(ta, tb);
    tb *= 1.0000004f;  // 1 + 2*gamma(3): conservative far bound
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
    if (!(t0 <= t1)) return false;
  }
  tEntry = t0;
  return true;
}

static void traverseTree(const std::vector<Node>& nodes, const std::vector<Tri>& tris, RayHit& r, bool anyHit,
                         long long* nodeVisits = nullptr);

// Closest or any hit.  In a two-level scene every instance is tried in turn: the ray moves into
// its object space (t keeps its meaning) and the mesh's tree answers; ties in t go to the
// smaller (instance, triangle) pair (the order the kernels' tie rule gives).
static void traverse(const Scene& s, RayHit& r, bool anyHit, long long* nodeVisits = nullptr)
{
  r.inst = -1;
  if (s.inst.empty())
  {
    traverseTree(s.nodes, s.tris, r, anyHit, nodeVisits);
    return;
  }
  int best = -1, bestInst = -1;
  float bu = 0.f, bv = 0.f, tfar = r.tfar;
  for (size_t i = 0; i < s.inst.size(); ++i)
  {
    const Scene::Instance& I = s.inst[i];
    RayHit q;
    q.o = xfmPointInv(I, r.o);
    q.d = xfmVectorInv(I, r.d);
    q.tnear = r.tnear;
    q.tfar = tfar;
    q.u = q.v = 0.f;
    traverseTree(s.meshTrees[size_t(I.mesh)].nodes, s.meshTrees[size_t(I.mesh)].tris, q, anyHit, nodeVisits);
    if (q.prim < 0) continue;
    if (anyHit) { r.prim = q.prim; r.tfar = q.tfar; r.inst = int(i); return; }
    // instances in increasing order: an equal t keeps the earlier (smaller) instance
    if (best < 0 || q.tfar < tfar) { tfar = q.tfar; best = q.prim; bestInst = int(i); bu = q.u; bv = q.v; }
  }
  r.prim = best;
  if (best >= 0) { r.tfar = tfar; r.u = bu; r.v = bv; r.inst = bestInst; }
}

static void traverseTree(const std::vector<Node>& nodes, const std::vector<Tri>& tris, RayHit& r, bool anyHit, long long* nodeVisits)
{
  V3 inv(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  r.prim = -1;
  float tfar = r.tfar;
  long long visits = 0;
  while (sp)
  {
    const Node& n = nodes[stack[--sp]];
    float te;
    ++visits;
    if (!boxHit(n, r.o, inv, r.tnear, tfar, te)) continue;
    if (n.count > 0)
    {
      for (int i = 0; i < n.count; ++i)
      {
        const Tri& tr = tris[n.first + i];
        float t, u, v;
        if (intersectTri(tr, r.o, r.d, r.tnear, tfar, t, u, v))
        {
          if (anyHit) { r.prim = tr.prim; r.tfar = t; if (nodeVisits) *nodeVisits += visits; return; }
          if (t < tfar || r.prim < 0 || (t == tfar && tr.prim < r.prim))
          {
            tfar = t; r.prim = tr.prim; r.u = u; r.v = v;
          }
        }
      }
    }
    else
    {
      // push the far child first so the near child is popped first
      const Node& a = nodes[n.first];
      const Node& b = nodes[n.first + 1];
      float ta, tb;
      bool ha = boxHit(a, r.o, inv, r.tnear, tfar, ta), hb = boxHit(b, r.o, inv, r.tnear, tfar, tb);
      if (ha && hb)
      {
        if (ta <= tb) { stack[sp++] = n.first + 1; stack[sp++] = n.first; }
        else { stack[sp++] = n.first; stack[sp++] = n.first + 1; }
      }
      else if (ha) stack[sp++] = n.first;
      else if (hb) stack[sp++] = n.first + 1;
    }
  }
  if (r.prim >= 0) r.tfar = tfar;
  if (nodeVisits) *nodeVisits += visits;
}

static void bruteForce(const Scene& s, RayHit& r, bool anyHit)
{
  r.prim = -1;
  r.inst = -1;
  float tfar = r.tfar;
  const size_t nInst = s.inst.empty() ? 1 : s.inst.size();
  for (size_t k = 0; k < nInst; ++k)
  {
    // one level: every triangle; two levels: each instance's mesh in its object space
    const std::vector<Tri>& tris = s.inst.empty() ? s.tris : s.meshTrees[size_t(s.inst[k].mesh)].tris;
    const V3 o = s.inst.empty() ? r.o : xfmPointInv(s.inst[k], r.o), d = s.inst.empty() ? r.d : xfmVectorInv(s.inst[k], r.d);
    const int inst = s.inst.empty() ? -1 : int(k);
    for (size_t i = 0; i < tris.size(); ++i)
    {
      float t, u, v;
      if (intersectTri(tris[i], o, d, r.tnear, tfar, t, u, v))
      {
        if (anyHit) { r.prim = tris[i].prim; r.tfar = t; r.inst = inst; return; }
        if (r.prim < 0 || t < tfar || (t == tfar && (inst < r.inst || (inst == r.inst && tris[i].prim < r.prim))))
        {
          tfar = t; r.prim = tris[i].prim; r.u = u; r.v = v; r.inst = inst;
        }
      }
    }
  }
  if (r.prim >= 0) r.tfar = tfar;
}

// ---------------------------------------------------------------------------------------
// Hit geometry: geometry/triangle_mesh.cpp:152-213 + render/scene.h:61-72
// ---------------------------------------------------------------------------------------
// inst >= 0: the mesh's post-intersection in object space, then the tangents to world space by
// the instance's linear part and the normals by its inverse transpose, renormalised
// (geometry/instance.h:17-26), before the orientation flip.
static Ctx postIntersect(const Scene& s, V3 o, V3 d, float t, int prim, float u, float v, int inst = -1)
{
  Ctx c;
  c.p = o + t * d;  // core/ray_simd.h:58-68
  c.eps = std::max(t, reduceMax(V3(std::fabs(c.p.x), std::fabs(c.p.y), std::fabs(c.p.z)))) * 0x1.fp-18f;
  int i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
  V3 p0 = s.pos[i0], p1 = s.pos[i1], p2 = s.pos[i2];
  V3 Ng = normalize(cross(p1 - p0, p2 - p0));
  V3 N = Ng;
  if (!s.nrm.empty())
  {
    // zero vertex normals (a geometry without normals in a mixed file) keep N = Ng, as a
    // TriangleMesh without normals does (geometry/triangle_mesh.cpp:176-189)
    float b0 = 1.f - u - v;
    V3 Ns = b0 * s.nrm[i0] + u * s.nrm[i1] + v * s.nrm[i2];
    if (dot(Ns, Ns) > 0.f)
    {
      N = normalize(Ns);
      if (dot(N, Ng) < 0.f) N = -N;
    }
  }
  c.Ng = Ng;
  c.f.N = N;
  makeFrame(c.f.U, c.f.V, N);  // no texcoords: triangle_mesh.cpp:235-242
  if (inst >= 0)
  {
    const Scene::Instance& I = s.inst[size_t(inst)];
    c.f.U = normalize(xfmRows(I.lin, c.f.U));
    c.f.V = normalize(xfmRows(I.lin, c.f.V));
    c.f.N = normalize(xfmRows(I.nrm, c.f.N));
    c.Ng = normalize(xfmRows(I.nrm, c.Ng));
  }
  c.backfacing = dot(c.Ng, d) > 0.f;
  if (c.backfacing) { c.Ng = -c.Ng; c.f.N = -c.f.N; }
  return c;
}

// Any-hit traversal with the transparent-shadow filter (render/transparent_filter.h:42-93):
// each hit on a transparent material multiplies T by the material's transparency along -d and
// the ray goes on until T falls below 1e-5; a hit on any other material occludes (T = 0).
// One tree of the filtered any-hit traversal: triangles tested with the ray (o, d) of the tree's
// space, hits shaded with the world ray of r on instance inst.  Returns true when occluded.
static bool traverseShadowTree(const Scene& s, const std::vector<Node>& nodes, const std::vector<Tri>& tris, V3 o, V3 d,
                               RayHit& r, V3& T, uint32_t& crossed, int inst)
{
  V3 inv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp)
  {
    const Node& n = nodes[stack[--sp]];
    float te;
    if (!boxHit(n, o, inv, r.tnear, r.tfar, te)) continue;
    if (n.count > 0)
    {
      for (int i = 0; i < n.count; ++i)
      {
        const Tri& tr = tris[n.first + i];
        float t, u, v;
        if (!intersectTri(tr, o, d, r.tnear, r.tfar, t, u, v)) continue;
        const Material& m = s.mats[s.triMat[tr.prim]];
        if (isTransparent(m))
        {
          Ctx c = postIntersect(s, r.o, r.d, t, tr.prim, u, v, inst);
          Bsdf b = makeBsdf(s.luts, m, c, -r.d, 0, -1);
          T *= bsdfTransparency(s.luts, b, -r.d);
          ++crossed;
          if (!(reduceMax(T) < 1e-5f)) continue;
        }
        else T = V3(0.f);
        r.prim = tr.prim;
        r.tfar = t;
        r.inst = inst;
        return true;
      }
    }
    else
    {
      stack[sp++] = n.first + 1;
      stack[sp++] = n.first;
    }
  }
  return false;
}

static void traverseShadow(const Scene& s, RayHit& r, V3& T, uint32_t& crossed)
{
  T = V3(1.f);
  crossed = 0;
  r.prim = -1;
  r.inst = -1;
  if (s.inst.empty())
  {
    traverseShadowTree(s, s.nodes, s.tris, r.o, r.d, r, T, crossed, -1);
    return;
  }
  for (size_t i = 0; i < s.inst.size(); ++i)
  {
    const Scene::Instance& I = s.inst[i];
    const Scene::MeshTree& M = s.meshTrees[size_t(I.mesh)];
    if (traverseShadowTree(s, M.nodes, M.tris, xfmPointInv(I, r.o), xfmVectorInv(I, r.d), r, T, crossed, int(i))) return;
  }
}

// ---------------------------------------------------------------------------------------
// Integrator
// ---------------------------------------------------------------------------------------
struct Camera
{
  float pos[3], U[3], V[3], N[3];
  float fov, nearClip, lensRadius, focalDistance;
};

struct RenderParams
{
  int W, H;
  uint32_t seed;
  int maxDepth, rrDepth;
  float minWeight, maxRadiance;
};

struct Accum  // sums are FP32, added in sample order (image/accum_buffer.h:439-486)
{
  float* color;   // [3][W*H] planar
  float* weight;  // [W*H]
  float* albedo; float* albedoW;
  float* albedo1; float* albedo1W;
  float* normal; float* normalW;
  float* normal1; float* normal1W;
  float* depth; float* depthW;
  float* color2;  // [3][W*H] sums of squares (AccumFlag::SampleVar) or null
  // per-sample debug outputs of the last pass (nullable)
  float* sampleL;      // [3][W*H]
  uint32_t* pathHash;  // [W*H]
  int* firstPrim;      // [W*H]
};

static inline void add3(float* buf, float* w, int npix, int pix, V3 c)
{
  if (!buf) return;
  buf[pix] += c.x; buf[npix + pix] += c.y; buf[2 * npix + pix] += c.z;
  w[pix] += 1.f;
}

// AccumBuffer3f::addSerial (image/accum_buffer.h:439-486): sum, weight and, with SampleVar, the
// sum of squares
static inline void addColor(Accum& A, int npix, int pix, V3 c)
{
  add3(A.color, A.weight, npix, pix, c);
  if (A.color2)
  {
    A.color2[pix] += c.x * c.x; A.color2[npix + pix] += c.y * c.y; A.color2[2 * npix + pix] += c.z * c.z;
  }
}

// integrator.h:27-46
static inline V3 clampL(V3 L, float maxRadiance)
{
  V3 c(0.f < L.x ? L.x : 0.f, 0.f < L.y ? L.y : 0.f, 0.f < L.z ? L.z : 0.f);
  float y = luminance(c);
  if (y > maxRadiance) c *= maxRadiance * (1.0f / y);
  return c;
}

static inline uint32_t hashStep(uint32_t h, uint32_t x) { return murmurMix(h, x); }
// Coarse direction cell (1/32 per axis): two paths whose sampled directions fall in different
// cells are not the same path even when they hit the same primitive.
static inline uint32_t dirCell(V3 d)
{
  auto q = [](float x) { return uint32_t(int(std::floor(x * 32.f)) + 64) & 0xffu; };
  return q(d.x) | (q(d.y) << 8) | (q(d.z) << 16);
}

struct Stats { long long rays = 0; long long nodes = 0; };

// The pixel sample (path_renderer_stream.h:228-236) from the sampler's dimension-set 0 value:
// the jitter when one is set, else the value warped by the pixel filter (PixelFilter::sample +
// 0.5, image/pixel_filter.h:57-63, Distribution1D::sampleContinuous, distribution1d.h:104-108).
static void pixelSample(const Scene& s, float& psx, float& psy)
{
  if (s.jitter)
  {
    psx = s.jitterXY[0];
    psy = s.jitterXY[1];
    return;
  }
  if (s.filterCdf.empty()) return;
  const int size = int(s.filterCdf.size()) - 1;
  auto warp = [&](float u) {
    float pdf;
    int i = dist1dSampleReuse(s.filterCdf.data(), size, pdf, u);
    float x = (float(i) + u) * (1.0f / float(size));
    return (x - 0.5f) * s.filterWidth + 0.5f;
  };
  psx = warp(psx);
  psy = warp(psy);
}

// One path of one sample (path_renderer_stream.h:209-268 init, :276-1071 bounce loop)
// ---- FP64 parameter derivatives of the BSDF values (the adjoint's checker) --------------
// value = f * cos as bsdfEval returns it; dA[c] = d value_c / d albedo_c; dAl = d value / d alpha
// (alpha = roughness^2).  Diffuse: rho / pi cos.  Coated diffuse (rough_coating.h:297-378):
// F s D G2 / (4 cos_o) [cos_i > 0] + rho / pi cos_i absorb (1 - Ro)(1 - Ri) / (1 - Ravg), with
// the reflection-albedo tables interpolated linearly (their slope along the roughness axis is
// the derivative).  Every operation after the float inputs is in double.
struct GradD { double v[3] = {0, 0, 0}, dA[3] = {0, 0, 0}, dAl[3] = {0, 0, 0}; };

static double lerpd(double a, double b, double t) { return a + t * (b - a); }
static double clampd(double x, double lo, double hi) { return std::min(std::max(x, lo), hi); }
// interp3D (and its slope along z) in double over a float table
static double interp3Dd(double px, double py, double pz, const float* f, int n, bool slopeZ)
{
  if (slopeZ && !(pz > 0.0 && pz < double(n - 1))) return 0.0;
  double xc = clampd(px, 0, n - 1), yc = clampd(py, 0, n - 1), zc = clampd(pz, 0, n - 1);
  double sx = xc - std::floor(xc), sy = yc - std::floor(yc), sz = zc - std::floor(zc);
  int x0 = std::min(int(xc), n - 1), x1 = std::min(x0 + 1, n - 1);
  int y0 = std::min(int(yc), n - 1), y1 = std::min(y0 + 1, n - 1);
  int z0 = std::min(int(zc), n - 1), z1 = std::min(z0 + 1, n - 1);
  auto at = [&](int x, int y, int z) { return double(f[x + y * n + z * n * n]); };
  double f0 = lerpd(lerpd(at(x0, y0, z0), at(x1, y0, z0), sx), lerpd(at(x0, y1, z0), at(x1, y1, z0), sx), sy);
  double f1 = lerpd(lerpd(at(x0, y0, z1), at(x1, y0, z1), sx), lerpd(at(x0, y1, z1), at(x1, y1, z1), sx), sy);
  return slopeZ ? f1 - f0 : lerpd(f0, f1, sz);
}
static double interp2Dd(double px, double py, const float* f, int n, bool slopeY)
{
  if (slopeY && !(py > 0.0 && py < double(n - 1))) return 0.0;
  double xc = clampd(px, 0, n - 1), yc = clampd(py, 0, n - 1);
  double sx = xc - std::floor(xc), sy = yc - std::floor(yc);
  int x0 = std::min(int(xc), n - 1), x1 = std::min(x0 + 1, n - 1);
  int y0 = std::min(int(yc), n - 1), y1 = std::min(y0 + 1, n - 1);
  double f0 = lerpd(f[x0 + y0 * n], f[x1 + y0 * n], sx), f1 = lerpd(f[x0 + y1 * n], f[x1 + y1 * n], sx);
  return slopeY ? f1 - f0 : lerpd(f0, f1, sy);
}
struct D3 { double x, y, z; };
static D3 d3(V3 a) { return D3{a.x, a.y, a.z}; }
static double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static D3 normd(D3 a) { double l = std::sqrt(dotd(a, a)); return D3{a.x / l, a.y / l, a.z / l}; }

static double interp1Dd(double x, const float* f, int n, bool slope)
{
  if (slope && !(x > 0.0 && x < double(n - 1))) return 0.0;
  double xc = clampd(x, 0, n - 1);
  int x0 = std::min(int(xc), n - 1), x1 = std::min(x0 + 1, n - 1);
  return slope ? double(f[x1]) - f[x0] : lerpd(f[x0], f[x1], xc - std::floor(xc));
}

// GGX D and the Smith masking term in FP64, with their alpha derivatives (ggx_distribution.h:24-102).
struct GgxD { double D, G, dD, dG; };
static GgxD ggxD(double a, D3 o0, D3 i0, D3 h0)
{
  double sh = h0.x * h0.x + h0.y * h0.y, z2 = h0.z * h0.z, q = a * a * z2 + sh;
  auto lam = [&](D3 w) { double t2 = (w.x * w.x + w.y * w.y) / (w.z * w.z); return 0.5 * (-1 + std::sqrt(1 + a * a * t2)); };
  auto dlam = [&](D3 w) { double t2 = (w.x * w.x + w.y * w.y) / (w.z * w.z); return 0.5 * a * t2 / std::sqrt(1 + a * a * t2); };
  GgxD r;
  r.D = a * a / (M_PI * q * q);
  r.G = 1.0 / (1.0 + lam(o0) + lam(i0));
  r.dD = a > 0.001 ? (2 * a / (M_PI * q * q)) * (1 - 2 * a * a * z2 / q) : 0.0;
  r.dG = a > 0.001 ? -r.G * r.G * (dlam(o0) + dlam(i0)) : 0.0;
  return r;
}
static double fresnelDielD(double cosI, double eta)  // optics.h:84-99: cosI clamped at 0
{
  double ci = std::max(cosI, 0.0), kk = 1.0 - eta * eta * (1.0 - ci * ci);
  if (kk < 0.0) return 1.0;
  double ct = std::sqrt(kk);
  double Rper = (eta * ci - ct) / (eta * ci + ct), Rpar = (ci - eta * ct) / (ci + eta * ct);
  return 0.5 * (Rpar * Rpar + Rper * Rper);
}
static double fresnelCondD(double cosI, double eta, double k)  // optics.h:179-190
{
  double c = std::fabs(cosI), c2 = c * c, tmp = eta * eta + k * k, e2 = 2 * eta * c;
  return 0.5 * ((tmp * c2 - e2 + 1) / (tmp * c2 + e2 + 1) + (tmp - e2 + c2) / (tmp + e2 + c2));
}
static double artisticFresnelD(double cosI, double r, double g)  // optics.h:214-226
{
  r = std::min(r, 0.99);
  double nMin = (1 - r) / (1 + r), nMax = (1 + std::sqrt(r)) / (1 - std::sqrt(r)), n = g * nMin + (1 - g) * nMax;
  double k2 = ((n + 1) * (n + 1) * r - (n - 1) * (n - 1)) / (1 - r);
  return fresnelCondD(cosI, n, std::sqrt(std::max(k2, 0.0)));
}
static double artisticAvgD(double r, double g)  // optics.h:228-240
{
  return 0.087237 + 0.0230685 * g - 0.0864902 * g * g + 0.0774594 * g * g * g + 0.782654 * r - 0.136432 * r * r +
         0.278708 * r * r * r + 0.19744 * g * r + 0.0360605 * g * g * r - 0.2586 * g * r * r;
}
static double fresnelDielAvgD(double eta)  // optics.h:144-175
{
  double re = 1.0 / eta;
  if (re >= 1.0) return (re - 1.0) / (4.08567 + 1.00071 * re);
  return 0.997118 + 0.1014 * re - 0.965241 * re * re - 0.130607 * re * re * re;
}
// dielectric albedo tables and their roughness slopes (albedo_tables.h:82-154)
static double dielDirD(const Luts& L, double cosO, double eta, double r, bool slope)
{
  if (eta <= 1.0) return interp3Dd(cosO * 15, (eta - 1.0 / 3.0) / (2.0 / 3.0) * 15, r * 15, L.dielEtaDir, 16, slope) * (slope ? 15 : 1);
  return interp3Dd(cosO * 15, (eta - 1.0) / 2.0 * 15, r * 15, L.dielRcpEtaDir, 16, slope) * (slope ? 15 : 1);
}
static double dielAvgD(const Luts& L, double eta, double r, bool slope)
{
  if (eta <= 1.0) return interp2Dd((eta - 1.0 / 3.0) / (2.0 / 3.0) * 15, r * 15, L.dielEtaAvg, 16, slope) * (slope ? 15 : 1);
  return interp2Dd((eta - 1.0) / 2.0 * 15, r * 15, L.dielRcpEtaAvg, 16, slope) * (slope ? 15 : 1);
}

// Rough conductor: value and derivatives per channel, albedo = reflectivity r (the mapping to n, k
// by a central difference of width 2e-3, as the device does), alpha through D, G2 and the tables.
static GradD conductorGradD(const Luts& L, const Bsdf& b, D3 N, D3 U, D3 Vv, D3 o, D3 i)
{
  GradD g;
  const double cosO = dotd(o, N), cosI = dotd(i, N);
  if (!(cosO > 0.0 && cosI > 0.0)) return g;
  const D3 h = normd(D3{o.x + i.x, o.y + i.y, o.z + i.z});
  const double cosOH = dotd(o, h);
  const D3 o0{dotd(o, U), dotd(o, Vv), dotd(o, N)}, i0{dotd(i, U), dotd(i, Vv), dotd(i, N)}, h0{dotd(h, U), dotd(h, Vv), dotd(h, N)};
  const double r = b.roughness, a = std::max(r * r, 0.001);
  const GgxD m = ggxD(a, o0, i0, h0);
  const double kF = m.D * m.G / (4 * cosO), dkF = (m.dD * m.G + m.D * m.dG) / (4 * cosO);
  const double Eo = interp2Dd(cosO * 31, r * 31, L.ggxDir, 32, false), Ei = interp2Dd(cosI * 31, r * 31, L.ggxDir, 32, false);
  const double dEo = 31 * interp2Dd(cosO * 31, r * 31, L.ggxDir, 32, true), dEi = 31 * interp2Dd(cosI * 31, r * 31, L.ggxDir, 32, true);
  const double E = interp1Dd(r * 31, L.ggxAvg, 32, false), dE = 31 * interp1Dd(r * 31, L.ggxAvg, 32, true);
  const double A = (1 - Eo) * (1 - Ei), B = 1 / (1 - E), dA = -dEo * (1 - Ei) - (1 - Eo) * dEi, dB = dE * B * B;
  const double drda = (a > 0.001 && r > 0) ? 0.5 / r : 0.0, w = cosI / M_PI;
  const double rr[3] = {b.rr.x, b.rr.y, b.rr.z}, gg[3] = {b.gg.x, b.gg.y, b.gg.z};
  for (int k = 0; k < 3; ++k)
  {
    const double F = artisticFresnelD(cosOH, rr[k], gg[k]);
    const double Fa = artisticAvgD(rr[k], gg[k]), den = 1 - Fa * (1 - E);
    const double S = Fa * Fa * E / den, dSdE = Fa * Fa * (1 - Fa) / (den * den);
    const double dSdF = (2 * Fa * E * den + Fa * Fa * E * (1 - E)) / (den * den);
    g.v[k] = F * kF + S * A * B * w;
    g.dAl[k] = F * dkF + (dSdE * dE * A * B + S * dA * B + S * A * dB) * w * drda;
    double dF = 0.0;
    if (rr[k] < 0.99)
    {
      const double lo = std::max(rr[k] - 1e-3, 0.0), hi = std::min(rr[k] + 1e-3, 0.99);
      dF = (artisticFresnelD(cosOH, hi, gg[k]) - artisticFresnelD(cosOH, lo, gg[k])) / (hi - lo);
    }
    const double dFa = 0.782654 - 2 * 0.136432 * rr[k] + 3 * 0.278708 * rr[k] * rr[k] + 0.19744 * gg[k] +
                       0.0360605 * gg[k] * gg[k] - 2 * 0.2586 * gg[k] * rr[k];
    g.dA[k] = dF * kF + dSdF * dFa * A * B * w;
  }
  return g;
}

// Rough dielectric: value and alpha derivative (no albedo), single-scattering lobe through D, G2,
// multiple-scattering lobe through the tables and fmsRatio.
static GradD dielectricGradD(const Luts& L, const Bsdf& b, const Ctx& ctx, D3 N, D3 U, D3 Vv, D3 o, D3 i)
{
  GradD g;
  const double cosO = dotd(o, N);
  if (!(cosO > 0.0)) return g;
  const double cosI = dotd(i, N);
  if (!(cosI * dotd(i, d3(ctx.Ng)) > 0.0)) return g;
  const double eta = b.deta, r = b.roughness, a = std::max(r * r, 0.001);
  const bool refl = cosI > 0.0;
  D3 h = refl ? D3{o.x + i.x, o.y + i.y, o.z + i.z} : D3{eta * o.x + i.x, eta * o.y + i.y, eta * o.z + i.z};
  h = normd(h);
  if (dotd(h, N) < 0.0) h = D3{-h.x, -h.y, -h.z};
  const double cosOH = dotd(o, h), cosIH = dotd(i, h);
  const double F = fresnelDielD(cosOH, eta);
  const D3 o0{dotd(o, U), dotd(o, Vv), dotd(o, N)}, i0{dotd(i, U), dotd(i, Vv), dotd(i, N)}, h0{dotd(h, U), dotd(h, Vv), dotd(h, N)};
  const GgxD m = ggxD(a, o0, i0, h0);
  double pre;
  if (refl) pre = F / (4 * cosO);
  else
  {
    const double den = cosOH + cosIH / eta;
    pre = (1 - F) / (eta * eta) * std::fabs(cosIH) / (den * den) * std::fabs(cosOH) / cosO;
  }
  const double single = pre * m.D * m.G, dSingle = pre * (m.dD * m.G + m.D * m.dG);
  const double Fe = fresnelDielAvgD(eta), Fr = fresnelDielAvgD(1 / eta);
  const double Ee = dielAvgD(L, eta, r, false), Er = dielAvgD(L, 1 / eta, r, false);
  const double dEe = dielAvgD(L, eta, r, true), dEr = dielAvgD(L, 1 / eta, r, true);
  const double aa = (1 - Fe) / (1 - Er), bb = (1 - Fr) / (1 - Ee), x = bb / (aa + bb), ratio = 1 - x * (1 - Fe);
  const double daa = (1 - Fe) / ((1 - Er) * (1 - Er)) * dEr, dbb = (1 - Fr) / ((1 - Ee) * (1 - Ee)) * dEe;
  const double dRatio = -(1 - Fe) * (dbb * aa - bb * daa) / ((aa + bb) * (aa + bb));
  const double Eo = dielDirD(L, cosO, eta, r, false), dEo = dielDirD(L, cosO, eta, r, true);
  double fms, dFms;
  if (refl)
  {
    const double Ei = dielDirD(L, cosI, eta, r, false), dEi = dielDirD(L, cosI, eta, r, true);
    const double A = (1 - Eo) * (1 - Ei), dA = -dEo * (1 - Ei) - (1 - Eo) * dEi, B = 1 / (1 - Ee), dB = dEe * B * B;
    fms = ratio * A * B * cosI / M_PI;
    dFms = (dRatio * A * B + ratio * dA * B + ratio * A * dB) * cosI / M_PI;
  }
  else
  {
    const double ac = std::fabs(cosI);
    const double Ei = dielDirD(L, ac, 1 / eta, r, false), dEi = dielDirD(L, ac, 1 / eta, r, true);
    const double A = (1 - Eo) * (1 - Ei), dA = -dEo * (1 - Ei) - (1 - Eo) * dEi, B = 1 / (1 - Er), dB = dEr * B * B;
    fms = (1 - ratio) * A * B * ac / M_PI;
    dFms = (-dRatio * A * B + (1 - ratio) * dA * B + (1 - ratio) * A * dB) * ac / M_PI;
  }
  const double drda = (a > 0.001 && r > 0) ? 0.5 / r : 0.0;
  for (int k = 0; k < 3; ++k)
  {
    g.v[k] = single + fms;
    g.dAl[k] = dSingle + dFms * drda;
  }
  return g;
}

static GradD lobeEvalGradD(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi)
{
  GradD g;
  if (!(b.importance > 0.f)) return g;
  const D3 N = d3(b.fr.N), U = d3(b.fr.U), Vv = d3(b.fr.V), o = d3(wo), i = d3(wi);
  // the diffuse base per unit albedo: Lambert, or Oren-Nayar in Fujii's form (rough_diffuse.h:65-75)
  double kBase;
  {
    const double c = std::max(dotd(i, N), 0.0);
    if (!(b.diffRough > 0.f)) kBase = c / M_PI;
    else
    {
      const double cO = std::max(dotd(o, N), 0.0), sr = b.diffRough;
      const double sdot = dotd(o, i) - cO * c;
      const double sInvt = sdot > 0.0 ? sdot / std::max(cO, c) : sdot;
      kBase = c * (1.0 + sr * sInvt) / (M_PI + (M_PI / 2.0 - 2.0 / 3.0) * sr);
    }
  }
  if (b.kind == kDiffuse)
  {
    double rho[3] = {b.color.x, b.color.y, b.color.z};
    for (int k = 0; k < 3; ++k) { g.v[k] = rho[k] * kBase; g.dA[k] = kBase; }
    return g;
  }
  if (b.kind == kConductor) return conductorGradD(L, b, N, U, Vv, o, i);
  if (b.kind == kDielectric) return dielectricGradD(L, b, ctx, N, U, Vv, o, i);
  if (b.kind != kCoated) return g;
  const double cosO = dotd(o, N);
  if (!(cosO > 0.0)) return g;
  const double cosI = dotd(i, N);
  if (!(cosI * dotd(i, d3(ctx.Ng)) > 0.0)) return g;
  const double eta = b.ceta, s = b.scale, r = b.roughness, a = std::max(double(r) * r, 0.001);
  auto refr = [&](double c) { return std::sqrt(std::max(1.0 - eta * eta * (1.0 - c * c), 0.0)); };
  double length = 1.0 / refr(cosO) + 1.0 / refr(std::fabs(cosI));
  if (cosI <= 0.0) length *= 0.5;
  const double t = double(b.thickness) * length;
  const double coat[3] = {b.coatColor.x, b.coatColor.y, b.coatColor.z}, rho[3] = {b.color.x, b.color.y, b.color.z};
  const double e = (eta - 1.0 / 3.0) / (1.0 - 1.0 / 3.0);
  const double Ro = s * interp3Dd(cosO * 15, e * 15, r * 15, L.dielReflDir, 16, false);
  const double Ri = s * interp3Dd(std::fabs(cosI) * 15, e * 15, r * 15, L.dielReflDir, 16, false);
  const double Ravg = s * interp2Dd(e * 15, r * 15, L.dielReflAvg, 16, false);
  const double dRo = 15 * s * interp3Dd(cosO * 15, e * 15, r * 15, L.dielReflDir, 16, true);
  const double dRi = 15 * s * interp3Dd(std::fabs(cosI) * 15, e * 15, r * 15, L.dielReflDir, 16, true);
  const double dRavg = 15 * s * interp2Dd(e * 15, r * 15, L.dielReflAvg, 16, true);
  const double inv = 1.0 / (1.0 - Ravg);
  const double T = (1 - Ro) * (1 - Ri) * inv;
  const double dTdr = (-dRo * (1 - Ri) - (1 - Ro) * dRi) * inv + (1 - Ro) * (1 - Ri) * dRavg * inv * inv;
  const double drda = (a > 0.001 && r > 0) ? 0.5 / r : 0.0;
  for (int k = 0; k < 3; ++k)
  {
    double unit = (1.0 + s * (std::pow(coat[k], t) - 1.0)) * kBase;
    g.v[k] = rho[k] * unit * T;
    g.dA[k] = unit * T;
    g.dAl[k] = rho[k] * unit * dTdr * drda;
  }
  if (cosI > 0.0)
  {
    D3 h = normd(D3{o.x + i.x, o.y + i.y, o.z + i.z});
    double cosOH = dotd(o, h);
    double ci = std::max(cosOH, 0.0), kk = 1.0 - eta * eta * (1.0 - ci * ci), F = 1.0;
    if (kk >= 0.0)
    {
      double ct = std::sqrt(kk);
      double Rper = (eta * ci - ct) / (eta * ci + ct), Rpar = (ci - eta * ct) / (ci + eta * ct);
      F = 0.5 * (Rpar * Rpar + Rper * Rper);
    }
    F *= s;
    D3 o0{dotd(o, U), dotd(o, Vv), dotd(o, N)}, i0{dotd(i, U), dotd(i, Vv), dotd(i, N)}, h0{dotd(h, U), dotd(h, Vv), dotd(h, N)};
    double sh = h0.x * h0.x + h0.y * h0.y, z2 = h0.z * h0.z, q = a * a * z2 + sh;
    double D = a * a / (M_PI * q * q);
    double dD = a > 0.001 ? (2 * a / (M_PI * q * q)) * (1 - 2 * a * a * z2 / q) : 0.0;
    auto lam = [&](D3 w) { double t2 = (w.x * w.x + w.y * w.y) / (w.z * w.z); return 0.5 * (-1 + std::sqrt(1 + a * a * t2)); };
    auto dlam = [&](D3 w) { double t2 = (w.x * w.x + w.y * w.y) / (w.z * w.z); return 0.5 * a * t2 / std::sqrt(1 + a * a * t2); };
    double G = 1.0 / (1.0 + lam(o0) + lam(i0));
    double dG = a > 0.001 ? -G * G * (dlam(o0) + dlam(i0)) : 0.0;
    double kf = F / (4 * cosO);
    for (int k = 0; k < 3; ++k) { g.v[k] += kf * D * G; g.dAl[k] += kf * (dD * G + D * dG); }
  }
  return g;
}

// Under a sheen layer (sheen.h:218-279): the lobe attenuated by T = min(1 - Ro, 1 - Ri) plus the
// sheen reflection, whose own parameters are not differentiated.
static GradD bsdfEvalGradD(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi)
{
  GradD g = lobeEvalGradD(L, b, ctx, wo, wi);
  if (!b.sheen) return g;
  const float cosO = dot(wo, b.fr.N), cosI = dot(wi, b.fr.N);
  if (!(cosO > 0.f) || !(cosI * dot(wi, ctx.Ng) > 0.f)) return GradD();
  const double Ro = double(sheenAlbedo(L, cosO, b.sheenRough)) * b.sheenScale;
  const double Ri = double(sheenAlbedo(L, std::fabs(cosI), b.sheenRough)) * b.sheenScale;
  const double T = std::min(1.0 - Ro, 1.0 - Ri);
  V3 sr = cosI > 0.f ? sheenReflection(b, wo, wi, cosO, cosI) : V3(0.f);
  const double srv[3] = {sr.x, sr.y, sr.z};
  for (int k = 0; k < 3; ++k)
  {
    g.v[k] = g.v[k] * T + srv[k];
    g.dA[k] *= T;
    g.dAl[k] *= T;
  }
  return g;
}

// One path vertex as the adjoint sees it: the vertex's material, its unoccluded next-event
// contribution with the BSDF derivatives at the light direction, and the radiance gathered up
// to and including that estimate with the derivatives at the sampled continuation.
struct AdjVertex { int mat = 0; bool nee = false, ext = false; V3 dirL, Lrun; GradD gNee, gExt; };
struct AdjPath { std::vector<AdjVertex> v; V3 L; };

static void tracePath(const Scene& s, const Camera& cam, const RenderParams& P, int x, int y, uint32_t sampleIndex,
                      Accum& A, Stats& st, bool recordSample, AdjPath* adj = nullptr)
{
  const int npix = P.W * P.H;
  const int pix = y * P.W + x;
  const uint32_t pixelId = uint32_t(x) | (uint32_t(y) << 16);
  Sampler smp;
  smp.set(P.seed, sampleIndex, pixelId);
  float psx, psy, lsx, lsy;
  smp.get2D(0, psx, psy);  // sampleDimPixel
  pixelSample(s, psx, psy);
  float imx = (float(x) + psx) * (1.0f / float(P.W));
  float imy = (float(y) + psy) * (1.0f / float(P.H));
  smp.get2D(2, lsx, lsy);  // sampleDimLens

  // camera/pinhole_camera.h:72-133, thin_lens_camera.h:155-182
  V3 U(cam.U[0], cam.U[1], cam.U[2]), Vb(cam.V[0], cam.V[1], cam.V[2]), Nb(cam.N[0], cam.N[1], cam.N[2]);
  float aspect = float(P.W) / float(P.H);
  float hh = std::tan(0.5f * cam.fov), hw = hh * aspect;
  V3 imageO = toGlobal(Basis{U, Vb, Nb}, V3(-hw, hh, -1.f));
  V3 imageDx = toGlobal(Basis{U, Vb, Nb}, V3(2.f * hw, 0.f, 0.f));
  V3 imageDy = toGlobal(Basis{U, Vb, Nb}, V3(0.f, -2.f * hh, 0.f));
  V3 org(cam.pos[0], cam.pos[1], cam.pos[2]);
  V3 dir;
  float tnear;
  if (cam.lensRadius == 0.f)  // render/device_cpu.cpp:117-128 picks the pinhole camera
  {
    dir = imageO + imx * imageDx + imy * imageDy;
    float len = length(dir);
    dir = dir * (1.0f / len);
    tnear = cam.nearClip * len;
  }
  else
  {
    float lx, ly;
    uniformSampleDisk(lsx, lsy, lx, ly);
    lx *= cam.lensRadius; ly *= cam.lensRadius;
    V3 lensOffset = toGlobal(Basis{U, Vb, Nb}, V3(lx, ly, 0.f));
    dir = cam.focalDistance * (imageO + imx * imageDx + imy * imageDy) - lensOffset;
    float len = length(dir);
    org = org + lensOffset;
    dir = dir * (1.0f / len);
    tnear = (cam.nearClip / cam.focalDistance) * len;
  }

  V3 T(1.f), L(0.f);
  int mediumId = 0;
  float lastPdf = 1.f;
  int lastType = BsdfType::SpecularReflection;
  bool aov1Done = false, aov2Done = false;
  uint32_t hash = 0x9e3779b9u;
  int firstPrim = -2;
  V3 camN(Nb);

  for (int depth = 0;; ++depth)
  {
    const int sampleDim = 4 + 7 * depth;
    RayHit r;
    r.o = org; r.d = dir; r.tnear = tnear; r.tfar = kPosMax;
    traverse(s, r, false, nullptr);
    st.rays++;
    if (depth == 0) firstPrim = r.prim;
    hash = hashStep(hash, uint32_t(r.prim + 1));
    if (r.inst >= 0) hash = hashStep(hash, 0x40000000u | uint32_t(r.inst));  // the instance hit (two-level scenes)
    int matId = r.prim >= 0 ? s.triMat[r.prim] : 0;
    if (matId == 0)
    {
      // miss shading (:285-415)
      float lightPdf;
      uint32_t envMask;
      V3 Le = evalEnv(s.ls, dir, lightPdf, &envMask);
      hash = hashStep(hash, 0x100u | envMask);  // which environment lights the ray sees
      float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
      L += T * w * Le;
      V3 c = clampL(L, P.maxRadiance);
      addColor(A, npix, pix, c);
      if (!aov1Done)
      {
        add3(A.albedo1, A.albedo1W, npix, pix, V3(0.f));
        add3(A.normal1, A.normal1W, npix, pix, V3(0.f));
        if (A.depth) { A.depth[pix] += kPosMax * 2.f; A.depthW[pix] += 1.f; }  // +inf
      }
      if (!aov2Done)
      {
        add3(A.albedo, A.albedoW, npix, pix, V3(0.f));
        add3(A.normal, A.normalW, npix, pix, V3(0.f));
      }
      if (recordSample)
      {
        if (A.sampleL) { A.sampleL[pix] = c.x; A.sampleL[npix + pix] = c.y; A.sampleL[2 * npix + pix] = c.z; }
        if (A.pathHash) A.pathHash[pix] = hash;
        if (A.firstPrim) A.firstPrim[pix] = firstPrim;
      }
      if (adj) adj->L = L;
      return;
    }
    if (isEmissive(s.mats[matId]))
    {
      // area-light hit (:419-582), at any depth: emission with MIS, then the path ends
      float lightPdf;
      V3 Le = evalArea(s.ls, dir, r.tfar, r.prim, lightPdf);
      float w = ((lastType & BsdfType::Smooth) == 0) ? 1.f : misWeight(lastPdf, lightPdf);
      L += T * w * Le;
      V3 c = clampL(L, P.maxRadiance);
      addColor(A, npix, pix, c);
      V3 Nv = normalize(-dir);
      if (!aov1Done)
      {
        add3(A.albedo1, A.albedo1W, npix, pix, V3(1.f));
        add3(A.normal1, A.normal1W, npix, pix, Nv);
        V3 ph = org + dir * r.tfar;
        if (A.depth) { A.depth[pix] += dot(ph - V3(cam.pos[0], cam.pos[1], cam.pos[2]), -camN); A.depthW[pix] += 1.f; }
      }
      if (!aov2Done)
      {
        add3(A.albedo, A.albedoW, npix, pix, V3(1.f));
        add3(A.normal, A.normalW, npix, pix, Nv);
      }
      if (recordSample)
      {
        if (A.sampleL) { A.sampleL[pix] = c.x; A.sampleL[npix + pix] = c.y; A.sampleL[2 * npix + pix] = c.z; }
        if (A.pathHash) A.pathHash[pix] = hash;
        if (A.firstPrim) A.firstPrim[pix] = firstPrim;
      }
      if (adj) adj->L = L;
      return;
    }

    bool ext = false;
    bool dirOk = false;
    AdjVertex av;
    bool haveAv = false;
    V3 dirL(0.f);
    V3 shadowO, shadowD;
    float shadowNear = 0.f, shadowFar = 0.f;
    V3 nextO, nextD;
    float nextNear = 0.f;
    if (depth < P.maxDepth)
    {
      const Material& m = s.mats[matId];
      Ctx ctx = postIntersect(s, org, dir, r.tfar, r.prim, r.u, r.v, r.inst);
      V3 wo = -dir;
      const int mediumIntId = s.matMedium[matId];
      V3 texBase;
      const bool textured = !s.baseTex.empty() && s.baseTex[size_t(matId)] >= 0;
      if (textured) texBase = texturedBase(s, m, r.prim, r.u, r.v, s.baseTex[size_t(matId)]);
      Bsdf b = makeBsdf(s.luts, m, ctx, wo, mediumId, mediumIntId, textured ? &texBase : nullptr);
      // sampler.resetSample with the same index and pixel (:641-642) leaves the state unchanged
      bool mAov1 = !aov1Done, mAov2 = !aov2Done;
      bool mRR = !(mAov1 || mAov2);
      // volume transmission (:652-657): a segment that ends here inside a medium is attenuated
      // by the medium's colour to the power of its length
      if (mediumId > 0)
      {
        const V3 mc = s.mediumColor[mediumId];
        T = T * V3(std::pow(mc.x, r.tfar), std::pow(mc.y, r.tfar), std::pow(mc.z, r.tfar));
      }

      // direct light (:659-718)
      if (b.flags & BsdfType::Smooth)
      {
        float a0, a1, a2;
        smp.get3D(sampleDim + 3, a0, a1, a2);
        V3 wiL(0.f);
        float lightPdf, dist;
        int picked;
        V3 lw = sampleLight(s.ls, ctx.p, wiL, lightPdf, dist, a0, a1, a2, &picked);
        hash = hashStep(hash, 0x200u | uint32_t(picked));  // which light NEE sampled
        hash = hashStep(hash, dirCell(wiL));
        if (reduceMax(lw) > 0.f && lightPdf > 0.f)
        {
          float bsdfPdf;
          V3 f = bsdfEval(s.luts, b, ctx, wo, wiL, bsdfPdf);
          if (reduceMax(f) > 0.f)
          {
            shadowO = ctx.p; shadowD = wiL; shadowNear = ctx.eps; shadowFar = dist - ctx.eps;
            V3 dirT = T * f;
            if (depth >= P.rrDepth)
            {
              float contProb = mRR ? std::min(luminance(dirT * (1.0f / bsdfPdf)), 0.95f) : 1.f;
              bsdfPdf *= contProb;
            }
            float w = misWeight(lightPdf, bsdfPdf);
            dirL = dirT * w * lw;
            dirOk = true;
            if (adj) av.gNee = bsdfEvalGradD(s.luts, b, ctx, wo, wiL);
          }
        }
      }

      // BSDF sample (:720-730)
      float b0, b1, b2;
      smp.get3D(sampleDim + 0, b0, b1, b2);
      V3 wi(0.f);
      float pdf = 0.f;
      int type = 0;
      V3 bw = bsdfSample(s.luts, b, ctx, wo, wi, pdf, type, b0, b1, b2);
      ext = reduceMax(bw) > 0.f && pdf > 0.f;
      int extType = ext ? type : 0;
      hash = hashStep(hash, uint32_t(extType));
      if (ext) hash = hashStep(hash, dirCell(wi));

      if (mAov1)  // first-hit AOVs (:739-819)
      {
        add3(A.albedo1, A.albedo1W, npix, pix, bsdfAlbedo(b));
        add3(A.normal1, A.normal1W, npix, pix, bsdfNormal(b));
        if (A.depth) { A.depth[pix] += dot(ctx.p - V3(cam.pos[0], cam.pos[1], cam.pos[2]), -camN); A.depthW[pix] += 1.f; }
        aov1Done = true;
      }
      mAov2 = mAov2 && (extType & BsdfType::Delta) == 0;  // (:828-832)
      if (mAov2)
      {
        add3(A.albedo, A.albedoW, npix, pix, bsdfAlbedo(b));
        add3(A.normal, A.normalW, npix, pix, bsdfNormal(b));
        aov2Done = true;
      }

      T *= bw;
      ext = ext && reduceMax(T) >= P.minWeight;
      if (depth >= P.rrDepth)  // (:880-887)
      {
        float contProb = mRR ? std::min(luminance(T), 0.95f) : 1.f;
        ext = ext && smp.get1D(sampleDim + 6) <= contProb;
        T *= 1.0f / contProb;
        pdf *= contProb;
      }
      if (ext)
      {
        nextO = ctx.p; nextD = wi; nextNear = ctx.eps;
        if ((type & BsdfType::Continuation) == 0) { lastPdf = pdf; lastType = type; }
        // material.h:84-95: crossing a media interface swaps interior/exterior
        if (isMediaInterface(m) && dot(wo, ctx.Ng) * dot(wi, ctx.Ng) < 0.f)
          mediumId = (mediumId == mediumIntId) ? 0 : mediumIntId;
      }
      if (adj)
      {
        haveAv = true;
        av.mat = matId;
        av.dirL = dirL;
        // a delta (continuation) sample carries no differentiated BSDF value
        av.ext = ext && !(type & BsdfType::Delta);
        if (av.ext) av.gExt = bsdfEvalGradD(s.luts, b, ctx, wo, wi);
      }
    }

    // shadow ray + final pass (:951-1065)
    if (dirOk)
    {
      RayHit sr;
      sr.o = shadowO; sr.d = shadowD; sr.tnear = shadowNear; sr.tfar = shadowFar;
      V3 Tsh(1.f);
      uint32_t crossed = 0;
      if (s.anyTransparent) traverseShadow(s, sr, Tsh, crossed);
      else traverse(s, sr, true, nullptr);
      st.rays++;
      // the visibility event, with the transparent surfaces crossed, is part of the path
      hash = hashStep(hash, sr.prim < 0 ? 1u + 4u * crossed : 2u);
      if (sr.prim < 0)
      {
        if (s.anyTransparent) { dirL *= Tsh; av.dirL = dirL; }  // the shadow ray's transmittance
        L += dirL;
      }
      av.nee = sr.prim < 0;
    }
    if (haveAv)
    {
      av.nee = av.nee && dirOk;
      av.Lrun = L;
      adj->v.push_back(av);
    }
    if (!ext)
    {
      V3 c = clampL(L, P.maxRadiance);
      addColor(A, npix, pix, c);
      if (recordSample)
      {
        if (A.sampleL) { A.sampleL[pix] = c.x; A.sampleL[npix + pix] = c.y; A.sampleL[2 * npix + pix] = c.z; }
        if (A.pathHash) A.pathHash[pix] = hash;
        if (A.firstPrim) A.firstPrim[pix] = firstPrim;
      }
      if (adj) adj->L = L;
      return;
    }
    org = nextO; dir = nextD; tnear = nextNear;
  }
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------------------
extern "C" {

void* orc_scene_create(int nverts, const float* pos, const float* nrm, int ntris, const int* idx, const int* triMat,
                       int nmats, const void* mats)
{
  Scene* s = new Scene();
  s->pos.resize(nverts);
  for (int i = 0; i < nverts; ++i) s->pos[i] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  if (nrm)
  {
    s->nrm.resize(nverts);
    for (int i = 0; i < nverts; ++i) s->nrm[i] = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  s->idx.assign(idx, idx + 3 * size_t(ntris));
  s->triMat.assign(triMat, triMat + ntris);
  s->mats.assign((const Material*)mats, (const Material*)mats + nmats);
  computeBounds(*s);
  assignMedia(*s);
  precomputeLuts(s->luts);
  buildBvh(*s);
  return s;
}

void orc_scene_destroy(void* p) { delete (Scene*)p; }

// Two-level scene: meshes [meshBegin[m], meshBegin[m + 1]) of the triangles and n instances,
// each an int mesh followed by the 12 floats of the rows of [L | t] (x_world = L x + t); n = 0:
// one level.  The inverse, the linear part and its inverse transpose are formed in double and
// rounded to float, as the product's host does.
void orc_scene_set_instances(void* p, int nMeshes, const int* meshBegin, int n, const void* inst)
{
  Scene& s = *(Scene*)p;
  s.inst.clear();
  s.meshTrees.clear();
  s.meshBegin.clear();
  if (n <= 0) return;
  s.meshBegin.assign(meshBegin, meshBegin + nMeshes + 1);
  s.meshTrees.resize(size_t(nMeshes));
  for (int m = 0; m < nMeshes; ++m) buildTree(s, meshBegin[m], meshBegin[m + 1], s.meshTrees[size_t(m)].nodes, s.meshTrees[size_t(m)].tris);
  const char* q = (const char*)inst;
  for (int i = 0; i < n; ++i, q += 52)
  {
    Scene::Instance I;
    std::memcpy(&I.mesh, q, 4);
    float x[12];
    std::memcpy(x, q + 4, 48);
    const double L[9] = {x[0], x[1], x[2], x[4], x[5], x[6], x[8], x[9], x[10]}, Tt[3] = {x[3], x[7], x[11]};
    // inverse by cofactors / determinant
    const double c00 = L[4] * L[8] - L[5] * L[7], c01 = L[5] * L[6] - L[3] * L[8], c02 = L[3] * L[7] - L[4] * L[6];
    const double det = L[0] * c00 + L[1] * c01 + L[2] * c02;
    double A[9];
    A[0] = c00 / det; A[1] = (L[2] * L[7] - L[1] * L[8]) / det; A[2] = (L[1] * L[5] - L[2] * L[4]) / det;
    A[3] = c01 / det; A[4] = (L[0] * L[8] - L[2] * L[6]) / det; A[5] = (L[2] * L[3] - L[0] * L[5]) / det;
    A[6] = c02 / det; A[7] = (L[1] * L[6] - L[0] * L[7]) / det; A[8] = (L[0] * L[4] - L[1] * L[3]) / det;
    for (int r = 0; r < 3; ++r)
    {
      for (int k = 0; k < 3; ++k)
      {
        I.inv[4 * r + k] = float(A[3 * r + k]);
        I.lin[3 * r + k] = float(L[3 * r + k]);
        I.nrm[3 * r + k] = float(A[3 * k + r]);
      }
      I.inv[4 * r + 3] = float(-(A[3 * r] * Tt[0] + A[3 * r + 1] * Tt[1] + A[3 * r + 2] * Tt[2]));
    }
    s.inst.push_back(I);
  }
  // scene bounds and radius (render/scene.cpp:367): each instanced mesh's vertex box, its
  // corners through x' = L x + t in double, rounded to float
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  bool any = false;
  for (int i = 0; i < n; ++i)
  {
    const int m = s.inst[size_t(i)].mesh;
    if (meshBegin[m + 1] == meshBegin[m]) continue;
    double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
    for (int t = meshBegin[m]; t < meshBegin[m + 1]; ++t)
      for (int j = 0; j < 3; ++j)
      {
        const V3 p = s.pos[s.idx[3 * t + j]];
        const double pv[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) { blo[k] = std::min(blo[k], pv[k]); bhi[k] = std::max(bhi[k], pv[k]); }
      }
    float x[12];
    std::memcpy(x, (const char*)inst + 52 * size_t(i) + 4, 48);
    for (int corner = 0; corner < 8; ++corner)
    {
      const double pc[3] = {(corner & 1) ? bhi[0] : blo[0], (corner & 2) ? bhi[1] : blo[1], (corner & 4) ? bhi[2] : blo[2]};
      for (int k = 0; k < 3; ++k)
      {
        const double w = double(x[4 * k]) * pc[0] + double(x[4 * k + 1]) * pc[1] + double(x[4 * k + 2]) * pc[2] + double(x[4 * k + 3]);
        lo[k] = std::min(lo[k], w);
        hi[k] = std::max(hi[k], w);
      }
    }
    any = true;
  }
  if (any)
  {
    s.bmin = V3(float(lo[0]), float(lo[1]), float(lo[2]));
    s.bmax = V3(float(hi[0]), float(hi[1]), float(hi[2]));
    s.sceneRadius = length(s.bmax - s.bmin) * 0.5f;
  }
}

// Closest hit with the instance of each hit (-1: miss or one level); mode as orc_intersect.
void orc_intersect_inst(void* p, int n, const float* rays, int mode, int* prim, float* t, float* u, float* v, int* inst)
{
  Scene& s = *(Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    const float* q = rays + 8 * i;
    RayHit r;
    r.o = V3(q[0], q[1], q[2]); r.d = V3(q[3], q[4], q[5]); r.tnear = q[6]; r.tfar = q[7];
    r.u = r.v = 0.f;
    const bool any = mode >= 2;
    if (mode == 0 || mode == 2) traverse(s, r, any);
    else bruteForce(s, r, any);
    prim[i] = r.prim;
    t[i] = r.prim >= 0 ? r.tfar : q[7];
    u[i] = r.u;
    v[i] = r.v;
    inst[i] = r.prim >= 0 ? r.inst : -1;
  }
}

// UV set 0 for the scene's vertices (null: none).
void orc_scene_set_uvs(void* p, const float* uv, int nverts)
{
  Scene& s = *(Scene*)p;
  if (uv) s.uv.assign(uv, uv + 2 * size_t(nverts));
  else s.uv.clear();
}

// Appends a texture: w x h texels as image rows top to bottom, format 0: RGBA8 sRGB (the 2.2
// gamma table of image/pixel.cpp:21, pow(i / 255, 2.2) in float), 1: RGBA8 linear (/ 255,
// image/pixel.h:137-145), 2: RGB FP32; address 0 repeat, 1 clamp; uvParams: scale u, v, offset
// u, v.  Returns its index.
int orc_scene_add_texture(void* p, int w, int h, int format, int addrU, int addrV, const float* uvParams, const void* texels)
{
  Scene& s = *(Scene*)p;
  Texture t;
  t.w = w;
  t.h = h;
  t.repeatU = addrU == 0;
  t.repeatV = addrV == 0;
  t.scale[0] = uvParams[0];
  t.scale[1] = uvParams[1];
  t.offset[0] = uvParams[2];
  t.offset[1] = uvParams[3];
  t.texels.resize(size_t(w) * h);
  float gamma[256];
  for (int i = 0; i < 256; ++i) gamma[i] = std::pow((float)i / 255.0f, 2.2f);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
    {
      const size_t src = size_t(h - 1 - y) * w + x;  // bottom row first (image_texture.h:352-353)
      V3 c;
      if (format == 2)
      {
        const float* f = (const float*)texels + 3 * src;
        c = V3(f[0], f[1], f[2]);
      }
      else
      {
        const unsigned char* b = (const unsigned char*)texels + 4 * src;
        c = format == 0 ? V3(gamma[b[0]], gamma[b[1]], gamma[b[2]])
                        : V3((float)b[0] * (1.f / 255.f), (float)b[1] * (1.f / 255.f), (float)b[2] * (1.f / 255.f));
      }
      t.texels[size_t(y) * w + x] = c;
    }
  s.textures.push_back(t);
  return int(s.textures.size()) - 1;
}

void orc_material_set_texture(void* p, int material, int texture)
{
  Scene& s = *(Scene*)p;
  if (s.baseTex.size() != s.mats.size()) s.baseTex.assign(s.mats.size(), -1);
  s.baseTex[size_t(material)] = texture;
}

// UV set 0 at n hits (prim, u, v) into out (n x 2).
void orc_hit_uv(void* p, int n, const int* prim, const float* uv, float* out)
{
  const Scene& s = *(const Scene*)p;
  for (int i = 0; i < n; ++i) hitUV(s, prim[i], uv[2 * i], uv[2 * i + 1], out[2 * i], out[2 * i + 1]);
}

// Bilinear lookup of texture `tex` at n (u, v) pairs into out (n x 3): the lookup alone, for
// the texture tests.
void orc_texture_lookup(void* p, int tex, int n, const float* uv, float* out)
{
  const Scene& s = *(const Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    const V3 c = textureGet(s.textures[size_t(tex)], uv[2 * i], uv[2 * i + 1]);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}

float orc_scene_radius(void* p) { return ((Scene*)p)->sceneRadius; }

// lights: records of orc::Light; env image in texture row order (row 0 = nadir), lutMax as
// the reference's lutSize prop (light/image_light.h:25-40, 2x bucketing while too large)
int orc_scene_set_lights(void* p, int nlights, const void* lights, int envW, int envH, const float* envRgb, int lutMax)
{
  Scene& s = *(Scene*)p;
  if (nlights < 0 || nlights > 8) return -1;
  buildAreaLight(s);
  s.lights.assign((const Light*)lights, (const Light*)lights + nlights);
  if (nlights == 0 && s.area.v0.empty())
  {
    // render/scene.cpp:432-438: no lights at all -> a default ambient light (Ke = 1)
    Light amb{};
    amb.kind = kAmbient; amb.r = amb.g = amb.b = 1.f;
    s.lights.push_back(amb);
    nlights = 1;
  }
  for (Light& l : s.lights)
    if (l.kind == kDistant)  // light/distant_light.h:34 normalizes the direction
    {
      V3 d = normalize(V3(l.dx, l.dy, l.dz));
      l.dx = d.x; l.dy = d.y; l.dz = d.z;
    }
  if (envRgb && envW > 0 && envH > 0)
  {
    s.envRgb.assign(envRgb, envRgb + size_t(envW) * envH * 3);
    s.env.W = envW; s.env.H = envH; s.env.rgb = s.envRgb.data();
    int lw = envW, lh = envH, bucket = 1;
    if (lutMax > 0)
      while (std::max(lw, lh) > lutMax && lw % 2 == 0 && lh % 2 == 0) { lw /= 2; lh /= 2; bucket *= 2; }
    std::vector<float> w(size_t(lw) * lh);
    for (int y = 0; y < lh; ++y)
      for (int x = 0; x < lw; ++x)
      {
        float sum = 0.f;
        for (int y2 = 0; y2 < bucket; ++y2)
          for (int x2 = 0; x2 < bucket; ++x2) sum += luminance(envTexel(s.env, x * bucket + x2, y * bucket + y2));
        w[size_t(y) * lw + x] = sum * std::sin((float(y) + 0.5f) / float(lh) * kPi);
      }
    s.envX.resize(size_t(lh) * (lw + 1));
    s.envY.resize(lh + 1);
    dist2dInit(s.envX.data(), s.envY.data(), w.data(), lw, lh);
    s.env.LW = lw; s.env.LH = lh; s.env.xDists = s.envX.data(); s.env.yDist = s.envY.data();
  }
  s.ls.count = nlights;
  s.ls.lights = s.lights.data();
  s.ls.area = s.area.v0.empty() ? nullptr : &s.area;
  s.ls.offset = s.ls.area ? 1 : 0;
  s.ls.env = s.env.rgb ? &s.env : nullptr;
  s.ls.sceneRadius = s.sceneRadius;
  float pw[9];
  if (s.ls.area) pw[0] = s.area.power;
  for (int i = 0; i < nlights; ++i) pw[s.ls.offset + i] = lightPower(s.lights[i], s.ls.env, s.sceneRadius);
  dist1dInit(s.ls.cdf, pw, s.ls.offset + nlights);
  return 0;
}

int orc_light_count(void* p)
{
  Scene& s = *(Scene*)p;
  return s.ls.offset + s.ls.count;
}

void orc_get_light_cdf(void* p, float* out)
{
  Scene& s = *(Scene*)p;
  for (int i = 0; i <= s.ls.offset + s.ls.count; ++i) out[i] = s.ls.cdf[i];
}

void orc_get_luts(void* p, float* out)  // ggxDir, ggxAvg, dielEtaDir, dielEtaAvg, dielRcpEtaDir, dielRcpEtaAvg
{
  std::memcpy(out, &((Scene*)p)->luts, sizeof(Luts));
}

void orc_compute_luts(float* out)
{
  Luts* L = new Luts();
  precomputeLuts(*L);
  std::memcpy(out, L, sizeof(Luts));
  delete L;
}

// rays: n x [ox oy oz dx dy dz near far]; mode 0 = BVH closest, 1 = brute-force closest,
// 2 = BVH any-hit, 3 = brute-force any-hit
void orc_intersect(void* p, int n, const float* rays, int mode, int* prim, float* t, float* u, float* v)
{
  Scene& s = *(Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    const float* q = rays + 8 * i;
    RayHit r;
    r.o = V3(q[0], q[1], q[2]); r.d = V3(q[3], q[4], q[5]); r.tnear = q[6]; r.tfar = q[7];
    r.u = r.v = 0.f;
    bool any = mode >= 2;
    if (mode == 0 || mode == 2) traverse(s, r, any);
    else bruteForce(s, r, any);
    prim[i] = r.prim;
    if (t) t[i] = r.prim >= 0 ? r.tfar : q[7];
    if (u) u[i] = r.u;
    if (v) v[i] = r.v;
  }
}

// One sample pass per sample index in [sppStart, sppStart+sppCount), pixels split over
// `threads` std::threads by 32x32 tiles (the reference's Tasking::run over tiles,
// sys/tasking.h:22-41).  rowBegin/rowEnd bound the rendered rows (the CPU baseline renders a
// bounded band of the frame).  Returns the number of rays traced.
long long orc_render(void* p, const void* camp, int W, int H, int rowBegin, int rowEnd, int sppStart, int sppCount,
                     uint32_t seed, int maxDepth, int rrDepth, float minWeight, float maxRadiance, int threads,
                     float** buffers /* color, weight, albedo, albedoW, albedo1, albedo1W, normal, normalW, normal1, normal1W, depth, depthW, sampleL, color2 */,
                     uint32_t* pathHash, int* firstPrim)
{
  Scene& s = *(Scene*)p;
  const Camera& cam = *(const Camera*)camp;
  RenderParams P{W, H, seed, maxDepth, rrDepth, minWeight, maxRadiance};
  Accum A;
  A.color = buffers[0]; A.weight = buffers[1];
  A.albedo = buffers[2]; A.albedoW = buffers[3];
  A.albedo1 = buffers[4]; A.albedo1W = buffers[5];
  A.normal = buffers[6]; A.normalW = buffers[7];
  A.normal1 = buffers[8]; A.normal1W = buffers[9];
  A.depth = buffers[10]; A.depthW = buffers[11];
  A.sampleL = buffers[12];
  A.color2 = buffers[13];
  A.pathHash = pathHash; A.firstPrim = firstPrim;
  if (threads < 1) threads = 1;
  rowBegin = std::max(rowBegin, 0);
  rowEnd = std::min(rowEnd, H);
  const int tilesX = (W + 31) / 32, tilesY = (rowEnd - rowBegin + 31) / 32;
  std::atomic<long long> rays{0};
  for (int sIdx = sppStart; sIdx < sppStart + sppCount; ++sIdx)
  {
    std::atomic<int> next{0};
    const bool last = sIdx == sppStart + sppCount - 1;
    auto worker = [&]() {
      Stats st;
      for (;;)
      {
        int tile = next.fetch_add(1);
        if (tile >= tilesX * tilesY) break;
        int tx = tile % tilesX, ty = tile / tilesX;
        for (int y = rowBegin + ty * 32; y < std::min(rowBegin + ty * 32 + 32, rowEnd); ++y)
          for (int x = tx * 32; x < std::min(tx * 32 + 32, W); ++x)
            tracePath(s, cam, P, x, y, uint32_t(sIdx), A, st, last);
      }
      rays += st.rays;
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; ++i) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
  }
  return rays.load();
}

// The pixel sample of later renders: jitter (enable, jx, jy) or the pixel filter of
// DeviceCpu::initFilter (render/device_cpu.cpp:235-282; type 0 box, 1 triangle, 2 Gaussian,
// 3 Blackman-Harris; width <= 0 and sigma <= 0 take the reference defaults), its lookup table
// restated from image/pixel_filter.h:30-44, 77-139 and distribution1d.cpp:20-35.
void orc_set_pixel_sample(void* p, int jitter, float jx, float jy, int type, float width, float sigma, int bias)
{
  Scene& s = *(Scene*)p;
  s.jitter = jitter != 0;
  s.jitterXY[0] = jx;
  s.jitterXY[1] = jy;
  s.filterCdf.clear();
  s.filterWidth = 0.f;
  if (type < 1 || type > 3) return;
  if (!(width > 0.f)) width = type == 1 ? 2.f : 3.f;
  if (type == 2 && !(sigma > 0.f)) sigma = width / 6.f;
  const float radius = width / 2.f;
  auto eval = [&](float x) -> float {
    switch (type)
    {
      case 1: return std::max(radius - std::fabs(x), 0.f);
      case 2:
      {
        const float a = -1.f / (2.f * sigma * sigma);
        const float b = bias ? std::exp(a * radius * radius) : 0.f;
        return std::max(std::exp(a * x * x) - b, 0.f);
      }
      default:
      {
        const float u = x / radius;
        if (u < -1.f || u > 1.f) return 0.f;
        const float t = (u + 1.f) * kPi;
        return 0.35875f - 0.48829f * std::cos(t) + 0.14128f * std::cos(2.f * t) - 0.01168f * std::cos(3.f * t);
      }
    }
  };
  const int n = 2048, size = 2 * n + 1;
  std::vector<float> lut(size);
  for (int i = 0; i < size; ++i) lut[i] = eval((float(i - n) / float(n)) * radius);
  s.filterCdf.resize(size + 1);
  dist1dInit(s.filterCdf.data(), lut.data(), size);
  s.filterWidth = width;
}

// The pixel-sample positions the renderer takes for sampler values s (n x 2) with the scene's
// jitter / pixel filter: out n x 2.
void orc_pixel_sample(void* p, int n, const float* sv, float* out)
{
  const Scene& s = *(const Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    float x = sv[2 * i], y = sv[2 * i + 1];
    pixelSample(s, x, y);
    out[2 * i] = x;
    out[2 * i + 1] = y;
  }
}

// FP64 gradient of the loss mean((I - target)^2) over pixels and channels with respect to
// (albedo r, g, b, alpha) of every material, for the image of sample passes 0 .. spp-1; the
// estimator of the HIP adjoint (detached sampling, suffix radiance times df/f at each vertex),
// evaluated along the oracle's own paths with per-vertex records.  Returns the loss.
double orc_adjoint(void* p, const void* camp, int W, int H, int spp, uint32_t seed, int maxDepth, int rrDepth,
                   float minWeight, const float* target, int threads, double* grad /* nMat x 4 */)
{
  Scene& s = *(Scene*)p;
  const Camera& cam = *(const Camera*)camp;
  const int npix = W * H;
  std::vector<float> color(3 * size_t(npix), 0.f), weight(npix, 0.f);
  float* bufs[14] = {color.data(), weight.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr};
  orc_render(p, camp, W, H, 0, H, 0, spp, seed, maxDepth, rrDepth, minWeight, 3.402823466e38f, threads, bufs, nullptr, nullptr);
  const double norm = 1.0 / (3.0 * npix);
  std::vector<double> delta(3 * size_t(npix));
  double loss = 0.0;
  for (int i = 0; i < npix; ++i)
    for (int k = 0; k < 3; ++k)
    {
      double I = double(color[size_t(k) * npix + i]) / double(weight[i]);
      double r = I - double(target[3 * size_t(i) + k]);
      loss += r * r;
      delta[size_t(k) * npix + i] = 2.0 * r * norm;
    }
  const int nMat = int(s.mats.size());
  RenderParams P{W, H, seed, maxDepth, rrDepth, minWeight, 3.402823466e38f};
  if (threads < 1) threads = 1;
  std::vector<std::vector<double>> part(threads, std::vector<double>(4 * size_t(nMat), 0.0));
  std::atomic<int> next{0};
  auto worker = [&](int tid) {
    Accum A{};
    Stats st;
    std::vector<double>& g = part[tid];
    for (;;)
    {
      int row = next.fetch_add(1);
      if (row >= H) break;
      for (int x = 0; x < W; ++x)
      {
        const int pix = row * W + x;
        const double dl[3] = {delta[pix] / spp, delta[size_t(npix) + pix] / spp, delta[2 * size_t(npix) + pix] / spp};
        for (int smp = 0; smp < spp; ++smp)
        {
          AdjPath ap;
          tracePath(s, cam, P, x, row, uint32_t(smp), A, st, false, &ap);
          const double Lt[3] = {ap.L.x, ap.L.y, ap.L.z};
          for (const AdjVertex& v : ap.v)
          {
            double* gm = &g[4 * size_t(v.mat)];
            const double dL[3] = {v.dirL.x, v.dirL.y, v.dirL.z}, Lr[3] = {v.Lrun.x, v.Lrun.y, v.Lrun.z};
            for (int c = 0; c < 3; ++c)
            {
              if (v.nee && v.gNee.v[c] > 0.0)
              {
                double w = dl[c] * dL[c] / v.gNee.v[c];
                gm[c] += w * v.gNee.dA[c];
                gm[3] += w * v.gNee.dAl[c];
              }
              if (v.ext && v.gExt.v[c] > 0.0)
              {
                double w = dl[c] * (Lt[c] - Lr[c]) / v.gExt.v[c];
                gm[c] += w * v.gExt.dA[c];
                gm[3] += w * v.gExt.dAl[c];
              }
            }
          }
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < threads; ++i) pool.emplace_back(worker, i);
  worker(0);
  for (auto& t : pool) t.join();
  for (size_t k = 0; k < 4 * size_t(nMat); ++k)
  {
    double sum = 0.0;
    for (int t = 0; t < threads; ++t) sum += part[t][k];
    grad[k] = sum;
  }
  return loss * norm;
}

// Value and parameter derivatives (FP64) of the BSDF on given frames and directions, with the
// parameter rows of orc_bsdf: n x [value rgb, d value / d albedo rgb, d value / d alpha rgb].
void orc_bsdf_grad(int n, const float* luts, const float* params, const float* geo, double* out)
{
  const Luts& L = *(const Luts*)luts;
  for (int i = 0; i < n; ++i)
  {
    const float* pr = params + 11 * i;
    const float* gg = geo + 18 * i;
    Ctx ctx;
    ctx.f.U = V3(gg[0], gg[1], gg[2]); ctx.f.V = V3(gg[3], gg[4], gg[5]); ctx.f.N = V3(gg[6], gg[7], gg[8]);
    ctx.Ng = V3(gg[9], gg[10], gg[11]);
    V3 wo(gg[12], gg[13], gg[14]), wi(gg[15], gg[16], gg[17]);
    Material m{};
    m.kind = int(pr[0]); m.r = pr[1]; m.g = pr[2]; m.b = pr[3]; m.er = pr[4]; m.eg = pr[5]; m.eb = pr[6]; m.roughness = pr[7];
    m.ior = 1.0f / pr[8]; m.thickness = pr[9]; m.weight = pr[10];
    Bsdf b = makeBsdf(L, m, ctx, wo, 0, -1);
    b.fr = ctx.f;
    if (b.kind == kDielectric) b.deta = pr[8];  // the supplied eta as is, as orc_bsdf does
    GradD g = bsdfEvalGradD(L, b, ctx, wo, wi);
    double* o = out + 9 * size_t(i);
    for (int k = 0; k < 3; ++k) { o[k] = g.v[k]; o[3 + k] = g.dA[k]; o[6 + k] = g.dAl[k]; }
  }
}

// ---- unit-level entry points used by the golden-vector tests --------------------------

void orc_sampler(int n, const uint32_t* seedPidIdx, int ndim, float* out1, float* out2, float* out3)
{
  for (int i = 0; i < n; ++i)
  {
    Sampler s;
    s.set(seedPidIdx[3 * i], seedPidIdx[3 * i + 2], seedPidIdx[3 * i + 1]);
    for (int d = 0; d < ndim; ++d)
    {
      out1[i * ndim + d] = s.get1D(d);
      s.get2D(d, out2[(i * ndim + d) * 2], out2[(i * ndim + d) * 2 + 1]);
      s.get3D(d, out3[(i * ndim + d) * 3], out3[(i * ndim + d) * 3 + 1], out3[(i * ndim + d) * 3 + 2]);
    }
  }
}

void orc_morton(int n, int tileSizeX, int* xy)
{
  for (int i = 0; i < n; ++i) morton8x8(i, tileSizeX, xy[2 * i], xy[2 * i + 1]);
}

// cfg: [w h fov nearClip lensRadius focal px py pz]; s: n x 4; out: n x 2 x 7 (pinhole, thin lens)
void orc_camera_rays(const float* cfg, int n, const float* s, float* out)
{
  V3 U(-1.f, 0.f, 0.f), Vb(0.f, 1.f, 0.f), Nb(0.f, 0.f, -1.f);
  Basis B{U, Vb, Nb};
  float W = cfg[0], H = cfg[1], fov = cfg[2], nearClip = cfg[3], lensRadius = cfg[4], focal = cfg[5];
  V3 pos(cfg[6], cfg[7], cfg[8]);
  float hh = std::tan(0.5f * fov), hw = hh * (W / H);
  V3 imageO = toGlobal(B, V3(-hw, hh, -1.f)), imageDx = toGlobal(B, V3(2.f * hw, 0.f, 0.f)), imageDy = toGlobal(B, V3(0.f, -2.f * hh, 0.f));
  for (int i = 0; i < n; ++i)
  {
    const float* q = s + 4 * i;
    float* o = out + 14 * i;
    V3 d = imageO + q[0] * imageDx + q[1] * imageDy;
    float len = length(d);
    V3 dn = d * (1.0f / len);
    o[0] = pos.x; o[1] = pos.y; o[2] = pos.z; o[3] = dn.x; o[4] = dn.y; o[5] = dn.z; o[6] = nearClip * len;
    float lx, ly;
    uniformSampleDisk(q[2], q[3], lx, ly);
    lx *= lensRadius; ly *= lensRadius;
    V3 off = toGlobal(B, V3(lx, ly, 0.f));
    V3 d2 = focal * (imageO + q[0] * imageDx + q[1] * imageDy) - off;
    float len2 = length(d2);
    V3 o2 = pos + off;
    V3 dn2 = d2 * (1.0f / len2);
    o[7] = o2.x; o[8] = o2.y; o[9] = o2.z; o[10] = dn2.x; o[11] = dn2.y; o[12] = dn2.z; o[13] = (nearClip / focal) * len2;
  }
}

void orc_warps(int n, const float* s, float* cosOut, float* sphOut, float* diskOut, float* coneOut /* n x 2 x 4 */)
{
  const float angles[2] = {0.1f, 0.00445f};
  for (int i = 0; i < n; ++i)
  {
    float pdf;
    V3 c = cosineSampleHemisphere(pdf, s[2 * i], s[2 * i + 1]);
    cosOut[4 * i] = c.x; cosOut[4 * i + 1] = c.y; cosOut[4 * i + 2] = c.z; cosOut[4 * i + 3] = pdf;
    V3 sp = uniformSampleSphere(pdf, s[2 * i], s[2 * i + 1]);
    sphOut[4 * i] = sp.x; sphOut[4 * i + 1] = sp.y; sphOut[4 * i + 2] = sp.z; sphOut[4 * i + 3] = pdf;
    uniformSampleDisk(s[2 * i], s[2 * i + 1], diskOut[2 * i], diskOut[2 * i + 1]);
    for (int a = 0; a < 2; ++a)
    {
      V3 cn = uniformSampleCone(pdf, s[2 * i], s[2 * i + 1], angles[a]);
      float* o = coneOut + (i * 2 + a) * 4;
      o[0] = cn.x; o[1] = cn.y; o[2] = cn.z; o[3] = pdf;
    }
  }
}

void orc_dist1d(const float* w, int size, int n, const float* s, int* idx, float* pdfSr, float* pdfI /* size+2 */)
{
  std::vector<float> cdf(size + 1);
  dist1dInit(cdf.data(), w, size);
  for (int i = 0; i < n; ++i)
  {
    float pdf, sr = s[i];
    idx[i] = dist1dSampleReuse(cdf.data(), size, pdf, sr);
    pdfSr[2 * i] = pdf; pdfSr[2 * i + 1] = sr;
  }
  for (int i = -1; i <= size; ++i) pdfI[i + 1] = dist1dPdf(cdf.data(), size, i);
}

void orc_dist2d(const float* w, int W, int H, int n, const float* s, int* idx, float* pdfSr, float* pdfCells)
{
  std::vector<float> xd(size_t(H) * (W + 1)), yd(H + 1);
  dist2dInit(xd.data(), yd.data(), w, W, H);
  for (int i = 0; i < n; ++i)
  {
    float pdf, sx = s[2 * i], sy = s[2 * i + 1];
    int ix, iy;
    dist2dSampleReuse(xd.data(), yd.data(), W, H, pdf, sx, sy, ix, iy);
    idx[2 * i] = ix; idx[2 * i + 1] = iy;
    pdfSr[3 * i] = pdf; pdfSr[3 * i + 1] = sx; pdfSr[3 * i + 2] = sy;
  }
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) pdfCells[y * W + x] = dist2dPdf(xd.data(), yd.data(), W, H, x, y);
}

// params: n x [kind r g b er eg eb roughness eta thickness weight]; geo: n x [U V N Ng wo wi] (18);
// s: n x 3; out: sample n x [weight wi pdf], type n, eval n x [value pdf albedo].  Coated rows: r g b = diffuse color, er eg eb = coat color, eta = ambient/coat IOR.
// outSample: n x [weight3 wi3 pdf]; outType: n; outEval: n x [value3 pdf]
// Uses precomputed LUTs `luts` (layout of orc::Luts).
void orc_bsdf(int n, const float* luts, const float* params, const float* geo, const float* s, float* outSample, int* outType, float* outEval)
{
  const Luts& L = *(const Luts*)luts;
  for (int i = 0; i < n; ++i)
  {
    const float* pr = params + 11 * i;
    const float* g = geo + 18 * i;
    Ctx ctx;
    ctx.f.U = V3(g[0], g[1], g[2]); ctx.f.V = V3(g[3], g[4], g[5]); ctx.f.N = V3(g[6], g[7], g[8]);
    ctx.Ng = V3(g[9], g[10], g[11]);
    V3 wo(g[12], g[13], g[14]), wi(g[15], g[16], g[17]);
    Bsdf b;
    b.kind = int(pr[0]);
    b.fr = ctx.f;  // golden vectors construct the BSDF on the given frame directly
    b.roughness = pr[7];
    b.ggx = GGX(roughnessToAlpha(pr[7]));
    Material m{};
    m.kind = b.kind; m.r = pr[1]; m.g = pr[2]; m.b = pr[3]; m.er = pr[4]; m.eg = pr[5]; m.eb = pr[6]; m.roughness = pr[7];
    if (b.kind == kDielectric || b.kind == kCoated) m.ior = 1.0f / pr[8];  // eta = 1/ior seen from outside
    m.thickness = pr[9]; m.weight = pr[10];
    // build parameters exactly as makeBsdf does, then keep the supplied frame
    Ctx c2 = ctx;
    Bsdf built = makeBsdf(L, m, c2, wo, 0, b.kind == kDielectric ? (pr[8] > 1.f ? 0 : 1) : -1);
    built.fr = ctx.f;
    if (b.kind == kDielectric)
    {
      // supplied eta is used as is (the reference constructs RoughDielectricBsdf(eta) directly)
      float eta = pr[8];
      built.deta = eta;
      built.EavgEta = dielAlbedoAvg(L, eta, built.roughness);
      built.EavgRcpEta = dielAlbedoAvg(L, 1.0f / eta, built.roughness);
      float FavgEta = fresnelDielectricAvg(eta), FavgRcpEta = fresnelDielectricAvg(1.0f / eta);
      float a = (1.f - FavgEta) / (1.f - built.EavgRcpEta), bb = (1.f - FavgRcpEta) / (1.f - built.EavgEta);
      float x = bb / (a + bb);
      built.fmsRatio = 1.f - x * (1.f - FavgEta);
    }
    V3 swi(0.f);
    float spdf = 0.f;
    int stype = 0;
    V3 w = bsdfSample(L, built, ctx, wo, swi, spdf, stype, s[3 * i], s[3 * i + 1], s[3 * i + 2]);
    float* o = outSample + 7 * i;
    o[0] = w.x; o[1] = w.y; o[2] = w.z; o[3] = swi.x; o[4] = swi.y; o[5] = swi.z; o[6] = spdf;
    outType[i] = stype;
    float epdf;
    V3 ev = bsdfEval(L, built, ctx, wo, wi, epdf);
    V3 alb = bsdfAlbedo(built);
    float* e = outEval + 7 * i;
    e[0] = ev.x; e[1] = ev.y; e[2] = ev.z; e[3] = epdf; e[4] = alb.x; e[5] = alb.y; e[6] = alb.z;
  }
}

// The whole BSDF of a material record as the renderer builds it (makeBsdf from outside, the
// opacity mix included) on a supplied frame: mats n x Material, geo n x 18, s n x 3.
// outSample n x 7 [weight wi pdf], outType n, outEval n x 4 [value pdf], outAux n x 9
// [albedo, normal, transparency].
void orc_bsdf_mat(int n, const float* luts, const void* mats, const float* geo, const float* s, float* outSample, int* outType,
                  float* outEval, float* outAux, const int* inside)
{
  const Luts& L = *(const Luts*)luts;
  const Material* ms = (const Material*)mats;
  for (int i = 0; i < n; ++i)
  {
    const float* g = geo + 18 * i;
    Ctx ctx;
    ctx.f.U = V3(g[0], g[1], g[2]); ctx.f.V = V3(g[3], g[4], g[5]); ctx.f.N = V3(g[6], g[7], g[8]);
    ctx.Ng = V3(g[9], g[10], g[11]);
    V3 wo(g[12], g[13], g[14]), wi(g[15], g[16], g[17]);
    // inside[i]: the path is in the material's interior medium (a media interface is backfacing)
    Bsdf b = makeBsdf(L, ms[i], ctx, wo, (inside && inside[i]) ? 1 : 0, 1);
    b.fr = ctx.f;
    if (b.mixLobes)
    {
      auto lobes = std::make_shared<StdLobes>(*b.mixLobes);
      for (int k = 0; k < lobes->n; ++k) lobes->child[k].fr = ctx.f;
      b.mixLobes = lobes;
    }
    V3 swi(0.f);
    float spdf = 0.f;
    int stype = 0;
    V3 w = bsdfSample(L, b, ctx, wo, swi, spdf, stype, s[3 * i], s[3 * i + 1], s[3 * i + 2]);
    float* o = outSample + 7 * i;
    o[0] = w.x; o[1] = w.y; o[2] = w.z; o[3] = swi.x; o[4] = swi.y; o[5] = swi.z; o[6] = spdf;
    outType[i] = stype;
    float epdf;
    V3 ev = bsdfEval(L, b, ctx, wo, wi, epdf);
    float* e = outEval + 4 * i;
    e[0] = ev.x; e[1] = ev.y; e[2] = ev.z; e[3] = epdf;
    V3 alb = bsdfAlbedo(b), nrm = bsdfNormal(b);
    V3 tr = isTransparent(ms[i]) ? bsdfTransparency(L, b, wo) : V3(0.f);
    float* x = outAux + 9 * i;
    x[0] = alb.x; x[1] = alb.y; x[2] = alb.z; x[3] = nrm.x; x[4] = nrm.y; x[5] = nrm.z; x[6] = tr.x; x[7] = tr.y; x[8] = tr.z;
  }
}

// rays: n x [o d t u v]; prim: n; out ctx: n x 16 [p eps Ng U V N] before the orientation flip,
// frame: n x 9 after flip + makeFrame(wo = -d)
void orc_post_intersect(int nverts, const float* pos, const float* nrm, const int* idx, int n, const float* rays,
                        const int* prim, float* outCtx, float* outFrame)
{
  Scene s;
  s.pos.resize(nverts);
  for (int i = 0; i < nverts; ++i) s.pos[i] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  if (nrm)
  {
    s.nrm.resize(nverts);
    for (int i = 0; i < nverts; ++i) s.nrm[i] = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  }
  int maxIdx = 0;
  for (int i = 0; i < n; ++i) maxIdx = std::max(maxIdx, prim[i]);
  s.idx.assign(idx, idx + 3 * size_t(maxIdx + 1));
  for (int i = 0; i < n; ++i)
  {
    const float* q = rays + 9 * i;
    V3 o(q[0], q[1], q[2]), d(q[3], q[4], q[5]);
    Ctx c = postIntersect(s, o, d, q[6], prim[i], q[7], q[8]);
    // undo the flip for the pre-flip output
    V3 Ng = c.backfacing ? -c.Ng : c.Ng, N = c.backfacing ? -c.f.N : c.f.N;
    float* oc = outCtx + 16 * i;
    float vals[16] = {c.p.x, c.p.y, c.p.z, c.eps, Ng.x, Ng.y, Ng.z, c.f.U.x, c.f.U.y, c.f.U.z, c.f.V.x, c.f.V.y, c.f.V.z, N.x, N.y, N.z};
    std::memcpy(oc, vals, sizeof(vals));
    Basis fr = makeShadingFrame(c, -d);
    float* of = outFrame + 9 * i;
    float fv[9] = {fr.U.x, fr.U.y, fr.U.z, fr.V.x, fr.V.y, fr.V.z, fr.N.x, fr.N.y, fr.N.z};
    std::memcpy(of, fv, sizeof(fv));
  }
}

void orc_env_eval(void* p, int n, const float* dirs, float* out /* n x 4: value3, pdf */)
{
  Scene& s = *(Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    float pdf;
    V3 v = evalEnv(s.ls, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), pdf);
    out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = pdf;
  }
}

// samples the light list from the point org (the area light's samples depend on it)
void orc_light_sample(void* p, int n, const float* s3, const float* org, float* out /* n x 8: wi3 weight3 pdf dist */)
{
  Scene& s = *(Scene*)p;
  for (int i = 0; i < n; ++i)
  {
    V3 wi(0.f);
    float pdf, dist;
    V3 w = sampleLight(s.ls, V3(org[0], org[1], org[2]), wi, pdf, dist, s3[3 * i], s3[3 * i + 1], s3[3 * i + 2]);
    float* o = out + 8 * i;
    o[0] = wi.x; o[1] = wi.y; o[2] = wi.z; o[3] = w.x; o[4] = w.y; o[5] = w.z; o[6] = pdf; o[7] = dist;
  }
}

}  // extern "C"
