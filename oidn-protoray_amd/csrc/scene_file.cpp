// Scene files: the loaders behind the reference's Scene(path, props) (render/scene.cpp:64-372).
//   .rkscene  JSON description + one .rkbin buffer (rkscene/rkscene.h:847-903)
//   .gltf     glTF 2.0 (JSON + external or data-URI buffers; rkscene/rkscene_gltf.h)
//   .mesh     TriangleMesh blob (geometry/triangle_mesh.cpp:362-380, sys/blob.h:13-37) with
//             the materials of a .mtl file of the same base name (render/scene.cpp:136-172)
// A loaded file becomes one triangle mesh in world space: instanced objects are flattened
// (each instance's copy transformed, as the reference does with the `flatten` prop or for
// emissive objects, render/scene.cpp:343-363), with per-triangle material IDs in the
// reference's numbering (ID 0 = environment, emissive materials first, the default material
// last) and material records in this library's layout.  Host code only: no GPU is touched.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/prt_hip.h"

namespace prt {
int setLastError(int code, const std::string& msg);
}

namespace {

struct LoadError
{
  std::string msg;
};
[[noreturn]] void fail(const std::string& m) { throw LoadError{m}; }

// ---------------------------------------------------------------------------------------
// A small JSON reader (objects, arrays, strings with escapes, numbers, true/false/null).
// ---------------------------------------------------------------------------------------
struct Json
{
  enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<Json> a;
  std::vector<std::pair<std::string, Json>> o;

  const Json* get(const std::string& k) const
  {
    if (type != Obj) return nullptr;
    for (const auto& kv : o)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  double num() const
  {
    if (type == Num) return n;
    if (type == Bool) return b ? 1.0 : 0.0;
    fail("JSON value is not a number");
  }
  const std::string& str() const
  {
    if (type != Str) fail("JSON value is not a string");
    return s;
  }
  const std::vector<Json>& arr() const
  {
    if (type != Arr) fail("JSON value is not an array");
    return a;
  }
};

class JsonReader
{
public:
  explicit JsonReader(const std::string& text) : t(text) {}
  Json parse()
  {
    Json v = value();
    ws();
    if (i != t.size()) fail("trailing characters after JSON document");
    return v;
  }

private:
  const std::string& t;
  size_t i = 0;

  void ws()
  {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\t' || t[i] == '\n' || t[i] == '\r')) ++i;
  }
  char peek()
  {
    ws();
    if (i >= t.size()) fail("unexpected end of JSON");
    return t[i];
  }
  void expect(char c)
  {
    if (peek() != c) fail(std::string("JSON: expected '") + c + "'");
    ++i;
  }
  Json value()
  {
    const char c = peek();
    Json v;
    if (c == '{')
    {
      v.type = Json::Obj;
      ++i;
      if (peek() == '}') { ++i; return v; }
      for (;;)
      {
        Json k = value();
        if (k.type != Json::Str) fail("JSON: object key is not a string");
        expect(':');
        v.o.emplace_back(k.s, value());
        const char d = peek();
        ++i;
        if (d == '}') break;
        if (d != ',') fail("JSON: expected ',' or '}'");
      }
    }
    else if (c == '[')
    {
      v.type = Json::Arr;
      ++i;
      if (peek() == ']') { ++i; return v; }
      for (;;)
      {
        v.a.push_back(value());
        const char d = peek();
        ++i;
        if (d == ']') break;
        if (d != ',') fail("JSON: expected ',' or ']'");
      }
    }
    else if (c == '"')
    {
      v.type = Json::Str;
      ++i;
      while (i < t.size() && t[i] != '"')
      {
        char ch = t[i++];
        if (ch == '\\')
        {
          if (i >= t.size()) fail("JSON: bad escape");
          const char e = t[i++];
          switch (e)
          {
          case 'n': v.s += '\n'; break;
          case 't': v.s += '\t'; break;
          case 'r': v.s += '\r'; break;
          case 'b': v.s += '\b'; break;
          case 'f': v.s += '\f'; break;
          case 'u':
          {
            if (i + 4 > t.size()) fail("JSON: bad \\u escape");
            const unsigned cp = unsigned(std::stoul(t.substr(i, 4), nullptr, 16));
            i += 4;
            // UTF-8 encoding of a BMP code point (surrogate pairs are kept as two code points)
            if (cp < 0x80) v.s += char(cp);
            else if (cp < 0x800) { v.s += char(0xC0 | (cp >> 6)); v.s += char(0x80 | (cp & 0x3F)); }
            else { v.s += char(0xE0 | (cp >> 12)); v.s += char(0x80 | ((cp >> 6) & 0x3F)); v.s += char(0x80 | (cp & 0x3F)); }
            break;
          }
          default: v.s += e;
          }
        }
        else
          v.s += ch;
      }
      if (i >= t.size()) fail("JSON: unterminated string");
      ++i;
    }
    else if (c == 't' || c == 'f' || c == 'n')
    {
      const char* words[3] = {"true", "false", "null"};
      for (int w = 0; w < 3; ++w)
        if (t.compare(i, std::strlen(words[w]), words[w]) == 0)
        {
          i += std::strlen(words[w]);
          if (w < 2) { v.type = Json::Bool; v.b = w == 0; }
          return v;
        }
      fail("JSON: bad literal");
    }
    else
    {
      const char* b = t.c_str() + i;
      char* e = nullptr;
      v.type = Json::Num;
      v.n = std::strtod(b, &e);
      if (e == b) fail("JSON: bad number");
      i += size_t(e - b);
    }
    return v;
  }
};

std::string readText(const std::string& path)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) fail("failed to open file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string dirOf(const std::string& path)
{
  const size_t p = path.find_last_of('/');
  return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

std::string extOf(const std::string& path)
{
  const size_t s = path.find_last_of('/'), p = path.find_last_of('.');
  if (p == std::string::npos || (s != std::string::npos && p < s)) return std::string();
  std::string e = path.substr(p + 1);
  for (char& c : e) c = char(std::tolower(static_cast<unsigned char>(c)));
  return e;
}

// ---------------------------------------------------------------------------------------
// Scene description shared by the loaders (the reference's rkscene::Scene)
// ---------------------------------------------------------------------------------------
struct F3
{
  float x, y, z;
};

// Affine transform: columns v0, v1, v2 (linear part) and v3 (translation), p' = v0 x + v1 y + v2 z + v3.
struct Xfm
{
  F3 v[4] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}, {0.f, 0.f, 0.f}};
};

Xfm mul(const Xfm& a, const Xfm& b)
{
  Xfm c;
  for (int i = 0; i < 4; ++i)
  {
    const F3& q = b.v[i];
    c.v[i] = {a.v[0].x * q.x + a.v[1].x * q.y + a.v[2].x * q.z, a.v[0].y * q.x + a.v[1].y * q.y + a.v[2].y * q.z,
              a.v[0].z * q.x + a.v[1].z * q.y + a.v[2].z * q.z};
    if (i == 3)
    {
      c.v[3].x += a.v[3].x;
      c.v[3].y += a.v[3].y;
      c.v[3].z += a.v[3].z;
    }
  }
  return c;
}

// Standard material parameters (material/standard.cpp:27-81) as uniform values; a texture
// reference keeps its factor and is counted as approximated.
struct StdMat
{
  std::string name;
  float baseWeight = 1.f, baseColor[3] = {0.8f, 0.8f, 0.8f}, metalness = 0.f, diffuseRoughness = 0.f;
  float specularWeight = 1.f, specularColor[3] = {1.f, 1.f, 1.f}, specularRoughness = 0.3f, specularIor = 1.5f;
  float transmissionWeight = 0.f, transmissionColor[3] = {1.f, 1.f, 1.f}, transmissionDepth = 1.f;
  float subsurfaceWeight = 0.f;
  float coatWeight = 0.f, coatColor[3] = {1.f, 1.f, 1.f}, coatRoughness = 0.f, coatIor = 1.6f;
  float sheenWeight = 0.f, sheenColor[3] = {1.f, 1.f, 1.f}, sheenRoughness = 0.3f;
  float emissionLuminance = 0.f, emissionColor[3] = {1.f, 1.f, 1.f};
  bool thinWalled = false;
  float opacity = 1.f;
  bool textured = false;
  bool isDefault = false;  // the factory default (Principled with default props)
  bool luminous = false;   // the Luminous type of .mtl files
  float luminousColor[3] = {0.f, 0.f, 0.f};
};

struct Geom
{
  std::vector<F3> pos, nrm;
  std::vector<uint32_t> idx;
  int material = -1;
};

struct Obj
{
  std::vector<int> geoms;
  std::vector<Xfm> instances;  // empty: the identity, not instanced
};

struct Desc
{
  std::vector<StdMat> mats;
  std::vector<Geom> geoms;
  std::vector<Obj> objs;
  bool hasSun = false;
  F3 sunDir{0.f, 0.f, 0.f};
};

void readF3(const Json& j, float* out)
{
  const auto& a = j.arr();
  if (a.size() < 3) fail("expected a 3-vector");
  for (int k = 0; k < 3; ++k) out[k] = float(a[k].num());
}

// rkscene material parameter: a number / array, an image name (string) or {factor, image, ...}
template <int N>
void matParam(const Json& jm, const char* key, float* out, bool& textured)
{
  const Json* j = jm.get(key);
  if (!j) return;
  auto set = [&](const Json& v) {
    if (N == 1) out[0] = float(v.num());
    else readF3(v, out);
  };
  if (j->type == Json::Str)
  {
    for (int k = 0; k < N; ++k) out[k] = 1.f;
    textured = true;
  }
  else if (j->type == Json::Num || j->type == Json::Arr)
    set(*j);
  else if (j->type == Json::Obj)
  {
    if (j->get("image"))
    {
      for (int k = 0; k < N; ++k) out[k] = 1.f;
      textured = true;
    }
    if (const Json* f = j->get("factor")) set(*f);
    if (const Json* ty = j->get("type"))
      if (ty->str() == "bump") textured = true;
  }
}

StdMat rkMaterial(const Json& jm)
{
  StdMat m;
  bool& tx = m.textured;
  if (const Json* n = jm.get("name")) m.name = n->str();
  matParam<1>(jm, "baseWeight", &m.baseWeight, tx);
  matParam<3>(jm, "baseColor", m.baseColor, tx);
  matParam<1>(jm, "metalness", &m.metalness, tx);
  matParam<1>(jm, "diffuseRoughness", &m.diffuseRoughness, tx);
  bool normalMap = false;
  float dummy[3];
  matParam<1>(jm, "normal", dummy, normalMap);
  matParam<1>(jm, "specularWeight", &m.specularWeight, tx);
  matParam<3>(jm, "specularColor", m.specularColor, tx);
  matParam<1>(jm, "specularRoughness", &m.specularRoughness, tx);
  if (const Json* v = jm.get("specularIOR")) m.specularIor = float(v->num());
  matParam<1>(jm, "transmissionWeight", &m.transmissionWeight, tx);
  matParam<3>(jm, "transmissionColor", m.transmissionColor, tx);
  if (const Json* v = jm.get("transmissionDepth")) m.transmissionDepth = float(v->num());
  matParam<1>(jm, "subsurfaceWeight", &m.subsurfaceWeight, tx);
  matParam<1>(jm, "coatWeight", &m.coatWeight, tx);
  matParam<3>(jm, "coatColor", m.coatColor, tx);
  matParam<1>(jm, "coatRoughness", &m.coatRoughness, tx);
  if (const Json* v = jm.get("coatIOR")) m.coatIor = float(v->num());
  matParam<1>(jm, "sheenWeight", &m.sheenWeight, tx);
  matParam<3>(jm, "sheenColor", m.sheenColor, tx);
  matParam<1>(jm, "sheenRoughness", &m.sheenRoughness, tx);
  matParam<1>(jm, "emissionLuminance", &m.emissionLuminance, tx);
  matParam<3>(jm, "emissionColor", m.emissionColor, tx);
  if (const Json* v = jm.get("thinWalled")) m.thinWalled = v->num() != 0.0;
  matParam<1>(jm, "opacity", &m.opacity, tx);
  m.textured = tx || normalMap;
  return m;
}

// ---------------------------------------------------------------------------------------
// .rkscene (rkscene/rkscene.h:847-903): buffers[0] = the .rkbin file; typed views
// {type, count, byteOffset} into it.
// ---------------------------------------------------------------------------------------
struct Bin
{
  std::vector<char> data;
  template <class T>
  void read(const Json& jb, const char* type, std::vector<T>& out) const
  {
    if (!jb.get("type") || jb.get("type")->str() != type) fail(std::string("unexpected buffer type (expected ") + type + ")");
    const size_t count = size_t(jb.get("count")->num()), off = size_t(jb.get("byteOffset")->num());
    if (off + count * sizeof(T) > data.size()) fail("buffer access out of bounds");
    out.resize(count);
    if (count) std::memcpy(out.data(), data.data() + off, count * sizeof(T));
  }
};

Desc loadRkscene(const std::string& path)
{
  const Json js = JsonReader(readText(path)).parse();
  const Json* jbufs = js.get("buffers");
  if (!jbufs || jbufs->arr().size() != 1) fail("an .rkscene needs exactly one buffer");
  std::string uri = jbufs->a[0].get("uri")->str();
  const size_t byteLength = size_t(jbufs->a[0].get("byteLength")->num());
  if (uri.empty() || uri[0] != '/') uri = dirOf(path) + uri;
  Bin bin;
  {
    std::ifstream f(uri, std::ios::binary);
    if (!f) fail("failed to open file: " + uri);
    bin.data.resize(byteLength);
    f.read(bin.data.data(), std::streamsize(byteLength));
    if (size_t(f.gcount()) != byteLength) fail("the .rkbin file is shorter than its byteLength");
  }
  Desc d;
  if (const Json* jm = js.get("materials"))
    for (const Json& m : jm->arr()) d.mats.push_back(rkMaterial(m));
  if (const Json* jg = js.get("geometries"))
    for (const Json& g : jg->arr())
    {
      if (!g.get("type") || g.get("type")->str() != "triangle") fail("unsupported geometry type");
      Geom geo;
      if (const Json* p = g.get("positions")) bin.read(*p, "float3", geo.pos);
      if (const Json* n = g.get("normals")) bin.read(*n, "float3", geo.nrm);
      if (const Json* ix = g.get("indices")) bin.read(*ix, "uint", geo.idx);
      if (const Json* m = g.get("material")) geo.material = int(m->num());
      d.geoms.push_back(std::move(geo));
    }
  if (const Json* jo = js.get("objects"))
    for (const Json& o : jo->arr())
    {
      Obj ob;
      if (const Json* gs = o.get("geometries"))
        for (const Json& g : gs->arr()) ob.geoms.push_back(int(g.num()));
      if (const Json* in = o.get("instances"))
      {
        std::vector<float> m12;
        if (!in->get("type") || in->get("type")->str() != "float3x4") fail("unexpected buffer type (expected float3x4)");
        const size_t count = size_t(in->get("count")->num()), off = size_t(in->get("byteOffset")->num());
        if (off + count * 48 > bin.data.size()) fail("buffer access out of bounds");
        m12.resize(count * 12);
        if (count) std::memcpy(m12.data(), bin.data.data() + off, count * 48);
        for (size_t k = 0; k < count; ++k)
        {
          Xfm x;
          for (int c = 0; c < 4; ++c) x.v[c] = {m12[k * 12 + 3 * c], m12[k * 12 + 3 * c + 1], m12[k * 12 + 3 * c + 2]};
          ob.instances.push_back(x);
        }
      }
      d.objs.push_back(std::move(ob));
    }
  if (const Json* jl = js.get("lights"))
    for (const Json& l : jl->arr())
    {
      if (!l.get("type") || l.get("type")->str() != "sunSky") fail("unsupported light type");
      d.hasSun = true;
      if (const Json* dir = l.get("direction"))
      {
        float v[3];
        readF3(*dir, v);
        d.sunDir = {v[0], v[1], v[2]};
      }
    }
  return d;
}

// ---------------------------------------------------------------------------------------
// glTF 2.0 (rkscene/rkscene_gltf.h): buffers (file or base64 data URI), views, accessors,
// meshes of triangle primitives, the default scene's node tree (TRS or matrix), PBR
// materials with the KHR extensions the reference reads.
// ---------------------------------------------------------------------------------------
std::vector<char> base64(const std::string& s)
{
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
  };
  std::vector<char> out;
  uint32_t acc = 0;
  int bits = 0;
  for (char c : s)
  {
    const int v = val(c);
    if (v < 0) continue;  // padding and whitespace
    acc = (acc << 6) | uint32_t(v);
    bits += 6;
    if (bits >= 8)
    {
      bits -= 8;
      out.push_back(char((acc >> bits) & 0xFF));
    }
  }
  return out;
}

struct Gltf
{
  Json doc;
  std::vector<std::vector<char>> buffers;

  const Json& at(const char* key, int i) const
  {
    const Json* a = doc.get(key);
    if (!a || i < 0 || size_t(i) >= a->arr().size()) fail(std::string("invalid glTF ") + key + " index");
    return a->a[size_t(i)];
  }

  // accessor -> floats (VEC2 / VEC3, FLOAT) or indices (SCALAR, integer types)
  template <class Out>
  void accessor(int ai, int wantComps, bool indices, std::vector<Out>& out) const
  {
    const Json& acc = at("accessors", ai);
    const std::string type = acc.get("type")->str();
    const int comps = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4 : 0;
    if (comps != wantComps) fail("glTF type mismatch");
    const int ct = int(acc.get("componentType")->num());
    const size_t count = size_t(acc.get("count")->num());
    const size_t accOff = acc.get("byteOffset") ? size_t(acc.get("byteOffset")->num()) : 0;
    if (!acc.get("bufferView")) fail("glTF accessors without a buffer view are not supported");
    const Json& view = at("bufferViews", int(acc.get("bufferView")->num()));
    const int bi = int(view.get("buffer")->num());
    if (bi < 0 || size_t(bi) >= buffers.size()) fail("invalid glTF buffer index");
    const std::vector<char>& buf = buffers[size_t(bi)];
    const size_t viewOff = view.get("byteOffset") ? size_t(view.get("byteOffset")->num()) : 0;
    int compSize = 0;
    switch (ct)
    {
    case 5120: case 5121: compSize = 1; break;
    case 5122: case 5123: compSize = 2; break;
    case 5125: case 5126: compSize = 4; break;
    default: fail("glTF component type mismatch");
    }
    if (indices == (ct == 5126)) fail("glTF component type mismatch");
    const size_t elem = size_t(compSize) * comps;
    size_t stride = view.get("byteStride") ? size_t(view.get("byteStride")->num()) : 0;
    if (stride == 0) stride = elem;
    const size_t base = viewOff + accOff;
    if (count > 0 && base + (count - 1) * stride + elem > buf.size()) fail("glTF buffer access out of bounds");
    out.resize(count * comps);
    for (size_t k = 0; k < count; ++k)
      for (int c = 0; c < comps; ++c)
      {
        const char* p = buf.data() + base + k * stride + size_t(c) * compSize;
        double v = 0.0;
        switch (ct)
        {
        case 5120: v = double(*reinterpret_cast<const int8_t*>(p)); break;
        case 5121: v = double(*reinterpret_cast<const uint8_t*>(p)); break;
        case 5122: { int16_t t; std::memcpy(&t, p, 2); v = t; break; }
        case 5123: { uint16_t t; std::memcpy(&t, p, 2); v = t; break; }
        case 5125: { uint32_t t; std::memcpy(&t, p, 4); v = t; break; }
        case 5126: { float t; std::memcpy(&t, p, 4); out[k * comps + c] = Out(t); continue; }
        }
        out[k * comps + c] = Out(v);
      }
  }
};

float jnum(const Json& j, const char* k, float def)
{
  const Json* v = j.get(k);
  return v ? float(v->num()) : def;
}

Xfm nodeTransform(const Json& n)
{
  Xfm x;
  if (const Json* m = n.get("matrix"))
  {
    const auto& a = m->arr();
    if (a.size() != 16) fail("glTF node matrix needs 16 values");
    // column-major 4x4
    for (int c = 0; c < 4; ++c) x.v[c] = {float(a[c * 4].num()), float(a[c * 4 + 1].num()), float(a[c * 4 + 2].num())};
    return x;
  }
  float t[3] = {0.f, 0.f, 0.f}, q[4] = {0.f, 0.f, 0.f, 1.f}, s[3] = {1.f, 1.f, 1.f};
  if (const Json* v = n.get("translation")) readF3(*v, t);
  if (const Json* v = n.get("scale")) readF3(*v, s);
  if (const Json* v = n.get("rotation"))
  {
    const auto& a = v->arr();
    if (a.size() != 4) fail("glTF rotation needs 4 values");
    for (int k = 0; k < 4; ++k) q[k] = float(a[k].num());
  }
  // rotation matrix of the unit quaternion (x, y, z, w), columns scaled by s
  const float xx = q[0] * q[0], yy = q[1] * q[1], zz = q[2] * q[2];
  const float xy = q[0] * q[1], xz = q[0] * q[2], yz = q[1] * q[2];
  const float wx = q[3] * q[0], wy = q[3] * q[1], wz = q[3] * q[2];
  x.v[0] = {(1.f - 2.f * (yy + zz)) * s[0], 2.f * (xy + wz) * s[0], 2.f * (xz - wy) * s[0]};
  x.v[1] = {2.f * (xy - wz) * s[1], (1.f - 2.f * (xx + zz)) * s[1], 2.f * (yz + wx) * s[1]};
  x.v[2] = {2.f * (xz + wy) * s[2], 2.f * (yz - wx) * s[2], (1.f - 2.f * (xx + yy)) * s[2]};
  x.v[3] = {t[0], t[1], t[2]};
  return x;
}

StdMat gltfMaterial(const Json& gm)
{
  // glTF PBR -> Standard parameters (rkscene/rkscene_gltf.h, loadGLTFScene's material loop)
  StdMat m;
  if (const Json* n = gm.get("name")) m.name = n->str();
  float base[4] = {1.f, 1.f, 1.f, 1.f}, metallic = 1.f, rough = 1.f;
  if (const Json* pbr = gm.get("pbrMetallicRoughness"))
  {
    if (const Json* b = pbr->get("baseColorFactor"))
      for (int k = 0; k < 4; ++k) base[k] = float(b->arr().at(size_t(k)).num());
    metallic = jnum(*pbr, "metallicFactor", 1.f);
    rough = jnum(*pbr, "roughnessFactor", 1.f);
    if (pbr->get("baseColorTexture") || pbr->get("metallicRoughnessTexture")) m.textured = true;
  }
  if (gm.get("normalTexture") || gm.get("emissiveTexture")) m.textured = true;
  std::copy(base, base + 3, m.baseColor);
  m.metalness = metallic;
  m.specularRoughness = rough;
  float emissive[3] = {0.f, 0.f, 0.f}, emissiveStrength = 1.f;
  if (const Json* e = gm.get("emissiveFactor")) readF3(*e, emissive);
  float specular = 1.f, specColor[3] = {1.f, 1.f, 1.f}, ior = 1.5f, transmission = 0.f, thickness = 0.f;
  float sheenColor[3] = {0.f, 0.f, 0.f}, sheenRough = 0.f, attDist = INFINITY, attColor[3] = {1.f, 1.f, 1.f}, clearcoat = 0.f, clearcoatRough = 0.f;
  float diffTrans = 0.f;
  if (const Json* ext = gm.get("extensions"))
  {
    if (const Json* e = ext->get("KHR_materials_specular"))
    {
      specular = jnum(*e, "specularFactor", 1.f);
      if (const Json* c = e->get("specularColorFactor")) readF3(*c, specColor);
    }
    if (const Json* e = ext->get("KHR_materials_ior")) ior = jnum(*e, "ior", 1.5f);
    if (const Json* e = ext->get("KHR_materials_transmission")) transmission = jnum(*e, "transmissionFactor", 0.f);
    if (const Json* e = ext->get("KHR_materials_sheen"))
    {
      if (const Json* c = e->get("sheenColorFactor")) readF3(*c, sheenColor);
      sheenRough = jnum(*e, "sheenRoughnessFactor", 0.f);
      if (e->get("sheenColorTexture") || e->get("sheenRoughnessTexture")) m.textured = true;
    }
    if (const Json* e = ext->get("KHR_materials_diffuse_transmission")) diffTrans = jnum(*e, "diffuseTransmissionFactor", 0.f);
    if (const Json* e = ext->get("KHR_materials_volume"))
    {
      thickness = jnum(*e, "thicknessFactor", 0.f);
      attDist = jnum(*e, "attenuationDistance", INFINITY);
      if (const Json* c = e->get("attenuationColor")) readF3(*c, attColor);
    }
    if (const Json* e = ext->get("KHR_materials_clearcoat"))
    {
      clearcoat = jnum(*e, "clearcoatFactor", 0.f);
      clearcoatRough = jnum(*e, "clearcoatRoughnessFactor", 0.f);
    }
    if (const Json* e = ext->get("KHR_materials_emissive_strength")) emissiveStrength = jnum(*e, "emissiveStrength", 1.f);
  }
  m.specularWeight = specular;
  std::copy(specColor, specColor + 3, m.specularColor);
  m.specularIor = ior;
  m.transmissionWeight = transmission;
  if (transmission != 0.f) std::copy(base, base + 3, m.transmissionColor);
  // KHR_materials_sheen (rkscene/rkscene_gltf.h:975-978): weight 1 for a non-black colour
  if (sheenColor[0] != 0.f || sheenColor[1] != 0.f || sheenColor[2] != 0.f) m.sheenWeight = 1.f;
  std::copy(sheenColor, sheenColor + 3, m.sheenColor);
  m.sheenRoughness = sheenRough;
  m.subsurfaceWeight = diffTrans;
  if (thickness != 0.f && std::isfinite(attDist))
  {
    m.transmissionDepth = attDist;
    std::copy(attColor, attColor + 3, m.transmissionColor);
  }
  m.coatWeight = clearcoat;
  m.coatRoughness = clearcoatRough;
  m.coatIor = 1.5f;
  if (emissive[0] != 0.f || emissive[1] != 0.f || emissive[2] != 0.f)
  {
    std::copy(emissive, emissive + 3, m.emissionColor);
    m.emissionLuminance = emissiveStrength;
  }
  m.thinWalled = thickness == 0.f;
  if (const Json* am = gm.get("alphaMode"))
    if (am->str() == "MASK" || am->str() == "BLEND") m.opacity = base[3];
  return m;
}

Desc loadGltf(const std::string& path)
{
  Gltf g;
  g.doc = JsonReader(readText(path)).parse();
  if (const Json* bufs = g.doc.get("buffers"))
    for (const Json& b : bufs->arr())
    {
      const std::string uri = b.get("uri") ? b.get("uri")->str() : std::string();
      const size_t len = size_t(b.get("byteLength")->num());
      std::vector<char> data;
      if (uri.compare(0, 5, "data:") == 0)
      {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.find(";base64") == std::string::npos) fail("unsupported glTF data URI");
        data = base64(uri.substr(comma + 1));
      }
      else if (!uri.empty())
      {
        const std::string file = uri[0] == '/' ? uri : dirOf(path) + uri;
        std::ifstream f(file, std::ios::binary);
        if (!f) fail("failed to open file: " + file);
        data.resize(len);
        f.read(data.data(), std::streamsize(len));
        data.resize(size_t(f.gcount()));
      }
      else
        fail("glTF buffers without a URI (GLB) are not supported");
      if (data.size() < len) fail("glTF buffer is shorter than its byteLength");
      g.buffers.push_back(std::move(data));
    }
  Desc d;
  if (const Json* mats = g.doc.get("materials"))
    for (const Json& m : mats->arr()) d.mats.push_back(gltfMaterial(m));
  // meshes become objects (one per mesh, instanced by the nodes that use it)
  std::map<int, int> meshObj;
  auto meshToObject = [&](int mi) {
    auto it = meshObj.find(mi);
    if (it != meshObj.end()) return it->second;
    const Json& mesh = g.at("meshes", mi);
    Obj ob;
    for (const Json& prim : mesh.get("primitives")->arr())
    {
      const int mode = prim.get("mode") ? int(prim.get("mode")->num()) : 4;
      if (mode != 4) fail("only triangle glTF primitives are supported");
      const Json* attrs = prim.get("attributes");
      Geom geo;
      if (const Json* p = attrs->get("POSITION"))
      {
        std::vector<float> f;
        g.accessor(int(p->num()), 3, false, f);
        for (size_t k = 0; k + 2 < f.size(); k += 3) geo.pos.push_back({f[k], f[k + 1], f[k + 2]});
      }
      if (const Json* n = attrs->get("NORMAL"))
      {
        std::vector<float> f;
        g.accessor(int(n->num()), 3, false, f);
        for (size_t k = 0; k + 2 < f.size(); k += 3) geo.nrm.push_back({f[k], f[k + 1], f[k + 2]});
      }
      if (const Json* ix = prim.get("indices"))
        g.accessor(int(ix->num()), 1, true, geo.idx);
      else
        for (uint32_t k = 0; k < uint32_t(geo.pos.size()); ++k) geo.idx.push_back(k);
      if (const Json* m = prim.get("material")) geo.material = int(m->num());
      ob.geoms.push_back(int(d.geoms.size()));
      d.geoms.push_back(std::move(geo));
    }
    d.objs.push_back(std::move(ob));
    meshObj[mi] = int(d.objs.size()) - 1;
    return int(d.objs.size()) - 1;
  };
  int sceneIdx = g.doc.get("scene") ? int(g.doc.get("scene")->num()) : 0;
  const Json& scene = g.at("scenes", sceneIdx);
  // pre-order walk, roots and children in document order (objects and their instances are
  // numbered in the order the walk meets them)
  std::vector<std::pair<int, Xfm>> stack;
  if (const Json* ns = scene.get("nodes"))
    for (auto it = ns->arr().rbegin(); it != ns->arr().rend(); ++it) stack.push_back({int(it->num()), Xfm()});
  int guard = 0;
  while (!stack.empty())
  {
    if (++guard > 10000000) fail("glTF node hierarchy too deep or cyclic");
    auto [ni, parent] = stack.back();
    stack.pop_back();
    const Json& node = g.at("nodes", ni);
    const Xfm x = mul(parent, nodeTransform(node));
    if (const Json* m = node.get("mesh")) d.objs[size_t(meshToObject(int(m->num())))].instances.push_back(x);
    if (const Json* ch = node.get("children"))
      for (auto it = ch->arr().rbegin(); it != ch->arr().rend(); ++it) stack.push_back({int(it->num()), x});
  }
  return d;
}

// ---------------------------------------------------------------------------------------
// .mesh blob + .mtl (render/scene.cpp:136-172, material/mtl_loader.cpp,
// material/material_factory.cpp:29-101)
// ---------------------------------------------------------------------------------------
constexpr int kMeshBlobId = 0x04df8666;

struct MtlProps
{
  std::map<std::string, std::vector<float>> num;
  std::map<std::string, std::string> str;
  bool has(const std::string& k) const { return num.count(k) || str.count(k); }
  float f(const std::string& k, float def) const
  {
    auto it = num.find(k);
    return it != num.end() && !it->second.empty() ? it->second[0] : def;
  }
  void f3(const std::string& k, float* out) const
  {
    auto it = num.find(k);
    if (it == num.end() || it->second.empty()) return;
    for (int c = 0; c < 3; ++c) out[c] = it->second[std::min<size_t>(size_t(c), it->second.size() - 1)];
  }
};

std::map<std::string, MtlProps> loadMtl(const std::string& path)
{
  std::map<std::string, MtlProps> lib;
  std::ifstream f(path);
  if (!f) return lib;
  std::string line, cur;
  while (std::getline(f, line))
  {
    std::istringstream ls(line);
    std::string head;
    if (!(ls >> head) || head[0] == '#') continue;
    std::string rest;
    std::getline(ls, rest);
    const size_t b = rest.find_first_not_of(" \t");
    rest = b == std::string::npos ? std::string() : rest.substr(b);
    while (!rest.empty() && (rest.back() == '\r' || rest.back() == ' ' || rest.back() == '\t')) rest.pop_back();
    if (head == "newmtl")
    {
      cur = rest;
      lib[cur];
      continue;
    }
    std::istringstream vs(rest);
    std::vector<float> v;
    float x;
    while (v.size() < 3 && (vs >> x)) v.push_back(x);
    if (v.empty()) lib[cur].str[head] = rest;
    else lib[cur].num[head] = v;
  }
  return lib;
}

StdMat mtlMaterial(const MtlProps& p)
{
  StdMat m;
  const std::string type = p.str.count("type") ? p.str.at("type") : std::string("OBJ");
  if (type == "Standard")
  {
    p.f3("baseColor", m.baseColor);
    m.baseWeight = p.f("baseWeight", 1.f);
    m.metalness = p.f("metalness", 0.f);
    m.specularWeight = p.f("specularWeight", 1.f);
    p.f3("specularColor", m.specularColor);
    m.specularRoughness = p.f("specularRoughness", 0.3f);
    m.specularIor = p.f("specularIOR", 1.5f);
    m.transmissionWeight = p.f("transmissionWeight", 0.f);
    p.f3("transmissionColor", m.transmissionColor);
    m.transmissionDepth = p.f("transmissionDepth", 1.f);
    m.coatWeight = p.f("coatWeight", 0.f);
    p.f3("coatColor", m.coatColor);
    m.coatRoughness = p.f("coatRoughness", 0.f);
    m.coatIor = p.f("coatIOR", 1.6f);
    m.sheenWeight = p.f("sheenWeight", 0.f);
    p.f3("sheenColor", m.sheenColor);
    m.sheenRoughness = p.f("sheenRoughness", 0.3f);
    m.emissionLuminance = p.f("emissionLuminance", 0.f);
    p.f3("emissionColor", m.emissionColor);
    m.thinWalled = p.f("thinWalled", 0.f) != 0.f;
    m.opacity = p.f("opacity", 1.f);
    for (const char* k : {"baseColorMap", "normalMap", "bumpMap"})
      if (p.has(k)) m.textured = true;
    return m;
  }
  if (type == "Luminous")
  {
    m.luminous = true;
    p.f3("color", m.luminousColor);
    if (!p.has("color")) p.f3("Ke", m.luminousColor);
    return m;
  }
  // OBJ and Principled (the legacy StandardMaterial parameter names)
  m.isDefault = false;
  float kd[3] = {0.8f, 0.8f, 0.8f};
  if (p.has("baseColor")) p.f3("baseColor", kd);
  else p.f3("Kd", kd);
  std::copy(kd, kd + 3, m.baseColor);
  m.baseWeight = p.f("diffuse", 1.f);
  m.metalness = p.f("metallic", 0.f);
  m.specularWeight = p.f("specular", 1.f);
  p.f3("edgeColor", m.specularColor);
  float rough = p.f("roughness", 0.f);
  if (!p.has("roughness") && p.has("Ns")) rough = std::sqrt(std::sqrt(2.f / (p.f("Ns", 0.f) + 2.f)));
  m.specularRoughness = rough;
  float ior = p.f("ior", 1.f);
  if (!p.has("ior") && p.has("Ks"))
  {
    float ks[3] = {0.f, 0.f, 0.f};
    p.f3("Ks", ks);
    const float f0 = (ks[0] + ks[1] + ks[2]) / 3.f * 0.08f;
    ior = (1.f + std::sqrt(f0)) / (1.f - std::sqrt(f0));
  }
  m.specularIor = ior;
  m.transmissionWeight = p.f("transmission", 0.f);
  if (!p.has("transmission") && p.has("Tf"))
  {
    float tf[3] = {1.f, 1.f, 1.f};
    p.f3("Tf", tf);
    if (std::min({tf[0], tf[1], tf[2]}) < 1.f)
    {
      m.transmissionWeight = 1.f;
      for (int c = 0; c < 3; ++c) m.transmissionColor[c] = 1.f - tf[c];
    }
  }
  m.coatWeight = p.f("coat", 0.f);
  // legacy sheen (material/standard.cpp:133-137, :605-607): roughness 0.2, colour lerped to the
  // base colour by sheenTint
  m.sheenWeight = p.f("sheen", 0.f);
  p.f3("sheenColor", m.sheenColor);
  m.sheenRoughness = p.f("sheenRoughness", 0.2f);
  const float tint = std::min(std::max(p.f("sheenTint", 0.f), 0.f), 1.f);
  for (int c = 0; c < 3; ++c) m.sheenColor[c] += tint * (m.baseColor[c] - m.sheenColor[c]);
  m.thinWalled = p.f("thin", 0.f) != 0.f;
  m.opacity = p.f("opacity", 1.f);
  for (const char* k : {"map_Kd", "baseColorMap", "map_Bump", "map_bump", "bumpMap", "normalMap", "map_d"})
    if (p.has(k)) m.textured = true;
  return m;
}

StdMat defaultMaterial()
{
  StdMat m;
  m.isDefault = true;
  return m;
}

template <class T>
void rd(std::ifstream& f, T& v)
{
  f.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!f) fail("unexpected end of .mesh stream");
}

// Loads the blob straight into a one-object description whose triangles carry their own
// material IDs (already in the reference's numbering: 0 = environment, lights first).
struct MeshFile
{
  std::vector<F3> pos, nrm;
  std::vector<int32_t> idx, mat;
  std::vector<StdMat> mats;
};

MeshFile loadMesh(const std::string& path)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) fail("failed to open file: " + path);
  int32_t id = 0;
  rd(f, id);
  if (id != kMeshBlobId) fail("file has wrong blob ID: " + path);
  float bounds[6];
  f.read(reinterpret_cast<char*>(bounds), sizeof(bounds));
  int32_t nt = 0, nv = 0;
  uint8_t hasNormals = 0;
  int8_t ntc = 0;
  rd(f, nt);
  rd(f, nv);
  rd(f, hasNormals);
  rd(f, ntc);
  if (nt < 0 || nv < 0 || ntc < 0) fail("invalid triangle mesh data in stream");
  MeshFile m;
  m.idx.resize(size_t(nt) * 3);
  m.mat.resize(size_t(nt));
  f.read(reinterpret_cast<char*>(m.idx.data()), std::streamsize(m.idx.size() * 4));
  f.read(reinterpret_cast<char*>(m.mat.data()), std::streamsize(m.mat.size() * 4));
  const int stride = 3 + (hasNormals ? 3 : 0) + 2 * ntc;
  std::vector<float> attr(size_t(nv) * stride);
  f.read(reinterpret_cast<char*>(attr.data()), std::streamsize(attr.size() * 4));
  if (!f) fail("unexpected end of .mesh stream");
  m.pos.resize(size_t(nv));
  if (hasNormals) m.nrm.resize(size_t(nv));
  for (int i = 0; i < nv; ++i)
  {
    const float* a = attr.data() + size_t(i) * stride;
    m.pos[size_t(i)] = {a[0], a[1], a[2]};
    if (hasNormals) m.nrm[size_t(i)] = {a[3], a[4], a[5]};
  }
  int32_t nn = 0;
  rd(f, nn);
  if (nn < 0) fail("invalid array size in stream");
  std::vector<std::string> names(static_cast<size_t>(nn));
  for (auto& s : names)
  {
    int32_t len = 0;
    rd(f, len);
    if (len < 0) fail("invalid string size in stream");
    s.resize(size_t(len));
    f.read(s.data(), len);
    if (!f) fail("unexpected end of .mesh stream");
  }
  std::string base = path.substr(0, path.size() - 5);  // strip ".mesh"
  const auto lib = loadMtl(base + ".mtl");
  if (names.empty())
    m.mats.push_back(defaultMaterial());
  else
    for (const auto& n : names)
    {
      auto it = lib.find(n);
      StdMat sm = it != lib.end() ? mtlMaterial(it->second) : defaultMaterial();
      sm.name = n;
      m.mats.push_back(sm);
    }
  for (int32_t& t : m.mat)
    if (t < 0 || size_t(t) >= m.mats.size()) t = 0;
  for (int32_t v : m.idx)
    if (v < 0 || v >= nv) fail("triangle index out of range in .mesh");
  return m;
}

// ---------------------------------------------------------------------------------------
// Standard parameters -> material record: a single-lobe kind where StandardMaterial's composition
// reduces to one lobe of weight 1, kind 5 (the weighted bases and the coat) otherwise.  Textures
// (their factors are kept), subsurface and zero roughness (the smooth lobes, modelled by the
// rough ones at the minimum roughness) count the material as approximated.
// ---------------------------------------------------------------------------------------
bool emissive(const StdMat& m)
{
  if (m.luminous) return m.luminousColor[0] > 0.f || m.luminousColor[1] > 0.f || m.luminousColor[2] > 0.f;
  return m.emissionLuminance > 0.f;
}

prt_material toRecord(const StdMat& m, bool& approx)
{
  prt_material r{};
  r.thickness = 1.f;
  r.weight = 1.f;
  r.ior = 1.5f;
  if (m.luminous)
  {
    r.kind = 0;
    r.r = r.g = r.b = 0.f;
    for (int c = 0; c < 3; ++c) r.emission[c] = m.luminousColor[c];
    return r;
  }
  if (m.isDefault)
  {
    // the factory default (Principled, ior = outside ior 1: no specular layer): diffuse 0.8
    r.kind = 0;
    r.r = r.g = r.b = 0.8f;
    return r;
  }
  if (m.emissionLuminance > 0.f)
    for (int c = 0; c < 3; ++c) r.emission[c] = m.emissionColor[c] * m.emissionLuminance;
  const float ior = m.specularIor > 0.f && m.specularIor < 1.f ? 1.f / m.specularIor : m.specularIor;
  const float coatIor = m.coatIor > 0.f && m.coatIor < 1.f ? 1.f / m.coatIor : m.coatIor;
  auto cl = [](float x) { return std::min(std::max(x, 0.f), 1.f); };
  const float eps = 1e-5f;  // material/standard.h:79
  const float base = cl(m.baseWeight), metal = cl(m.metalness), spec = cl(m.specularWeight);
  const float trans = cl(m.transmissionWeight), coat = cl(m.coatWeight);
  const bool specIor = std::fabs(ior - 1.f) > eps, hasCoat = coat > eps && std::fabs(coatIor - 1.f) > eps;
  // textures keep their factors and subsurface is not modelled; roughness 0 (the smooth lobes) is
  // modelled by the rough lobes at the minimum roughness
  approx = m.textured || m.subsurfaceWeight > 0.f;
  // the sheen layer and opacity over any lobe set (material/standard.cpp:599-627)
  r.sheen_weight = m.sheenWeight;
  for (int c = 0; c < 3; ++c) r.sheen_color[c] = m.sheenColor[c];
  r.sheen_roughness = m.sheenRoughness;
  r.transparency = 1.f - cl(m.opacity);
  r.diffuse_roughness = cl(m.diffuseRoughness);
  for (int c = 0; c < 3; ++c)
  {
    r.transmission_color[c] = m.transmissionColor[c];
    r.coat_color[c] = m.coatColor[c];
  }
  r.transmission_depth = m.transmissionDepth;
  // the single-lobe kinds where the composition reduces to one lobe of weight 1 (their kernels
  // are the leanest), kind 5 for every other parameter set
  if (metal >= 1.f && spec >= 1.f && !hasCoat)
  {
    r.kind = 1;  // metal base alone (:533-546)
    r.r = m.baseColor[0] * base; r.g = m.baseColor[1] * base; r.b = m.baseColor[2] * base;
    r.er = m.specularColor[0]; r.eg = m.specularColor[1]; r.eb = m.specularColor[2];
    r.roughness = m.specularRoughness;
    approx = approx || m.specularRoughness <= 0.f;
    return r;
  }
  if (metal <= 0.f && trans >= 1.f && spec >= 1.f && specIor && !hasCoat)
  {
    r.ior = ior;
    r.roughness = m.specularRoughness;
    approx = approx || m.specularRoughness <= 0.f;
    if (m.thinWalled)
    {
      // thin-walled transmission: ThinRoughDielectricBsdf with transmissionColor and thickness /
      // transmissionDepth, thickness 1 (material/standard.cpp:534-543, :48-52, :79-80)
      r.kind = 4;
      r.r = m.transmissionColor[0]; r.g = m.transmissionColor[1]; r.b = m.transmissionColor[2];
      r.thickness = m.transmissionDepth > 0.f ? 1.f / m.transmissionDepth : 3.4e38f;
      return r;
    }
    // the medium inside: transmissionColor ^ transmissionDepth (material/standard.cpp:169-173)
    r.kind = 2;
    r.r = r.g = r.b = 1.f;
    return r;
  }
  if (metal <= 0.f && trans <= 0.f && !(specIor && spec > eps && hasCoat))
  {
    r.r = m.baseColor[0] * base; r.g = m.baseColor[1] * base; r.b = m.baseColor[2] * base;
    if (specIor && spec > eps)
    {
      // the default Standard material: a rough dielectric layer of weight specularWeight over the
      // diffuse base (:499-503)
      r.kind = 3;
      r.er = r.eg = r.eb = 1.f;
      r.thickness = 1.f;
      r.weight = spec;
      r.ior = ior;
      r.roughness = m.specularRoughness;
      approx = approx || m.specularRoughness <= 0.f;
      return r;
    }
    if (hasCoat)
    {
      // the coat over a diffuse base (:551-566), StandardMaterial's coat thickness 0.5 (:66)
      r.kind = 3;
      r.er = m.coatColor[0]; r.eg = m.coatColor[1]; r.eb = m.coatColor[2];
      r.thickness = 0.5f;
      r.weight = coat;
      r.ior = coatIor;
      r.roughness = m.coatRoughness;
      approx = approx || m.coatRoughness <= 0.f;
      return r;
    }
    r.kind = 0;
    return r;
  }
  // kind 5: the weighted bases and the coat (material/standard.cpp:409-611)
  r.kind = PRT_KIND_STANDARD;
  r.r = m.baseColor[0]; r.g = m.baseColor[1]; r.b = m.baseColor[2];
  r.er = m.specularColor[0]; r.eg = m.specularColor[1]; r.eb = m.specularColor[2];
  r.roughness = m.specularRoughness;
  r.ior = m.specularIor > 0.f ? m.specularIor : 1.5f;
  r.base_weight = m.baseWeight;
  r.metalness = m.metalness;
  r.specular_weight = m.specularWeight;
  r.transmission_weight = m.transmissionWeight;
  r.thin_walled = m.thinWalled ? 1 : 0;
  r.thickness = 1.f;  // thicknessPar.scale (:79)
  r.coat_weight = m.coatWeight;
  r.coat_roughness = m.coatRoughness;
  r.coat_ior = m.coatIor > 0.f ? m.coatIor : 1.6f;
  r.coat_thickness = 0.5f;  // coatThicknessPar.scale (:66)
  approx = approx || (m.specularRoughness <= 0.f && spec > eps) || (hasCoat && m.coatRoughness <= 0.f);
  return r;
}

}  // namespace

struct prt_scene_file
{
  std::vector<float> pos, nrm;
  std::vector<int32_t> idx, triMat;
  std::vector<prt_material> mats;
  std::vector<std::string> names;
  prt_scene_file_info info{};
  // two-level load: meshes as triangle ranges and the instances that place them (empty: flat)
  std::vector<int32_t> meshBegin;
  std::vector<prt_instance> inst;
};

namespace {

// Material IDs as the reference's loadScene numbers them (render/scene.cpp:200-262):
// 0 = environment placeholder, the file's materials, the default material last; then the
// emissive materials are swapped to the front (IDs 1..).  Returns file material -> ID.
void numberMaterials(const std::vector<StdMat>& in, prt_scene_file& out, std::vector<int>& fileToId, int& defaultId)
{
  std::vector<StdMat> mats;
  std::vector<int> fileIdx;
  mats.push_back(defaultMaterial());
  fileIdx.push_back(-1);
  for (size_t i = 0; i < in.size(); ++i)
  {
    mats.push_back(in[i]);
    fileIdx.push_back(int(i));
  }
  mats.push_back(defaultMaterial());
  fileIdx.push_back(-1);
  int next = 1;
  for (size_t i = 1; i < mats.size(); ++i)
    if (emissive(mats[i]))
    {
      if (int(i) != next)
      {
        std::swap(mats[i], mats[size_t(next)]);
        std::swap(fileIdx[i], fileIdx[size_t(next)]);
      }
      ++next;
    }
  fileToId.assign(in.size(), -1);
  defaultId = -1;
  for (size_t i = 1; i < mats.size(); ++i)
  {
    if (fileIdx[i] >= 0) fileToId[size_t(fileIdx[i])] = int(i);
    else defaultId = int(i);  // the reference's materialMap[-1]: the last placeholder wins
  }
  out.mats.clear();
  out.names.clear();
  int approx = 0;
  for (size_t i = 0; i < mats.size(); ++i)
  {
    bool a = false;
    out.mats.push_back(toRecord(mats[i], a));
    out.names.push_back(mats[i].name);
    if (a && i > 0) ++approx;
  }
  out.info.approximated_materials = approx;
}

void appendGeometry(prt_scene_file& out, const Geom& g, const Xfm* x, int matId, bool withNormals)
{
  const size_t base = out.pos.size() / 3;
  // normals transform with the inverse transpose of the linear part
  float n[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (x)
  {
    const F3 *a = &x->v[0], *b = &x->v[1], *c = &x->v[2];
    // columns a, b, c of M; cofactor matrix = det * inverse transpose
    const F3 bc = {b->y * c->z - b->z * c->y, b->z * c->x - b->x * c->z, b->x * c->y - b->y * c->x};
    const F3 ca = {c->y * a->z - c->z * a->y, c->z * a->x - c->x * a->z, c->x * a->y - c->y * a->x};
    const F3 ab = {a->y * b->z - a->z * b->y, a->z * b->x - a->x * b->z, a->x * b->y - a->y * b->x};
    // inverse transpose has columns (b x c, c x a, a x b) / det
    const float det = a->x * bc.x + a->y * bc.y + a->z * bc.z;
    const float inv = det != 0.f ? 1.f / det : 0.f;
    const F3 cols[3] = {bc, ca, ab};
    for (int k = 0; k < 3; ++k)
    {
      n[k][0] = cols[k].x * inv;
      n[k][1] = cols[k].y * inv;
      n[k][2] = cols[k].z * inv;
    }
  }
  for (size_t v = 0; v < g.pos.size(); ++v)
  {
    const F3 p = g.pos[v];
    if (x)
    {
      out.pos.push_back(x->v[0].x * p.x + x->v[1].x * p.y + x->v[2].x * p.z + x->v[3].x);
      out.pos.push_back(x->v[0].y * p.x + x->v[1].y * p.y + x->v[2].y * p.z + x->v[3].y);
      out.pos.push_back(x->v[0].z * p.x + x->v[1].z * p.y + x->v[2].z * p.z + x->v[3].z);
    }
    else
    {
      out.pos.push_back(p.x);
      out.pos.push_back(p.y);
      out.pos.push_back(p.z);
    }
    if (withNormals)
    {
      F3 q = v < g.nrm.size() ? g.nrm[v] : F3{0.f, 0.f, 0.f};
      if (x) q = {n[0][0] * q.x + n[1][0] * q.y + n[2][0] * q.z, n[0][1] * q.x + n[1][1] * q.y + n[2][1] * q.z,
                  n[0][2] * q.x + n[1][2] * q.y + n[2][2] * q.z};
      out.nrm.push_back(q.x);
      out.nrm.push_back(q.y);
      out.nrm.push_back(q.z);
    }
  }
  if (g.idx.size() % 3) fail("triangle index count is not a multiple of 3");
  for (size_t t = 0; t + 2 < g.idx.size(); t += 3)
  {
    for (int k = 0; k < 3; ++k)
    {
      if (g.idx[t + k] >= g.pos.size()) fail("triangle index out of range");
      out.idx.push_back(int32_t(base + g.idx[t + k]));
    }
    out.triMat.push_back(matId);
  }
}

void flatten(const Desc& d, prt_scene_file& out)
{
  std::vector<int> fileToId;
  int defaultId = 0;
  numberMaterials(d.mats, out, fileToId, defaultId);
  bool withNormals = false;
  for (const Geom& g : d.geoms)
    if (!g.nrm.empty()) withNormals = true;
  int instances = 0;
  for (const Obj& o : d.objs)
  {
    const std::vector<Xfm>& xs = o.instances;
    const int copies = xs.empty() ? 1 : int(xs.size());
    instances += xs.empty() ? 0 : int(xs.size());
    for (int c = 0; c < copies; ++c)
      for (int gi : o.geoms)
      {
        if (gi < 0 || size_t(gi) >= d.geoms.size()) fail("object references a missing geometry");
        const Geom& g = d.geoms[size_t(gi)];
        const int mid = g.material >= 0 && size_t(g.material) < fileToId.size() ? fileToId[size_t(g.material)] : defaultId;
        appendGeometry(out, g, xs.empty() ? nullptr : &xs[size_t(c)], mid, withNormals);
      }
  }
  if (!withNormals) out.nrm.clear();
  out.info.instance_count = instances;
  out.info.has_sun_sky = d.hasSun ? 1 : 0;
  out.info.sun_direction[0] = d.sunDir.x;
  out.info.sun_direction[1] = d.sunDir.y;
  out.info.sun_direction[2] = d.sunDir.z;
}

// Two levels (render/scene.cpp:325-363 without `flatten`): each instanced object becomes a mesh
// in its object space, placed by one instance per transform; objects without instances, and
// emissive objects (which the reference always flattens, :343-363), go to one world-space mesh
// placed once by the identity.
void twoLevel(const Desc& d, prt_scene_file& out)
{
  std::vector<int> fileToId;
  int defaultId = 0;
  numberMaterials(d.mats, out, fileToId, defaultId);
  bool withNormals = false;
  for (const Geom& g : d.geoms)
    if (!g.nrm.empty()) withNormals = true;
  auto matOf = [&](const Geom& g) {
    return g.material >= 0 && size_t(g.material) < fileToId.size() ? fileToId[size_t(g.material)] : defaultId;
  };
  auto emissive = [&](const Obj& o) {
    for (int gi : o.geoms)
    {
      if (gi < 0 || size_t(gi) >= d.geoms.size()) fail("object references a missing geometry");
      const prt_material& m = out.mats[size_t(matOf(d.geoms[size_t(gi)]))];
      if (m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f) return true;
    }
    return false;
  };
  auto twoLevelObj = [&](const Obj& o) { return !o.instances.empty() && !emissive(o); };
  out.meshBegin.assign(1, 0);
  // the world-space mesh
  for (const Obj& o : d.objs)
  {
    if (twoLevelObj(o)) continue;
    const int copies = o.instances.empty() ? 1 : int(o.instances.size());
    for (int c = 0; c < copies; ++c)
      for (int gi : o.geoms)
        appendGeometry(out, d.geoms[size_t(gi)], o.instances.empty() ? nullptr : &o.instances[size_t(c)],
                       matOf(d.geoms[size_t(gi)]), withNormals);
  }
  const prt_instance identity{0, {1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f}};
  if (!out.idx.empty())
  {
    out.meshBegin.push_back(int32_t(out.idx.size() / 3));
    out.inst.push_back(identity);
  }
  // one mesh per instanced object, in object space
  for (const Obj& o : d.objs)
  {
    if (!twoLevelObj(o)) continue;
    for (int gi : o.geoms) appendGeometry(out, d.geoms[size_t(gi)], nullptr, matOf(d.geoms[size_t(gi)]), withNormals);
    const int mesh = int(out.meshBegin.size()) - 1;
    out.meshBegin.push_back(int32_t(out.idx.size() / 3));
    for (const Xfm& x : o.instances)  // columns v[0..2] of L, v[3] = t -> rows of [L | t]
      out.inst.push_back(prt_instance{mesh, {x.v[0].x, x.v[1].x, x.v[2].x, x.v[3].x, x.v[0].y, x.v[1].y, x.v[2].y, x.v[3].y,
                                             x.v[0].z, x.v[1].z, x.v[2].z, x.v[3].z}});
  }
  if (!withNormals) out.nrm.clear();
  out.info.instance_count = int32_t(out.inst.size());
  out.info.has_sun_sky = d.hasSun ? 1 : 0;
  out.info.sun_direction[0] = d.sunDir.x;
  out.info.sun_direction[1] = d.sunDir.y;
  out.info.sun_direction[2] = d.sunDir.z;
}

void fromMesh(const MeshFile& m, prt_scene_file& out)
{
  // .mesh material IDs are stored already numbered (render/scene.cpp:136-172): no
  // environment placeholder is added and no reordering happens here
  out.mats.clear();
  int approx = 0;
  for (size_t i = 0; i < m.mats.size(); ++i)
  {
    bool a = false;
    out.mats.push_back(toRecord(m.mats[i], a));
    out.names.push_back(m.mats[i].name);
    if (a && i > 0) ++approx;
  }
  out.info.approximated_materials = approx;
  for (const F3& p : m.pos)
  {
    out.pos.push_back(p.x);
    out.pos.push_back(p.y);
    out.pos.push_back(p.z);
  }
  for (const F3& n : m.nrm)
  {
    out.nrm.push_back(n.x);
    out.nrm.push_back(n.y);
    out.nrm.push_back(n.z);
  }
  out.idx = m.idx;
  out.triMat = m.mat;
}

}  // namespace

extern "C" {

static int loadFile(const char* path, prt_scene_file** out, prt_scene_file_info* info, bool instanced)
{
  if (!path || !out) return prt::setLastError(PRT_ERR_INVALID, "null argument");
  *out = nullptr;
  auto f = std::make_unique<prt_scene_file>();
  try
  {
    const std::string p(path), ext = extOf(p);
    if (ext == "rkscene")
    {
      if (instanced) twoLevel(loadRkscene(p), *f);
      else flatten(loadRkscene(p), *f);
    }
    else if (ext == "gltf")
    {
      if (instanced) twoLevel(loadGltf(p), *f);
      else flatten(loadGltf(p), *f);
    }
    else if (ext == "mesh") fromMesh(loadMesh(p), *f);
    else return prt::setLastError(PRT_ERR_INVALID, "unknown scene format: " + p);
  }
  catch (const LoadError& e)
  {
    return prt::setLastError(PRT_ERR_INVALID, e.msg);
  }
  catch (const std::exception& e)
  {
    return prt::setLastError(PRT_ERR_INVALID, std::string("scene file: ") + e.what());
  }
  if (f->idx.size() / 3 > size_t(INT32_MAX) || f->pos.size() / 3 > size_t(INT32_MAX))
    return prt::setLastError(PRT_ERR_INVALID, "scene too large for 32-bit indices");
  f->info.vertex_count = int64_t(f->pos.size() / 3);
  f->info.triangle_count = int64_t(f->idx.size() / 3);
  f->info.material_count = int32_t(f->mats.size());
  f->info.has_normals = f->nrm.empty() ? 0 : 1;
  if (info) *info = f->info;
  *out = f.release();
  return PRT_OK;
}

int prt_scene_file_load(const char* path, prt_scene_file** out, prt_scene_file_info* info)
{
  return loadFile(path, out, info, false);
}

int prt_scene_file_load_instanced(const char* path, prt_scene_file** out, prt_scene_file_info* info)
{
  return loadFile(path, out, info, true);
}

int prt_scene_file_get_instances(prt_scene_file* f, int32_t* n_meshes, int32_t* mesh_begin, int32_t* n_instances,
                                 prt_instance* instances)
{
  if (!f) return prt::setLastError(PRT_ERR_INVALID, "null scene file");
  if (n_meshes) *n_meshes = f->meshBegin.empty() ? 0 : int32_t(f->meshBegin.size()) - 1;
  if (n_instances) *n_instances = int32_t(f->inst.size());
  if (mesh_begin && !f->meshBegin.empty()) std::memcpy(mesh_begin, f->meshBegin.data(), f->meshBegin.size() * sizeof(int32_t));
  if (instances && !f->inst.empty()) std::memcpy(instances, f->inst.data(), f->inst.size() * sizeof(prt_instance));
  return PRT_OK;
}

int prt_scene_file_get(prt_scene_file* f, float* positions, float* normals, int32_t* indices, int32_t* tri_material,
                       prt_material* materials)
{
  if (!f) return prt::setLastError(PRT_ERR_INVALID, "null scene file");
  if (positions && !f->pos.empty()) std::memcpy(positions, f->pos.data(), f->pos.size() * sizeof(float));
  if (normals && !f->nrm.empty()) std::memcpy(normals, f->nrm.data(), f->nrm.size() * sizeof(float));
  if (indices && !f->idx.empty()) std::memcpy(indices, f->idx.data(), f->idx.size() * sizeof(int32_t));
  if (tri_material && !f->triMat.empty()) std::memcpy(tri_material, f->triMat.data(), f->triMat.size() * sizeof(int32_t));
  if (materials && !f->mats.empty()) std::memcpy(materials, f->mats.data(), f->mats.size() * sizeof(prt_material));
  return PRT_OK;
}

int prt_scene_file_material_name(prt_scene_file* f, int32_t id, char* buf, int32_t len)
{
  if (!f || !buf || len < 1) return prt::setLastError(PRT_ERR_INVALID, "null argument");
  if (id < 0 || size_t(id) >= f->names.size()) return prt::setLastError(PRT_ERR_INVALID, "material ID out of range");
  std::snprintf(buf, size_t(len), "%s", f->names[size_t(id)].c_str());
  return PRT_OK;
}

int prt_scene_file_destroy(prt_scene_file* f)
{
  delete f;
  return PRT_OK;
}

}  // extern "C"
