// Scene-file harness: TEST INFRASTRUCTURE ONLY (built into oracle/_ref by Makefile.ref from the
// reference's own sources, never linked into the product).
//
//   ref_scene dump FILE            parse an .rkscene / .gltf with the reference's loaders
//                                  (rkscene/rkscene.h:847-903, rkscene/rkscene_gltf.h) and print
//                                  the parsed scene as JSON
//   ref_scene mesh DESC OUT        write a .mesh blob with the reference's TriangleMesh stream
//                                  operator and saveBlob (geometry/triangle_mesh.cpp:362-397,
//                                  sys/blob.h:27-37)
//   ref_scene clone DESC m0..m11   copy a mesh through TriangleMesh::clone(Affine3f)
//                                  (geometry/triangle_mesh.cpp:326-359; the flattening of an
//                                  instance) and print its positions and normals as JSON
//
// DESC is a text mesh: "nv nt hasNormals nnames", nv lines "x y z [nx ny nz]", nt lines
// "i j k material", nnames lines of material names ("-" for the empty name).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "rkscene/rkscene.h"
#include "rkscene/rkscene_gltf.h"
#include "geometry/triangle_mesh.h"
#include "math/affine3.h"
#include "sys/blob.h"

namespace {

void f(float v) { std::printf("%.9g", v); }

void f3(const rkscene::float3& v)
{
  std::printf("[");
  f(v.x); std::printf(","); f(v.y); std::printf(","); f(v.z);
  std::printf("]");
}

template <class T>
void tex1(const char* name, const rkscene::Texture<T>& t)
{
  std::printf("\"%s\":", name);
  f(t.factor);
  std::printf(",\"%s_image\":%d,", name, t.image.empty() ? 0 : 1);
}

void tex3(const char* name, const rkscene::Texture3f& t)
{
  std::printf("\"%s\":", name);
  f3(t.factor);
  std::printf(",\"%s_image\":%d,", name, t.image.empty() ? 0 : 1);
}

int dump(const std::string& path)
{
  const std::string ext = path.substr(path.find_last_of('.') + 1);
  std::shared_ptr<rkscene::Scene> s = ext == "gltf" ? rkscene::loadGLTF(path) : rkscene::load(path);
  std::printf("{\"materials\":[");
  for (size_t i = 0; i < s->materials.size(); ++i)
  {
    const auto& m = *s->materials[i];
    std::printf("%s{\"name\":\"%s\",", i ? "," : "", m.name.c_str());
    tex1("baseWeight", m.baseWeight);
    tex3("baseColor", m.baseColor);
    tex1("metalness", m.metalness);
    tex1("diffuseRoughness", m.diffuseRoughness);
    tex1("specularWeight", m.specularWeight);
    tex3("specularColor", m.specularColor);
    tex1("specularRoughness", m.specularRoughness);
    std::printf("\"specularIor\":"); f(m.specularIor); std::printf(",");
    tex1("transmissionWeight", m.transmissionWeight);
    tex3("transmissionColor", m.transmissionColor);
    std::printf("\"transmissionDepth\":"); f(m.transmissionDepth); std::printf(",");
    tex1("subsurfaceWeight", m.subsurfaceWeight);
    tex1("coatWeight", m.coatWeight);
    tex3("coatColor", m.coatColor);
    tex1("coatRoughness", m.coatRoughness);
    std::printf("\"coatIor\":"); f(m.coatIor); std::printf(",");
    tex1("sheenWeight", m.sheenWeight);
    tex1("emissionLuminance", m.emissionLuminance);
    tex3("emissionColor", m.emissionColor);
    tex1("opacity", m.opacity);
    std::printf("\"normal_image\":%d,\"thinWalled\":%d}", m.normal.image.empty() ? 0 : 1, m.thinWalled ? 1 : 0);
  }
  std::printf("],\"geometries\":[");
  for (size_t i = 0; i < s->geometries.size(); ++i)
  {
    const auto& g = *s->geometries[i];
    std::printf("%s{\"material\":%d,\"positions\":[", i ? "," : "", g.material);
    for (size_t k = 0; k < g.positions.size(); ++k)
    {
      std::printf("%s", k ? "," : "");
      f(g.positions[k].x); std::printf(","); f(g.positions[k].y); std::printf(","); f(g.positions[k].z);
    }
    std::printf("],\"normals\":[");
    for (size_t k = 0; k < g.normals.size(); ++k)
    {
      std::printf("%s", k ? "," : "");
      f(g.normals[k].x); std::printf(","); f(g.normals[k].y); std::printf(","); f(g.normals[k].z);
    }
    std::printf("],\"indices\":[");
    for (size_t k = 0; k < g.indices.size(); ++k) std::printf("%s%u", k ? "," : "", g.indices[k]);
    std::printf("]}");
  }
  std::printf("],\"objects\":[");
  for (size_t i = 0; i < s->objects.size(); ++i)
  {
    const auto& o = *s->objects[i];
    std::printf("%s{\"geometries\":[", i ? "," : "");
    for (size_t k = 0; k < o.geometries.size(); ++k) std::printf("%s%d", k ? "," : "", o.geometries[k]);
    std::printf("],\"instances\":[");
    for (size_t k = 0; k < o.instances.size(); ++k)
    {
      std::printf("%s[", k ? "," : "");
      for (int c = 0; c < 4; ++c)
      {
        std::printf("%s", c ? "," : "");
        f(o.instances[k].v[c].x); std::printf(","); f(o.instances[k].v[c].y); std::printf(","); f(o.instances[k].v[c].z);
      }
      std::printf("]");
    }
    std::printf("]}");
  }
  std::printf("],\"lights\":[");
  for (size_t i = 0; i < s->lights.size(); ++i)
  {
    std::printf("%s{\"direction\":", i ? "," : "");
    f3(s->lights[i]->direction);
    std::printf("}");
  }
  std::printf("]}\n");
  return 0;
}

void readDesc(const std::string& path, prt::TriangleMesh& mesh)
{
  std::ifstream in(path);
  int nv, nt, hasN, nn;
  in >> nv >> nt >> hasN >> nn;
  mesh.alloc(nt, nv, hasN != 0, 0);
  for (int i = 0; i < nv; ++i)
  {
    prt::Vec3f p;
    in >> p.x >> p.y >> p.z;
    mesh.setPosition(i, p);
    if (hasN)
    {
      prt::Vec3f n;
      in >> n.x >> n.y >> n.z;
      mesh.setNormal(i, n);
    }
  }
  for (int i = 0; i < nt; ++i)
  {
    int a, b, c, m;
    in >> a >> b >> c >> m;
    mesh.indices[i] = prt::Vec3i(a, b, c);
    mesh.materialIds[i] = m;
  }
  std::string line;
  std::getline(in, line);
  for (int i = 0; i < nn; ++i)
  {
    std::getline(in, line);
    mesh.materialNames.pushBack(line == "-" ? std::string() : line);
  }
  mesh.getBounds();
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 3)
  {
    std::fprintf(stderr, "usage: ref_scene dump FILE | mesh DESC OUT | clone DESC m0..m11\n");
    return 2;
  }
  const std::string mode = argv[1];
  try
  {
    if (mode == "dump") return dump(argv[2]);
    prt::TriangleMesh mesh;
    readDesc(argv[2], mesh);
    if (mode == "mesh" && argc >= 4)
    {
      prt::saveBlob(argv[3], mesh);
      return 0;
    }
    if (mode == "clone" && argc >= 15)
    {
      float m[12];
      for (int i = 0; i < 12; ++i) m[i] = std::stof(argv[3 + i]);
      const prt::Affine3f xf(prt::Vec3f(m[0], m[1], m[2]), prt::Vec3f(m[3], m[4], m[5]), prt::Vec3f(m[6], m[7], m[8]),
                             prt::Vec3f(m[9], m[10], m[11]));
      auto c = std::static_pointer_cast<prt::TriangleMesh>(mesh.clone(xf));
      std::printf("{\"positions\":[");
      for (int i = 0; i < c->getVertexCount(); ++i)
      {
        const prt::Vec3f p = c->getPosition(i);
        std::printf("%s%.9g,%.9g,%.9g", i ? "," : "", p.x, p.y, p.z);
      }
      std::printf("],\"normals\":[");
      if (c->normals)
        for (int i = 0; i < c->getVertexCount(); ++i)
        {
          const prt::Vec3f n = c->getNormal(i);
          std::printf("%s%.9g,%.9g,%.9g", i ? "," : "", n.x, n.y, n.z);
        }
      std::printf("]}\n");
      return 0;
    }
  }
  catch (const std::exception& e)
  {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "bad arguments\n");
  return 2;
}
