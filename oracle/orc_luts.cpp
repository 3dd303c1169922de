// Oracle: energy-compensation LUT precompute.  TEST INFRASTRUCTURE ONLY.
// Restates bsdf/albedo.h:16-120 (stratified 32x32 midpoint quadrature of the GGX and
// dielectric single-scattering albedo) and bsdf/albedo_tables.cpp:23-107 (table layout and
// trapezoid average), and albedo.h:287-330 / albedo_tables.cpp:155-170 for the sheen table.
// Only the tables the stream path's BSDFs read are built: GGX, dielectric (both eta ranges),
// dielectric reflection (the coating's and the thin dielectric's) and sheen.
#include "orc_core.h"

namespace orc {

// albedo.h:18-40
static float ggxAlbedoSample(float cosThetaO, const GGX& m, float sx, float sy)
{
  cosThetaO = std::max(cosThetaO, 1e-6f);
  V3 wo(cos2sin(cosThetaO), 0.f, cosThetaO);
  float whPdf;
  V3 wh = m.sampleVisible(wo, whPdf, sx, sy);
  float cosOH = dot(wo, wh);
  V3 wi = reflect(wo, wh, cosOH);
  if (wi.z <= 0.f) return 0.f;
  return m.G2(wo, wi) * rcpSafe(m.G1(wo));
}

// albedo.h:63-103
static float dielAlbedoSample(float cosThetaO, float eta, const GGX& m, float sx, float sy)
{
  cosThetaO = std::max(cosThetaO, 1e-6f);
  V3 wo(cos2sin(cosThetaO), 0.f, cosThetaO);
  float whPdf;
  V3 wh = m.sampleVisible(wo, whPdf, sx, sy);
  float cosOH = dot(wo, wh);
  float cosTH;
  float F = fresnelDielectricEx(cosOH, cosTH, eta);
  float w = 0.f;
  V3 wi = reflect(wo, wh, cosOH);
  if (wi.z > 0.f) w += F * (m.G2(wo, wi) * rcpSafe(m.G1(wo)));
  wi = refract(wo, wh, cosOH, cosTH, eta);
  if (wi.z < 0.f) w += (1.f - F) * (m.G2(wo, wi) * rcpSafe(m.G1(wo)));
  return w;
}

// albedo.h:124-151 (reflection only)
static float dielReflAlbedoSample(float cosThetaO, float eta, const GGX& m, float sx, float sy)
{
  cosThetaO = std::max(cosThetaO, 1e-6f);
  V3 wo(cos2sin(cosThetaO), 0.f, cosThetaO);
  float whPdf;
  V3 wh = m.sampleVisible(wo, whPdf, sx, sy);
  float cosOH = dot(wo, wh);
  float F = fresnelDielectric(cosOH, eta);
  V3 wi = reflect(wo, wh, cosOH);
  if (wi.z <= 0.f) return 0.f;
  return F * (m.G2(wo, wi) * rcpSafe(m.G1(wo)));
}

template <class F>
static float integrateStratified(F&& f)  // albedo.h:42-58 (n = sqrt(1024) = 32)
{
  const int n = 32;
  float sum = 0.f;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
    {
      float sx = std::min((float(i) + 0.5f) / float(n), 1.f - 1e-6f);
      float sy = std::min((float(j) + 0.5f) / float(n), 1.f - 1e-6f);
      sum += f(sx, sy);
    }
  return std::min(sum / float(n * n), 1.f);
}

static float integrateAvg(const float* f, int size)  // albedo_tables.cpp:23-36
{
  const int n = 1024 + 1;
  float sum = 0.f;
  for (int i = 0; i < n; ++i)
  {
    const float c = float(i) / float(n - 1);
    sum += interp1D(c * float(size - 1), f, size) * c * ((i == 0 || i == n - 1) ? 0.5f : 1.f);
  }
  return std::min(2.f * (sum / float(n - 1)), 1.f);
}

template <bool kReflOnly>
static void dielTable(float minEta, float maxEta, float* dir, float* avg)  // albedo_tables.cpp:61-92
{
  const int size = 16;
  for (int k = 0; k < size; k++)
  {
    const float roughness = float(k) / float(size - 1);
    GGX m(roughnessToAlpha(roughness));
    for (int j = 0; j < size; j++)
    {
      const float eta = lerpf(minEta, maxEta, float(j) / float(size - 1));
      float* d = dir + (k * size + j) * size;
      for (int i = 0; i < size; ++i)
      {
        const float c = float(i) / float(size - 1);
        d[i] = integrateStratified([&](float sx, float sy) {
          return kReflOnly ? dielReflAlbedoSample(c, eta, m, sx, sy) : dielAlbedoSample(c, eta, m, sx, sy);
        });
      }
      avg[k * size + j] = integrateAvg(d, size);
    }
  }
}

// albedo.h:287-330: the sheen lobe's albedo, uniform hemisphere samples
static float sheenAlbedoSample(float cosThetaO, const SheenDist& m, float sx, float sy)
{
  cosThetaO = std::max(cosThetaO, 1e-6f);
  V3 wo(cos2sin(cosThetaO), 0.f, cosThetaO);
  float phi = 2.f * kPi * sx;
  V3 wi(std::cos(phi) * cos2sin(sy), std::sin(phi) * cos2sin(sy), sy);
  V3 wh = normalize(wo + wi);
  return m.D(wh.z) * m.G2(cosThetaO, wi.z) * (1.0f / (4.f * cosThetaO * (1.f / (2.f * kPi))));
}

void precomputeLuts(Luts& L)
{
  // albedo_tables.cpp:155-170
  for (int j = 0; j < 16; j++)
  {
    SheenDist m{roughnessToAlpha(float(j) / 15.f)};
    for (int i = 0; i < 16; ++i)
    {
      const float c = float(i) / 15.f;
      L.sheenDir[j * 16 + i] = integrateStratified([&](float sx, float sy) { return sheenAlbedoSample(c, m, sx, sy); });
    }
  }
  // albedo_tables.cpp:38-59
  for (int j = 0; j < 32; j++)
  {
    const float roughness = float(j) / 31.f;
    GGX m(roughnessToAlpha(roughness));
    float* d = L.ggxDir + j * 32;
    for (int i = 0; i < 32; ++i)
    {
      const float c = float(i) / 31.f;
      d[i] = integrateStratified([&](float sx, float sy) { return ggxAlbedoSample(c, m, sx, sy); });
    }
    L.ggxAvg[j] = integrateAvg(d, 32);
  }
  // albedo_tables.cpp:94-107 (rcp(maxIor) = 1/3, rcp(minIor) = 1)
  dielTable<false>(1.0f / 3.0f, 1.0f, L.dielEtaDir, L.dielEtaAvg);
  dielTable<false>(1.0f, 3.0f, L.dielRcpEtaDir, L.dielRcpEtaAvg);
  // albedo_tables.cpp:109-116
  dielTable<true>(1.0f / 3.0f, 1.0f, L.dielReflDir, L.dielReflAvg);
}

}  // namespace orc
