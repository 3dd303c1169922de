// Analytic parameter derivatives of the BSDF values the path-replay adjoint needs.  For a hit's
// BSDF b and directions (wo, wi), bsdfEvalGrad returns the value bsdfEval returns (f * cos)
// and its partial derivatives with respect to the material's albedo (per colour channel: the
// diffuse base colour) and its GGX alpha (alpha = roughness^2).  Differentiated lobes:
//   diffuse (bsdf/diffuse.h:55-68): value = rho / pi * cos.
//   coated diffuse (bsdf/rough_coating.h:297-378 over bsdf/diffuse.h): value =
//     F s D G2 / (4 cos_o) [cos_i > 0]  +  rho / pi * cos_i * absorb * (1 - Ro)(1 - Ri) / (1 - Ravg)
//     with Ro, Ri, Ravg from the dielectric reflection albedo tables (linear in roughness
//     between table rows), D = alpha^2 / (pi (alpha^2 z^2 + x^2 + y^2)^2) and
//     G2 = 1 / (1 + L(wo) + L(wi)), L(w) = (sqrt(1 + alpha^2 t^2) - 1) / 2, t^2 = (x^2 + y^2) / z^2.
// The other lobes have no differentiated parameters (zero derivatives).
#pragma once

#include "dev_math.h"

namespace prt {

struct BsdfGrad
{
  f3 value;    // f * cos, as bsdfEval
  f3 dAlbedo;  // d value_c / d albedo_c
  f3 dAlpha;   // d value / d alpha
};

// Derivative of interp3D along its third coordinate (grid units); zero where the coordinate is
// clamped to the table.
PRT_DEV float interp3DdZ(float px, float py, float pz, const float* f, int n)
{
  if (!(pz > 0.f && pz < float(n - 1))) return 0.f;
  float xc = clampf(px, 0.f, float(n - 1)), yc = clampf(py, 0.f, float(n - 1));
  float sx = xc - floorf(xc), sy = yc - floorf(yc);
  int x0 = min(int(xc), n - 1), x1 = min(x0 + 1, n - 1);
  int y0 = min(int(yc), n - 1), y1 = min(y0 + 1, n - 1);
  int z0 = min(int(pz), n - 1), z1 = min(z0 + 1, n - 1);
  int ny = n, nz = n * n;
  float f00 = lerpf(f[x0 + y0 * ny + z0 * nz], f[x1 + y0 * ny + z0 * nz], sx);
  float f01 = lerpf(f[x0 + y1 * ny + z0 * nz], f[x1 + y1 * ny + z0 * nz], sx);
  float f10 = lerpf(f[x0 + y0 * ny + z1 * nz], f[x1 + y0 * ny + z1 * nz], sx);
  float f11 = lerpf(f[x0 + y1 * ny + z1 * nz], f[x1 + y1 * ny + z1 * nz], sx);
  return lerpf(f10, f11, sy) - lerpf(f00, f01, sy);
}
// Derivative of interp2D along its second coordinate (grid units).
PRT_DEV float interp2DdY(float px, float py, const float* f, int n)
{
  if (!(py > 0.f && py < float(n - 1))) return 0.f;
  float xc = clampf(px, 0.f, float(n - 1));
  float sx = xc - floorf(xc);
  int x0 = min(int(xc), n - 1), x1 = min(x0 + 1, n - 1);
  int y0 = min(int(py), n - 1), y1 = min(y0 + 1, n - 1);
  return lerpf(f[x0 + y1 * n], f[x1 + y1 * n], sx) - lerpf(f[x0 + y0 * n], f[x1 + y0 * n], sx);
}

// GGX distribution and Smith masking derivatives with respect to alpha.
PRT_DEV float ggxDdAlpha(float a, f3 wh)
{
  float s = wh.x * wh.x + wh.y * wh.y, z2 = wh.z * wh.z;
  float q = a * a * z2 + s;
  return (2.f * a / (kPi * q * q)) * (1.f - 2.f * a * a * z2 / q);
}
PRT_DEV float ggxLambdaDAlpha(float a, f3 w)
{
  float t2 = (w.x * w.x + w.y * w.y) / (w.z * w.z);
  return 0.5f * a * t2 / sqrtf(1.f + a * a * t2);
}

PRT_DEV BsdfGrad bsdfEvalGrad(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi)
{
  BsdfGrad g;
  g.value = g.dAlbedo = g.dAlpha = f3(0.f);
  if (!(b.importance > 0.f)) return g;
  if (b.kind == PRT_KIND_DIFFUSE)
  {
    float cosI = fmaxf(dot(wi, b.fr.N), 0.f);
    g.value = b.c0 * (1.f / kPi) * cosI;
    g.dAlbedo = f3((1.f / kPi) * cosI);
    return g;
  }
  if (b.kind != PRT_KIND_COATED)
  {
    float pdf;
    g.value = bsdfEval(lut, b, c, wo, wi, pdf);
    return g;
  }
  const float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return g;
  const float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return g;
  const float scale = b.c2.y, r = b.roughness, a = b.ggx.a;
  // absorption along the refracted path (rough_coating.h:319-327)
  float length = 1.0f / refractCos(cosO, b.eta) + 1.0f / refractCos(fabsf(cosI), b.eta);
  if (cosI <= 0.f) length *= 0.5f;
  const float t = b.c2.x * length;
  f3 absorb(1.f + scale * (powf(b.c1.x, t) - 1.f), 1.f + scale * (powf(b.c1.y, t) - 1.f), 1.f + scale * (powf(b.c1.z, t) - 1.f));
  // energy conservation T(roughness) and its slope
  const float e = (b.eta - (1.0f / 3.0f)) / (1.0f - (1.0f / 3.0f));
  const float* dir = lut + PRT_LUT_DIEL_REFL_DIR;
  const float Ro = dielReflAlbedo(lut, cosO, b.eta, r) * scale;
  const float Ri = dielReflAlbedo(lut, fabsf(cosI), b.eta, r) * scale;
  const float dRo = 15.f * scale * interp3DdZ(cosO * 15.f, e * 15.f, r * 15.f, dir, 16);
  const float dRi = 15.f * scale * interp3DdZ(fabsf(cosI) * 15.f, e * 15.f, r * 15.f, dir, 16);
  const float dRavg = 15.f * scale * interp2DdY(e * 15.f, r * 15.f, lut + PRT_LUT_DIEL_REFL_AVG, 16);
  const float inv = 1.0f / (1.f - b.Eavg);
  const float T = (1.f - Ro) * (1.f - Ri) * inv;
  const float dTdr = (-dRo * (1.f - Ri) - (1.f - Ro) * dRi) * inv + (1.f - Ro) * (1.f - Ri) * dRavg * inv * inv;
  // alpha = roughness^2 (alpha above its 0.001 floor): d/dalpha = d/droughness / (2 roughness)
  const float drda = (a > 0.001f && r > 0.f) ? 0.5f / r : 0.f;
  const f3 unit = absorb * ((1.f / kPi) * fmaxf(cosI, 0.f));
  g.value = b.c0 * unit * T;
  g.dAlbedo = unit * T;
  g.dAlpha = b.c0 * unit * (dTdr * drda);
  if (cosI > 0.f)
  {
    const f3 wh = normalize(wo + wi);
    const float cosOH = dot(wo, wh);
    const float F = fresnelDielectric(cosOH, b.eta) * scale;
    const f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
    const float D = b.ggx.D(wh0), G = b.ggx.G2(wo0, wi0);
    const float dD = a > 0.001f ? ggxDdAlpha(a, wh0) : 0.f;
    const float dG = a > 0.001f ? -G * G * (ggxLambdaDAlpha(a, wo0) + ggxLambdaDAlpha(a, wi0)) : 0.f;
    const float k = F * (1.0f / (4.f * cosO));
    g.value = g.value + f3(k * D * G);
    g.dAlpha = g.dAlpha + f3(k * (dD * G + D * dG));
  }
  return g;
}

}  // namespace prt
