// Analytic parameter derivatives of the BSDF values the path-replay adjoint needs.  For a hit's
// BSDF b and directions (wo, wi), bsdfEvalGrad returns the value bsdfEval returns (f * cos)
// and its partial derivatives with respect to the material's albedo (per colour channel: the
// diffuse base colour) and its GGX alpha (alpha = roughness^2).  Differentiated lobes:
//   diffuse (bsdf/diffuse.h:55-68): value = rho / pi * cos, or with a diffuse roughness the
//     Oren-Nayar value (bsdf/rough_diffuse.h:65-75), also rho times a factor of the directions.
//   coated diffuse (bsdf/rough_coating.h:297-378 over bsdf/diffuse.h): value =
//     F s D G2 / (4 cos_o) [cos_i > 0]  +  rho / pi * cos_i * absorb * (1 - Ro)(1 - Ri) / (1 - Ravg)
//     with Ro, Ri, Ravg from the dielectric reflection albedo tables (linear in roughness
//     between table rows), D = alpha^2 / (pi (alpha^2 z^2 + x^2 + y^2)^2) and
//     G2 = 1 / (1 + L(wo) + L(wi)), L(w) = (sqrt(1 + alpha^2 t^2) - 1) / 2, t^2 = (x^2 + y^2) / z^2.
//   rough conductor (bsdf/rough_conductor.h:132-202 with optics.h:214-261): value =
//     F(cos_oh; n(r, g), k(r, g)) D G2 / (4 cos_o) + fmsScale (1 - Eo)(1 - Ei) / (pi (1 - Eavg)) cos_i,
//     fmsScale = Favg^2 Eavg / (1 - Favg (1 - Eavg)); albedo = the reflectivity r (Gulbrandsen's
//     mapping to n, k is differentiated by a central difference of width 2e-3 in r, the Favg fit
//     analytically), alpha through D, G2 and the albedo tables Eo, Ei, Eavg.
//   rough dielectric (bsdf/rough_dielectric.h:296-526): alpha only (no albedo), through D, G2 of
//     the single-scattering lobe and the tables and fmsRatio of the multiple-scattering lobe.
//   thin dielectric: not differentiated (zero derivatives).
//   a sheen layer attenuates the lobe's derivatives by its T = min(1 - Ro, 1 - Ri); its own
//     parameters are not differentiated.
// Table derivatives are those of the linear interpolation in roughness (zero where clamped).
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

// Derivative of interp1D (grid units).
PRT_DEV float interp1DdX(float x, const float* f, int size)
{
  if (!(x > 0.f && x < float(size - 1))) return 0.f;
  const int x0 = min(int(x), size - 1), x1 = min(x0 + 1, size - 1);
  return f[x1] - f[x0];
}

// d/droughness of the dielectric albedo tables (the lookups of dielAlbedo and of the averages).
PRT_DEV float dielAlbedoDr(const float* lut, float cosO, float eta, float r)
{
  if (eta <= 1.f)
    return 15.f * interp3DdZ(cosO * 15.f, (eta - (1.0f / 3.0f)) / (1.0f - (1.0f / 3.0f)) * 15.f, r * 15.f,
                             lut + PRT_LUT_DIEL_ETA_DIR, 16);
  return 15.f * interp3DdZ(cosO * 15.f, (eta - 1.f) / 2.f * 15.f, r * 15.f, lut + PRT_LUT_DIEL_RCPETA_DIR, 16);
}
PRT_DEV float dielAvgDr(const float* lut, float eta, float r)
{
  if (eta <= 1.f)
    return 15.f * interp2DdY((eta - (1.0f / 3.0f)) / (1.0f - (1.0f / 3.0f)) * 15.f, r * 15.f, lut + PRT_LUT_DIEL_ETA_AVG, 16);
  return 15.f * interp2DdY((eta - 1.f) / 2.f * 15.f, r * 15.f, lut + PRT_LUT_DIEL_RCPETA_AVG, 16);
}

// Conductor Fresnel through the artistic mapping (optics.h:214-226).
PRT_DEV float artisticFresnel(float cosI, float r, float g)
{
  r = fminf(r, 0.99f);
  const float nMin = (1.f - r) / (1.f + r), nMax = (1.f + sqrtf(r)) / (1.f - sqrtf(r));
  const float n = g * nMin + (1.f - g) * nMax;
  const float k2 = (sqr(n + 1.f) * r - sqr(n - 1.f)) / (1.f - r);
  return fresnelConductor(cosI, n, sqrtf(fmaxf(k2, 0.f)));
}
// d Favg / d r of the Kulla-Conty fit (optics.h:228-240).
PRT_DEV float artisticAvgDr(float r, float g)
{
  return 0.782654f - 2.f * 0.136432f * r + 3.f * 0.278708f * r * r + 0.19744f * g + 0.0360605f * g * g -
         2.f * 0.2586f * g * r;
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

PRT_DEV BsdfGrad conductorEvalGrad(const float* lut, const Bsdf& b, const prt_material_constants& m, f3 wo, f3 wi)
{
  BsdfGrad g;
  g.value = g.dAlbedo = g.dAlpha = f3(0.f);
  const float cosO = dot(wo, b.fr.N), cosI = dot(wi, b.fr.N);
  if (!(cosO > 0.f && cosI > 0.f)) return g;
  const f3 wh = normalize(wi + wo);
  const float cosOH = dot(wo, wh);
  const f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  const float r = b.roughness, a = b.ggx.a;
  const float D = b.ggx.D(wh0), G = b.ggx.G2(wo0, wi0);
  const float dD = a > 0.001f ? ggxDdAlpha(a, wh0) : 0.f;
  const float dG = a > 0.001f ? -G * G * (ggxLambdaDAlpha(a, wo0) + ggxLambdaDAlpha(a, wi0)) : 0.f;
  const float kF = D * G / (4.f * cosO), dkF = (dD * G + D * dG) / (4.f * cosO);
  const float Eo = ggxAlbedo(lut, cosO, r), Ei = ggxAlbedo(lut, cosI, r);
  const float dEo = 31.f * interp2DdY(cosO * 31.f, r * 31.f, lut + PRT_LUT_GGX_DIR, 32);
  const float dEi = 31.f * interp2DdY(cosI * 31.f, r * 31.f, lut + PRT_LUT_GGX_DIR, 32);
  const float dEavg = 31.f * interp1DdX(r * 31.f, lut + PRT_LUT_GGX_AVG, 32);
  const float E = b.Eavg, A = (1.f - Eo) * (1.f - Ei), B = 1.0f / (1.f - E);
  const float dA = -dEo * (1.f - Ei) - (1.f - Eo) * dEi, dB = dEavg * B * B;
  const float drda = (a > 0.001f && r > 0.f) ? 0.5f / r : 0.f;  // alpha = roughness^2 above its floor
  const float w = cosI * (1.f / kPi);
  const float F[3] = {fresnelConductor(cosOH, b.c0.x, b.c1.x), fresnelConductor(cosOH, b.c0.y, b.c1.y),
                      fresnelConductor(cosOH, b.c0.z, b.c1.z)};
  const float S[3] = {b.c2.x, b.c2.y, b.c2.z}, rr[3] = {b.rr.x, b.rr.y, b.rr.z};
  float v[3], dAl[3], dAb[3];
  for (int j = 0; j < 3; ++j)
  {
    const float Fa = m.Favg[j], den = 1.f - Fa * (1.f - E);
    const float dSdE = Fa * Fa * (1.f - Fa) / (den * den);
    const float dSdF = (2.f * Fa * E * den + Fa * Fa * E * (1.f - E)) / (den * den);
    v[j] = F[j] * kF + S[j] * A * B * w;
    dAl[j] = F[j] * dkF + (dSdE * dEavg * A * B + S[j] * dA * B + S[j] * A * dB) * w * drda;
    // reflectivity: the Fresnel term through n(r), k(r) (clamped at 0.99), and Favg(r)
    float dF = 0.f;
    if (rr[j] < 0.99f)
    {
      const float lo = fmaxf(rr[j] - 1e-3f, 0.f), hi = fminf(rr[j] + 1e-3f, 0.99f);
      dF = (artisticFresnel(cosOH, hi, m.edge[j]) - artisticFresnel(cosOH, lo, m.edge[j])) / (hi - lo);
    }
    dAb[j] = dF * kF + dSdF * artisticAvgDr(rr[j], m.edge[j]) * A * B * w;
  }
  g.value = f3(v[0], v[1], v[2]);
  g.dAlpha = f3(dAl[0], dAl[1], dAl[2]);
  g.dAlbedo = f3(dAb[0], dAb[1], dAb[2]);
  return g;
}

PRT_DEV BsdfGrad dielectricEvalGrad(const float* lut, const Bsdf& b, const prt_material_constants& m, const Ctx& c, f3 wo, f3 wi)
{
  BsdfGrad g;
  g.value = g.dAlbedo = g.dAlpha = f3(0.f);
  const float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return g;
  const float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return g;
  const float r = b.roughness, a = b.ggx.a;
  const float drda = (a > 0.001f && r > 0.f) ? 0.5f / r : 0.f;
  // single scattering: a factor times D G2 (dielSingle)
  float sPdf;
  const float sValue = dielSingle(b, c, wo, wi, sPdf);
  const bool refl = cosI > 0.f;
  f3 wh = normalize(refl ? wi + wo : b.eta * wo + wi);
  if (dot(wh, b.fr.N) < 0.f) wh = -wh;
  const f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  const float D = b.ggx.D(wh0), G = b.ggx.G2(wo0, wi0);
  const float dD = a > 0.001f ? ggxDdAlpha(a, wh0) : 0.f;
  const float dG = a > 0.001f ? -G * G * (ggxLambdaDAlpha(a, wo0) + ggxLambdaDAlpha(a, wi0)) : 0.f;
  const float dSingle = D * G > 0.f ? sValue / (D * G) * (dD * G + D * dG) : 0.f;
  // multiple scattering (dielFms) with fmsRatio(EavgEta, EavgRcpEta) (rough_dielectric.h:279-291)
  const int o = (b.eta == m.eta[1] && m.eta[1] != m.eta[0]) ? 1 : 0;
  const float Fe = m.FavgEta[o], Fr = m.FavgRcpEta[o];
  const float Ee = b.EavgEta, Er = b.EavgRcpEta;
  const float dEe = dielAvgDr(lut, b.eta, r), dEr = dielAvgDr(lut, 1.0f / b.eta, r);
  const float aa = (1.f - Fe) / (1.f - Er), bb = (1.f - Fr) / (1.f - Ee);
  const float daa = (1.f - Fe) / sqr(1.f - Er) * dEr, dbb = (1.f - Fr) / sqr(1.f - Ee) * dEe;
  const float dRatio = -(1.f - Fe) * (dbb * aa - bb * daa) / sqr(aa + bb);
  const float Eo = dielAlbedo(lut, cosO, b.eta, r), dEo = dielAlbedoDr(lut, cosO, b.eta, r);
  float fms, dFms;
  if (refl)
  {
    const float Ei = dielAlbedo(lut, cosI, b.eta, r), dEi = dielAlbedoDr(lut, cosI, b.eta, r);
    const float A = (1.f - Eo) * (1.f - Ei), dA = -dEo * (1.f - Ei) - (1.f - Eo) * dEi;
    const float B = 1.0f / (1.f - Ee), dB = dEe * B * B, w = cosI * (1.f / kPi);
    fms = b.fmsRatio * A * B * w;
    dFms = (dRatio * A * B + b.fmsRatio * dA * B + b.fmsRatio * A * dB) * w;
  }
  else
  {
    const float ac = fabsf(cosI);
    const float Ei = dielAlbedo(lut, ac, 1.0f / b.eta, r), dEi = dielAlbedoDr(lut, ac, 1.0f / b.eta, r);
    const float A = (1.f - Eo) * (1.f - Ei), dA = -dEo * (1.f - Ei) - (1.f - Eo) * dEi;
    const float B = 1.0f / (1.f - Er), dB = dEr * B * B, w = ac * (1.f / kPi);
    fms = (1.f - b.fmsRatio) * A * B * w;
    dFms = (-dRatio * A * B + (1.f - b.fmsRatio) * dA * B + (1.f - b.fmsRatio) * A * dB) * w;
  }
  g.value = f3(sValue + fms);
  g.dAlpha = f3(dSingle + dFms * drda);
  return g;
}

PRT_DEV BsdfGrad lobeEvalGrad(const float* lut, const Bsdf& b, const prt_material_constants& m, const Ctx& c, f3 wo, f3 wi)
{
  BsdfGrad g;
  g.value = g.dAlbedo = g.dAlpha = f3(0.f);
  if (!(b.importance > 0.f)) return g;
  // the diffuse base per unit albedo: Lambert, or Oren-Nayar (linear in the albedo too)
  Bsdf unitBase = b;
  unitBase.c0 = f3(1.f);
  float basePdf;
  const float kBase = diffuseEval(unitBase, wo, wi, basePdf).x;
  if (b.kind == PRT_KIND_DIFFUSE)
  {
    g.value = b.c0 * kBase;
    g.dAlbedo = f3(kBase);
    return g;
  }
  if (b.kind == PRT_KIND_CONDUCTOR) return conductorEvalGrad(lut, b, m, wo, wi);
  if (b.kind == PRT_KIND_DIELECTRIC) return dielectricEvalGrad(lut, b, m, c, wo, wi);
  if (b.kind != PRT_KIND_COATED) return g;  // thin dielectric: not differentiated
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
  const f3 unit = absorb * kBase;
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

// Under a sheen layer (bsdf/sheen.h:218-279) the lobe's value and derivatives are attenuated by
// T = min(1 - Ro, 1 - Ri) and the sheen reflection (not differentiated) joins the value.
PRT_DEV BsdfGrad bsdfEvalGrad(const float* lut, const Bsdf& b, const prt_material_constants& m, const Ctx& c, f3 wo, f3 wi)
{
  BsdfGrad g = lobeEvalGrad(lut, b, m, c, wo, wi);
  if (!b.sheen) return g;
  const float cosO = dot(wo, b.fr.N), cosI = dot(wi, b.fr.N);
  if (!(cosO > 0.f) || !(cosI * dot(wi, c.Ng) > 0.f))
  {
    g.value = g.dAlbedo = g.dAlpha = f3(0.f);
    return g;
  }
  const float Ro = sheenAlbedo(lut, cosO, b.sheenRough) * b.sheenScale;
  const float Ri = sheenAlbedo(lut, fabsf(cosI), b.sheenRough) * b.sheenScale;
  const float T = fminf(1.f - Ro, 1.f - Ri);
  g.value = g.value * T;
  if (cosI > 0.f) g.value = g.value + sheenReflection(b, wo, wi, cosO, cosI);
  g.dAlbedo = g.dAlbedo * T;
  g.dAlpha = g.dAlpha * T;
  return g;
}

}  // namespace prt
