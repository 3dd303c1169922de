// Device-side math of the stream path tracer for gfx950: vectors, the Sobol-Burley sampler,
// sample warps, discrete distributions, GGX / Fresnel, the energy-compensation LUT lookups,
// the three BSDFs of the stream path and the environment lights.  Every function follows
// the reference file:line it cites; one GPU lane evaluates what one SIMD lane of the
// reference evaluates.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "prt_types.h"

#define PRT_DEV __device__ __forceinline__

namespace prt {

// Loads and stores through a pointer known to address device (global) memory.  The kernels
// receive their buffers inside a struct in device memory, so the compiler cannot infer the
// address space of those pointers and would use flat instructions, which also count against
// the LDS counter (every wait on a node load then waits on the traversal stack as well) and
// take 64-bit VGPR addresses.
template <class T>
PRT_DEV T ldg(const T* p)
{
  return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
PRT_DEV void stg(T* p, const T& v)
{
  *(__attribute__((address_space(1))) T*)p = v;
}
// float4 / int4 go through the native vector types (copying HIP's vector structs would pass a
// generic reference to their copy constructor).
typedef float prtV4f __attribute__((ext_vector_type(4)));
typedef int prtV4i __attribute__((ext_vector_type(4)));
template <>
PRT_DEV float4 ldg<float4>(const float4* p)
{
  const prtV4f v = *(const __attribute__((address_space(1))) prtV4f*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
template <>
PRT_DEV int4 ldg<int4>(const int4* p)
{
  const prtV4i v = *(const __attribute__((address_space(1))) prtV4i*)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
template <>
PRT_DEV void stg<float4>(float4* p, const float4& v)
{
  *(__attribute__((address_space(1))) prtV4f*)p = prtV4f{v.x, v.y, v.z, v.w};
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPosMax = 3.402823466e38f;

struct f3
{
  float x, y, z;
  PRT_DEV f3() {}
  PRT_DEV f3(float a) : x(a), y(a), z(a) {}
  PRT_DEV f3(float a, float b, float c) : x(a), y(b), z(c) {}
};
PRT_DEV f3 operator+(f3 a, f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
PRT_DEV f3 operator-(f3 a, f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
PRT_DEV f3 operator-(f3 a) { return f3(-a.x, -a.y, -a.z); }
PRT_DEV f3 operator*(f3 a, f3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
PRT_DEV f3 operator*(f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
PRT_DEV f3 operator*(float s, f3 a) { return f3(a.x * s, a.y * s, a.z * s); }
PRT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PRT_DEV f3 cross(f3 a, f3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
PRT_DEV float lengthSqr(f3 a) { return dot(a, a); }
PRT_DEV float length(f3 a) { return sqrtf(dot(a, a)); }
PRT_DEV f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }  // math/vec3.h:399-402
PRT_DEV float reduceMax(f3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
PRT_DEV float sqr(float a) { return a * a; }
PRT_DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
PRT_DEV float lerpf(float a, float b, float s) { return a + s * (b - a); }          // math/math.h:89-92
PRT_DEV float rcpSafe(float x) { return x == 0.f ? kPosMax : 1.0f / x; }            // math/simd.h:107-111
PRT_DEV float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }  // image/color.h:17-22
PRT_DEV float cos2sin(float x) { return sqrtf(fmaxf(1.f - x * x, 0.f)); }           // math/math.h:57-66

struct Frame { f3 U, V, N; };
PRT_DEV f3 toGlobal(const Frame& b, f3 a) { return a.x * b.U + a.y * b.V + a.z * b.N; }           // math/basis3.h:72-75
PRT_DEV f3 toLocal(const Frame& b, f3 a) { return f3(dot(a, b.U), dot(a, b.V), dot(a, b.N)); }    // math/basis3.h:77-80

// math/basis3.h:173-180
PRT_DEV void makeFrame(f3& U, f3& V, f3 N)
{
  f3 U0(0.f, N.z, -N.y), U1(-N.z, 0.f, N.x);
  U = normalize(lengthSqr(U0) > lengthSqr(U1) ? U0 : U1);
  V = normalize(cross(N, U));
}

// ---------------------------------------------------------------------------------------
// Sobol-Burley sampler (sampling/sobol_burley.h, sobol_burley_sampler_simd.h) -- bit-exact
// ---------------------------------------------------------------------------------------
PRT_DEV uint32_t reverseBits(uint32_t x) { return __builtin_bitreverse32(x); }  // math/simd.h:140-147
PRT_DEV uint32_t hashHP(uint32_t i)  // sobol_burley.h:31-44
{
  i ^= i >> 16; i *= 0x21f0aaadu; i ^= i >> 15; i *= 0xd35a2d97u; i ^= i >> 15;
  return i ^ 0xe6fe3bebu;
}
PRT_DEV uint32_t reversedBitOwen(uint32_t n, uint32_t seed)  // sobol_burley.h:52-62
{
  n ^= n * 0x3d20adeau; n += seed; n *= (seed >> 16) | 1u; n ^= n * 0x05526c56u; n ^= n * 0x53a22864u;
  return n;
}
// math/simd.h:76-84, with the default AVX2 host build's two-step uint32 -> float conversion
// (math/simd/vfloat8_avx2.h:171-178): low 31 bits converted, then 2^31 added.
PRT_DEV float toFloatUnormExcl(uint32_t x)
{
  float lo = (float)(int32_t)(x & 0x7fffffffu);
  float hi = (x & 0x80000000u) ? 2147483648.0f : 0.0f;
  return (lo + hi) * (1.0f / 4294967808.0f);
}
// Sobol dimension 2 direction numbers with reversed bits (the third column of
// sampling/sobol_burley.cpp:9-51), generated at compile time from the primitive polynomial
// x^2 + x + 1 with initial m = {1, 3}: v_k = v_{k-1} ^ v_{k-2} ^ (v_{k-2} >> 2), MSB-first,
// then bit-reversed.
struct SobolColumn { uint32_t v[32]; };
constexpr SobolColumn makeSobolDim2()
{
  SobolColumn c{};
  uint32_t v[33] = {};
  v[1] = 1u << 31;
  v[2] = 3u << 30;
  for (int k = 3; k <= 32; ++k) v[k] = v[k - 1] ^ v[k - 2] ^ (v[k - 2] >> 2);
  for (int k = 1; k <= 32; ++k)
  {
    uint32_t r = 0;
    for (int b = 0; b < 32; ++b) r |= ((v[k] >> b) & 1u) << (31 - b);
    c.v[k - 1] = r;
  }
  return c;
}
__constant__ static const SobolColumn kSobolDim2 = makeSobolDim2();

PRT_DEV uint32_t sobolDim1(uint32_t v)  // sobol_burley.h:117-128 (Ahmed 2024)
{
  v ^= v << 16;
  v ^= (v & 0x00FF00FFu) << 8;
  v ^= (v & 0x0F0F0F0Fu) << 4;
  v ^= (v & 0x33333333u) << 2;
  v ^= (v & 0x55555555u) << 1;
  return reverseBits(v);
}
PRT_DEV uint32_t sobolDim2(uint32_t rev)  // sobol_burley.h:129-143: XOR of columns for set bits
{
  uint32_t r = 0;
  // bit k of the reversed index (counting from the MSB) selects column k
  while (rev)
  {
    int j = __builtin_clz(rev);
    r ^= kSobolDim2.v[j];
    rev &= ~(0x80000000u >> j);
  }
  return r;
}
PRT_DEV float owenFinish(uint32_t r, uint32_t scramble) { return toFloatUnormExcl(reverseBits(reversedBitOwen(r, scramble))); }

struct Sampler
{
  uint32_t index, scramble;
  // setSample (sobol_burley_sampler_simd.h:36-62, math/hash.h:12-42)
  PRT_DEV void set(uint32_t seed, uint32_t sampleIndex, uint32_t pixelId)
  {
    index = sampleIndex;
    uint32_t k = pixelId * 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    uint32_t h = seed ^ k;
    h = ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    scramble = h;
  }
  PRT_DEV float get1D(uint32_t dim) const  // sobol_burley.h:222-238
  {
    uint32_t s = scramble ^ hashHP(dim);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xbff95bfeu);
    return owenFinish(reverseBits(i), s ^ 0x635c77bdu);
  }
  PRT_DEV void get2D(uint32_t set, float& a, float& b) const  // sobol_burley.h:250-267
  {
    uint32_t s = scramble ^ hashHP(set);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xf8ade99au);
    a = owenFinish(reverseBits(i), s ^ 0xe0aaaf76u);
    b = owenFinish(sobolDim1(i), s ^ 0x94964d4eu);
  }
  PRT_DEV void get3D(uint32_t set, float& a, float& b, float& c) const  // sobol_burley.h:279-297
  {
    uint32_t s = scramble ^ hashHP(set);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xcaa726acu);
    a = owenFinish(reverseBits(i), s ^ 0x9e78e391u);
    b = owenFinish(sobolDim1(i), s ^ 0x67c33241u);
    c = owenFinish(sobolDim2(i), s ^ 0x78c395c5u);
  }
};

// ---------------------------------------------------------------------------------------
// Warps (sampling/shape_sampler.h)
// ---------------------------------------------------------------------------------------
PRT_DEV f3 cosineSampleHemisphere(float& pdf, float sx, float sy)  // :69-86
{
  float ct = sqrtf(sx), st = sqrtf(1.f - sx);
  float sp, cp;
  sincosf(2.f * kPi * sy, &sp, &cp);
  pdf = ct * (1.f / kPi);
  return f3(cp * st, sp * st, ct);
}
PRT_DEV f3 uniformSampleSphere(float& pdf, float sx, float sy)  // :23-40
{
  float ct = 1.f - 2.f * sx, st = 2.f * sqrtf(sx * (1.f - sx));
  float sp, cp;
  sincosf(2.f * kPi * sy, &sp, &cp);
  pdf = 1.f / (4.f * kPi);
  return f3(cp * st, sp * st, ct);
}
PRT_DEV f3 uniformSampleCone(float& pdf, float sx, float sy, float cosAngle, float conePdf)  // :100-117
{
  float ct = 1.f - sx * (1.f - cosAngle);
  float st = cos2sin(ct);
  float sp, cp;
  sincosf(2.f * kPi * sy, &sp, &cp);
  pdf = conePdf;
  return f3(cp * st, sp * st, ct);
}

// ---------------------------------------------------------------------------------------
// Discrete distributions (sampling/distribution1d.h:76-101, distribution1d.cpp:35-55)
// ---------------------------------------------------------------------------------------
PRT_DEV int dist1dLookup(const float* cdf, int size, float s)
{
  int first = 0, len = size;
  while (len > 0)
  {
    int half = len >> 1, mid = first + half;
    if (s < cdf[mid]) len = half;
    else { first = mid + 1; len = len - half - 1; }
  }
  return first - 1 > 0 ? first - 1 : 0;
}
PRT_DEV int dist1dSampleReuse(const float* cdf, int size, float& pdf, float& s)
{
  int i = dist1dLookup(cdf, size, s);
  float c0 = cdf[i], c1 = cdf[i + 1];
  pdf = c1 - c0;
  s = (s - c0) * (1.0f / pdf);
  return i;
}
PRT_DEV float dist1dPdf(const float* cdf, int size, int i) { return (i < 0 || i >= size) ? 0.f : cdf[i + 1] - cdf[i]; }
// dist1dLookup with a guide table (the cutpoint method): guide[g] is the index the full search
// returns for s = g / PRT_ENV_GUIDE.  The full search's result does not decrease with s, so for
// s in bin g it lies in [guide[g], guide[g + 1]]: the last i there with cdf[i] <= s, found by
// searching cdf[guide[g] + 1 .. guide[g + 1]] only -- the same index in a few steps.
PRT_DEV int dist1dLookupGuided(const float* cdf, const int* guide, float s)
{
  const int g = min(int(s * float(PRT_ENV_GUIDE)), PRT_ENV_GUIDE - 1);
  const int lo = guide[g];
  int first = lo + 1, len = guide[g + 1] - lo;
  while (len > 0)
  {
    int half = len >> 1, mid = first + half;
    if (s < cdf[mid]) len = half;
    else { first = mid + 1; len = len - half - 1; }
  }
  return first - 1;
}
PRT_DEV int dist1dSampleReuseGuided(const float* cdf, const int* guide, float& pdf, float& s)
{
  int i = dist1dLookupGuided(cdf, guide, s);
  float c0 = cdf[i], c1 = cdf[i + 1];
  pdf = c1 - c0;
  s = (s - c0) * (1.0f / pdf);
  return i;
}

// ---------------------------------------------------------------------------------------
// GGX (bsdf/ggx_distribution.h) -- isotropic alpha = max(roughness^2, 0.001)
// ---------------------------------------------------------------------------------------
struct GGX
{
  float a;
  PRT_DEV float D(f3 wh) const  // :24-31
  {
    float c2 = sqr(wh.z);
    float e = (sqr(wh.x / a) + sqr(wh.y / a)) / c2;
    return 1.0f / (kPi * a * a * sqr(c2 * (1.f + e)));
  }
  PRT_DEV float lambda(f3 w) const  // :84-90
  {
    float c2 = sqr(w.z);
    float invA2 = (sqr(w.x * a) + sqr(w.y * a)) / c2;
    return 0.5f * (-1.f + sqrtf(1.f + invA2));
  }
  PRT_DEV float G1(f3 w) const { return 1.0f / (1.f + lambda(w)); }
  PRT_DEV float G2(f3 wo, f3 wi) const { return 1.0f / (1.f + lambda(wo) + lambda(wi)); }
  PRT_DEV float evalVisible(f3 wh, f3 wo, float cosOH, float& pdf) const  // :42-48
  {
    float d = D(wh);
    pdf = G1(wo) * fabsf(cosOH) * d / fabsf(wo.z);
    return d;
  }
  PRT_DEV f3 sampleVisible(f3 wo, float& pdf, float sx, float sy) const  // :52-78 (Heitz 2017)
  {
    f3 V = normalize(f3(a * wo.x, a * wo.y, wo.z));
    f3 T1 = (V.z < 0.9999f) ? normalize(cross(V, f3(0.f, 0.f, 1.f))) : f3(1.f, 0.f, 0.f);
    f3 T2 = cross(T1, V);
    float aa = 1.f / (1.f + V.z);
    float r = sqrtf(sx);
    float phi = (sy < aa) ? sy / aa * kPi : kPi + (sy - aa) / (1.f - aa) * kPi;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float P1 = r * cp;
    float P2 = r * sp * ((sy < aa) ? 1.f : V.z);
    f3 wh = P1 * T1 + P2 * T2 + sqrtf(fmaxf(0.f, 1.f - P1 * P1 - P2 * P2)) * V;
    wh = normalize(f3(a * wh.x, a * wh.y, fmaxf(0.f, wh.z)));
    pdf = G1(wo) * fabsf(dot(wo, wh)) * D(wh) / fabsf(wo.z);
    return wh;
  }
};

// bsdf/optics.h
PRT_DEV f3 reflect(f3 I, f3 N) { return 2.f * dot(I, N) * N - I; }
PRT_DEV f3 reflect(f3 I, f3 N, float cosI) { return 2.f * cosI * N - I; }
PRT_DEV f3 refract(f3 I, f3 N, float cosI, float cosT, float eta) { return eta * (cosI * N - I) - cosT * N; }
PRT_DEV float fresnelDielectric(float cosI, float cosT, float eta)  // :61-67
{
  float Rper = (eta * cosI - cosT) / (eta * cosI + cosT);
  float Rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
  return 0.5f * (Rpar * Rpar + Rper * Rper);
}
PRT_DEV float fresnelDielectric(float cosI, float eta)  // :84-99
{
  cosI = fmaxf(cosI, 0.f);
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  return (k >= 0.f) ? fresnelDielectric(cosI, sqrtf(k), eta) : 1.f;
}
PRT_DEV float fresnelDielectricEx(float cosI, float& cosT, float eta)  // :122-139
{
  cosI = fabsf(cosI);
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  if (!(k >= 0.f)) { cosT = 0.f; return 1.f; }
  cosT = sqrtf(k);
  return fresnelDielectric(cosI, cosT, eta);
}
PRT_DEV float fresnelConductor(float cosI, float eta, float k)  // :179-190
{
  cosI = fabsf(cosI);
  float c2 = sqr(cosI), tmp = sqr(eta) + sqr(k), e2 = 2.f * eta * cosI;
  float Rpar = (tmp * c2 - e2 + 1.f) / (tmp * c2 + e2 + 1.f);
  float Rper = (tmp - e2 + c2) / (tmp + e2 + c2);
  return 0.5f * (Rpar + Rper);
}

// ---------------------------------------------------------------------------------------
// LUT lookups (math/interpolation.h:26-227, bsdf/albedo_tables.h:20-154)
// ---------------------------------------------------------------------------------------
PRT_DEV float interp1D(float x, const float* f, int size)
{
  float xc = clampf(x, 0.f, float(size - 1));
  float s = xc - floorf(xc);
  int x0 = min(int(xc), size - 1), x1 = min(x0 + 1, size - 1);
  return lerpf(f[x0], f[x1], s);
}
PRT_DEV float interp2D(float px, float py, const float* f, int n)
{
  float xc = clampf(px, 0.f, float(n - 1)), yc = clampf(py, 0.f, float(n - 1));
  float sx = xc - floorf(xc), sy = yc - floorf(yc);
  int x0 = min(int(xc), n - 1), x1 = min(x0 + 1, n - 1);
  int y0 = min(int(yc), n - 1), y1 = min(y0 + 1, n - 1);
  float f0 = lerpf(f[x0 + y0 * n], f[x1 + y0 * n], sx);
  float f1 = lerpf(f[x0 + y1 * n], f[x1 + y1 * n], sx);
  return lerpf(f0, f1, sy);
}
PRT_DEV float interp3D(float px, float py, float pz, const float* f, int n)
{
  float xc = clampf(px, 0.f, float(n - 1)), yc = clampf(py, 0.f, float(n - 1)), zc = clampf(pz, 0.f, float(n - 1));
  float sx = xc - floorf(xc), sy = yc - floorf(yc), sz = zc - floorf(zc);
  int x0 = min(int(xc), n - 1), x1 = min(x0 + 1, n - 1);
  int y0 = min(int(yc), n - 1), y1 = min(y0 + 1, n - 1);
  int z0 = min(int(zc), n - 1), z1 = min(z0 + 1, n - 1);
  int ny = n, nz = n * n;
  float f00 = lerpf(f[x0 + y0 * ny + z0 * nz], f[x1 + y0 * ny + z0 * nz], sx);
  float f01 = lerpf(f[x0 + y1 * ny + z0 * nz], f[x1 + y1 * ny + z0 * nz], sx);
  float f10 = lerpf(f[x0 + y0 * ny + z1 * nz], f[x1 + y0 * ny + z1 * nz], sx);
  float f11 = lerpf(f[x0 + y1 * ny + z1 * nz], f[x1 + y1 * ny + z1 * nz], sx);
  return lerpf(lerpf(f00, f01, sy), lerpf(f10, f11, sy), sz);
}
PRT_DEV float sheenAlbedo(const float* lut, float cosO, float roughness)  // MicrofacetSheenAlbedoTable::eval
{
  return interp2D(cosO * 15.f, roughness * 15.f, lut + PRT_LUT_SHEEN_DIR, 16);
}
// Sheen microfiber distribution (bsdf/sheen_distribution.h:10-78; Conty and Kulla 2017), r = alpha
struct SheenDist
{
  float r;
  PRT_DEV float D(float cosTheta) const
  {
    float inv = 1.0f / r;
    return (2.f + inv) * powf(cos2sin(cosTheta), inv) * (1.f / (2.f * kPi));
  }
  PRT_DEV float L(float x) const  // the fitted exponent of Lambda
  {
    float w0 = sqr(1.f - r), w1 = 1.f - w0;
    float a = w0 * 25.3245f + w1 * 21.5473f, b = w0 * 3.32435f + w1 * 3.82987f, c = w0 * 0.16801f + w1 * 0.19823f;
    float d = w0 * -1.27393f + w1 * -1.97760f, e = w0 * -4.85967f + w1 * -4.32054f;
    return a / (1.f + b * powf(x, c)) + d * x + e;
  }
  PRT_DEV float lambda(float c) const { return c < 0.5f ? expf(L(c)) : expf(2.f * L(0.5f) - L(1.f - c)); }
  PRT_DEV float lambdaI(float c) const { return powf(lambda(c), 1.f + 2.f * sqr(sqr(sqr(1.f - c)))); }
  PRT_DEV float G2(float cosO, float cosI) const { return 1.0f / (1.f + lambda(cosO) + lambdaI(cosI)); }
};
PRT_DEV float ggxAlbedo(const float* lut, float cosO, float roughness)  // MicrofacetAlbedoTable::eval
{
  return interp2D(cosO * 31.f, roughness * 31.f, lut + PRT_LUT_GGX_DIR, 32);
}
// MicrofacetDielectricAlbedoTable::eval (eta <= 1: etaDir over [1/3, 1], else rcpEtaDir over [1, 3])
PRT_DEV float dielAlbedo(const float* lut, float cosO, float eta, float roughness)
{
  if (eta <= 1.f)
  {
    float e = (eta - (1.0f / 3.0f)) / (1.0f - (1.0f / 3.0f));
    return interp3D(cosO * 15.f, e * 15.f, roughness * 15.f, lut + PRT_LUT_DIEL_ETA_DIR, 16);
  }
  float e = (eta - 1.f) / (3.f - 1.f);
  return interp3D(cosO * 15.f, e * 15.f, roughness * 15.f, lut + PRT_LUT_DIEL_RCPETA_DIR, 16);
}

// MicrofacetDielectricReflectionAlbedoTable::eval (eta in [1/3, 1])
PRT_DEV float dielReflAlbedo(const float* lut, float cosO, float eta, float roughness)
{
  float e = (eta - (1.0f / 3.0f)) / (1.0f - (1.0f / 3.0f));
  return interp3D(cosO * 15.f, e * 15.f, roughness * 15.f, lut + PRT_LUT_DIEL_REFL_DIR, 16);
}

// ---------------------------------------------------------------------------------------
// Shading context and BSDFs
// ---------------------------------------------------------------------------------------
namespace BsdfType {
enum : int {
  DiffuseReflection = 1 << 0, GlossyReflection = 1 << 1, SpecularReflection = 1 << 2,
  DiffuseTransmission = 1 << 3, GlossyTransmission = 1 << 4, SpecularTransmission = 1 << 5,
  UnbentTransmission = 1 << 6, Transparency = 1 << 7,
  Smooth = DiffuseReflection | DiffuseTransmission | GlossyReflection | GlossyTransmission,
  Delta = SpecularReflection | SpecularTransmission | UnbentTransmission | Transparency,
  Continuation = UnbentTransmission | Transparency,
};
}  // bsdf/bsdf.h:15-44

struct Ctx { f3 p, Ng; Frame f; float eps; };

// core/shading_context.h:144-191 (normal adaption + Gram-Schmidt)
PRT_DEV Frame shadingFrame(const Ctx& c, f3 wo)
{
  f3 N = c.f.N;
  f3 wi = reflect(wo, N);
  if (dot(wi, c.Ng) < 0.f)
  {
    float a = -dot(wi, c.Ng), b = dot(N, c.Ng);
    if (b < 0.f) { b = -b; wi = -wi; }
    f3 wi1 = normalize(wi + a * rcpSafe(b) * N + 1e-5f * c.Ng);
    N = normalize(wo + wi1);
  }
  Frame fr;
  fr.N = N;
  fr.U = c.f.U - dot(c.f.U, N) * N;
  float ul = lengthSqr(fr.U);
  fr.U = fr.U * (1.0f / sqrtf(ul));
  fr.V = c.f.V - dot(c.f.V, N) * N;
  fr.V = fr.V - dot(fr.V, fr.U) * fr.U;
  float vl = lengthSqr(fr.V);
  fr.V = fr.V * (1.0f / sqrtf(vl));
  if (ul == 0.f || vl == 0.f) makeFrame(fr.U, fr.V, fr.N);
  return fr;
}

// One BSDF instance for a hit: the material record plus per-material constants precomputed
// on the host at upload (prt_material_constants); the per-lane part is the frame.
struct Bsdf
{
  Frame fr;
  int kind;
  float roughness;
  GGX ggx;
  // diffuse: color; conductor: eta/k per channel + fms terms; dielectric: eta + fms terms;
  // coated: diffuse color, coat color, (thickness, scale, diffuse importance), Ravg in Eavg
  f3 c0, c1, c2;  // diffuse: color | conductor: eta, k, fmsScale | coated: see above
  f3 rr;          // conductor reflectivity (albedo AOV)
  float Eavg;     // conductor: average albedo
  float eta, EavgEta, EavgRcpEta, fmsRatio;  // dielectric (oriented)
  float importance;
  float diffRough;            // kinds 0 and 3: Oren-Nayar base when > 0
  int mix;                    // opacity mix: 0 lobe, 1 lobe + transparent BTDF, 2 transparent
  float opacity, transp;
  bool sheen;                 // a sheen layer over the lobe (bsdf/sheen.h)
  f3 sheenColor;
  float sheenScale, sheenRough, sheenA;
  // false: the scene has no thin dielectric, opacity, Oren-Nayar base, sheen or kind-5 material; a
  // compile-time constant in the kernels built for such scenes, where it folds those branches away
  bool ext;
  bool stdk;    // false: no kind-5 material in the scene (compile-time in the kernels built without them)
  bool smooth;  // a diffuse or glossy lobe: next-event estimation applies
  // kind 5: the mix of bases (the records kids[0 .. stdm->nKids) with stdm's scales and
  // importances) and whether the coat covers it (the coated fields hold the coat: c1 colour,
  // c2 = (thickness, scale, mix importance), Eavg = Ravg, eta, roughness, ggx); rr the albedo
  const prt_material_constants* stdm;
  const prt_material_constants* kids;
  bool coat;
};

// Importance of the surface: the lobe's, or under a sheen layer the larger of the sheen
// colour's and the lobe's (sheen.h:403-406).
PRT_DEV float surfImportance(const Bsdf& b)
{
  return b.sheen ? fmaxf(reduceMax(b.sheenColor), b.importance) : b.importance;
}
// The lobe's albedo (multi_bsdf.h:314-337 with one child).
PRT_DEV f3 lobeAlbedo(const Bsdf& b)
{
  if (b.kind == PRT_KIND_DIFFUSE) return b.c0;
  return (b.kind == PRT_KIND_DIELECTRIC) ? f3(1.f) : b.rr;  // conductor reflectivity | coated, thin albedo
}
// The albedo AOV of the whole BSDF: the opacity mix ignores the transparent BTDF's negative
// albedo and rescales (multi_bsdf.h:339-362); transparent alone returns -1 (transparent.h:52-55).
PRT_DEV f3 bsdfAlbedo(const Bsdf& b)
{
  if (b.mix == 2) return f3(-1.f);
  if (b.mix == 1 && !(surfImportance(b) * b.opacity > 0.f)) return f3(0.f);
  return lobeAlbedo(b);  // the sheen layer passes the base's albedo (sheen.h:408-411)
}
// The normal AOV: the first child with a non-zero importance (multi_bsdf.h:408-428); the
// transparent BTDF has no frame.
PRT_DEV f3 bsdfNormal(const Bsdf& b)
{
  bool lobe = b.importance > 0.f;
  if (b.stdk && b.kind == PRT_KIND_STANDARD)  // kind 5: a base with an importance and a frame
  {
    lobe = false;
    for (int t = 0; t < b.stdm->nKids; ++t)
      lobe = lobe || (b.stdm->kidImp[t] > 0.f && b.kids[t].kind != PRT_KIND_TRANSPARENT);
  }
  return (b.mix != 2 && lobe && (b.mix == 0 || surfImportance(b) * b.opacity > 0.f)) ? b.fr.N : f3(0.f);
}

// bsdf/diffuse.h:55-68, or with diffRough > 0 the Oren-Nayar BRDF in Fujii's form
// (bsdf/rough_diffuse.h:65-75)
PRT_DEV f3 diffuseEval(const Bsdf& b, f3 wo, f3 wi, float& pdf)
{
  float c = fmaxf(dot(wi, b.fr.N), 0.f);
  pdf = c * (1.f / kPi);
  if (!(b.diffRough > 0.f)) return b.c0 * (1.f / kPi) * c;
  float cO = fmaxf(dot(wo, b.fr.N), 0.f);
  float s = dot(wo, wi) - cO * c;
  float sInvt = s > 0.f ? s * rcpSafe(fmaxf(cO, c)) : s;
  return b.c0 * (c * (1.f + b.diffRough * sInvt) * (1.0f / (kPi + (kPi / 2.f - 2.f / 3.f) * b.diffRough)));
}
// Cosine sampling of the diffuse base (diffuse.h:61-67, rough_diffuse.h:77-87): the weight,
// zero below the geometric surface.
PRT_DEV f3 diffuseSample(const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy)
{
  wi = toGlobal(b.fr, cosineSampleHemisphere(pdf, sx, sy));
  if (!(dot(wi, c.Ng) > 0.f)) return f3(0.f);
  type = BsdfType::DiffuseReflection;
  if (!(b.diffRough > 0.f)) return b.c0;
  f3 v = diffuseEval(b, wo, wi, pdf);
  return v * (1.0f / pdf);
}
// bsdf/rough_conductor.h:132-202
PRT_DEV f3 conductorF(const Bsdf& b, float c)
{
  return f3(fresnelConductor(c, b.c0.x, b.c1.x), fresnelConductor(c, b.c0.y, b.c1.y), fresnelConductor(c, b.c0.z, b.c1.z));
}
PRT_DEV f3 conductorEval(const float* lut, const Bsdf& b, f3 wo, f3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N), cosI = dot(wi, b.fr.N);
  if (!(cosO > 0.f && cosI > 0.f)) return f3(0.f);
  f3 wh = normalize(wi + wo);
  float cosOH = dot(wo, wh);
  f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  f3 F = conductorF(b, cosOH);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  float Eo = ggxAlbedo(lut, cosO, b.roughness), Ei = ggxAlbedo(lut, cosI, b.roughness);
  f3 fms = b.c2 * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
  pdf = whPdf * (1.0f / (4.f * fabsf(cosOH)));
  return F * (D * G * (1.0f / (4.f * cosO))) + fms;
}
PRT_DEV f3 conductorSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  f3 wo0 = toLocal(b.fr, wo);
  float whPdf;
  f3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
  wi = reflect(wo, wh);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI > 0.f && dot(wi, c.Ng) > 0.f)) return f3(0.f);
  float cosOH = dot(wo, wh);
  f3 wi0 = toLocal(b.fr, wi);
  f3 F = conductorF(b, cosOH);
  float G = b.ggx.G2(wo0, wi0);
  float Eo = ggxAlbedo(lut, cosO, b.roughness), Ei = ggxAlbedo(lut, cosI, b.roughness);
  f3 fms = b.c2 * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
  type = BsdfType::GlossyReflection;
  pdf = whPdf * (1.0f / (4.f * fabsf(cosOH)));
  return F * (G * rcpSafe(b.ggx.G1(wo0))) + fms * (1.0f / pdf);
}
// bsdf/rough_dielectric.h:296-526
PRT_DEV float dielSingle(const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return 0.f;
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return 0.f;
  bool refl = cosI > 0.f;
  f3 wh = normalize(refl ? wi + wo : b.eta * wo + wi);
  if (dot(wh, b.fr.N) < 0.f) wh = -wh;
  float cosOH = dot(wo, wh), cosIH = dot(wi, wh);
  float F = fresnelDielectric(cosOH, b.eta);
  float value;
  if (refl)
  {
    pdf = F * (1.0f / (4.f * fabsf(cosOH)));
    value = F * (1.0f / (4.f * cosO));
  }
  else
  {
    pdf = (1.f - F) * (1.0f / sqr(b.eta)) * fabsf(cosIH) * (1.0f / sqr(cosOH + (1.0f / b.eta) * cosIH));
    value = pdf * fabsf(cosOH) * (1.0f / cosO);
  }
  f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  pdf *= whPdf;
  return value * D * G;
}
PRT_DEV void dielFms(const float* lut, const Bsdf& b, float Eo, float cosI, float& fmsPdf, float& fmsValue)
{
  if (cosI > 0.f)
  {
    float Ei = dielAlbedo(lut, cosI, b.eta, b.roughness);
    fmsPdf = b.fmsRatio * cosI * (1.f / kPi);
    fmsValue = b.fmsRatio * (1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.EavgEta))) * cosI;
  }
  else
  {
    float ac = fabsf(cosI);
    float Ei = dielAlbedo(lut, ac, 1.0f / b.eta, b.roughness);
    fmsPdf = (1.f - b.fmsRatio) * ac * (1.f / kPi);
    fmsValue = (1.f - b.fmsRatio) * (1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.EavgRcpEta))) * ac;
  }
}
PRT_DEV f3 dielEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  float Eo = dielAlbedo(lut, cosO, b.eta, b.roughness);
  float fmsPdf, fmsValue, sPdf;
  dielFms(lut, b, Eo, cosI, fmsPdf, fmsValue);
  float sValue = dielSingle(b, c, wo, wi, sPdf);
  pdf = Eo * sPdf + (1.f - Eo) * fmsPdf;
  return f3(sValue + fmsValue);
}
PRT_DEV f3 dielSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float Eo = dielAlbedo(lut, cosO, b.eta, b.roughness);
  bool ok;
  if (sc <= Eo)
  {
    f3 wo0 = toLocal(b.fr, wo);
    float whPdf;
    f3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    float cosOH = dot(wo, wh), cosTH;
    float F = fresnelDielectricEx(cosOH, cosTH, b.eta);
    if (sc * (1.0f / Eo) <= F) { wi = reflect(wo, wh, cosOH); ok = dot(wi, b.fr.N) > 0.f; }
    else { wi = refract(wo, wh, cosOH, cosTH, b.eta); ok = dot(wi, b.fr.N) < 0.f; }
  }
  else
  {
    float sc1 = (sc - Eo) * (1.0f / (1.f - Eo));
    float cp;
    f3 dir = toGlobal(b.fr, cosineSampleHemisphere(cp, sx, sy));
    wi = (sc1 <= b.fmsRatio) ? dir : -dir;
    ok = true;
  }
  if (!ok) return f3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  type = cosI > 0.f ? BsdfType::GlossyReflection : BsdfType::GlossyTransmission;
  float fmsPdf, fmsValue, sPdf;
  dielFms(lut, b, Eo, cosI, fmsPdf, fmsValue);
  float sValue = dielSingle(b, c, wo, wi, sPdf);
  pdf = Eo * sPdf + (1.f - Eo) * fmsPdf;
  return f3((sValue + fmsValue) * (1.0f / pdf));
}

// Coated diffuse: RoughCoatingBsdfSimd over DiffuseBrdfSimd (bsdf/rough_coating.h:297-593).
// c0 = diffuse color, c1 = coat color, c2 = (thickness, scale, diffuse importance),
// Eavg = Ravg, eta <= 1.
PRT_DEV float refractCos(float cosI, float eta)  // optics.h:50-56
{
  return sqrtf(fmaxf(1.f - eta * eta * (1.f - cosI * cosI), 0.f));
}
// absorption through the coat along the refracted path and energy conservation (:331-350): the
// term of Kelemen and Szirmay-Kalos 2001 for a base with diffuse lobes only, min(1 - Ro, 1 - Ri)
// for any other base (Kulla and Conty 2017)
PRT_DEV f3 coatBase(const float* lut, const Bsdf& b, f3 baseValue, float cosO, float cosI, float Ro, bool diffuseOnly)
{
  float length = 1.0f / refractCos(cosO, b.eta) + 1.0f / refractCos(fabsf(cosI), b.eta);
  if (cosI <= 0.f) length *= 0.5f;
  const float t = b.c2.x * length;
  f3 absorbed(baseValue.x * powf(b.c1.x, t), baseValue.y * powf(b.c1.y, t), baseValue.z * powf(b.c1.z, t));
  baseValue = baseValue + b.c2.y * (absorbed - baseValue);
  float Ri = dielReflAlbedo(lut, fabsf(cosI), b.eta, b.roughness) * b.c2.y;
  return baseValue * (diffuseOnly ? (1.f - Ro) * (1.f - Ri) * (1.0f / (1.f - b.Eavg)) : fminf(1.f - Ro, 1.f - Ri));
}
PRT_DEV float coatPickProb(const Bsdf& b, float Ro)
{
  float baseImportance = (1.f - Ro) * b.c2.z;
  return Ro * (1.0f / (Ro + baseImportance));
}
// specular reflection of the coat around the half vector wh (:365-378)
PRT_DEV float coatReflection(const Bsdf& b, f3 wo, f3 wi, f3 wh, float cosO, float& coatPdf)
{
  float cosOH = dot(wo, wh);
  float F = fresnelDielectric(cosOH, b.eta) * b.c2.y;
  f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  coatPdf = whPdf * (1.0f / (4.f * fabsf(cosOH)));
  return F * D * G * (1.0f / (4.f * cosO));
}
// RoughCoatingBsdfSimd over a base (:297-537).  Base: eval(wo, wi, pdf), sample(wo, wi, pdf, type,
// sx, sy, sc), diffuseOnly(), and kDelta (whether a base sample can be a delta).
template <class Base>
PRT_DEV f3 coatEvalOver(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf, const Base& base)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  float basePdf;
  f3 baseValue = base.eval(wo, wi, basePdf);
  float Ro = dielReflAlbedo(lut, cosO, b.eta, b.roughness) * b.c2.y;
  baseValue = coatBase(lut, b, baseValue, cosO, cosI, Ro, base.diffuseOnly());
  float coatPick = coatPickProb(b, Ro), basePick = 1.f - coatPick;
  if (cosI > 0.f)
  {
    float coatPdf;
    float coatValue = coatReflection(b, wo, wi, normalize(wo + wi), cosO, coatPdf);
    pdf = coatPick * coatPdf + basePick * basePdf;
    return f3(coatValue) + baseValue;
  }
  pdf = basePdf * basePick;
  return baseValue;
}
template <class Base>
PRT_DEV f3 coatSampleOver(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy,
                          float sc, const Base& base)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  f3 wo0 = toLocal(b.fr, wo);
  float Ro = dielReflAlbedo(lut, cosO, b.eta, b.roughness) * b.c2.y;
  f3 wh, baseValue;
  float basePdf;
  float coatPick = coatPickProb(b, Ro), basePick = 1.f - coatPick;
  if (sc <= coatPick)
  {
    float whPdf;
    wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    wi = reflect(wo, wh, dot(wo, wh));
    if (!(dot(wi, b.fr.N) > 0.f)) return f3(0.f);
    type = BsdfType::GlossyReflection;
    baseValue = base.eval(wo, wi, basePdf);
  }
  else
  {
    f3 bw = base.sample(wo, wi, basePdf, type, sx, sy, (sc - coatPick) * (1.0f / basePick));
    if (!(reduceMax(bw) > 0.f)) return f3(0.f);
    baseValue = (Base::kDelta && (type & BsdfType::Delta)) ? bw : bw * basePdf;
    wh = normalize(wo + wi);
  }
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  baseValue = coatBase(lut, b, baseValue, cosO, cosI, Ro, base.diffuseOnly());
  if (Base::kDelta && (type & BsdfType::Delta))  // a delta base sample: no coat reflection (:484-490)
  {
    pdf = basePdf;
    return baseValue * (1.0f / basePick);
  }
  if (cosI <= 0.f)
  {
    pdf = basePick * basePdf;
    return baseValue * (1.0f / pdf);
  }
  float coatPdf;
  float coatValue = coatReflection(b, wo, wi, wh, cosO, coatPdf);
  pdf = coatPick * coatPdf + basePick * basePdf;
  return (f3(coatValue) + baseValue) * (1.0f / pdf);
}
// the coated kind's diffuse base: it reads only the 2D part of the remapped sample (:424-427)
struct DiffuseBase
{
  static constexpr bool kDelta = false;
  const Bsdf& b;
  const Ctx& c;
  PRT_DEV f3 eval(f3 wo, f3 wi, float& pdf) const { return diffuseEval(b, wo, wi, pdf); }
  PRT_DEV f3 sample(f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float) const
  {
    return diffuseSample(b, c, wo, wi, pdf, type, sx, sy);
  }
  PRT_DEV bool diffuseOnly() const { return true; }
};
PRT_DEV f3 coatEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  return coatEvalOver(lut, b, c, wo, wi, pdf, DiffuseBase{b, c});
}
PRT_DEV f3 coatSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  return coatSampleOver(lut, b, c, wo, wi, pdf, type, sx, sy, sc, DiffuseBase{b, c});
}

// --- thin rough dielectric (bsdf/thin_rough_dielectric.h:170-345): c0 = transmission color,
// c1.x = fmsScale, c2.x = thickness, Eavg = GGX average albedo, eta <= 1.  Reflection is a GGX
// lobe with energy compensation; transmission goes straight through (unbent, a continuation
// lobe) with probability 1 - R.
PRT_DEV float thinReflectance(const float* lut, const Bsdf& b, float cosO, float& Eo)
{
  Eo = ggxAlbedo(lut, cosO, b.roughness);
  return dielReflAlbedo(lut, cosO, b.eta, b.roughness) + b.c1.x * (1.f - Eo);  // :262-264
}
PRT_DEV f3 thinTransmittance(const Bsdf& b, float cosO)  // pow(color, thickness / cosT) (:319)
{
  float e = (1.0f / refractCos(cosO, b.eta)) * b.c2.x;
  return f3(powf(b.c0.x, e), powf(b.c0.y, e), powf(b.c0.z, e));
}
PRT_DEV f3 thinEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)  // :194-240
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f) || !(cosI > 0.f)) return f3(0.f);
  f3 wh = normalize(wi + wo);
  float cosOH = dot(wo, wh);
  f3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float F = fresnelDielectric(cosOH, b.eta);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  float Eo, Ei = ggxAlbedo(lut, cosI, b.roughness);
  float R = thinReflectance(lut, b, cosO, Eo);
  float fms = b.c1.x * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
  pdf = whPdf * (1.0f / (4.f * fabsf(cosOH))) * R;
  return f3(F * D * G * (1.0f / (4.f * cosO)) + fms);
}
PRT_DEV f3 thinSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)  // :242-305
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float Eo;
  float R = thinReflectance(lut, b, cosO, Eo);
  pdf = 0.f;
  wi = f3(0.f);
  if (sc <= R)
  {
    f3 wo0 = toLocal(b.fr, wo);
    float whPdf;
    f3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    wi = reflect(wo, wh);
    float cosI = dot(wi, b.fr.N);
    if (!(cosI > 0.f && dot(wi, c.Ng) > 0.f)) return f3(0.f);
    float cosOH = dot(wo, wh);
    f3 wi0 = toLocal(b.fr, wi);
    float F = fresnelDielectric(cosOH, b.eta);
    float G = b.ggx.G2(wo0, wi0);
    float Ei = ggxAlbedo(lut, cosI, b.roughness);
    float fms = b.c1.x * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
    type = BsdfType::GlossyReflection;
    pdf = whPdf * (1.0f / (4.f * fabsf(cosOH))) * R;
    return f3(F * G * rcpSafe(b.ggx.G1(wo0)) * (1.0f / R) + fms * (1.0f / pdf));
  }
  wi = -wo;
  if (!(dot(wi, c.Ng) < 0.f)) return f3(0.f);
  type = BsdfType::UnbentTransmission;
  pdf = 1.f - R;
  return thinTransmittance(b, cosO);
}
// getTransparency (:307-321): what a shadow ray keeps through the surface
PRT_DEV f3 thinTransparency(const float* lut, const Bsdf& b, f3 wo)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float Eo;
  float R = thinReflectance(lut, b, cosO, Eo);
  return thinTransmittance(b, cosO) * (1.f - R);
}

// One lobe of a single kind, or the transparent BTDF (bsdf/transparent.h:32-71: evaluates to zero,
// samples -wo with pdf 1).
PRT_DEV f3 leafEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  pdf = 0.f;
  if (b.kind == PRT_KIND_DIFFUSE) return diffuseEval(b, wo, wi, pdf);
  if (b.kind == PRT_KIND_CONDUCTOR) return conductorEval(lut, b, wo, wi, pdf);
  if (b.kind == PRT_KIND_COATED) return coatEval(lut, b, c, wo, wi, pdf);
  if (b.ext && b.kind == PRT_KIND_THIN_DIELECTRIC) return thinEval(lut, b, c, wo, wi, pdf);
  if (b.ext && b.kind == PRT_KIND_TRANSPARENT) return f3(0.f);
  return dielEval(lut, b, c, wo, wi, pdf);
}
PRT_DEV f3 leafSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.kind == PRT_KIND_DIFFUSE) return diffuseSample(b, c, wo, wi, pdf, type, sx, sy);
  if (b.kind == PRT_KIND_CONDUCTOR) return conductorSample(lut, b, c, wo, wi, pdf, type, sx, sy);
  if (b.kind == PRT_KIND_COATED) return coatSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
  if (b.ext && b.kind == PRT_KIND_THIN_DIELECTRIC) return thinSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
  if (b.ext && b.kind == PRT_KIND_TRANSPARENT)
  {
    wi = -wo;
    type = BsdfType::Transparency;
    pdf = 1.f;
    return f3(1.f);
  }
  return dielSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
}

// Base t of a kind-5 mix as a lobe on the parent's frame.
PRT_DEV Bsdf kidBsdf(const Bsdf& parent, int t)
{
  const prt_material_constants& k = parent.kids[t];
  Bsdf b;
  b.fr = parent.fr;
  b.kind = k.kind;
  b.roughness = k.roughness;
  b.ggx.a = k.alpha;
  b.importance = k.importance;
  b.c0 = f3(k.c0[0], k.c0[1], k.c0[2]);
  b.c1 = f3(k.c1[0], k.c1[1], k.c1[2]);
  b.c2 = f3(k.c2[0], k.c2[1], k.c2[2]);
  b.rr = f3(k.rr[0], k.rr[1], k.rr[2]);
  b.Eavg = k.Eavg;
  b.eta = k.eta[0];  // each orientation has its own base records
  b.EavgEta = k.EavgEta[0];
  b.EavgRcpEta = k.EavgRcpEta[0];
  b.fmsRatio = k.fmsRatio[0];
  b.diffRough = k.diffuseRoughness;
  b.mix = 0;
  b.sheen = false;
  b.ext = true;
  b.stdk = false;
  b.smooth = k.smooth != 0;
  b.stdm = nullptr;
  b.kids = nullptr;
  b.coat = false;
  return b;
}

// kind 5's weighted mix of bases (multi_bsdf.h:166-287): eval sums the bases' values times their
// scales and their pdfs times their importances; sample picks a base by importance with the
// remapped sample, then (unless the pick is a delta) adds the other bases' values and pdfs.
PRT_DEV f3 mixEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  const prt_material_constants& m = *b.stdm;
  f3 value(0.f);
  pdf = 0.f;
  for (int t = 0; t < m.nKids; ++t)
  {
    if (!(m.kidImp[t] > 0.f)) continue;
    float cp;
    f3 cv = leafEval(lut, kidBsdf(b, t), c, wo, wi, cp);
    value = value + cv * m.kidScale[t];
    pdf += cp * m.kidImp[t];
  }
  pdf *= 1.0f / m.impSum;
  return value;
}
PRT_DEV f3 mixSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  const prt_material_constants& m = *b.stdm;
  if (!(m.impSum > 0.f)) return f3(0.f);
  if (m.nKids == 1) return leafSample(lut, kidBsdf(b, 0), c, wo, wi, pdf, type, sx, sy, sc) * m.kidScale[0];
  const float x = sc * m.impSum;
  int choice = 0;
  float prefix = m.kidImp[0], imp = m.kidImp[0], scale = m.kidScale[0];
  for (int t = 1; t < m.nKids; ++t)
  {
    if (!(x > prefix)) break;
    choice = t;
    prefix += m.kidImp[t];
    imp = m.kidImp[t];
    scale = m.kidScale[t];
  }
  f3 cwi(0.f);
  float cpdf = 0.f;
  int ctype = 0;
  f3 cw = leafSample(lut, kidBsdf(b, choice), c, wo, cwi, cpdf, ctype, sx, sy, (x + imp - prefix) * (1.0f / imp));
  f3 w(0.f);
  wi = f3(0.f);
  pdf = 0.f;
  type = 0;
  if (reduceMax(cw) > 0.f) { w = cw; wi = cwi; pdf = cpdf; type = ctype; }
  w = w * scale;
  if (!(reduceMax(w) > 0.f && pdf > 0.f)) return f3(0.f);
  if (type & BsdfType::Delta) return w * (m.impSum * (1.0f / imp));
  f3 value = w * pdf;
  pdf = pdf * imp;
  for (int t = 0; t < m.nKids; ++t)
  {
    if (t == choice || !(m.kidImp[t] > 0.f)) continue;
    float cp;
    f3 cv = leafEval(lut, kidBsdf(b, t), c, wo, wi, cp);
    value = value + cv * m.kidScale[t];
    pdf += cp * m.kidImp[t];
  }
  pdf *= 1.0f / m.impSum;
  return value * (1.0f / pdf);
}
struct MixBase
{
  static constexpr bool kDelta = true;
  const float* lut;
  const Bsdf& b;
  const Ctx& c;
  PRT_DEV f3 eval(f3 wo, f3 wi, float& pdf) const { return mixEval(lut, b, c, wo, wi, pdf); }
  PRT_DEV f3 sample(f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc) const
  {
    return mixSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
  }
  PRT_DEV bool diffuseOnly() const { return b.stdm->mixDiffuseOnly != 0; }  // rough_coating.h:349
};

// The material's lobe: a MultiBsdf with one child of weight 1 for kinds 0-4
// (bsdf/multi_bsdf.h:166-210), kind 5's mix of bases under its coat when it has one.
PRT_DEV f3 lobeEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  if (b.stdk && b.kind == PRT_KIND_STANDARD)
  {
    const MixBase base{lut, b, c};
    return b.coat ? coatEvalOver(lut, b, c, wo, wi, pdf, base) : base.eval(wo, wi, pdf);
  }
  f3 v(0.f);
  pdf = 0.f;
  if (b.importance > 0.f)
  {
    float cp;
    v = leafEval(lut, b, c, wo, wi, cp);
    pdf = cp * b.importance;
  }
  pdf *= 1.0f / b.importance;
  return v;
}
PRT_DEV f3 lobeSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.stdk && b.kind == PRT_KIND_STANDARD)
  {
    const MixBase base{lut, b, c};
    return b.coat ? coatSampleOver(lut, b, c, wo, wi, pdf, type, sx, sy, sc, base) : base.sample(wo, wi, pdf, type, sx, sy, sc);
  }
  if (!(b.importance > 0.f)) return f3(0.f);
  return leafSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
}

// --- sheen layer over the lobe (bsdf/sheen.h:199-451): a microfiber reflection picked by
// importance against the lobe, which it attenuates by min(1 - Ro, 1 - Ri)
PRT_DEV void sheenPicks(const Bsdf& b, float Ro, float& coatPick, float& basePick)
{
  float coatImp = Ro * reduceMax(b.sheenColor), baseImp = (1.f - Ro) * b.importance;
  coatPick = coatImp * (1.0f / (coatImp + baseImp));
  basePick = 1.f - coatPick;
}
PRT_DEV f3 sheenReflection(const Bsdf& b, f3 wo, f3 wi, float cosO, float cosI)  // :253-265, :372-384
{
  f3 wh = normalize(wo + wi);
  SheenDist sd{b.sheenA};
  float D = sd.D(dot(wh, b.fr.N)), G = sd.G2(cosO, cosI);
  return b.sheenColor * (D * G * (1.0f / (4.f * cosO)) * b.sheenScale);
}
PRT_DEV f3 sheenEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)  // :218-279
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  float basePdf;
  f3 baseValue = lobeEval(lut, b, c, wo, wi, basePdf);
  float Ro = sheenAlbedo(lut, cosO, b.sheenRough) * b.sheenScale;
  float Ri = sheenAlbedo(lut, fabsf(cosI), b.sheenRough) * b.sheenScale;
  baseValue = baseValue * fminf(1.f - Ro, 1.f - Ri);
  float coatPick, basePick;
  sheenPicks(b, Ro, coatPick, basePick);
  if (cosI > 0.f)
  {
    pdf = coatPick * (1.f / (2.f * kPi)) + basePick * basePdf;
    return sheenReflection(b, wo, wi, cosO, cosI) + baseValue;
  }
  pdf = basePdf * basePick;
  return baseValue;
}
PRT_DEV f3 sheenSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)  // :281-391
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return f3(0.f);
  float Ro = sheenAlbedo(lut, cosO, b.sheenRough) * b.sheenScale;
  float coatPick, basePick;
  sheenPicks(b, Ro, coatPick, basePick);
  float basePdf = 0.f;
  f3 baseValue(0.f);
  wi = f3(0.f);
  pdf = 0.f;
  type = 0;
  if (sc <= coatPick)
  {
    type = BsdfType::DiffuseReflection;
    float phi = 2.f * kPi * sx, sp, cp;
    sincosf(phi, &sp, &cp);
    const float st = cos2sin(sy);
    wi = toGlobal(b.fr, f3(cp * st, sp * st, sy));  // uniformSampleHemisphere (shape_sampler.h:54-61)
    baseValue = lobeEval(lut, b, c, wo, wi, basePdf);
  }
  else
  {
    f3 bwi(0.f);
    float bpdf = 0.f;
    int btype = 0;
    f3 bw = lobeSample(lut, b, c, wo, bwi, bpdf, btype, sx, sy, (sc - coatPick) * (1.0f / basePick));
    if (!(reduceMax(bw) > 0.f)) return f3(0.f);
    wi = bwi;
    basePdf = bpdf;
    type = btype;
    baseValue = bw * ((type & BsdfType::Delta) ? 1.f : basePdf);
  }
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, c.Ng) > 0.f)) return f3(0.f);
  float Ri = sheenAlbedo(lut, fabsf(cosI), b.sheenRough) * b.sheenScale;
  baseValue = baseValue * fminf(1.f - Ro, 1.f - Ri);
  if (type & BsdfType::Delta)
  {
    pdf = basePdf;
    return baseValue * (1.0f / basePick);
  }
  if (cosI <= 0.f)
  {
    pdf = basePick * basePdf;
    return baseValue * (1.0f / pdf);
  }
  pdf = coatPick * (1.f / (2.f * kPi)) + basePick * basePdf;
  return (sheenReflection(b, wo, wi, cosO, cosI) + baseValue) * (1.0f / pdf);
}

// The surface: the lobe, under the sheen layer where there is one.
PRT_DEV f3 surfEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  return b.sheen ? sheenEval(lut, b, c, wo, wi, pdf) : lobeEval(lut, b, c, wo, wi, pdf);
}
PRT_DEV f3 surfSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  return b.sheen ? sheenSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc) : lobeSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
}

// The whole BSDF with the opacity mix (material/standard.cpp:613-627): MultiBsdf{lobe x opacity,
// transparent BTDF x transp} (multi_bsdf.h:166-287, transparent.h:40-45).  The transparent
// BTDF evaluates to zero and has importance 1.
PRT_DEV f3 bsdfEval(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3 wi, float& pdf)
{
  if (b.mix == 0) return surfEval(lut, b, c, wo, wi, pdf);
  pdf = 0.f;
  if (b.mix == 2) return f3(0.f);
  const float imp0 = surfImportance(b) * b.opacity, sum = imp0 + b.transp;
  f3 v(0.f);
  if (imp0 > 0.f)
  {
    float cp;
    v = surfEval(lut, b, c, wo, wi, cp) * b.opacity;
    pdf = cp * imp0;
  }
  pdf *= 1.0f / sum;
  return v;
}
PRT_DEV f3 bsdfSample(const float* lut, const Bsdf& b, const Ctx& c, f3 wo, f3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.mix == 0) return surfSample(lut, b, c, wo, wi, pdf, type, sx, sy, sc);
  if (b.mix == 2)
  {
    wi = -wo;
    type = BsdfType::Transparency;
    pdf = 1.f;
    return f3(1.f);
  }
  const float imp0 = surfImportance(b) * b.opacity, sum = imp0 + b.transp;
  if (!(sum > 0.f)) return f3(0.f);
  // choose a child by importance and remap the sample (:219-245)
  const float x = sc * sum;
  const bool lobe = !(x > imp0);
  const float imp = lobe ? imp0 : b.transp, prefix = lobe ? imp0 : sum;
  const float sc1 = (x + imp - prefix) * (1.0f / imp);
  f3 w(0.f);
  wi = f3(0.f);
  pdf = 0.f;
  type = 0;
  if (lobe)
  {
    f3 cwi(0.f);
    float cpdf = 0.f;
    int ctype = 0;
    f3 cw = surfSample(lut, b, c, wo, cwi, cpdf, ctype, sx, sy, sc1);
    if (reduceMax(cw) > 0.f) { w = cw * b.opacity; wi = cwi; pdf = cpdf; type = ctype; }
  }
  else
  {
    w = f3(b.transp);
    wi = -wo;
    pdf = 1.f;
    type = BsdfType::Transparency;
  }
  if (!(reduceMax(w) > 0.f && pdf > 0.f)) return f3(0.f);
  if (type & BsdfType::Delta) return w * (sum * (1.0f / imp));  // (:262-272)
  // a smooth sample of the lobe: the transparent BTDF adds nothing to value or pdf (:274-287)
  f3 value = w * pdf;
  pdf = pdf * imp * (1.0f / sum);
  return value * (1.0f / pdf);
}
// getTransparency of a material whose shadow rays pass through (material/standard.cpp:633-661):
// a thin dielectric's lobe transmits, every other lobe is opaque, the transparent BTDF passes 1.
PRT_DEV f3 bsdfTransparency(const float* lut, const Bsdf& b, f3 wo)
{
  if (b.mix == 2) return f3(1.f);
  f3 t = (b.kind == PRT_KIND_THIN_DIELECTRIC && b.importance > 0.f) ? thinTransparency(lut, b, wo) : f3(0.f);
  if (b.stdk && b.kind == PRT_KIND_STANDARD)
  {
    // the mix (multi_bsdf.h:289-305): thin dielectrics transmit, the transparent BTDF passes 1
    const prt_material_constants& m = *b.stdm;
    t = f3(0.f);
    for (int k = 0; k < m.nKids; ++k)
    {
      if (!(m.kidImp[k] > 0.f)) continue;
      const int kind = b.kids[k].kind;
      if (kind == PRT_KIND_THIN_DIELECTRIC) t = t + thinTransparency(lut, kidBsdf(b, k), wo) * m.kidScale[k];
      else if (kind == PRT_KIND_TRANSPARENT) t = t + f3(m.kidScale[k]);
    }
    if (b.coat)  // the coat's absorption along the refracted direction and its reflectance (rough_coating.h:539-560)
    {
      const float cosO = dot(wo, b.fr.N);
      if (!(cosO > 0.f)) t = f3(0.f);
      else
      {
        const float e = b.c2.x * (1.0f / refractCos(cosO, b.eta));
        f3 absorbed(t.x * powf(b.c1.x, e), t.y * powf(b.c1.y, e), t.z * powf(b.c1.z, e));
        t = (t + b.c2.y * (absorbed - t)) * (1.f - dielReflAlbedo(lut, cosO, b.eta, b.roughness) * b.c2.y);
      }
    }
  }
  if (b.sheen)  // sheen.h:393-401
  {
    float cosO = dot(wo, b.fr.N);
    t = cosO > 0.f ? t * (1.f - sheenAlbedo(lut, cosO, b.sheenRough) * b.sheenScale) : f3(0.f);
  }
  if (b.mix == 0) return t;
  return t * b.opacity + f3(b.transp);
}

// Builds the BSDF for a hit from the material constants (prt_material_constants, computed on
// the host from the formulas of rough_conductor.h:121-130 and rough_dielectric.h:279-291).
// ext = false (for scenes without thin dielectrics, opacity or Oren-Nayar bases) drops those
// paths at compile time where the argument is a constant; stdk = false likewise drops kind 5.
PRT_DEV Bsdf makeBsdf(const prt_material_constants& m, const Ctx& c, f3 wo, bool inside, bool ext = true, bool stdk = true)
{
  Bsdf b;
  b.fr = shadingFrame(c, wo);
  b.kind = m.kind;
  b.roughness = m.roughness;
  b.ggx.a = m.alpha;
  b.importance = m.importance;
  b.c0 = f3(m.c0[0], m.c0[1], m.c0[2]);
  b.c1 = f3(m.c1[0], m.c1[1], m.c1[2]);
  b.c2 = f3(m.c2[0], m.c2[1], m.c2[2]);
  b.rr = f3(m.rr[0], m.rr[1], m.rr[2]);
  b.Eavg = m.Eavg;
  int o = inside ? 1 : 0;  // material/material.h:66-76: inside the medium = backfacing
  b.eta = m.eta[o];
  b.EavgEta = m.EavgEta[o];
  b.EavgRcpEta = m.EavgRcpEta[o];
  b.fmsRatio = m.fmsRatio[o];
  b.ext = ext;
  b.diffRough = ext ? m.diffuseRoughness : 0.f;
  b.mix = ext ? m.mix : 0;
  b.opacity = m.opacity;
  b.transp = m.transp;
  b.sheen = ext && m.sheen;
  b.sheenColor = f3(m.sheenColor[0], m.sheenColor[1], m.sheenColor[2]);
  b.sheenScale = m.sheenScale;
  b.sheenRough = m.sheenRoughness;
  b.sheenA = m.sheenAlpha;
  b.stdk = ext && stdk;
  b.smooth = !b.stdk || m.smooth != 0;
  b.stdm = nullptr;
  b.kids = nullptr;
  b.coat = false;
  if (b.stdk && m.kind == PRT_KIND_STANDARD)
  {
    // the mix of bases of this orientation, and the coat over it
    b.stdm = &m;
    b.kids = &m + m.kidBase[o];
    b.coat = m.coat != 0;
    b.c1 = f3(m.coatColor[0], m.coatColor[1], m.coatColor[2]);
    b.c2 = f3(m.coatThickness, m.coatScale, m.impSum);
    b.Eavg = m.coatRavg[o];
    b.eta = m.coatEta[o];
    b.roughness = m.coatRoughness;
    b.ggx.a = m.coatAlpha;
  }
  return b;
}

// ---------------------------------------------------------------------------------------
// Environment lights (light/*.h, render/scene.h:137-250)
// ---------------------------------------------------------------------------------------
struct EnvView
{
  const float4* texels;  // [H][W] rgb + pad, texture row order
  const float* xDists;   // [LH][LW+1]
  const float* yDist;    // [LH+1]
  const int* xGuide;     // [LH][PRT_ENV_GUIDE+1] guide tables of the row CDFs
  const int* yGuide;     // [PRT_ENV_GUIDE+1] guide table of the row-selection CDF
  int W, H, LW, LH;
};

// image/image_texture.h:252-283: bilinear, u repeat, v clamp
PRT_DEV f3 envLookup(const EnvView& e, float u, float v)
{
  float x = u - 0.5f / float(e.W), y = v - 0.5f / float(e.H);
  float xr = (x - floorf(x)) * float(e.W);
  float yr = clampf(y * float(e.H), 0.f, float(e.H - 1));
  float sx = xr - floorf(xr), sy = yr - floorf(yr);
  int x0 = min(max(int(xr), 0), e.W - 1), y0 = min(max(int(yr), 0), e.H - 1);
  int x1 = (x0 + 1 == e.W) ? 0 : x0 + 1, y1 = min(y0 + 1, e.H - 1);
  float4 c00 = e.texels[y0 * e.W + x0], c10 = e.texels[y0 * e.W + x1];
  float4 c01 = e.texels[y1 * e.W + x0], c11 = e.texels[y1 * e.W + x1];
  f3 a(c00.x, c00.y, c00.z), b(c10.x, c10.y, c10.z), c(c01.x, c01.y, c01.z), d(c11.x, c11.y, c11.z);
  f3 c0 = a + sx * (b - a), c1 = c + sx * (d - c);
  return c0 + sy * (c1 - c0);
}

PRT_DEV f3 envEvalOne(const prt_light_dev& l, const EnvView& e, f3 wi, float& pdf)
{
  pdf = 0.f;
  if (l.kind == PRT_LIGHT_AMBIENT)  // ambient_light.h:33-37
  {
    pdf = 1.f / (4.f * kPi);
    return f3(l.r, l.g, l.b);
  }
  if (l.kind == PRT_LIGHT_DISTANT)  // distant_light.h:48-57
  {
    float c = dot(wi, f3(l.dx, l.dy, l.dz));
    if (!(c >= l.cosAngle)) return f3(0.f);
    pdf = l.conePdf;
    return f3(l.r, l.g, l.b);
  }
  // image_light.h:79-91
  float theta = acosf(clampf(wi.y, -1.f, 1.f));
  float phi = atan2f(wi.x, -wi.z);
  float u = 0.5f + phi * (1.0f / (2.0f * kPi));
  float v = 1.0f - theta * (1.0f / kPi);
  int ix = min(max(int(u * float(e.LW)), 0), e.LW - 1);
  int iy = min(max(int(v * float(e.LH)), 0), e.LH - 1);
  float p = dist1dPdf(e.xDists + iy * (e.LW + 1), e.LW, ix) * dist1dPdf(e.yDist, e.LH, iy);
  pdf = p * float(e.LW) * float(e.LH) * (1.0f / (2.0f * kPi * kPi * sinf(theta)));
  return envLookup(e, u, v) * f3(l.r, l.g, l.b);
}

PRT_DEV f3 envSampleOne(const prt_light_dev& l, const EnvView& e, f3& wi, float& pdf, float sx, float sy)
{
  if (l.kind == PRT_LIGHT_AMBIENT)  // ambient_light.h:46-51
  {
    wi = uniformSampleSphere(pdf, sx, sy);
    return f3(l.r, l.g, l.b) * (1.0f / pdf);
  }
  if (l.kind == PRT_LIGHT_DISTANT)  // distant_light.h:66-71
  {
    Frame fr;
    fr.U = f3(l.fu[0], l.fu[1], l.fu[2]);
    fr.V = f3(l.fv[0], l.fv[1], l.fv[2]);
    fr.N = f3(l.dx, l.dy, l.dz);
    wi = toGlobal(fr, uniformSampleCone(pdf, sx, sy, l.cosAngle, l.conePdf));
    return f3(l.r, l.g, l.b) * (1.0f / pdf);
  }
  // image_light.h:112-134 (Distribution2D::sampleReuse, distribution2d.h:53-64)
  float yPdf, xPdf;
  int iy = dist1dSampleReuseGuided(e.yDist, e.yGuide, yPdf, sy);
  int ix = dist1dSampleReuseGuided(e.xDists + iy * (e.LW + 1), e.xGuide + iy * (PRT_ENV_GUIDE + 1), xPdf, sx);
  pdf = xPdf * yPdf;
  float u = (float(ix) + sx) * (1.0f / float(e.LW));
  float v = (float(iy) + sy) * (1.0f / float(e.LH));
  float theta = (1.0f - v) * kPi, phi = (u - 0.5f) * (2.0f * kPi);
  float sp, cp, st, ct;
  sincosf(phi, &sp, &cp);
  sincosf(theta, &st, &ct);
  wi = f3(sp * st, ct, -cp * st);
  pdf *= float(e.LW) * float(e.LH) * (1.0f / (2.0f * kPi * kPi * st));
  return envLookup(e, u, v) * f3(l.r, l.g, l.b) * (1.0f / pdf);
}

// MIS power heuristic (render/path_renderer_stream.h:1077-1080)
PRT_DEV float misWeight(float a, float b) { return sqr(a) * rcpSafe(sqr(a) + sqr(b)); }

}  // namespace prt
