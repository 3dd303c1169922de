// Oracle core: TEST INFRASTRUCTURE ONLY.
//
// A scalar, clean-room C++ restatement of the math ProtoRay's stream path tracer runs per
// path (sampler, camera, warps, distributions, microfacet LUTs, BSDFs, lights, hit
// geometry).  It is the checker the HIP path is compared against and the timed CPU
// baseline; it is never linked into the product library.  Each function cites the
// reference file:line it follows.  Arithmetic is IEEE FP32 (the reference's SIMD rcp /
// rsqrt are approximation + one Newton step, math/simd/vfloat8_avx2.h:227-232,327-342,
// so float results agree with the reference to a few ulp, not bit for bit; the sampler is
// integer work and is bit-exact).
#pragma once

#include <cmath>
#include <memory>
#include <cstdint>
#include <cfloat>
#include <algorithm>

namespace orc {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPosMax = FLT_MAX;
constexpr float kEps = FLT_EPSILON;  // sys/constants.h:71 (numeric_limits<float>::epsilon)

struct V3
{
  float x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(float a) : x(a), y(a), z(a) {}
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
inline V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
inline V3& operator*=(V3& a, float s) { a = a * s; return a; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline float lengthSqr(V3 a) { return dot(a, a); }
// math/vec3.h:399-402 (a * rsqrt(dot(a,a)))
inline V3 normalize(V3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }
inline float reduceMax(V3 a) { return std::max(std::max(a.x, a.y), a.z); }
inline V3 vmax(V3 a, V3 b) { return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
inline V3 vmin(V3 a, V3 b) { return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
inline float sqr(float a) { return a * a; }
inline float clampf(float x, float lo, float hi) { return std::min(std::max(x, lo), hi); }
inline float lerpf(float a, float b, float s) { return a + s * (b - a); }  // math/math.h:89-92
inline float rcpSafe(float x) { return x == 0.f ? kPosMax : 1.0f / x; }  // math/simd.h:107-111
// image/color.h:17-22
inline float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
inline float cos2sin(float x) { return std::sqrt(std::max(1.f - x * x, 0.f)); }  // math/math.h:57-66

struct Basis { V3 U, V, N; };
inline V3 toGlobal(const Basis& b, V3 a) { return a.x * b.U + a.y * b.V + a.z * b.N; }      // math/basis3.h:72-75,139-142
inline V3 toLocal(const Basis& b, V3 a) { return V3(dot(a, b.U), dot(a, b.V), dot(a, b.N)); }  // math/basis3.h:77-80

// math/basis3.h:173-180
inline void makeFrame(V3& U, V3& V, V3 N)
{
  V3 U0(0.f, N.z, -N.y);
  V3 U1(-N.z, 0.f, N.x);
  U = normalize(lengthSqr(U0) > lengthSqr(U1) ? U0 : U1);
  V = normalize(cross(N, U));
}
inline Basis makeBasis(V3 N) { Basis b; makeFrame(b.U, b.V, N); b.N = N; return b; }

// ---------------------------------------------------------------------------------------
// Sobol-Burley sampler (bit-exact integer part)
// ---------------------------------------------------------------------------------------

// sampling/sobol_burley.cpp:9-51 holds the reversed-bit direction numbers of the first four
// Sobol dimensions.  Dimensions 0 and 1 have closed forms (sobol_burley.h:110-128) and only
// dimension 2 is used by a 3D sample, so this restatement generates the dimension-2 column
// from its primitive polynomial (x^2 + x + 1, m = {1, 3}; Joe-Kuo) instead of keeping a table.
inline const uint32_t* sobolDim2Reversed()
{
  static uint32_t t[32];
  static bool init = false;
  if (!init)
  {
    // standard (MSB-first) direction numbers v_k = m_k << (32-k), then bit-reverse
    uint32_t v[33];
    uint32_t m1 = 1, m2 = 3;
    v[1] = m1 << 31;
    v[2] = m2 << 30;
    for (int k = 3; k <= 32; ++k)
      v[k] = v[k - 2] ^ (v[k - 2] >> 2) ^ v[k - 1];  // s = 2, a = 1: m_k = 2 m_{k-1} ^ 4 m_{k-2} ^ m_{k-2}
    for (int k = 1; k <= 32; ++k)
    {
      uint32_t x = v[k], r = 0;
      for (int b = 0; b < 32; ++b) r |= ((x >> b) & 1u) << (31 - b);
      t[k - 1] = r;
    }
    init = true;
  }
  return t;
}

inline uint32_t reverseBits(uint32_t x)  // math/simd.h:140-147
{
  x = ((x & 0xaaaaaaaau) >> 1) | ((x & 0x55555555u) << 1);
  x = ((x & 0xccccccccu) >> 2) | ((x & 0x33333333u) << 2);
  x = ((x & 0xf0f0f0f0u) >> 4) | ((x & 0x0f0f0f0fu) << 4);
  x = ((x & 0xff00ff00u) >> 8) | ((x & 0x00ff00ffu) << 8);
  return (x >> 16) | (x << 16);
}
inline uint32_t hashHP(uint32_t i)  // sampling/sobol_burley.h:31-44
{
  i ^= i >> 16; i *= 0x21f0aaadu; i ^= i >> 15; i *= 0xd35a2d97u; i ^= i >> 15;
  return i ^ 0xe6fe3bebu;
}
inline uint32_t reversedBitOwen(uint32_t n, uint32_t seed)  // sampling/sobol_burley.h:52-62
{
  n ^= n * 0x3d20adeau; n += seed; n *= (seed >> 16) | 1u; n ^= n * 0x05526c56u; n ^= n * 0x53a22864u;
  return n;
}
// math/simd.h:76-84 with the uint32 -> float conversion of the default AVX2 host build
// (math/simd/vfloat8_avx2.h:171-178): the low 31 bits are converted, then 2^31 is added as a
// float, i.e. two roundings.  (The AVX-512 and NEON builds convert in one rounding and can
// differ by one ulp.)
inline float uintToFloatAvx2(uint32_t x)
{
  float lo = float(int32_t(x & 0x7fffffffu));
  float hi = (x & 0x80000000u) ? 2147483648.0f : 0.0f;
  return lo + hi;
}
inline float toFloatUnormExcl(uint32_t x) { return uintToFloatAvx2(x) * (1.0f / 4294967808.0f); }

// sampling/sobol_burley.h:150-190 (SIMD form: bit-serial table walk for dims >= 2)
inline float sobolBurley(uint32_t revBitIndex, uint32_t dim, uint32_t scramble)
{
  uint32_t r = 0;
  if (dim == 0)
    r = reverseBits(revBitIndex);
  else if (dim == 1)
  {
    uint32_t v = revBitIndex;
    v ^= v << 16;
    v ^= (v & 0x00FF00FFu) << 8;
    v ^= (v & 0x0F0F0F0Fu) << 4;
    v ^= (v & 0x33333333u) << 2;
    v ^= (v & 0x55555555u) << 1;
    r = reverseBits(v);
  }
  else
  {
    const uint32_t* tab = sobolDim2Reversed();  // only dimension 2 is reached (3D samples)
    for (int i = 0; revBitIndex != 0; ++i, revBitIndex <<= 1)
      if (revBitIndex & 0x80000000u) r ^= tab[i];
  }
  return toFloatUnormExcl(reverseBits(reversedBitOwen(r, scramble)));
}

// math/hash.h:12-42
inline uint32_t murmurMix(uint32_t h, uint32_t k)
{
  k *= 0xcc9e2d51u; k = (k << 15) | (k >> 17); k *= 0x1b873593u;
  h ^= k; h = ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
  return h;
}
inline uint32_t murmurFinalize(uint32_t h)
{
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

// SobolBurleySamplerSimd<false> (sampling/sobol_burley_sampler_simd.h:36-85)
struct Sampler
{
  uint32_t index = 0, scramble = 0;
  void set(uint32_t seed, uint32_t sampleIndex, uint32_t pixelId)
  {
    index = sampleIndex;
    scramble = murmurFinalize(murmurMix(seed, pixelId));
  }
  float get1D(uint32_t dim) const  // sobol_burley.h:222-238
  {
    uint32_t s = scramble ^ hashHP(dim);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xbff95bfeu);
    return sobolBurley(i, 0, s ^ 0x635c77bdu);
  }
  void get2D(uint32_t set, float& a, float& b) const  // sobol_burley.h:250-267
  {
    uint32_t s = scramble ^ hashHP(set);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xf8ade99au);
    a = sobolBurley(i, 0, s ^ 0xe0aaaf76u);
    b = sobolBurley(i, 1, s ^ 0x94964d4eu);
  }
  void get3D(uint32_t set, float& a, float& b, float& c) const  // sobol_burley.h:279-297
  {
    uint32_t s = scramble ^ hashHP(set);
    uint32_t i = reversedBitOwen(reverseBits(index), s ^ 0xcaa726acu);
    a = sobolBurley(i, 0, s ^ 0x9e78e391u);
    b = sobolBurley(i, 1, s ^ 0x67c33241u);
    c = sobolBurley(i, 2, s ^ 0x78c395c5u);
  }
};

// image/morton.h:14-21: pixel offset of task i inside a tile of width tileSizeX (8x8 Morton blocks)
inline void morton8x8(int i, int tileSizeX, int& x, int& y)
{
  int logTile = 0;
  while ((1 << logTile) < tileSizeX) ++logTile;
  int m = i & 63;
  // de-interleave the 6-bit Morton code (x in even bits, y in odd bits)
  int mx = (m & 1) | ((m >> 1) & 2) | ((m >> 2) & 4);
  int my = ((m >> 1) & 1) | ((m >> 2) & 2) | ((m >> 3) & 4);
  x = mx + ((i >> 3) & (tileSizeX - 8));
  y = my + ((i >> (logTile + 3)) << 3);
}

// ---------------------------------------------------------------------------------------
// Shape warps (sampling/shape_sampler.h)
// ---------------------------------------------------------------------------------------
inline V3 cosineSampleHemisphere(float& pdf, float sx, float sy)  // :69-86
{
  float cosTheta = std::sqrt(sx);
  float sinTheta = std::sqrt(1.f - sx);
  float phi = 2.f * kPi * sy;
  pdf = cosTheta * (1.f / kPi);
  return V3(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
}
inline float cosineSampleHemispherePdf(float c) { return c * (1.f / kPi); }  // :94-98
inline V3 uniformSampleSphere(float& pdf, float sx, float sy)  // :23-40
{
  float cosTheta = 1.f - 2.f * sx;
  float sinTheta = 2.f * std::sqrt(sx * (1.f - sx));
  float phi = 2.f * kPi * sy;
  pdf = 1.f / (4.f * kPi);
  return V3(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
}
inline V3 uniformSampleCone(float& pdf, float sx, float sy, float angle)  // :100-117
{
  float cosTheta = 1.f - sx * (1.f - std::cos(angle));
  float sinTheta = cos2sin(cosTheta);
  float phi = 2.f * kPi * sy;
  pdf = 1.f / (4.f * kPi * sqr(std::sin(0.5f * angle)));
  return V3(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
}
inline float uniformSampleConePdf(float cosTheta, float angle)  // :119-123
{
  return cosTheta < std::cos(angle) ? 0.f : 1.f / (4.f * kPi * sqr(std::sin(0.5f * angle)));
}
inline void uniformSampleDisk(float sx, float sy, float& x, float& y)  // :169-179
{
  float r = std::sqrt(sx);
  float theta = 2.f * kPi * sy;
  x = r * std::cos(theta);
  y = r * std::sin(theta);
}

// ---------------------------------------------------------------------------------------
// Discrete distributions (sampling/distribution1d.{h,cpp}, distribution2d.{h,cpp})
// ---------------------------------------------------------------------------------------
// distribution1d.cpp:18-32: cdf has size+1 entries
inline void dist1dInit(float* cdf, const float* f, int size)
{
  cdf[0] = 0.f;
  for (int i = 1; i < size + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1];
  float norm = (cdf[size] == 0.f) ? 0.f : 1.0f / cdf[size];
  for (int i = 1; i < size + 1; ++i) cdf[i] *= norm;
  cdf[size] = 1.f;
}
// distribution1d.cpp:35-55 (upper bound over cdf[0..size-1], minus one, clamped at 0)
inline int dist1dLookup(const float* cdf, int size, float s)
{
  int first = 0, len = size;
  while (len > 0)
  {
    int half = len >> 1, mid = first + half;
    if (s < cdf[mid]) len = half;
    else { first = mid + 1; len = len - half - 1; }
  }
  return std::max(first - 1, 0);
}
// distribution1d.h:76-83
inline int dist1dSampleReuse(const float* cdf, int size, float& pdf, float& s)
{
  int i = dist1dLookup(cdf, size, s);
  pdf = cdf[i + 1] - cdf[i];
  s = (s - cdf[i]) * (1.0f / pdf);
  return i;
}
inline float dist1dPdf(const float* cdf, int size, int i)  // distribution1d.h:96-101
{
  if (i < 0 || i >= size) return 0.f;
  return cdf[i + 1] - cdf[i];
}
// distribution2d.cpp:16-34: xDists[H][W+1], yDist[H+1]
inline void dist2dInit(float* xDists, float* yDist, const float* f, int W, int H)
{
  float* fy = new float[H];
  for (int y = 0; y < H; ++y)
  {
    fy[y] = 0.f;
    for (int x = 0; x < W; ++x) fy[y] += f[y * W + x];
    dist1dInit(xDists + y * (W + 1), f + y * W, W);
  }
  dist1dInit(yDist, fy, H);
  delete[] fy;
}
// distribution2d.h:53-64
inline void dist2dSampleReuse(const float* xDists, const float* yDist, int W, int H, float& pdf, float& sx, float& sy, int& ix, int& iy)
{
  float yPdf, xPdf;
  iy = dist1dSampleReuse(yDist, H, yPdf, sy);
  ix = dist1dSampleReuse(xDists + iy * (W + 1), W, xPdf, sx);
  pdf = xPdf * yPdf;
}
inline float dist2dPdf(const float* xDists, const float* yDist, int W, int H, int ix, int iy)  // distribution2d.h:79-86
{
  if (iy < 0 || iy >= H) return 0.f;
  return dist1dPdf(xDists + iy * (W + 1), W, ix) * dist1dPdf(yDist, H, iy);
}

// ---------------------------------------------------------------------------------------
// GGX microfacet distribution (bsdf/ggx_distribution.h) and roughness mapping
// ---------------------------------------------------------------------------------------
inline float roughnessToAlpha(float r) { return std::max(sqr(r), 0.001f); }  // bsdf/microfacet.h:10-24 (isotropic)

struct GGX
{
  float ax, ay;
  explicit GGX(float a) : ax(a), ay(a) {}
  float eval(V3 wh) const  // :24-31
  {
    float c2 = sqr(wh.z);
    float e = (sqr(wh.x / ax) + sqr(wh.y / ay)) / c2;
    return 1.0f / (kPi * ax * ay * sqr(c2 * (1.f + e)));
  }
  float lambda(V3 w) const  // :84-90
  {
    float c2 = sqr(w.z);
    float invA2 = (sqr(w.x * ax) + sqr(w.y * ay)) / c2;
    return 0.5f * (-1.f + std::sqrt(1.f + invA2));
  }
  float G1(V3 w) const { return 1.0f / (1.f + lambda(w)); }                  // :92-95
  float G2(V3 wo, V3 wi) const { return 1.0f / (1.f + lambda(wo) + lambda(wi)); }  // :99-102
  float evalVisible(V3 wh, V3 wo, float cosThetaOH, float& pdf) const  // :42-48
  {
    float D = eval(wh);
    pdf = G1(wo) * std::fabs(cosThetaOH) * D / std::fabs(wo.z);
    return D;
  }
  // Heitz 2017 VNDF sampling (:52-78)
  V3 sampleVisible(V3 wo, float& pdf, float sx, float sy) const
  {
    V3 V = normalize(V3(ax * wo.x, ay * wo.y, wo.z));
    V3 T1 = (V.z < 0.9999f) ? normalize(cross(V, V3(0, 0, 1))) : V3(1, 0, 0);
    V3 T2 = cross(T1, V);
    float a = 1.f / (1.f + V.z);
    float r = std::sqrt(sx);
    float phi = (sy < a) ? sy / a * kPi : kPi + (sy - a) / (1.f - a) * kPi;
    float P1 = r * std::cos(phi);
    float P2 = r * std::sin(phi) * ((sy < a) ? 1.f : V.z);
    V3 wh = P1 * T1 + P2 * T2 + std::sqrt(std::max(0.f, 1.f - P1 * P1 - P2 * P2)) * V;
    wh = normalize(V3(ax * wh.x, ay * wh.y, std::max(0.f, wh.z)));
    pdf = G1(wo) * std::fabs(dot(wo, wh)) * eval(wh) / std::fabs(wo.z);
    return wh;
  }
};

// bsdf/optics.h
inline V3 reflect(V3 I, V3 N) { return 2.f * dot(I, N) * N - I; }                // :12-16
inline V3 reflect(V3 I, V3 N, float cosI) { return 2.f * cosI * N - I; }        // :20-24
inline V3 refract(V3 I, V3 N, float cosI, float cosT, float eta) { return eta * (cosI * N - I) - cosT * N; }  // :30-34
inline float fresnelDielectric(float cosI, float cosT, float eta)  // :61-67
{
  float Rper = (eta * cosI - cosT) / (eta * cosI + cosT);
  float Rpar = (cosI - eta * cosT) / (cosI + eta * cosT);
  return 0.5f * (Rpar * Rpar + Rper * Rper);
}
inline float fresnelDielectric(float cosI, float eta)  // :84-99 (SIMD form)
{
  cosI = std::max(cosI, 0.f);
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  if (!(k >= 0.f)) return 1.f;
  return fresnelDielectric(cosI, std::sqrt(k), eta);
}
inline float fresnelDielectricEx(float cosI, float& cosT, float eta)  // :122-139 (SIMD form)
{
  cosT = 0.f;
  cosI = std::fabs(cosI);
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  if (!(k >= 0.f)) return 1.f;
  cosT = std::sqrt(k);
  return fresnelDielectric(cosI, cosT, eta);
}
inline float fresnelDielectricAvg(float eta)  // :144-175
{
  float re = 1.0f / eta;
  if (re >= 1.f) return (re - 1.f) / (4.08567f + 1.00071f * re);
  float re2 = sqr(re), re3 = re2 * re;
  return 0.997118f + 0.1014f * re - 0.965241f * re2 - 0.130607f * re3;
}
inline float fresnelConductor(float cosI, float eta, float k)  // :179-190
{
  cosI = std::fabs(cosI);
  float tmp = sqr(eta) + sqr(k);
  float Rpar = (tmp * sqr(cosI) - 2.f * eta * cosI + 1.f) / (tmp * sqr(cosI) + 2.f * eta * cosI + 1.f);
  float Rper = (tmp - 2.f * eta * cosI + sqr(cosI)) / (tmp + 2.f * eta * cosI + sqr(cosI));
  return 0.5f * (Rpar + Rper);
}
// :214-226 (Gulbrandsen 2014)
inline void artisticToIor(float r, float g, float& eta, float& k)
{
  r = std::min(r, 0.99f);
  float nMin = (1.f - r) / (1.f + r);
  float nMax = (1.f + std::sqrt(r)) / (1.f - std::sqrt(r));
  float n = g * nMin + (1.f - g) * nMax;
  float k2 = (sqr(n + 1.f) * r - sqr(n - 1.f)) / (1.f - r);
  eta = n;
  k = std::sqrt(std::max(k2, 0.f));
}
inline float artisticAvg(float r, float g)  // :228-240 (Kulla-Conty fit)
{
  float r2 = r * r, r3 = r2 * r, g2 = g * g, g3 = g2 * g;
  return 0.087237f + 0.0230685f * g - 0.0864902f * g2 + 0.0774594f * g3
       + 0.782654f * r - 0.136432f * r2 + 0.278708f * r3
       + 0.19744f * g * r + 0.0360605f * g2 * r - 0.2586f * g * r2;
}

// ---------------------------------------------------------------------------------------
// Energy-compensation LUTs (bsdf/albedo_tables.{h,cpp}, bsdf/albedo.h) and lookups
// (math/interpolation.h)
// ---------------------------------------------------------------------------------------
inline float interp1D(float x, const float* f, int size)  // interpolation.h:26-36
{
  float xc = clampf(x, 0.f, float(size - 1));
  float s = xc - std::floor(xc);
  int x0 = std::min(int(xc), size - 1), x1 = std::min(x0 + 1, size - 1);
  return lerpf(f[x0], f[x1], s);
}
inline float interp2D(float px, float py, const float* f, int sx, int sy)  // :77-98
{
  float xc = clampf(px, 0.f, float(sx - 1)), yc = clampf(py, 0.f, float(sy - 1));
  float fx = xc - std::floor(xc), fy = yc - std::floor(yc);
  int x0 = std::min(int(xc), sx - 1), x1 = std::min(x0 + 1, sx - 1);
  int y0 = std::min(int(yc), sy - 1), y1 = std::min(y0 + 1, sy - 1);
  float f0 = lerpf(f[x0 + y0 * sx], f[x1 + y0 * sx], fx);
  float f1 = lerpf(f[x0 + y1 * sx], f[x1 + y1 * sx], fx);
  return lerpf(f0, f1, fy);
}
inline float interp3D(float px, float py, float pz, const float* f, int n)  // :197-227 (cube n^3)
{
  float xc = clampf(px, 0.f, float(n - 1)), yc = clampf(py, 0.f, float(n - 1)), zc = clampf(pz, 0.f, float(n - 1));
  float sx = xc - std::floor(xc), sy = yc - std::floor(yc), sz = zc - std::floor(zc);
  int x0 = std::min(int(xc), n - 1), x1 = std::min(x0 + 1, n - 1);
  int y0 = std::min(int(yc), n - 1), y1 = std::min(y0 + 1, n - 1);
  int z0 = std::min(int(zc), n - 1), z1 = std::min(z0 + 1, n - 1);
  int ny = n, nz = n * n;
  float f00 = lerpf(f[x0 + y0 * ny + z0 * nz], f[x1 + y0 * ny + z0 * nz], sx);
  float f01 = lerpf(f[x0 + y1 * ny + z0 * nz], f[x1 + y1 * ny + z0 * nz], sx);
  float f10 = lerpf(f[x0 + y0 * ny + z1 * nz], f[x1 + y0 * ny + z1 * nz], sx);
  float f11 = lerpf(f[x0 + y1 * ny + z1 * nz], f[x1 + y1 * ny + z1 * nz], sx);
  return lerpf(lerpf(f00, f01, sy), lerpf(f10, f11, sy), sz);
}

struct Luts
{
  float ggxDir[32 * 32];            // MicrofacetAlbedoTable::dir  [roughness][cosThetaO]
  float ggxAvg[32];                 // MicrofacetAlbedoTable::avg  [roughness]
  float dielEtaDir[16 * 16 * 16];   // MicrofacetDielectricAlbedoTable::etaDir [rough][eta][cos], eta in [1/3,1]
  float dielEtaAvg[16 * 16];
  float dielRcpEtaDir[16 * 16 * 16];  // eta in [1,3]
  float dielRcpEtaAvg[16 * 16];
  float dielReflDir[16 * 16 * 16];  // MicrofacetDielectricReflectionAlbedoTable::etaDir, eta in [1/3,1]
  float dielReflAvg[16 * 16];
  float sheenDir[16 * 16];          // MicrofacetSheenAlbedoTable::dir [roughness][cosThetaO]
};

// albedo_tables.h:20-42
inline float ggxAlbedo(const Luts& L, float cosThetaO, float roughness)
{
  return interp2D(cosThetaO * 31.f, roughness * 31.f, L.ggxDir, 32, 32);
}
inline float ggxAlbedoAvg(const Luts& L, float roughness) { return interp1D(roughness * 31.f, L.ggxAvg, 32); }
// albedo_tables.h:82-154 (minIor 1, maxIor 3; rcp(maxIor) evaluated as 1/3)
inline float dielAlbedo(const Luts& L, float cosThetaO, float eta, float roughness)
{
  if (eta <= 1.f)
  {
    const float minEta = 1.0f / 3.0f, maxEta = 1.0f;
    float e = (eta - minEta) / (maxEta - minEta);
    return interp3D(cosThetaO * 15.f, e * 15.f, roughness * 15.f, L.dielEtaDir, 16);
  }
  float e = (eta - 1.f) / (3.f - 1.f);
  return interp3D(cosThetaO * 15.f, e * 15.f, roughness * 15.f, L.dielRcpEtaDir, 16);
}
inline float dielAlbedoAvg(const Luts& L, float eta, float roughness)
{
  if (eta <= 1.f)
  {
    const float minEta = 1.0f / 3.0f, maxEta = 1.0f;
    float e = (eta - minEta) / (maxEta - minEta);
    return interp2D(e * 15.f, roughness * 15.f, L.dielEtaAvg, 16, 16);
  }
  float e = (eta - 1.f) / (3.f - 1.f);
  return interp2D(e * 15.f, roughness * 15.f, L.dielRcpEtaAvg, 16, 16);
}

// MicrofacetDielectricReflectionAlbedoTable::eval / evalAvg (albedo_tables.h:160-205)
inline float dielReflAlbedo(const Luts& L, float cosThetaO, float eta, float roughness)
{
  const float minEta = 1.0f / 3.0f, maxEta = 1.0f;
  float e = (eta - minEta) / (maxEta - minEta);
  return interp3D(cosThetaO * 15.f, e * 15.f, roughness * 15.f, L.dielReflDir, 16);
}
inline float dielReflAlbedoAvg(const Luts& L, float eta, float roughness)
{
  const float minEta = 1.0f / 3.0f, maxEta = 1.0f;
  float e = (eta - minEta) / (maxEta - minEta);
  return interp2D(e * 15.f, roughness * 15.f, L.dielReflAvg, 16, 16);
}

// MicrofacetSheenAlbedoTable::eval (albedo_tables.h:287-291)
inline float sheenAlbedo(const Luts& L, float cosThetaO, float roughness)
{
  return interp2D(cosThetaO * 15.f, roughness * 15.f, L.sheenDir, 16, 16);
}

// Sheen microfiber distribution (bsdf/sheen_distribution.h:10-78; Conty and Kulla 2017), r = alpha
struct SheenDist
{
  float r;
  float D(float cosTheta) const
  {
    float inv = 1.0f / r;
    return (2.f + inv) * std::pow(cos2sin(cosTheta), inv) * (1.f / (2.f * kPi));
  }
  float L(float x) const  // the fitted exponent of Lambda
  {
    float w0 = sqr(1.f - r), w1 = 1.f - w0;
    float a = w0 * 25.3245f + w1 * 21.5473f, b = w0 * 3.32435f + w1 * 3.82987f, c = w0 * 0.16801f + w1 * 0.19823f;
    float d = w0 * -1.27393f + w1 * -1.97760f, e = w0 * -4.85967f + w1 * -4.32054f;
    return a / (1.f + b * std::pow(x, c)) + d * x + e;
  }
  float lambda(float c) const { return c < 0.5f ? std::exp(L(c)) : std::exp(2.f * L(0.5f) - L(1.f - c)); }
  float lambdaI(float c) const { return std::pow(lambda(c), 1.f + 2.f * sqr(sqr(sqr(1.f - c)))); }  // softer terminator
  float G2(float cosO, float cosI) const { return 1.0f / (1.f + lambda(cosO) + lambdaI(cosI)); }
};

void precomputeLuts(Luts& L);  // orc_luts.cpp

// ---------------------------------------------------------------------------------------
// Materials and BSDFs
// ---------------------------------------------------------------------------------------
namespace BsdfType {
enum : int {
  DiffuseReflection = 1 << 0, GlossyReflection = 1 << 1, SpecularReflection = 1 << 2,
  DiffuseTransmission = 1 << 3, GlossyTransmission = 1 << 4, SpecularTransmission = 1 << 5,
  UnbentTransmission = 1 << 6, Transparency = 1 << 7,
  Diffuse = DiffuseReflection | DiffuseTransmission, Glossy = GlossyReflection | GlossyTransmission,
  Specular = SpecularReflection | SpecularTransmission, Continuation = UnbentTransmission | Transparency,
  Smooth = Diffuse | Glossy, Delta = Specular | Continuation,
};
}  // bsdf/bsdf.h:15-44

// Material kinds supported on the stream path (uniform StandardMaterial parameters,
// material/standard.cpp:409-573, each producing a MultiBsdf with one child of weight 1):
//   0 diffuse    : baseColor, specularWeight = 0            -> DiffuseBrdf
//   1 conductor  : metalness = 1, baseColor = r, specularColor = g, specularRoughness
//                                                           -> RoughConductorBrdf<FresnelConductorArtistic>
//   2 dielectric : transmissionWeight = 1, specularIOR, specularRoughness
//                                                           -> RoughDielectricBsdf + media interface
//   4 thin dielectric : thinWalled, transmissionWeight = 1, specularIOR, specularRoughness,
//                      transmissionColor = r, thickness / transmissionDepth = thickness
//                                                           -> ThinRoughDielectricBsdf
// Any kind: transparency = 1 - opacity mixes in a TransparentBtdf; diffuseRoughness (kinds 0
// and 3) selects the Oren-Nayar base.
enum MatKind : int { kDiffuse = 0, kConductor = 1, kDielectric = 2, kCoated = 3, kThinDielectric = 4, kStandard = 5 };
constexpr int kTransparentLobe = 6;  // the transparent BTDF as a base of kind 5 (never a material kind)

struct Material  // 148-byte record, the prt_material layout of the HIP path
{
  int kind;
  float r, g, b;     // albedo (diffuse) / reflectivity (conductor)
  float er, eg, eb;  // edge tint (conductor)
  float roughness;   // specularRoughness (alpha = max(roughness^2, 0.001))
  float ior;         // specularIOR (dielectric) / coat IOR (coated), relative to an ambient IOR of 1
  float thickness;   // coated: coat thickness (absorption path scale)
  float weight;      // coated: coat weight (specularWeight / coatWeight)
  float emission[3]; // emissionColor * emissionLuminance; non-zero makes the material an area light
  float transparency;      // 1 - opacity
  float diffuseRoughness;  // Oren-Nayar base of kinds 0 and 3 when > 0
  float sheenWeight, sheenColor[3], sheenRoughness;  // the sheen layer
  // kinds 2 and 5: the interior medium's colour = transmissionColor ^ transmissionDepth
  // (material/standard.cpp:169-173); depth 0 leaves it clear
  float transmissionColor[3], transmissionDepth;
  // kind 5: the StandardMaterial parameters no single-lobe kind reads (standard.cpp:27-81)
  float baseWeight, metalness, specularWeight, transmissionWeight;
  float coatWeight, coatColor[3], coatRoughness, coatIor, coatThickness;
  int thinWalled;
};
static_assert(sizeof(Material) == 148, "the prt_material layout");
// MaterialType::Transparent (material/standard.cpp:152-166): shadow rays pass through
inline bool isTransparent(const Material& m)
{
  if (m.transparency != 0.f || m.kind == kThinDielectric) return true;
  return m.kind == kStandard && m.transmissionWeight != 0.f && m.thinWalled;
}

// MaterialType::MediaInterface (material/standard.cpp:159-175): a transmissive surface that is not
// thin-walled separates the ambient medium from its interior one.
inline bool isMediaInterface(const Material& m)
{
  if (m.kind == kDielectric) return true;
  return m.kind == kStandard && m.transmissionWeight != 0.f && !m.thinWalled;
}
// The IOR of that interface, as the medium records it (>= 1).
inline float interfaceIor(const Material& m) { return m.ior < 1.f ? 1.0f / m.ior : m.ior; }

inline bool isEmissive(const Material& m) { return m.emission[0] > 0.f || m.emission[1] > 0.f || m.emission[2] > 0.f; }

struct Ctx  // core/shading_context.h:113-124 subset
{
  V3 p;
  Basis f;  // geometric shading frame from postIntersect (+ orientation flip)
  V3 Ng;
  bool backfacing;
  float eps;
};

// core/shading_context.h:144-191: shading normal adaption + Gram-Schmidt frame
inline Basis makeShadingFrame(const Ctx& ctx, V3 wo)
{
  Basis fr;
  V3 N = ctx.f.N;
  V3 wi = reflect(wo, N);
  if (dot(wi, ctx.Ng) < 0.f)
  {
    float a = -dot(wi, ctx.Ng);
    float b = dot(N, ctx.Ng);
    if (b < 0.f) { b = -b; wi = -wi; }
    V3 wi1 = normalize(wi + a * rcpSafe(b) * N + 1e-5f * ctx.Ng);
    N = normalize(wo + wi1);
  }
  fr.N = N;
  fr.U = ctx.f.U - dot(ctx.f.U, N) * N;
  float Ulen2 = lengthSqr(fr.U);
  fr.U = fr.U * (1.0f / std::sqrt(Ulen2));
  fr.V = ctx.f.V - dot(ctx.f.V, N) * N;
  fr.V = fr.V - dot(fr.V, fr.U) * fr.U;
  float Vlen2 = lengthSqr(fr.V);
  fr.V = fr.V * (1.0f / std::sqrt(Vlen2));
  if (Ulen2 == 0.f || Vlen2 == 0.f) makeFrame(fr.U, fr.V, fr.N);
  return fr;
}

struct StdLobes;

// Per-hit BSDF instance (MultiBsdf{child, weight 1}, or for kind 5 the weighted mix of
// StandardMaterial's bases in `mix`, under a coat when `coat` is set).
struct Bsdf
{
  int kind;
  int flags;
  Basis fr;
  V3 color;            // diffuse
  // conductor
  V3 eta, k, rr, gg;
  float Eavg;
  V3 fmsScale;
  // dielectric
  float deta, EavgEta, EavgRcpEta, fmsRatio;
  // coated diffuse: base = color (diffuse), coat transmittance color, thickness, weight (scale),
  // eta <= 1, Ravg = average coat reflectance * scale, importance of the diffuse base
  V3 coatColor;
  float thickness, scale, ceta, Ravg, baseImportance;
  // thin dielectric: color = transmission color, eta <= 1, GGX Eavg, fmsScale, thickness
  float teta, tEavg, tfms, tthick;
  float diffRough;  // Oren-Nayar base when > 0
  int mix;          // 0 the lobe, 1 lobe x opacity + transparent BTDF x transp, 2 transparent only
  float opacity, transp;
  bool sheen;       // a sheen layer over the lobe
  V3 sheenColor;
  float sheenScale, sheenRough;
  SheenDist sheenDist{0.001f};
  float roughness;
  GGX ggx{0.001f};
  float importance;
  // kind 5: the bases (mix) and whether a coat (coatColor, thickness, scale, ceta, Ravg, roughness,
  // ggx above, baseImportance = the mix's importance) lies over them
  std::shared_ptr<const StdLobes> mixLobes;
  bool coat = false;
};

// kind 5: StandardMaterial's bases as one MultiBsdfSimd (multi_bsdf.h:151-187): up to three
// single-lobe children (the opaque base, the translucent base, the metal base) with their scales
// and importances (child importance x scale).
struct StdLobes
{
  int n = 0;
  Bsdf child[3];
  float scale[3] = {0.f, 0.f, 0.f}, imp[3] = {0.f, 0.f, 0.f};
  float impSum = 0.f;
  int flags = 0;
  void add(const Bsdf& c, float s)
  {
    child[n] = c;
    scale[n] = s;
    imp[n] = c.importance * s;
    impSum += imp[n];
    flags |= c.flags;
    ++n;
  }
};

// The single lobes (each the constructor of its reference class for the given parameters).
inline void initDiffuse(Bsdf& b, V3 color, float diffRough)
{
  b.kind = kDiffuse;
  b.color = color;
  b.diffRough = diffRough;
  b.flags = BsdfType::DiffuseReflection;
  b.importance = reduceMax(b.color);  // bsdf/diffuse.h:72-75
}
inline void initConductor(const Luts& L, Bsdf& b, V3 rr, V3 gg, float roughness)
{
  // rough_conductor.h:121-130 + optics.h:243-261
  b.kind = kConductor;
  b.roughness = roughness;
  b.ggx = GGX(roughnessToAlpha(roughness));
  b.rr = rr;
  b.gg = gg;
  artisticToIor(b.rr.x, b.gg.x, b.eta.x, b.k.x);
  artisticToIor(b.rr.y, b.gg.y, b.eta.y, b.k.y);
  artisticToIor(b.rr.z, b.gg.z, b.eta.z, b.k.z);
  b.Eavg = ggxAlbedoAvg(L, roughness);
  V3 Favg(artisticAvg(b.rr.x, b.gg.x), artisticAvg(b.rr.y, b.gg.y), artisticAvg(b.rr.z, b.gg.z));
  V3 one(1.f);
  V3 den = one - Favg * (one - V3(b.Eavg));
  b.fmsScale = V3(sqr(Favg.x) * b.Eavg / den.x, sqr(Favg.y) * b.Eavg / den.y, sqr(Favg.z) * b.Eavg / den.z);
  b.flags = BsdfType::GlossyReflection;
  b.importance = 1.f;
}
inline void initDielectric(const Luts& L, Bsdf& b, float eta, float roughness)
{
  // rough_dielectric.h:279-291
  b.kind = kDielectric;
  b.roughness = roughness;
  b.ggx = GGX(roughnessToAlpha(roughness));
  b.deta = eta;
  b.EavgEta = dielAlbedoAvg(L, eta, roughness);
  b.EavgRcpEta = dielAlbedoAvg(L, 1.0f / eta, roughness);
  float FavgEta = fresnelDielectricAvg(eta);
  float FavgRcpEta = fresnelDielectricAvg(1.0f / eta);
  float a = (1.f - FavgEta) / (1.f - b.EavgRcpEta);
  float bb = (1.f - FavgRcpEta) / (1.f - b.EavgEta);
  float x = bb / (a + bb);
  b.fmsRatio = 1.f - x * (1.f - FavgEta);
  b.flags = BsdfType::GlossyReflection | BsdfType::GlossyTransmission;
  b.importance = 1.f;
}
// RoughCoatingBsdf over a diffuse base; etaArg as the constructor receives it (rough_coating.h:
// 287-295: the member eta is folded to <= 1, Ravg uses the argument as given)
inline void initCoated(const Luts& L, Bsdf& b, V3 color, float diffRough, V3 coatColor, float thickness, float scale,
                       float etaArg, float roughness)
{
  b.kind = kCoated;
  b.roughness = roughness;
  b.ggx = GGX(roughnessToAlpha(roughness));
  b.color = color;
  b.diffRough = diffRough;
  b.coatColor = coatColor;
  b.thickness = thickness;
  b.scale = scale;
  b.Ravg = dielReflAlbedoAvg(L, etaArg, roughness) * b.scale;
  b.ceta = etaArg <= 1.f ? etaArg : 1.0f / etaArg;
  b.baseImportance = reduceMax(b.color);  // bsdf/diffuse.h:70-73
  b.flags = BsdfType::DiffuseReflection | BsdfType::GlossyReflection;
  b.importance = std::max(1.f, b.baseImportance);  // rough_coating.h:562-565
}
inline void initThin(const Luts& L, Bsdf& b, float eta, V3 color, float thickness, float roughness)
{
  // thin_rough_dielectric.h:180-192
  b.kind = kThinDielectric;
  b.roughness = roughness;
  b.ggx = GGX(roughnessToAlpha(roughness));
  b.teta = eta;
  b.color = color;
  b.tthick = thickness;
  b.tEavg = ggxAlbedoAvg(L, roughness);
  float Favg = fresnelDielectricAvg(b.teta);
  b.tfms = Favg * (1.f - b.tEavg) / (1.f - Favg * b.tEavg);
  b.flags = BsdfType::GlossyReflection | BsdfType::UnbentTransmission;
  b.importance = 1.f;
}
inline void initTransparent(Bsdf& b)  // bsdf/transparent.h:26-71
{
  b.kind = kTransparentLobe;
  b.flags = BsdfType::Transparency;
  b.importance = 1.f;
}

inline V3 clamp01(float r, float g, float b) { return V3(clampf(r, 0.f, 1.f), clampf(g, 0.f, 1.f), clampf(b, 0.f, 1.f)); }
inline float foldIor(float ior) { return ior < 1.f ? 1.0f / ior : ior; }

// StandardMaterial's lobes for kind 5 (material/standard.cpp:409-611, ambient IOR 1): the opaque,
// translucent and metal bases in one weighted mix, then the coat over the mix.
inline void initStandard(const Luts& L, Bsdf& b, const Material& m, bool fromOutside)
{
  const float eps = 1e-5f;  // material/standard.h:79
  auto lobes = std::make_shared<StdLobes>();
  const float base = clampf(m.baseWeight, 0.f, 1.f);
  const V3 baseColor = clamp01(m.r, m.g, m.b);
  const float metalness = clampf(m.metalness, 0.f, 1.f), specular = clampf(m.specularWeight, 0.f, 1.f);
  const float specRough = clampf(m.roughness, 0.f, 1.f), specIor = foldIor(std::max(m.ior, 0.f));
  const float diffRough = clampf(m.diffuseRoughness, 0.f, 1.f);
  const float dielectric = 1.f - metalness;
  if (dielectric > eps)
  {
    const float transmission = clampf(m.transmissionWeight, 0.f, 1.f);
    const float eta = clampf(fromOutside ? 1.0f / specIor : specIor, 1.f / 3.f, 3.f);
    const float opaque = dielectric * (1.f - transmission);
    if (opaque > eps)  // (:428-505): diffuse, under a specular coating where the IOR differs
    {
      Bsdf c;
      if (std::fabs(specIor - 1.f) > eps && specular > eps)
        initCoated(L, c, baseColor * base, diffRough, V3(1.f), 1.f, specular, eta, specRough);
      else
        initDiffuse(c, baseColor * base, diffRough);
      lobes->add(c, opaque);
    }
    const float translucent = dielectric * transmission * specular;
    if (translucent > eps)  // (:507-530)
    {
      Bsdf c;
      if (std::fabs(specIor - 1.f) > eps)
      {
        if (m.thinWalled)
          initThin(L, c, eta, clamp01(m.transmissionColor[0], m.transmissionColor[1], m.transmissionColor[2]),
                   std::max(m.thickness, 0.f) / std::max(m.transmissionDepth, 0.f), specRough);
        else
          initDielectric(L, c, eta, specRough);
      }
      else
        initTransparent(c);
      lobes->add(c, translucent);
    }
  }
  const float metal = metalness * specular;
  if (metal > eps)  // (:533-546)
  {
    Bsdf c;
    initConductor(L, c, baseColor * base, clamp01(m.er, m.eg, m.eb), specRough);
    lobes->add(c, metal);
  }
  for (int i = 0; i < lobes->n; ++i) lobes->child[i].fr = b.fr;
  b.kind = kStandard;
  b.flags = lobes->flags;
  b.importance = lobes->impSum;
  b.baseImportance = lobes->impSum;
  b.coat = false;
  const float coat = clampf(m.coatWeight, 0.f, 1.f), coatIor = foldIor(std::max(m.coatIor, 0.f));
  if (coat > eps && std::fabs(coatIor - 1.f) > eps)  // (:551-566)
  {
    const float etaArg = clampf(fromOutside ? 1.0f / coatIor : coatIor, 1.f / 3.f, 3.f);
    const float rough = clampf(m.coatRoughness, 0.f, 1.f);
    b.coat = true;
    b.roughness = rough;
    b.ggx = GGX(roughnessToAlpha(rough));
    b.coatColor = clamp01(m.coatColor[0], m.coatColor[1], m.coatColor[2]);
    b.thickness = std::max(m.coatThickness, 0.f);
    b.scale = coat;
    b.Ravg = dielReflAlbedoAvg(L, etaArg, rough) * b.scale;
    b.ceta = etaArg <= 1.f ? etaArg : 1.0f / etaArg;
    b.flags |= BsdfType::GlossyReflection;
    b.importance = std::max(1.f, lobes->impSum);  // rough_coating.h:562-565
  }
  b.mixLobes = lobes;
}

// baseColor: a textured baseColor parameter's value at the hit, clamp(texture(uv) * (r, g, b))
// (material/material_param.h:37-41, material/standard.cpp:424), for kinds 0 and 3; null: the
// record's uniform (r, g, b).
// ---------------------------------------------------------------------------------------
// Image textures (image/image_texture.h:148-227): texels decoded to linear RGB, stored bottom
// row first (image_texture.h:352-353); bilinear lookup with a half-texel offset and repeat or
// clamp addressing per axis.
// ---------------------------------------------------------------------------------------
struct Texture
{
  int w = 0, h = 0;
  bool repeatU = true, repeatV = true;
  float scale[2] = {1.f, 1.f}, offset[2] = {0.f, 0.f};
  std::vector<V3> texels;  // [y * w + x], y from the bottom
};

// clamp(x, lo, hi) = min(select(lo < x, x, lo), hi) (math/math.h:83-86)
inline float clampSel(float x, float lo, float hi) { return std::min(lo < x ? x : lo, hi); }

inline V3 textureGet(const Texture& t, float u, float v)
{
  // uv = ctx.uv * uvScale + uvOffset (image_texture.h:266-269), then get(uv) (:186-227)
  const float su = u * t.scale[0] + t.offset[0], sv = v * t.scale[1] + t.offset[1];
  const float x = su - 0.5f / float(t.w), y = sv - 0.5f / float(t.h);
  const float xr = t.repeatU ? (x - std::floor(x)) * float(t.w) : clampSel(x * float(t.w), 0.f, float(t.w - 1));
  const float yr = t.repeatV ? (y - std::floor(y)) * float(t.h) : clampSel(y * float(t.h), 0.f, float(t.h - 1));
  const float sx = xr - std::floor(xr), sy = yr - std::floor(yr);
  const int x0 = std::min(std::max(int(xr), 0), t.w - 1), y0 = std::min(std::max(int(yr), 0), t.h - 1);
  const int x1 = t.repeatU ? (x0 + 1 == t.w ? 0 : x0 + 1) : std::min(x0 + 1, t.w - 1);
  const int y1 = t.repeatV ? (y0 + 1 == t.h ? 0 : y0 + 1) : std::min(y0 + 1, t.h - 1);
  auto at = [&](int xx, int yy) { return t.texels[size_t(yy) * t.w + xx]; };
  auto lerp = [](V3 a, V3 b, float s) { return a + s * (b - a); };  // math/math.h:88-92
  const V3 c0 = lerp(at(x0, y0), at(x1, y0), sx);
  const V3 c1 = lerp(at(x0, y1), at(x1, y1), sx);
  return lerp(c0, c1, sy);
}

inline Bsdf makeBsdf(const Luts& L, const Material& m, const Ctx& ctx, V3 wo, int mediumId, int mediumIntId,
                     const V3* baseColor = nullptr)
{
  Bsdf b;
  b.kind = m.kind;
  b.fr = makeShadingFrame(ctx, wo);
  float roughness = clampf(m.roughness, 0.f, 1.f);
  b.roughness = roughness;
  b.ggx = GGX(roughnessToAlpha(roughness));
  b.diffRough = clampf(m.diffuseRoughness, 0.f, 1.f);
  // opacity (material/standard.cpp:412-420, :613-627; eps = 1e-5, standard.h:79)
  b.opacity = clampf(1.f - m.transparency, 0.f, 1.f);
  b.transp = 1.f - b.opacity;
  b.mix = b.transp > 1e-5f ? (b.opacity > 1e-5f ? 1 : 2) : 0;
  b.teta = b.tEavg = b.tfms = b.tthick = 0.f;
  // sheen (material/standard.cpp:599-611; sheenTint 0 keeps sheenColor)
  b.sheenScale = clampf(m.sheenWeight, 0.f, 1.f);
  b.sheen = b.sheenScale > 1e-5f;
  b.sheenColor = V3(clampf(m.sheenColor[0], 0.f, 1.f), clampf(m.sheenColor[1], 0.f, 1.f), clampf(m.sheenColor[2], 0.f, 1.f));
  b.sheenRough = clampf(m.sheenRoughness, 0.f, 1.f);
  b.sheenDist = SheenDist{roughnessToAlpha(b.sheenRough)};
  // the media interface's orientation: backfacing = (mediumId == mediumIntId), other materials
  // are always seen from outside (material/material.h:54-61); thin-walled surfaces too
  // (material/standard.cpp:429)
  const bool backfacing = isMediaInterface(m) && (mediumId == mediumIntId);
  const V3 base = baseColor ? *baseColor : clamp01(m.r, m.g, m.b);
  if (m.kind == kDiffuse)
    initDiffuse(b, base, b.diffRough);
  else if (m.kind == kConductor)
    initConductor(L, b, clamp01(m.r, m.g, m.b), clamp01(m.er, m.eg, m.eb), roughness);
  else if (m.kind == kDielectric)  // material/standard.cpp:436-437 (fromOutside = !backfacing)
    initDielectric(L, b, clampf(!backfacing ? 1.0f / foldIor(m.ior) : foldIor(m.ior), 1.f / 3.f, 3.f), roughness);
  else if (m.kind == kCoated)  // material/standard.cpp:512-518, 580-597: never a media interface
    initCoated(L, b, base, b.diffRough, clamp01(m.er, m.eg, m.eb), std::max(m.thickness, 0.f),
               clampf(m.weight, 0.f, 1.f), clampf(1.0f / foldIor(m.ior), 1.f / 3.f, 3.f), roughness);
  else if (m.kind == kThinDielectric)
    initThin(L, b, clampf(1.0f / foldIor(m.ior), 1.f / 3.f, 3.f), clamp01(m.r, m.g, m.b), std::max(m.thickness, 0.f),
             roughness);
  else
    initStandard(L, b, m, m.thinWalled || !backfacing);
  if (b.sheen) b.flags |= BsdfType::DiffuseReflection;  // sheen.h:210-211
  if (b.mix == 1) b.flags |= BsdfType::Transparency;
  if (b.mix == 2) b.flags = BsdfType::Transparency;
  return b;
}

// --- diffuse (bsdf/diffuse.h:55-68), or the Oren-Nayar BRDF in Fujii's form when diffRough > 0
// (bsdf/rough_diffuse.h:65-87)
inline V3 diffuseEval(const Bsdf& b, V3 wo, V3 wi, float& pdf)
{
  float c = std::max(dot(wi, b.fr.N), 0.f);
  pdf = cosineSampleHemispherePdf(c);
  if (!(b.diffRough > 0.f)) return b.color * (1.f / kPi) * c;
  float cO = std::max(dot(wo, b.fr.N), 0.f);
  float s = dot(wo, wi) - cO * c;
  float sInvt = s > 0.f ? s * rcpSafe(std::max(cO, c)) : s;
  return b.color * (c * (1.f + b.diffRough * sInvt) * (1.0f / (kPi + (kPi / 2.f - 2.f / 3.f) * b.diffRough)));
}
inline V3 diffuseSample(const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy)
{
  wi = toGlobal(b.fr, cosineSampleHemisphere(pdf, sx, sy));
  if (!(dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  type = BsdfType::DiffuseReflection;
  if (!(b.diffRough > 0.f)) return b.color;
  V3 v = diffuseEval(b, wo, wi, pdf);
  return v * (1.0f / pdf);
}

// --- rough conductor (bsdf/rough_conductor.h:132-202)
inline V3 fresnelArtistic(const Bsdf& b, float c)
{
  return V3(fresnelConductor(c, b.eta.x, b.k.x), fresnelConductor(c, b.eta.y, b.k.y), fresnelConductor(c, b.eta.z, b.k.z));
}
inline V3 conductorEval(const Luts& L, const Bsdf& b, V3 wo, V3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N), cosI = dot(wi, b.fr.N);
  if (!(cosO > 0.f && cosI > 0.f)) return V3(0.f);
  V3 wh = normalize(wi + wo);
  float cosOH = dot(wo, wh);
  V3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  V3 F = fresnelArtistic(b, cosOH);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  float Eo = ggxAlbedo(L, cosO, b.roughness), Ei = ggxAlbedo(L, cosI, b.roughness);
  V3 fms = b.fmsScale * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
  pdf = whPdf * (1.0f / (4.f * std::fabs(cosOH)));
  return F * (D * G * (1.0f / (4.f * cosO))) + fms;
}
inline V3 conductorSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  V3 wo0 = toLocal(b.fr, wo);
  float whPdf;
  V3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
  wi = reflect(wo, wh);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI > 0.f && dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  float cosOH = dot(wo, wh);
  V3 wi0 = toLocal(b.fr, wi);
  V3 F = fresnelArtistic(b, cosOH);
  float G = b.ggx.G2(wo0, wi0);
  float Eo = ggxAlbedo(L, cosO, b.roughness), Ei = ggxAlbedo(L, cosI, b.roughness);
  V3 fms = b.fmsScale * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.Eavg))) * cosI);
  type = BsdfType::GlossyReflection;
  pdf = whPdf * (1.0f / (4.f * std::fabs(cosOH)));
  return F * (G * rcpSafe(b.ggx.G1(wo0))) + fms * (1.0f / pdf);
}

// --- rough dielectric (bsdf/rough_dielectric.h:296-526)
inline float dielEvalSingle(const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return 0.f;
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return 0.f;
  bool refl = cosI > 0.f;
  V3 wh = refl ? wi + wo : b.deta * wo + wi;
  wh = normalize(wh);
  if (dot(wh, b.fr.N) < 0.f) wh = -wh;
  float cosOH = dot(wo, wh), cosIH = dot(wi, wh);
  float F = fresnelDielectric(cosOH, b.deta);
  float value;
  if (refl)
  {
    pdf = F * (1.0f / (4.f * std::fabs(cosOH)));
    value = F * (1.0f / (4.f * cosO));
  }
  else
  {
    pdf = (1.f - F) * (1.0f / sqr(b.deta)) * std::fabs(cosIH) * (1.0f / sqr(cosOH + (1.0f / b.deta) * cosIH));
    value = pdf * std::fabs(cosOH) * (1.0f / cosO);
  }
  V3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  pdf = pdf * whPdf;
  return value * D * G;
}
inline void dielFms(const Luts& L, const Bsdf& b, float Eo, float cosI, float& fmsPdf, float& fmsValue)
{
  if (cosI > 0.f)
  {
    float Ei = dielAlbedo(L, cosI, b.deta, b.roughness);
    fmsPdf = b.fmsRatio * cosineSampleHemispherePdf(cosI);
    fmsValue = b.fmsRatio * (1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.EavgEta))) * cosI;
  }
  else
  {
    float Ei = dielAlbedo(L, std::fabs(cosI), 1.0f / b.deta, b.roughness);
    fmsPdf = (1.f - b.fmsRatio) * cosineSampleHemispherePdf(std::fabs(cosI));
    fmsValue = (1.f - b.fmsRatio) * (1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.EavgRcpEta))) * std::fabs(cosI);
  }
}
inline V3 dielEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  float Eo = dielAlbedo(L, cosO, b.deta, b.roughness);
  float fmsPdf, fmsValue;
  dielFms(L, b, Eo, cosI, fmsPdf, fmsValue);
  float singlePdf;
  float singleValue = dielEvalSingle(b, ctx, wo, wi, singlePdf);
  pdf = Eo * singlePdf + (1.f - Eo) * fmsPdf;
  return V3(singleValue + fmsValue);
}
inline V3 dielSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float Eo = dielAlbedo(L, cosO, b.deta, b.roughness);
  bool ok;
  if (sc <= Eo)
  {
    V3 wo0 = toLocal(b.fr, wo);
    float whPdf;
    V3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    float cosOH = dot(wo, wh);
    float cosTH;
    float F = fresnelDielectricEx(cosOH, cosTH, b.deta);
    float sc1 = sc * (1.0f / Eo);
    if (sc1 <= F)
    {
      wi = reflect(wo, wh, cosOH);
      ok = dot(wi, b.fr.N) > 0.f;
    }
    else
    {
      wi = refract(wo, wh, cosOH, cosTH, b.deta);
      ok = dot(wi, b.fr.N) < 0.f;
    }
  }
  else
  {
    float sc1 = (sc - Eo) * (1.0f / (1.f - Eo));
    float cpdf;
    V3 dir = toGlobal(b.fr, cosineSampleHemisphere(cpdf, sx, sy));
    wi = (sc1 <= b.fmsRatio) ? dir : -dir;
    ok = true;
  }
  if (!ok) return V3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  type = cosI > 0.f ? BsdfType::GlossyReflection : BsdfType::GlossyTransmission;
  float fmsPdf, fmsValue;
  dielFms(L, b, Eo, cosI, fmsPdf, fmsValue);
  float singlePdf;
  float singleValue = dielEvalSingle(b, ctx, wo, wi, singlePdf);
  pdf = Eo * singlePdf + (1.f - Eo) * fmsPdf;
  return V3((singleValue + fmsValue) * (1.0f / pdf));
}

// --- coated diffuse (bsdf/rough_coating.h:297-593, RoughCoatingBsdfSimd over DiffuseBrdfSimd)
inline float refractCos(float cosI, float eta)  // optics.h:50-56
{
  float k = 1.f - eta * eta * (1.f - cosI * cosI);
  return std::sqrt(std::max(k, 0.f));
}
// absorption through the coat and energy conservation (:331-350): Kelemen's term for a base with
// diffuse lobes only, min(1 - Ro, 1 - Ri) for any other base [Kulla and Conty 2017]
inline V3 coatBase(const Luts& L, const Bsdf& b, V3 baseValue, float cosO, float cosI, float Ro, bool diffuseOnly = true)
{
  float length = 1.0f / refractCos(cosO, b.ceta) + 1.0f / refractCos(std::fabs(cosI), b.ceta);
  if (cosI <= 0.f) length *= 0.5f;
  const float t = b.thickness * length;
  V3 absorbed(baseValue.x * std::pow(b.coatColor.x, t), baseValue.y * std::pow(b.coatColor.y, t),
               baseValue.z * std::pow(b.coatColor.z, t));
  baseValue = baseValue + b.scale * (absorbed - baseValue);
  float Ri = dielReflAlbedo(L, std::fabs(cosI), b.ceta, b.roughness) * b.scale;
  float T = diffuseOnly ? (1.f - Ro) * (1.f - Ri) * (1.0f / (1.f - b.Ravg)) : std::min(1.f - Ro, 1.f - Ri);
  return baseValue * T;
}
inline float coatPickProb(const Bsdf& b, float Ro)
{
  float coatImportance = Ro;
  float baseImportance = (1.f - Ro) * b.baseImportance;
  return coatImportance * (1.0f / (coatImportance + baseImportance));
}
// the coat's specular reflection at (wo, wi) around the half vector wh (:365-378)
inline float coatReflection(const Bsdf& b, V3 wo, V3 wi, V3 wh, float cosO, float& coatPdf)
{
  float cosOH = dot(wo, wh);
  float F = fresnelDielectric(cosOH, b.ceta) * b.scale;
  V3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  coatPdf = whPdf * (1.0f / (4.f * std::fabs(cosOH)));
  return F * D * G * (1.0f / (4.f * cosO));
}
// RoughCoatingBsdfSimd::eval over a base (:297-378); Base: eval(wo, wi, pdf), sample(wo, wi, pdf,
// type, sx, sy, sc), diffuseOnly()
template <class Base>
inline V3 coatEvalOver(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf, const Base& base)
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  float basePdf;
  V3 baseValue = base.eval(wo, wi, basePdf);
  float Ro = dielReflAlbedo(L, cosO, b.ceta, b.roughness) * b.scale;
  baseValue = coatBase(L, b, baseValue, cosO, cosI, Ro, base.diffuseOnly());
  float coatPick = coatPickProb(b, Ro), basePick = 1.f - coatPick;
  if (cosI > 0.f)
  {
    float coatPdf;
    float coatValue = coatReflection(b, wo, wi, normalize(wo + wi), cosO, coatPdf);
    pdf = coatPick * coatPdf + basePick * basePdf;
    return V3(coatValue) + baseValue;
  }
  pdf = basePdf * basePick;
  return baseValue;
}
// RoughCoatingBsdfSimd::sample over a base (:380-537)
template <class Base>
inline V3 coatSampleOver(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy,
                         float sc, const Base& base)
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  V3 wo0 = toLocal(b.fr, wo);
  float Ro = dielReflAlbedo(L, cosO, b.ceta, b.roughness) * b.scale;
  wi = V3(0.f);
  pdf = 0.f;
  type = 0;
  V3 wh(0.f), baseValue(0.f);
  float basePdf = 0.f;
  float coatPick = coatPickProb(b, Ro), basePick = 1.f - coatPick;
  if (sc <= coatPick)
  {
    float whPdf;
    wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    wi = reflect(wo, wh, dot(wo, wh));
    if (!(dot(wi, b.fr.N) > 0.f)) return V3(0.f);
    type = BsdfType::GlossyReflection;
    baseValue = base.eval(wo, wi, basePdf);
  }
  else
  {
    V3 baseWi;
    float basePdf2 = 0.f;
    int baseType = 0;
    V3 baseWeight = base.sample(wo, baseWi, basePdf2, baseType, sx, sy, (sc - coatPick) * (1.0f / basePick));
    if (!(reduceMax(baseWeight) > 0.f)) return V3(0.f);
    wi = baseWi;
    basePdf = basePdf2;
    type = baseType;
    baseValue = baseWeight * ((type & BsdfType::Delta) ? 1.f : basePdf);
    wh = normalize(wo + wi);
  }
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  baseValue = coatBase(L, b, baseValue, cosO, cosI, Ro, base.diffuseOnly());
  if (type & BsdfType::Delta)  // a delta base sample: no coat reflection to add (:484-490)
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
  return (V3(coatValue) + baseValue) * (1.0f / pdf);
}
// the diffuse base of the coated kind (it reads only the 2D part of the remapped sample, :424-427)
struct DiffuseBase
{
  const Bsdf& b;
  const Ctx& ctx;
  V3 eval(V3 wo, V3 wi, float& pdf) const { return diffuseEval(b, wo, wi, pdf); }
  V3 sample(V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float) const
  {
    return diffuseSample(b, ctx, wo, wi, pdf, type, sx, sy);
  }
  bool diffuseOnly() const { return true; }
};
inline V3 coatEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  return coatEvalOver(L, b, ctx, wo, wi, pdf, DiffuseBase{b, ctx});
}
inline V3 coatSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  return coatSampleOver(L, b, ctx, wo, wi, pdf, type, sx, sy, sc, DiffuseBase{b, ctx});
}

// --- thin rough dielectric (bsdf/thin_rough_dielectric.h:170-345)
inline float thinReflectance(const Luts& L, const Bsdf& b, float cosO, float& Eo)  // :262-264
{
  Eo = ggxAlbedo(L, cosO, b.roughness);
  return dielReflAlbedo(L, cosO, b.teta, b.roughness) + b.tfms * (1.f - Eo);
}
inline V3 thinTransmittance(const Bsdf& b, float cosO)  // :319
{
  float e = (1.0f / refractCos(cosO, b.teta)) * b.tthick;
  return V3(std::pow(b.color.x, e), std::pow(b.color.y, e), std::pow(b.color.z, e));
}
inline V3 thinEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)  // :194-240
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f) || !(cosI > 0.f)) return V3(0.f);
  V3 wh = normalize(wi + wo);
  float cosOH = dot(wo, wh);
  V3 wo0 = toLocal(b.fr, wo), wi0 = toLocal(b.fr, wi), wh0 = toLocal(b.fr, wh);
  float F = fresnelDielectric(cosOH, b.teta);
  float whPdf;
  float D = b.ggx.evalVisible(wh0, wo0, cosOH, whPdf);
  float G = b.ggx.G2(wo0, wi0);
  float Eo, Ei = ggxAlbedo(L, cosI, b.roughness);
  float R = thinReflectance(L, b, cosO, Eo);
  float fms = b.tfms * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.tEavg))) * cosI);
  pdf = whPdf * (1.0f / (4.f * std::fabs(cosOH))) * R;
  return V3(F * D * G * (1.0f / (4.f * cosO)) + fms);
}
inline V3 thinSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)  // :242-305
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float Eo;
  float R = thinReflectance(L, b, cosO, Eo);
  pdf = 0.f;
  wi = V3(0.f);
  if (sc <= R)
  {
    V3 wo0 = toLocal(b.fr, wo);
    float whPdf;
    V3 wh = toGlobal(b.fr, b.ggx.sampleVisible(wo0, whPdf, sx, sy));
    wi = reflect(wo, wh);
    float cosI = dot(wi, b.fr.N);
    if (!(cosI > 0.f && dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
    float cosOH = dot(wo, wh);
    V3 wi0 = toLocal(b.fr, wi);
    float F = fresnelDielectric(cosOH, b.teta);
    float G = b.ggx.G2(wo0, wi0);
    float Ei = ggxAlbedo(L, cosI, b.roughness);
    float fms = b.tfms * ((1.f - Eo) * (1.f - Ei) * (1.0f / (kPi * (1.f - b.tEavg))) * cosI);
    type = BsdfType::GlossyReflection;
    pdf = whPdf * (1.0f / (4.f * std::fabs(cosOH))) * R;
    return V3(F * G * rcpSafe(b.ggx.G1(wo0)) * (1.0f / R) + fms * (1.0f / pdf));
  }
  wi = -wo;
  if (!(dot(wi, ctx.Ng) < 0.f)) return V3(0.f);
  type = BsdfType::UnbentTransmission;
  pdf = 1.f - R;
  return thinTransmittance(b, cosO);
}
inline V3 thinTransparency(const Luts& L, const Bsdf& b, V3 wo)  // :307-321
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float Eo;
  float R = thinReflectance(L, b, cosO, Eo);
  return thinTransmittance(b, cosO) * (1.f - R);
}

// One lobe of a single kind, or the transparent BTDF (bsdf/transparent.h:32-71: evaluates to
// zero, samples -wo with pdf 1).
inline V3 leafEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  pdf = 0.f;
  if (b.kind == kDiffuse) return diffuseEval(b, wo, wi, pdf);
  if (b.kind == kConductor) return conductorEval(L, b, wo, wi, pdf);
  if (b.kind == kCoated) return coatEval(L, b, ctx, wo, wi, pdf);
  if (b.kind == kThinDielectric) return thinEval(L, b, ctx, wo, wi, pdf);
  if (b.kind == kTransparentLobe) return V3(0.f);
  return dielEval(L, b, ctx, wo, wi, pdf);
}
inline V3 leafSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.kind == kDiffuse) return diffuseSample(b, ctx, wo, wi, pdf, type, sx, sy);
  if (b.kind == kConductor) return conductorSample(L, b, ctx, wo, wi, pdf, type, sx, sy);
  if (b.kind == kCoated) return coatSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
  if (b.kind == kThinDielectric) return thinSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
  if (b.kind == kTransparentLobe)
  {
    wi = -wo;
    type = BsdfType::Transparency;
    pdf = 1.f;
    return V3(1.f);
  }
  return dielSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
}
inline V3 leafAlbedo(const Bsdf& b)
{
  if (b.kind == kDiffuse) return b.color;
  if (b.kind == kConductor) return b.rr;
  if (b.kind == kCoated)  // rough_coating.h:567-571
  {
    V3 absorbed(b.color.x * std::pow(b.coatColor.x, b.thickness), b.color.y * std::pow(b.coatColor.y, b.thickness),
                 b.color.z * std::pow(b.coatColor.z, b.thickness));
    return b.color + b.scale * (absorbed - b.color);
  }
  if (b.kind == kThinDielectric)  // thin_rough_dielectric.h:324-327
    return V3(std::pow(b.color.x, b.tthick), std::pow(b.color.y, b.tthick), std::pow(b.color.z, b.tthick));
  if (b.kind == kTransparentLobe) return V3(-1.f);  // transparent.h:52-55
  return V3(1.f);
}
inline V3 leafTransparency(const Luts& L, const Bsdf& b, V3 wo)  // what a shadow ray keeps through the lobe
{
  if (b.kind == kThinDielectric) return thinTransparency(L, b, wo);
  return V3(b.kind == kTransparentLobe ? 1.f : 0.f);
}

// kind 5's weighted mix of bases (multi_bsdf.h:166-287): eval sums the children's values times
// their scales and their pdfs times their importances; sample picks a child by importance with
// the remapped sample, then (unless the pick is a delta) adds the other children's values and pdfs.
inline V3 mixEval(const Luts& L, const StdLobes& s, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  V3 value(0.f);
  pdf = 0.f;
  for (int i = 0; i < s.n; ++i)
  {
    if (!(s.imp[i] > 0.f)) continue;
    float cp;
    V3 cv = leafEval(L, s.child[i], ctx, wo, wi, cp);
    value += cv * s.scale[i];
    pdf += cp * s.imp[i];
  }
  pdf *= 1.0f / s.impSum;
  return value;
}
inline V3 mixSample(const Luts& L, const StdLobes& s, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (!(s.impSum > 0.f)) return V3(0.f);
  if (s.n == 1) return leafSample(L, s.child[0], ctx, wo, wi, pdf, type, sx, sy, sc) * s.scale[0];
  const float x = sc * s.impSum;
  int choice = 0;
  float prefix = s.imp[0], imp = s.imp[0], scale = s.scale[0];
  for (int i = 1; i < s.n; ++i)
  {
    if (!(x > prefix)) break;
    choice = i;
    prefix += s.imp[i];
    imp = s.imp[i];
    scale = s.scale[i];
  }
  const float sc1 = (x + imp - prefix) * (1.0f / imp);
  V3 cwi(0.f);
  float cpdf = 0.f;
  int ctype = 0;
  V3 cw = leafSample(L, s.child[choice], ctx, wo, cwi, cpdf, ctype, sx, sy, sc1);
  V3 w(0.f);
  wi = V3(0.f);
  pdf = 0.f;
  type = 0;
  if (reduceMax(cw) > 0.f) { w = cw; wi = cwi; pdf = cpdf; type = ctype; }
  w = w * scale;
  if (!(reduceMax(w) > 0.f && pdf > 0.f)) return V3(0.f);
  if (type & BsdfType::Delta) return w * (s.impSum * (1.0f / imp));
  V3 value = w * pdf;
  pdf = pdf * imp;
  for (int i = 0; i < s.n; ++i)
  {
    if (i == choice || !(s.imp[i] > 0.f)) continue;
    float cp;
    V3 cv = leafEval(L, s.child[i], ctx, wo, wi, cp);
    value += cv * s.scale[i];
    pdf += cp * s.imp[i];
  }
  pdf *= 1.0f / s.impSum;
  return value * (1.0f / pdf);
}
// getAlbedo (:314-337): children with a negative albedo (the transparent BTDF) are left out and
// the rest rescaled to the total scale
inline V3 mixAlbedo(const StdLobes& s)
{
  V3 value(0.f);
  float scaleSum = 0.f, total = 0.f;
  for (int i = 0; i < s.n; ++i)
  {
    if (!(s.imp[i] > 0.f)) continue;
    V3 cv = leafAlbedo(s.child[i]);
    total += s.scale[i];
    if (!(cv.x >= 0.f)) continue;
    value += cv * s.scale[i];
    scaleSum += s.scale[i];
  }
  return value * (total * rcpSafe(scaleSum));
}
// getNormal (:408-428): the first child with a non-zero importance and a frame
inline bool mixHasNormal(const StdLobes& s)
{
  for (int i = 0; i < s.n; ++i)
    if (s.imp[i] > 0.f && s.child[i].kind != kTransparentLobe) return true;
  return false;
}
inline V3 mixTransparency(const Luts& L, const StdLobes& s, V3 wo)  // :289-305
{
  V3 value(0.f);
  for (int i = 0; i < s.n; ++i)
    if (s.imp[i] > 0.f) value += leafTransparency(L, s.child[i], wo) * s.scale[i];
  return value;
}
struct MixBase
{
  const Luts& L;
  const StdLobes& s;
  const Ctx& ctx;
  V3 eval(V3 wo, V3 wi, float& pdf) const { return mixEval(L, s, ctx, wo, wi, pdf); }
  V3 sample(V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc) const
  {
    return mixSample(L, s, ctx, wo, wi, pdf, type, sx, sy, sc);
  }
  bool diffuseOnly() const { return (s.flags & ~BsdfType::Diffuse) == 0; }  // rough_coating.h:349
};

// The material's lobe: a MultiBsdf with one child of weight 1 for kinds 0-4 (eval scales the
// child's pdf by importance / importanceSum, sample returns the child's weight times its scale 1,
// a zero importance disables sampling), kind 5's mix of bases, under its coat when it has one.
inline V3 lobeEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  if (b.kind == kStandard)
  {
    const MixBase base{L, *b.mixLobes, ctx};
    return b.coat ? coatEvalOver(L, b, ctx, wo, wi, pdf, base) : base.eval(wo, wi, pdf);
  }
  V3 v(0.f);
  pdf = 0.f;
  if (b.importance > 0.f)
  {
    float cp = 0.f;
    v = leafEval(L, b, ctx, wo, wi, cp);
    pdf = cp * b.importance;
  }
  pdf *= 1.0f / b.importance;
  return v;
}
inline V3 lobeSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.kind == kStandard)
  {
    const MixBase base{L, *b.mixLobes, ctx};
    return b.coat ? coatSampleOver(L, b, ctx, wo, wi, pdf, type, sx, sy, sc, base) : base.sample(wo, wi, pdf, type, sx, sy, sc);
  }
  if (!(b.importance > 0.f)) return V3(0.f);
  return leafSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
}
inline V3 lobeAlbedo(const Bsdf& b)  // multi_bsdf.h:314-337 with one child; kind 5: the mix, coated
{
  if (b.kind != kStandard) return leafAlbedo(b);
  V3 a = mixAlbedo(*b.mixLobes);
  if (!b.coat) return a;
  V3 absorbed(a.x * std::pow(b.coatColor.x, b.thickness), a.y * std::pow(b.coatColor.y, b.thickness),
               a.z * std::pow(b.coatColor.z, b.thickness));
  return a + b.scale * (absorbed - a);  // rough_coating.h:567-571
}
// whether the lobe reports a normal (the first child with an importance and a frame)
inline bool lobeHasNormal(const Bsdf& b)
{
  return b.kind == kStandard ? mixHasNormal(*b.mixLobes) : b.importance > 0.f;
}
inline V3 lobeTransparency(const Luts& L, const Bsdf& b, V3 wo)
{
  if (b.kind != kStandard) return b.importance > 0.f ? leafTransparency(L, b, wo) : V3(0.f);
  V3 t = mixTransparency(L, *b.mixLobes, wo);
  if (!b.coat) return t;
  float cosO = dot(wo, b.fr.N);  // rough_coating.h:539-560
  if (!(cosO > 0.f)) return V3(0.f);
  const float e = b.thickness * (1.0f / refractCos(cosO, b.ceta));
  V3 absorbed(t.x * std::pow(b.coatColor.x, e), t.y * std::pow(b.coatColor.y, e), t.z * std::pow(b.coatColor.z, e));
  t = t + b.scale * (absorbed - t);
  return t * (1.f - dielReflAlbedo(L, cosO, b.ceta, b.roughness) * b.scale);
}

// --- sheen layer over the lobe (bsdf/sheen.h:199-451)
inline void sheenPicks(const Bsdf& b, float Ro, float& coatPick, float& basePick)
{
  float coatImp = Ro * reduceMax(b.sheenColor), baseImp = (1.f - Ro) * b.importance;
  coatPick = coatImp * (1.0f / (coatImp + baseImp));
  basePick = 1.f - coatPick;
}
inline V3 sheenReflection(const Bsdf& b, V3 wo, V3 wi, float cosO, float cosI)  // :253-265
{
  V3 wh = normalize(wo + wi);
  float D = b.sheenDist.D(dot(wh, b.fr.N)), G = b.sheenDist.G2(cosO, cosI);
  return b.sheenColor * (D * G * (1.0f / (4.f * cosO)) * b.sheenScale);
}
inline V3 sheenEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)  // :218-279
{
  pdf = 0.f;
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  float basePdf;
  V3 baseValue = lobeEval(L, b, ctx, wo, wi, basePdf);
  float Ro = sheenAlbedo(L, cosO, b.sheenRough) * b.sheenScale;
  float Ri = sheenAlbedo(L, std::fabs(cosI), b.sheenRough) * b.sheenScale;
  baseValue *= std::min(1.f - Ro, 1.f - Ri);
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
inline V3 sheenSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)  // :281-391
{
  float cosO = dot(wo, b.fr.N);
  if (!(cosO > 0.f)) return V3(0.f);
  float Ro = sheenAlbedo(L, cosO, b.sheenRough) * b.sheenScale;
  float coatPick, basePick;
  sheenPicks(b, Ro, coatPick, basePick);
  float basePdf = 0.f;
  V3 baseValue(0.f);
  wi = V3(0.f);
  pdf = 0.f;
  type = 0;
  if (sc <= coatPick)
  {
    type = BsdfType::DiffuseReflection;
    float phi = 2.f * kPi * sx;
    wi = toGlobal(b.fr, V3(std::cos(phi) * cos2sin(sy), std::sin(phi) * cos2sin(sy), sy));  // shape_sampler.h:54-61
    baseValue = lobeEval(L, b, ctx, wo, wi, basePdf);
  }
  else
  {
    V3 bwi(0.f);
    float bpdf = 0.f;
    int btype = 0;
    V3 bw = lobeSample(L, b, ctx, wo, bwi, bpdf, btype, sx, sy, (sc - coatPick) * (1.0f / basePick));
    if (!(reduceMax(bw) > 0.f)) return V3(0.f);
    wi = bwi;
    basePdf = bpdf;
    type = btype;
    baseValue = bw * ((type & BsdfType::Delta) ? 1.f : basePdf);
  }
  float cosI = dot(wi, b.fr.N);
  if (!(cosI * dot(wi, ctx.Ng) > 0.f)) return V3(0.f);
  float Ri = sheenAlbedo(L, std::fabs(cosI), b.sheenRough) * b.sheenScale;
  baseValue *= std::min(1.f - Ro, 1.f - Ri);
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
inline float surfImportance(const Bsdf& b)  // sheen.h:403-406
{
  return b.sheen ? std::max(reduceMax(b.sheenColor), b.importance) : b.importance;
}
inline V3 surfEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  return b.sheen ? sheenEval(L, b, ctx, wo, wi, pdf) : lobeEval(L, b, ctx, wo, wi, pdf);
}
inline V3 surfSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  return b.sheen ? sheenSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc) : lobeSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
}

// The whole BSDF with the opacity mix (material/standard.cpp:613-627): MultiBsdf{lobe x opacity,
// TransparentBtdf x transp} (multi_bsdf.h:166-287, transparent.h:40-45); the transparent BTDF
// evaluates to zero with importance 1; alone (opacity <= 1e-5) it passes every ray straight on.
inline V3 bsdfEval(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3 wi, float& pdf)
{
  if (b.mix == 0) return surfEval(L, b, ctx, wo, wi, pdf);
  pdf = 0.f;
  if (b.mix == 2) return V3(0.f);
  const float imp0 = surfImportance(b) * b.opacity, sum = imp0 + b.transp;
  V3 v(0.f);
  if (imp0 > 0.f)
  {
    float cp;
    v = surfEval(L, b, ctx, wo, wi, cp) * b.opacity;
    pdf = cp * imp0;
  }
  pdf *= 1.0f / sum;
  return v;
}
inline V3 bsdfSample(const Luts& L, const Bsdf& b, const Ctx& ctx, V3 wo, V3& wi, float& pdf, int& type, float sx, float sy, float sc)
{
  if (b.mix == 0) return surfSample(L, b, ctx, wo, wi, pdf, type, sx, sy, sc);
  if (b.mix == 2)
  {
    wi = -wo;
    type = BsdfType::Transparency;
    pdf = 1.f;
    return V3(1.f);
  }
  const float imp0 = surfImportance(b) * b.opacity, sum = imp0 + b.transp;
  if (!(sum > 0.f)) return V3(0.f);
  // pick a child by importance, remap the sample (:219-245)
  const float x = sc * sum;
  const bool lobe = !(x > imp0);
  const float imp = lobe ? imp0 : b.transp, prefix = lobe ? imp0 : sum;
  const float sc1 = (x + imp - prefix) * (1.0f / imp);
  V3 w(0.f);
  wi = V3(0.f);
  pdf = 0.f;
  type = 0;
  if (lobe)
  {
    V3 cwi(0.f);
    float cpdf = 0.f;
    int ctype = 0;
    V3 cw = surfSample(L, b, ctx, wo, cwi, cpdf, ctype, sx, sy, sc1);
    if (reduceMax(cw) > 0.f) { w = cw * b.opacity; wi = cwi; pdf = cpdf; type = ctype; }
  }
  else
  {
    w = V3(b.transp);
    wi = -wo;
    pdf = 1.f;
    type = BsdfType::Transparency;
  }
  if (!(reduceMax(w) > 0.f && pdf > 0.f)) return V3(0.f);
  if (type & BsdfType::Delta) return w * (sum * (1.0f / imp));  // (:262-272)
  V3 value = w * pdf;  // (:274-287): the transparent BTDF adds nothing
  pdf = pdf * imp * (1.0f / sum);
  return value * (1.0f / pdf);
}
// getAlbedo (multi_bsdf.h:339-362): negative albedos (the transparent BTDF) are left out
inline V3 bsdfAlbedo(const Bsdf& b)
{
  if (b.mix == 2) return V3(-1.f);
  if (b.mix == 1 && !(surfImportance(b) * b.opacity > 0.f)) return V3(0.f);
  return lobeAlbedo(b);  // the sheen layer passes the base's albedo (sheen.h:408-411)
}
// getNormal (multi_bsdf.h:408-428): the first child with a non-zero importance
inline V3 bsdfNormal(const Bsdf& b)
{
  return (b.mix != 2 && lobeHasNormal(b) && (b.mix == 0 || surfImportance(b) * b.opacity > 0.f)) ? b.fr.N : V3(0.f);
}
// getTransparency of a transparent material (material/standard.cpp:633-661)
inline V3 bsdfTransparency(const Luts& L, const Bsdf& b, V3 wo)
{
  if (b.mix == 2) return V3(1.f);
  V3 t = lobeTransparency(L, b, wo);
  if (b.sheen)  // sheen.h:393-401
  {
    float cosO = dot(wo, b.fr.N);
    t = cosO > 0.f ? t * (1.f - sheenAlbedo(L, cosO, b.sheenRough) * b.sheenScale) : V3(0.f);
  }
  if (b.mix == 0) return t;
  return t * b.opacity + V3(b.transp);
}

// ---------------------------------------------------------------------------------------
// Lights (light/ambient_light.h, distant_light.h, image_light.h; render/scene.h:137-250)
// ---------------------------------------------------------------------------------------
enum LightKind : int { kAmbient = 0, kDistant = 1, kImage = 2 };

struct Light  // 32-byte record shared with the HIP path
{
  int kind;
  float r, g, b;       // Ke (ambient / distant) or scale (image)
  float dx, dy, dz;    // direction (distant)
  float angle;         // half angle (distant), radians
};

struct EnvImage  // lat-long environment in texture row order (row 0 = nadir)
{
  int W = 0, H = 0;          // texel size
  const float* rgb = nullptr;  // [H][W][3]
  int LW = 0, LH = 0;        // importance LUT size (after 2x bucketing)
  float* xDists = nullptr;   // [LH][LW+1]
  float* yDist = nullptr;    // [LH+1]
};

// image/image_texture.h:252-283 bilinear fetch, addressU = Repeat, addressV = Clamp
inline V3 envTexel(const EnvImage& e, int x, int y)
{
  const float* p = e.rgb + (size_t(y) * e.W + x) * 3;
  return V3(p[0], p[1], p[2]);
}
inline V3 envLookup(const EnvImage& e, float u, float v)
{
  float x = u - 0.5f / float(e.W);
  float y = v - 0.5f / float(e.H);
  float xr = (x - std::floor(x)) * float(e.W);
  float yr = clampf(y * float(e.H), 0.f, float(e.H - 1));
  float sx = xr - std::floor(xr), sy = yr - std::floor(yr);
  int x0 = std::min(std::max(int(xr), 0), e.W - 1);
  int y0 = std::min(std::max(int(yr), 0), e.H - 1);
  int x1 = (x0 + 1 == e.W) ? 0 : x0 + 1;
  int y1 = std::min(y0 + 1, e.H - 1);
  V3 c00 = envTexel(e, x0, y0), c11 = envTexel(e, x1, y1), c01 = envTexel(e, x0, y1), c10 = envTexel(e, x1, y0);
  V3 c0 = c00 + sx * (c10 - c00);
  V3 c1 = c01 + sx * (c11 - c01);
  return c0 + sy * (c1 - c0);
}

// light/triangle_light.h: the emissive triangles of the mesh as one area light, triangles
// picked by power (|cross(e1, e2)| * luminance), uniform on the picked triangle.
struct AreaLight
{
  std::vector<V3> v0, v1, v2, n, color;  // n = cross(v1 - v0, v2 - v0), not normalized
  std::vector<float> cdf;                // Distribution1D over the emissive triangles
  std::vector<int> triMap;               // primitive ID -> emissive triangle index or -1
  float power = 0.f;                     // getPower (:152-162)
};

// light list of render/scene.cpp:365-450: the area light first (if any), then the
// environment lights; lightDistribution over their powers
struct LightSet
{
  int count = 0;                 // environment lights
  int offset = 0;                // 1 if an area light precedes them
  const Light* lights = nullptr;
  const AreaLight* area = nullptr;
  float cdf[10];                 // light power distribution (Distribution1D over <= 9 lights)
  const EnvImage* env = nullptr;
  float sceneRadius = 0.f;
};

// sampling/shape_sampler.h:135-152 [Heitz 2019, low-distortion map triangle <- square]
inline V3 uniformSampleTriangle(float sx, float sy, V3 v0, V3 v1, V3 v2)
{
  float u = sx, v = sy;
  if (v > u) { u *= 0.5f; v -= u; }
  else { v *= 0.5f; u -= v; }
  return v2 + u * (v0 - v2) + v * (v1 - v2);
}
// triangle_light.h:84-100 (eval on a hit of an emissive triangle at distance dist)
inline V3 areaLightEval(const AreaLight& a, V3 wi, float dist, int prim, float& pdf)
{
  int tri = a.triMap[prim];
  float pickProb = dist1dPdf(a.cdf.data(), int(a.v0.size()), tri);
  float dirDotN = dot(wi, a.n[tri]);
  if (dirDotN >= 0.f) { pdf = 0.f; return V3(0.f); }
  pdf = 2.0f * dist * dist * rcpSafe(std::fabs(dirDotN)) * pickProb;
  return a.color[tri];
}
// triangle_light.h:123-150 (wi is left unchanged when the sample faces away)
inline V3 areaLightSample(const AreaLight& a, V3 p, V3& wi, float& pdf, float& dist, float sx, float sy, float sc)
{
  pdf = 0.f;
  float pickProb;
  int tri = dist1dLookup(a.cdf.data(), int(a.v0.size()), sc);
  pickProb = a.cdf[tri + 1] - a.cdf[tri];
  V3 dir = uniformSampleTriangle(sx, sy, a.v0[tri], a.v1[tri], a.v2[tri]) - p;
  dist = length(dir);
  float dirDotN = dot(dir, a.n[tri]);
  if (!(dirDotN < 0.f)) return V3(0.f);
  wi = dir * (1.0f / dist);
  pdf = 2.0f * dist * dist * dist * rcpSafe(std::fabs(dirDotN)) * pickProb;
  return a.color[tri] * (1.0f / pdf);
}

// ambient_light.h:53-57, distant_light.h:73-77, image_light.h:137-148
inline float lightPower(const Light& l, const EnvImage* env, float sceneRadius)
{
  V3 c(l.r, l.g, l.b);
  if (l.kind == kAmbient) return 4.f * sqr(kPi) * sqr(sceneRadius) * luminance(c);
  if (l.kind == kDistant) return 4.f * sqr(kPi) * sqr(std::sin(0.5f * l.angle)) * sqr(sceneRadius) * luminance(c);
  float sumL = 0.f;  // over the image the light emits: the texels times its scale
  for (int y = 0; y < env->H; ++y)
    for (int x = 0; x < env->W; ++x) sumL += luminance(envTexel(*env, x, y) * c);
  return 4.f * sqr(kPi) * sqr(sceneRadius) * (sumL / float(env->W * env->H));
}

// Evaluate one environment light toward direction wi.
inline V3 envLightEval(const Light& l, const EnvImage* env, V3 wi, float& pdf)
{
  pdf = 0.f;
  if (l.kind == kAmbient)  // ambient_light.h:33-37
  {
    pdf = 1.f / (4.f * kPi);
    return V3(l.r, l.g, l.b);
  }
  if (l.kind == kDistant)  // distant_light.h:48-57
  {
    V3 dir(l.dx, l.dy, l.dz);
    float c = dot(wi, dir);
    if (!(c >= std::cos(l.angle))) return V3(0.f);
    pdf = uniformSampleConePdf(c, l.angle);
    return V3(l.r, l.g, l.b);
  }
  // image_light.h:79-91
  float theta = std::acos(clampf(wi.y, -1.f, 1.f));
  float phi = std::atan2(wi.x, -wi.z);
  float u = 0.5f + phi * (1.0f / (2.0f * kPi));
  float v = 1.0f - theta * (1.0f / kPi);
  int ix = std::min(std::max(int(u * float(env->LW)), 0), env->LW - 1);
  int iy = std::min(std::max(int(v * float(env->LH)), 0), env->LH - 1);
  pdf = dist2dPdf(env->xDists, env->yDist, env->LW, env->LH, ix, iy) * float(env->LW) * float(env->LH) *
        (1.0f / (2.0f * kPi * kPi * std::sin(theta)));
  return envLookup(*env, u, v) * V3(l.r, l.g, l.b);
}

// Sample one light from a 2D sample; returns Le / pdf (times the image scale).
inline V3 envLightSample(const Light& l, const EnvImage* env, V3& wi, float& pdf, float& dist, float sx, float sy)
{
  dist = kPosMax;
  if (l.kind == kAmbient)  // ambient_light.h:46-51
  {
    wi = uniformSampleSphere(pdf, sx, sy);
    return V3(l.r, l.g, l.b) * (1.0f / pdf);
  }
  if (l.kind == kDistant)  // distant_light.h:66-71
  {
    Basis fr = makeBasis(V3(l.dx, l.dy, l.dz));
    wi = toGlobal(fr, uniformSampleCone(pdf, sx, sy, l.angle));
    return V3(l.r, l.g, l.b) * (1.0f / pdf);
  }
  // image_light.h:112-134
  int ix, iy;
  dist2dSampleReuse(env->xDists, env->yDist, env->LW, env->LH, pdf, sx, sy, ix, iy);
  float u = (float(ix) + sx) * (1.0f / float(env->LW));
  float v = (float(iy) + sy) * (1.0f / float(env->LH));
  float theta = (1.0f - v) * kPi;
  float phi = (u - 0.5f) * (2.0f * kPi);
  float st = std::sin(theta);
  wi = V3(std::sin(phi) * st, std::cos(theta), -std::cos(phi) * st);
  pdf *= float(env->LW) * float(env->LH) * (1.0f / (2.0f * kPi * kPi * st));
  return envLookup(*env, u, v) * V3(l.r, l.g, l.b) * (1.0f / pdf);
}

// render/scene.h:137-157: sum of all environment lights, pdf weighted by pick probability
inline V3 evalEnv(const LightSet& ls, V3 wi, float& pdf, uint32_t* hitMask = nullptr)
{
  V3 value(0.f);
  pdf = 0.f;
  uint32_t mask = 0;
  for (int i = 0; i < ls.count; ++i)
  {
    float cp;
    V3 v = envLightEval(ls.lights[i], ls.env, wi, cp);
    if (cp > 0.f) mask |= 1u << i;
    value += v;
    pdf += cp * dist1dPdf(ls.cdf, ls.offset + ls.count, ls.offset + i);
  }
  if (hitMask) *hitMask = mask;
  return value;
}
// render/scene.h:159-195: the area light seen along wi at distance dist on primitive prim
inline V3 evalArea(const LightSet& ls, V3 wi, float dist, int prim, float& pdf)
{
  V3 v = areaLightEval(*ls.area, wi, dist, prim, pdf);
  pdf *= dist1dPdf(ls.cdf, ls.offset + ls.count, 0);
  return v;
}
// render/scene.h:215-250
inline V3 sampleLight(const LightSet& ls, V3 p, V3& wi, float& pdf, float& dist, float sx, float sy, float sc, int* picked = nullptr)
{
  if (picked) *picked = 0;
  const int total = ls.offset + ls.count;
  if (total == 1)
    return ls.area ? areaLightSample(*ls.area, p, wi, pdf, dist, sx, sy, sc) : envLightSample(ls.lights[0], ls.env, wi, pdf, dist, sx, sy);
  float pick;
  float c = sc;
  int id = dist1dSampleReuse(ls.cdf, total, pick, c);
  if (picked) *picked = id;
  V3 w = (ls.area && id == 0) ? areaLightSample(*ls.area, p, wi, pdf, dist, sx, sy, c)
                              : envLightSample(ls.lights[id - ls.offset], ls.env, wi, pdf, dist, sx, sy);
  pdf *= pick;
  return w * (1.0f / pick);
}

// Power-heuristic MIS (render/path_renderer_stream.h:1077-1080)
inline float misWeight(float a, float b) { return sqr(a) * rcpSafe(sqr(a) + sqr(b)); }

}  // namespace orc
