// Learned denoiser for gfx950: a U-Net (4 levels down and up, 3x3 convolutions with 32 / 48 /
// 64 / 96 / 128 channels, ReLU, 2x2 max pooling, nearest upsampling with skip concatenation)
// from 9 input channels (colour, albedo, normal) to 3, trained and run with this file's kernels.
// It plugs in behind the reference's Denoiser interface (image/denoiser.h:10-41): init with the
// colour / albedo / normal / output images, execute, version.
//
// Arithmetic: activations and activation gradients in BF16 (round to nearest even), every
// convolution on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation), FP32 master
// weights, weight gradients and Adam state.  Tensors are NHWC with channels padded to multiples
// of 16 (one MFMA K step).
//
// Convolutions are implicit GEMMs.  Forward and data gradient: M = pixels (64 consecutive pixels
// of one image row per wave, as 4 MFMA row blocks), N = output channels (16 per MFMA column
// block), K = 9 taps x input channels; the data gradient is the forward kernel run on the output
// gradient with the weights transposed and flipped, and its output is masked by the ReLU of the
// layer input.  Weight gradient: M = output channels, N = input channels, K = pixels (16 image
// rows per wave), reduced with FP32 atomics.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prt_hip.h"
#include "internal.h"

namespace prt {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t toBf16(float f)  // round to nearest even
{
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float fromBf16(uint16_t h)
{
  uint32_t u = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

constexpr int kBlock = 256;

struct ConvArgs
{
  const uint16_t* x;  // input NHWC bf16, channel stride xC (= cinPad)
  int xC;
  const uint16_t* w;  // [9][coutPad][cinPad] bf16
  const float* bias;  // [coutReal] or null
  int N, H, W, cinPad, coutPad, coutReal;
  int relu;
  uint16_t* y;  // output bf16 NHWC, channel stride yC, channel offset yOff (null: no bf16 output)
  int yC, yOff;
  float* yf;  // optional FP32 output [N*H*W][coutReal]
  const uint16_t* mask;  // optional: outputs where mask <= 0 are zeroed (mask has y's layout)
};

// One wave: 64 pixels of one row (4 MFMA row blocks) x 16*NB output channels.
template <int NB>
__global__ void __launch_bounds__(kBlock) kConv(ConvArgs a)
{
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int groups = a.coutPad / (16 * NB);
  const int segs = (a.W + 63) / 64;
  const int total = a.N * a.H * segs * groups;
  if (wave >= total) return;  // whole waves leave together
  int t = wave;
  const int g = t % groups; t /= groups;
  const int xs = t % segs; t /= segs;
  const int y = t % a.H;
  const int n = t / a.H;
  const int x0 = xs * 64, co0 = g * 16 * NB;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[4][NB];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[p][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x4 zero = {0, 0, 0, 0};
  for (int tap = 0; tap < 9; ++tap)
  {
    const int yy = y + tap / 3 - 1, dx = tap % 3 - 1;
    const bool rowOk = yy >= 0 && yy < a.H;
    const uint16_t* xrow = a.x + (size_t(n) * a.H + (rowOk ? yy : 0)) * a.W * size_t(a.xC);
    const uint16_t* wt = a.w + size_t(tap) * a.coutPad * a.cinPad;
    for (int c0 = 0; c0 < a.cinPad; c0 += 16)
    {
      bf16x4 A[4], B[NB];
#pragma unroll
      for (int p = 0; p < 4; ++p)
      {
        const int xx = x0 + 16 * p + li + dx;
        A[p] = (rowOk && xx >= 0 && xx < a.W) ? *reinterpret_cast<const bf16x4*>(xrow + size_t(xx) * a.xC + c0 + lk) : zero;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b)
        B[b] = *reinterpret_cast<const bf16x4*>(wt + size_t(co0 + 16 * b + li) * a.cinPad + c0 + lk);
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[p][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[p], B[b], acc[p][b], 0, 0, 0);
    }
  }
  // D of block (p, b): lane holds output channel co0 + 16 b + li for pixels x0 + 16 p + lk + j
#pragma unroll
  for (int b = 0; b < NB; ++b)
  {
    const int co = co0 + 16 * b + li;
    const float bias = (a.bias && co < a.coutReal) ? a.bias[co] : 0.f;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int x = x0 + 16 * p + lk + j;
        if (x >= a.W || co >= a.coutPad) continue;
        const size_t pix = (size_t(n) * a.H + y) * a.W + x;
        float v = acc[p][b][j] + bias;
        if (a.relu) v = fmaxf(v, 0.f);
        if (co >= a.coutReal) v = 0.f;
        if (a.mask && !(fromBf16(a.mask[pix * a.yC + a.yOff + co]) > 0.f)) v = 0.f;
        if (a.y) a.y[pix * a.yC + a.yOff + co] = toBf16(v);
        if (a.yf && co < a.coutReal) a.yf[pix * a.coutReal + co] = v;
      }
  }
}

struct WgradArgs
{
  const uint16_t* dy;  // [pixels][coutPad] (masked output gradient)
  const uint16_t* x;   // [pixels][cinPad] layer input
  int N, H, W, cinPad, coutPad, rows;
  float* dw;  // [9][coutPad][cinPad] FP32, accumulated
};

// One wave: one tap, 32 output x 32 input channels (2 x 2 MFMA blocks, or fewer), `rows` image
// rows of one image; K runs over the pixels 16 at a time.
__global__ void __launch_bounds__(kBlock) kWgrad(WgradArgs a)
{
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int mBlocks = (a.coutPad + 31) / 32, nBlocks = (a.cinPad + 31) / 32;
  const int chunks = (a.H + a.rows - 1) / a.rows;
  const int total = 9 * mBlocks * nBlocks * a.N * chunks;
  if (wave >= total) return;
  int t = wave;
  const int nb = t % nBlocks; t /= nBlocks;
  const int mb = t % mBlocks; t /= mBlocks;
  const int tap = t % 9; t /= 9;
  const int ch = t % chunks;
  const int n = t / chunks;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int y0 = ch * a.rows, y1 = min(a.H, y0 + a.rows);
  for (int y = y0; y < y1; ++y)
  {
    const int yy = y + dy;
    const bool rowOk = yy >= 0 && yy < a.H;
    const size_t rowOut = (size_t(n) * a.H + y) * a.W;
    const size_t rowIn = (size_t(n) * a.H + (rowOk ? yy : 0)) * a.W;
    for (int xb = 0; xb < a.W; xb += 16)
    {
      bf16x4 A[2], B[2];
      for (int i = 0; i < 2; ++i)
      {
        const int co = mb * 32 + 16 * i + li;
        for (int j = 0; j < 4; ++j)
        {
          const int x = xb + lk + j;
          A[i][j] = (co < a.coutPad && x < a.W) ? short(a.dy[(rowOut + x) * a.coutPad + co]) : short(0);
        }
        const int ci = nb * 32 + 16 * i + li;
        for (int j = 0; j < 4; ++j)
        {
          const int xx = xb + lk + j + dx;
          B[i][j] = (ci < a.cinPad && rowOk && xx >= 0 && xx < a.W && xb + lk + j < a.W)
                        ? short(a.x[(rowIn + xx) * a.cinPad + ci]) : short(0);
        }
      }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[i], B[j], acc[i][j], 0, 0, 0);
    }
  }
  // D of block (i, j): lane holds input channel nb*32 + 16 j + li for output channels mb*32 + 16 i + lk + r
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
    {
      const int ci = nb * 32 + 16 * j + li;
      for (int r = 0; r < 4; ++r)
      {
        const int co = mb * 32 + 16 * i + lk + r;
        if (co < a.coutPad && ci < a.cinPad && acc[i][j][r] != 0.f)
          atomicAdd(&a.dw[(size_t(tap) * a.coutPad + co) * a.cinPad + ci], acc[i][j][r]);
      }
    }
}

// Bias gradient: per output channel, the sum of the (masked) output gradient over the pixels.
__global__ void __launch_bounds__(kBlock) kBiasGrad(const uint16_t* __restrict__ dy, int npix, int cPad, int cReal,
                                                    float* __restrict__ db)
{
  const int chunk = 1024;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int c = t % cReal, k = t / cReal;
  const int p0 = k * chunk;
  if (p0 >= npix) return;
  float s = 0.f;
  for (int p = p0; p < min(npix, p0 + chunk); ++p) s += fromBf16(dy[size_t(p) * cPad + c]);
  atomicAdd(&db[c], s);
}

// 2x2 max pooling (NHWC, H and W even).
__global__ void __launch_bounds__(kBlock) kPool(const uint16_t* __restrict__ x, int N, int H, int W, int C, uint16_t* __restrict__ y)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  if (i >= size_t(N) * Ho * Wo * C) return;
  const int c = int(i % C);
  size_t p = i / C;
  const int xo = int(p % Wo); p /= Wo;
  const int yo = int(p % Ho);
  const int n = int(p / Ho);
  auto at = [&](int yy, int xx) { return fromBf16(x[((size_t(n) * H + yy) * W + xx) * C + c]); };
  float m = fmaxf(fmaxf(at(2 * yo, 2 * xo), at(2 * yo, 2 * xo + 1)), fmaxf(at(2 * yo + 1, 2 * xo), at(2 * yo + 1, 2 * xo + 1)));
  y[i] = toBf16(m);
}

// Encoder level backward: de = (pool gradient routed to the first maximum of each 2x2 window
// + the skip slice of the decoder's concat gradient), masked by ReLU(e) > 0.
__global__ void __launch_bounds__(kBlock) kPoolBwd(const uint16_t* __restrict__ e, const uint16_t* __restrict__ dp,
                                                   const uint16_t* __restrict__ dcat, int catC, int catOff, int N, int H, int W, int C,
                                                   uint16_t* __restrict__ de)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= size_t(N) * H * W * C) return;
  const int c = int(i % C);
  size_t p = i / C;
  const int x = int(p % W); p /= W;
  const int y = int(p % H);
  const int n = int(p / H);
  const float v = fromBf16(e[i]);
  float g = 0.f;
  if (v > 0.f)
  {
    const int xo = x / 2, yo = y / 2;
    // the window's first maximum in order (0,0), (0,1), (1,0), (1,1)
    int best = -1;
    float bv = -3.4e38f;
    for (int k = 0; k < 4; ++k)
    {
      const int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
      const float w = fromBf16(e[((size_t(n) * H + yy) * W + xx) * C + c]);
      if (w > bv) { bv = w; best = k; }
    }
    if (best == (y - 2 * yo) * 2 + (x - 2 * xo) && dp)
      g += fromBf16(dp[((size_t(n) * (H / 2) + yo) * (W / 2) + xo) * C + c]);
    if (dcat) g += fromBf16(dcat[((size_t(n) * H + y) * W + x) * catC + catOff + c]);
  }
  de[i] = toBf16(g);
}

// Decoder input: cat[.., 0:Cl] = nearest 2x upsampling of the lower level, cat[.., Cl:] = skip.
__global__ void __launch_bounds__(kBlock) kUpcat(const uint16_t* __restrict__ low, int Cl, const uint16_t* __restrict__ skip, int Cs,
                                                 int N, int H, int W, uint16_t* __restrict__ cat)
{
  const int C = Cl + Cs;
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= size_t(N) * H * W * C) return;
  const int c = int(i % C);
  size_t p = i / C;
  const int x = int(p % W); p /= W;
  const int y = int(p % H);
  const int n = int(p / H);
  cat[i] = c < Cl ? low[((size_t(n) * (H / 2) + y / 2) * (W / 2) + x / 2) * Cl + c]
                  : skip[((size_t(n) * H + y) * W + x) * Cs + (c - Cl)];
}

// Gradient of the upsampled part of a concat: the sum over each 2x2 block (already masked).
__global__ void __launch_bounds__(kBlock) kUpBwd(const uint16_t* __restrict__ dcat, int catC, int Cl, int N, int H, int W,
                                                 uint16_t* __restrict__ dlow)
{
  const int Ho = H / 2, Wo = W / 2;
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= size_t(N) * Ho * Wo * Cl) return;
  const int c = int(i % Cl);
  size_t p = i / Cl;
  const int xo = int(p % Wo); p /= Wo;
  const int yo = int(p % Ho);
  const int n = int(p / Ho);
  float s = 0.f;
  for (int k = 0; k < 4; ++k)
    s += fromBf16(dcat[((size_t(n) * H + 2 * yo + k / 2) * W + 2 * xo + k % 2) * catC + c]);
  dlow[i] = toBf16(s);
}

// Network input: log(1 + colour), albedo, normal -> 16 BF16 channels (9 used).
__global__ void __launch_bounds__(kBlock) kInput(const float* __restrict__ in9, size_t npix, float hdrScale, uint16_t* __restrict__ x)
{
  const size_t p = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= npix) return;
  for (int c = 0; c < 16; ++c)
  {
    float v = 0.f;
    if (c < 3) v = log1pf(fmaxf(in9[p * 9 + c] * hdrScale, 0.f));
    else if (c < 9) v = in9[p * 9 + c];
    x[p * 16 + c] = toBf16(v);
  }
}

// L1 loss in log(1 + colour) space: loss sum and the BF16 output gradient.
__global__ void __launch_bounds__(kBlock) kLoss(const float* __restrict__ pred, const float* __restrict__ target, size_t npix,
                                                float scale, uint16_t* __restrict__ dout, float* __restrict__ loss)
{
  const size_t p = size_t(blockIdx.x) * kBlock + threadIdx.x;
  float l = 0.f;
  if (p < npix)
  {
    for (int c = 0; c < 16; ++c)
    {
      float g = 0.f;
      if (c < 3)
      {
        const float d = pred[p * 3 + c] - log1pf(fmaxf(target[p * 3 + c], 0.f));
        l += fabsf(d);
        g = d > 0.f ? scale : (d < 0.f ? -scale : 0.f);
      }
      dout[p * 16 + c] = toBf16(g);
    }
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

__global__ void __launch_bounds__(kBlock) kOutput(const float* __restrict__ pred, size_t npix, float invScale, float* __restrict__ out)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < npix * 3) out[i] = expm1f(pred[i]) * invScale;
}

// Adam on the FP32 master parameters; then the layers' BF16 weights (forward layout and the
// transposed / flipped layout of the data gradient) are rewritten from them.
__global__ void __launch_bounds__(kBlock) kAdam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                                const float* __restrict__ g, size_t n, float lr, float b1, float b2, float c1,
                                                float c2, float eps)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Master layout of a layer: weights [coutReal][cinReal][3][3] then bias [coutReal].
__global__ void __launch_bounds__(kBlock) kPackWeights(const float* __restrict__ master, int cinReal, int coutReal, int cinPad,
                                                       int coutPad, uint16_t* __restrict__ wf, uint16_t* __restrict__ wt)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= 9 * coutPad * cinPad) return;
  const int ci = i % cinPad, co = (i / cinPad) % coutPad, tap = i / (cinPad * coutPad);
  const float v = (ci < cinReal && co < coutReal) ? master[(size_t(co) * cinReal + ci) * 9 + tap] : 0.f;
  const uint16_t h = toBf16(v);
  wf[i] = h;                                                      // [tap][co][ci]
  wt[(size_t(8 - tap) * cinPad + ci) * coutPad + co] = h;         // [8 - tap][ci][co]
}

// FP32 weight gradient [9][coutPad][cinPad] -> the master layout [coutReal][cinReal][3][3].
__global__ void __launch_bounds__(kBlock) kUnpackGrad(const float* __restrict__ dwp, int cinReal, int coutReal, int cinPad, int coutPad,
                                                      float* __restrict__ g)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= coutReal * cinReal * 9) return;
  const int tap = i % 9, ci = (i / 9) % cinReal, co = i / (9 * cinReal);
  g[i] = dwp[(size_t(tap) * coutPad + co) * cinPad + ci];
}

__global__ void kFill(float* p, size_t n, float v)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocksFor(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Network description and host driver
// ---------------------------------------------------------------------------------------
struct DenoiseLayer
{
  int cin, cout, cinPad, coutPad;
  size_t wOff;  // offset of the layer's weights in the master vector (bias follows)
};

// enc0..enc4, dec3..dec0, out
static const int kLayerCin[10] = {9, 32, 48, 64, 96, 224, 160, 112, 80, 32};
static const int kLayerCout[10] = {32, 48, 64, 96, 128, 96, 64, 48, 32, 3};

}  // namespace prt

using namespace prt;

struct prt_denoiser
{
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  DenoiseLayer L[10];
  size_t nParams = 0;
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master, Adam, gradient
  uint16_t* wf[10] = {};
  uint16_t* wt[10] = {};
  float* dwp[10] = {};  // padded FP32 weight gradients
  int step = 0;
  // activation buffers for the current (N, H, W)
  int N = 0, H = 0, W = 0;
  std::vector<void*> bufs;
  uint16_t *x = nullptr, *e[5] = {}, *p[4] = {}, *cat[4] = {}, *d[4] = {};
  uint16_t *dout = nullptr, *dd[4] = {}, *dcat[4] = {}, *de[5] = {}, *dp[4] = {};
  float *pred = nullptr, *in9 = nullptr, *tgt = nullptr, *loss = nullptr, *outF = nullptr;
  // Denoiser::init parameters
  prt_denoiser_params params{};
  bool inited = false;
};

namespace {

int dfail(int code, const std::string& msg) { return prt::setLastError(code, msg); }

#define DN_TRY(x)                                                                                        \
  do {                                                                                                   \
    hipError_t e_ = (x);                                                                                 \
    if (e_ != hipSuccess) return dfail(e_ == hipErrorOutOfMemory ? PRT_ERR_NOMEM : PRT_ERR_HIP,          \
                                       std::string(#x ": ") + hipGetErrorString(e_));                    \
  } while (0)

void freeActs(prt_denoiser* dn)
{
  for (void* b : dn->bufs) hipFree(b);
  dn->bufs.clear();
  dn->N = dn->H = dn->W = 0;
}

template <class T>
hipError_t actAlloc(prt_denoiser* dn, T*& p, size_t count)
{
  hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1));
  if (e == hipSuccess) dn->bufs.push_back(p);
  return e;
}

// Activation and gradient buffers for N images of H x W (multiples of 16).
int ensureActs(prt_denoiser* dn, int N, int H, int W)
{
  if (dn->N == N && dn->H == H && dn->W == W) return PRT_OK;
  freeActs(dn);
  const int encC[5] = {32, 48, 64, 96, 128};
  const int catC[4] = {80, 112, 160, 224};  // level 0..3: up(lower) + skip
  const int decC[4] = {32, 48, 64, 96};
  auto np = [&](int lvl) { return size_t(N) * (H >> lvl) * (W >> lvl); };
  DN_TRY(actAlloc(dn, dn->x, np(0) * 16));
  DN_TRY(actAlloc(dn, dn->in9, np(0) * 9));
  DN_TRY(actAlloc(dn, dn->tgt, np(0) * 3));
  DN_TRY(actAlloc(dn, dn->pred, np(0) * 3));
  DN_TRY(actAlloc(dn, dn->outF, np(0) * 3));
  DN_TRY(actAlloc(dn, dn->loss, 1));
  DN_TRY(actAlloc(dn, dn->dout, np(0) * 16));
  for (int l = 0; l < 5; ++l)
  {
    DN_TRY(actAlloc(dn, dn->e[l], np(l) * encC[l]));
    DN_TRY(actAlloc(dn, dn->de[l], np(l) * encC[l]));
  }
  for (int l = 0; l < 4; ++l)
  {
    DN_TRY(actAlloc(dn, dn->p[l], np(l + 1) * encC[l]));
    DN_TRY(actAlloc(dn, dn->dp[l], np(l + 1) * encC[l]));
    DN_TRY(actAlloc(dn, dn->cat[l], np(l) * catC[l]));
    DN_TRY(actAlloc(dn, dn->dcat[l], np(l) * catC[l]));
    DN_TRY(actAlloc(dn, dn->d[l], np(l) * decC[l]));
    DN_TRY(actAlloc(dn, dn->dd[l], np(l) * decC[l]));
  }
  dn->N = N; dn->H = H; dn->W = W;
  return PRT_OK;
}

void conv(prt_denoiser* dn, int layer, const uint16_t* x, int N, int H, int W, uint16_t* y, int yC, int yOff, bool relu, float* yf)
{
  const DenoiseLayer& l = dn->L[layer];
  ConvArgs a{};
  a.x = x; a.xC = l.cinPad; a.w = dn->wf[layer]; a.bias = dn->w + l.wOff + size_t(l.cout) * l.cin * 9;
  a.N = N; a.H = H; a.W = W; a.cinPad = l.cinPad; a.coutPad = l.coutPad; a.coutReal = l.cout;
  a.relu = relu ? 1 : 0; a.y = y; a.yC = yC; a.yOff = yOff; a.yf = yf; a.mask = nullptr;
  const size_t waves = size_t(N) * H * ((W + 63) / 64) * (l.coutPad / 16);
  kConv<1><<<blocksFor(waves * 64), kBlock, 0, dn->stream>>>(a);
}

// data gradient of layer `layer`: dx (layout of its input, cinPad channels) from dy (coutPad),
// masked by the layer input (a ReLU output) when mask is given
void dgrad(prt_denoiser* dn, int layer, const uint16_t* dy, int N, int H, int W, uint16_t* dx, const uint16_t* mask)
{
  const DenoiseLayer& l = dn->L[layer];
  ConvArgs a{};
  a.x = dy; a.xC = l.coutPad; a.w = dn->wt[layer]; a.bias = nullptr;
  a.N = N; a.H = H; a.W = W; a.cinPad = l.coutPad; a.coutPad = l.cinPad; a.coutReal = l.cinPad;
  a.relu = 0; a.y = dx; a.yC = l.cinPad; a.yOff = 0; a.yf = nullptr; a.mask = mask;
  const size_t waves = size_t(N) * H * ((W + 63) / 64) * (l.cinPad / 16);
  kConv<1><<<blocksFor(waves * 64), kBlock, 0, dn->stream>>>(a);
}

void wgrad(prt_denoiser* dn, int layer, const uint16_t* dy, const uint16_t* x, int N, int H, int W)
{
  const DenoiseLayer& l = dn->L[layer];
  WgradArgs a{dy, x, N, H, W, l.cinPad, l.coutPad, 16, dn->dwp[layer]};
  const size_t waves = size_t(9) * ((l.coutPad + 31) / 32) * ((l.cinPad + 31) / 32) * N * ((H + 15) / 16);
  kWgrad<<<blocksFor(waves * 64), kBlock, 0, dn->stream>>>(a);
  const size_t npix = size_t(N) * H * W;
  const size_t threads = ((npix + 1023) / 1024) * l.cout;
  kBiasGrad<<<blocksFor(threads), kBlock, 0, dn->stream>>>(dy, int(npix), l.coutPad, l.cout, dn->g + l.wOff + size_t(l.cout) * l.cin * 9);
}

void forward(prt_denoiser* dn, int N, int H, int W)
{
  const int encC[5] = {32, 48, 64, 96, 128};
  const uint16_t* in = dn->x;
  for (int l = 0; l < 5; ++l)
  {
    const int s = l;
    conv(dn, l, in, N, H >> s, W >> s, dn->e[l], encC[l], 0, true, nullptr);
    if (l < 4)
    {
      const size_t n = size_t(N) * (H >> (s + 1)) * (W >> (s + 1)) * encC[l];
      kPool<<<blocksFor(n), kBlock, 0, dn->stream>>>(dn->e[l], N, H >> s, W >> s, encC[l], dn->p[l]);
      in = dn->p[l];
    }
  }
  const int decC[4] = {32, 48, 64, 96};
  const uint16_t* low = dn->e[4];
  int lowC = 128;
  for (int l = 3; l >= 0; --l)
  {
    const int h = H >> l, w = W >> l;
    const size_t n = size_t(N) * h * w * (lowC + encC[l]);
    kUpcat<<<blocksFor(n), kBlock, 0, dn->stream>>>(low, lowC, dn->e[l], encC[l], N, h, w, dn->cat[l]);
    conv(dn, 8 - l, dn->cat[l], N, h, w, dn->d[l], decC[l], 0, true, nullptr);  // dec3 = layer 5 ... dec0 = layer 8
    low = dn->d[l];
    lowC = decC[l];
  }
  conv(dn, 9, dn->d[0], N, H, W, nullptr, 16, 0, false, dn->pred);
}

void backward(prt_denoiser* dn, int N, int H, int W)
{
  const int encC[5] = {32, 48, 64, 96, 128};
  const int decC[4] = {32, 48, 64, 96};
  // output layer
  wgrad(dn, 9, dn->dout, dn->d[0], N, H, W);
  dgrad(dn, 9, dn->dout, N, H, W, dn->dd[0], dn->d[0]);
  // decoder: dd[l] is the masked gradient of d[l]
  for (int l = 0; l < 4; ++l)
  {
    const int h = H >> l, w = W >> l, layer = 8 - l, catC = (l == 3 ? 128 : decC[l + 1]) + encC[l];
    wgrad(dn, layer, dn->dd[l], dn->cat[l], N, h, w);
    dgrad(dn, layer, dn->dd[l], N, h, w, dn->dcat[l], dn->cat[l]);
    const int lowC = l == 3 ? 128 : decC[l + 1];
    uint16_t* dlow = l == 3 ? dn->de[4] : dn->dd[l + 1];
    const size_t n = size_t(N) * (h / 2) * (w / 2) * lowC;
    kUpBwd<<<blocksFor(n), kBlock, 0, dn->stream>>>(dn->dcat[l], catC, lowC, N, h, w, dlow);
  }
  // encoder: de[4] is the masked gradient of e[4]
  for (int l = 4; l >= 0; --l)
  {
    const int h = H >> l, w = W >> l;
    const uint16_t* in = l == 0 ? dn->x : dn->p[l - 1];
    wgrad(dn, l, dn->de[l], in, N, h, w);
    if (l > 0)
    {
      dgrad(dn, l, dn->de[l], N, h, w, dn->dp[l - 1], nullptr);
      const int catC = (l - 1 == 3 ? 128 : decC[l]) + encC[l - 1];
      const int lowC = l - 1 == 3 ? 128 : decC[l];
      const size_t n = size_t(N) * (h * 2) * (w * 2) * encC[l - 1];
      kPoolBwd<<<blocksFor(n), kBlock, 0, dn->stream>>>(dn->e[l - 1], dn->dp[l - 1], dn->dcat[l - 1], catC, lowC, N, h * 2, w * 2,
                                                        encC[l - 1], dn->de[l - 1]);
    }
  }
}

int packWeights(prt_denoiser* dn)
{
  for (int i = 0; i < 10; ++i)
  {
    const DenoiseLayer& l = dn->L[i];
    kPackWeights<<<blocksFor(size_t(9) * l.coutPad * l.cinPad), kBlock, 0, dn->stream>>>(dn->w + l.wOff, l.cin, l.cout, l.cinPad,
                                                                                         l.coutPad, dn->wf[i], dn->wt[i]);
  }
  DN_TRY(hipGetLastError());
  return PRT_OK;
}

// Loss and parameter gradient of a batch (inputs already on the device); the gradient is left
// in dn->g (master layout).  Returns the mean L1 loss.
int lossAndGrad(prt_denoiser* dn, int N, int H, int W, double* lossOut)
{
  const size_t npix = size_t(N) * H * W;
  DN_TRY(hipMemsetAsync(dn->loss, 0, sizeof(float), dn->stream));
  DN_TRY(hipMemsetAsync(dn->g, 0, sizeof(float) * dn->nParams, dn->stream));
  for (int i = 0; i < 10; ++i)
    DN_TRY(hipMemsetAsync(dn->dwp[i], 0, sizeof(float) * 9 * dn->L[i].coutPad * dn->L[i].cinPad, dn->stream));
  kInput<<<blocksFor(npix), kBlock, 0, dn->stream>>>(dn->in9, npix, 1.f, dn->x);
  forward(dn, N, H, W);
  kLoss<<<blocksFor(npix), kBlock, 0, dn->stream>>>(dn->pred, dn->tgt, npix, 1.0f / float(npix * 3), dn->dout, dn->loss);
  backward(dn, N, H, W);
  for (int i = 0; i < 10; ++i)
  {
    const DenoiseLayer& l = dn->L[i];
    kUnpackGrad<<<blocksFor(size_t(l.cout) * l.cin * 9), kBlock, 0, dn->stream>>>(dn->dwp[i], l.cin, l.cout, l.cinPad, l.coutPad,
                                                                                  dn->g + l.wOff);
  }
  DN_TRY(hipGetLastError());
  float loss = 0.f;
  DN_TRY(hipMemcpyAsync(&loss, dn->loss, sizeof(float), hipMemcpyDeviceToHost, dn->stream));
  DN_TRY(hipStreamSynchronize(dn->stream));
  if (lossOut) *lossOut = double(loss) / double(npix * 3);
  return PRT_OK;
}

int uploadBatch(prt_denoiser* dn, const float* input, const float* target, int N, int H, int W)
{
  if (N < 1 || H < 16 || W < 16 || H % 16 || W % 16) return dfail(PRT_ERR_INVALID, "batch images must be multiples of 16 in size");
  int rc = ensureActs(dn, N, H, W);
  if (rc != PRT_OK) return rc;
  const size_t npix = size_t(N) * H * W;
  DN_TRY(hipMemcpyAsync(dn->in9, input, sizeof(float) * npix * 9, hipMemcpyHostToDevice, dn->stream));
  if (target) DN_TRY(hipMemcpyAsync(dn->tgt, target, sizeof(float) * npix * 3, hipMemcpyHostToDevice, dn->stream));
  return PRT_OK;
}

}  // namespace

extern "C" {

int prt_denoiser_create(int32_t device_ordinal, prt_denoiser** out)
{
  if (!out) return dfail(PRT_ERR_INVALID, "out is null");
  *out = nullptr;
  int n = 0;
  DN_TRY(hipGetDeviceCount(&n));
  if (device_ordinal < 0 || device_ordinal >= n) return dfail(PRT_ERR_INVALID, "invalid device ordinal");
  DN_TRY(hipSetDevice(device_ordinal));
  prt_denoiser* dn = new prt_denoiser();
  dn->device = device_ordinal;
  size_t off = 0;
  for (int i = 0; i < 10; ++i)
  {
    DenoiseLayer& l = dn->L[i];
    l.cin = kLayerCin[i];
    l.cout = kLayerCout[i];
    l.cinPad = (l.cin + 15) / 16 * 16;
    l.coutPad = (l.cout + 15) / 16 * 16;
    l.wOff = off;
    off += size_t(l.cout) * l.cin * 9 + l.cout;
  }
  dn->nParams = off;
  hipError_t e = hipStreamCreateWithFlags(&dn->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&dn->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&dn->ev[1]);
  if (e == hipSuccess) e = hipMalloc(&dn->w, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&dn->m, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&dn->v, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&dn->g, sizeof(float) * off);
  for (int i = 0; i < 10 && e == hipSuccess; ++i)
  {
    const size_t k = size_t(9) * dn->L[i].coutPad * dn->L[i].cinPad;
    e = hipMalloc(&dn->wf[i], sizeof(uint16_t) * k);
    if (e == hipSuccess) e = hipMalloc(&dn->wt[i], sizeof(uint16_t) * k);
    if (e == hipSuccess) e = hipMalloc(&dn->dwp[i], sizeof(float) * k);
  }
  if (e != hipSuccess)
  {
    prt_denoiser_destroy(dn);
    return dfail(PRT_ERR_NOMEM, hipGetErrorString(e));
  }
  // He-initialised weights from a fixed seed (a 32-bit LCG), zero biases, zero Adam state
  std::vector<float> w(off, 0.f);
  uint32_t s = 12345u;
  auto uni = [&]() { s = s * 1664525u + 1013904223u; return (float(s >> 8) + 0.5f) * (1.0f / 16777216.0f); };
  for (int i = 0; i < 10; ++i)
  {
    const DenoiseLayer& l = dn->L[i];
    const float bound = std::sqrt(6.0f / float(9 * l.cin));
    for (size_t k = 0; k < size_t(l.cout) * l.cin * 9; ++k) w[l.wOff + k] = (2.f * uni() - 1.f) * bound;
  }
  e = hipMemcpy(dn->w, w.data(), sizeof(float) * off, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dn->m, 0, sizeof(float) * off);
  if (e == hipSuccess) e = hipMemset(dn->v, 0, sizeof(float) * off);
  if (e != hipSuccess || packWeights(dn) != PRT_OK)
  {
    prt_denoiser_destroy(dn);
    return dfail(PRT_ERR_HIP, "denoiser initialisation failed");
  }
  *out = dn;
  return PRT_OK;
}

int prt_denoiser_destroy(prt_denoiser* dn)
{
  if (!dn) return PRT_OK;
  hipSetDevice(dn->device);
  if (dn->stream) (void)hipStreamSynchronize(dn->stream);
  freeActs(dn);
  for (float* p : {dn->w, dn->m, dn->v, dn->g})
    if (p) hipFree(p);
  for (int i = 0; i < 10; ++i)
  {
    if (dn->wf[i]) hipFree(dn->wf[i]);
    if (dn->wt[i]) hipFree(dn->wt[i]);
    if (dn->dwp[i]) hipFree(dn->dwp[i]);
  }
  for (auto& e : dn->ev)
    if (e) (void)hipEventDestroy(e);
  if (dn->stream) (void)hipStreamDestroy(dn->stream);
  delete dn;
  return PRT_OK;
}

int prt_denoiser_get_version(prt_denoiser* dn, char* buf, int32_t len)
{
  if (!dn || !buf || len < 1) return dfail(PRT_ERR_INVALID, "null argument");
  std::snprintf(buf, size_t(len), "prt-unet-bf16 1.0 (%zu parameters)", dn->nParams);
  return PRT_OK;
}

int prt_denoiser_param_count(prt_denoiser* dn, int64_t* n)
{
  if (!dn || !n) return dfail(PRT_ERR_INVALID, "null argument");
  *n = int64_t(dn->nParams);
  return PRT_OK;
}

int prt_denoiser_get_weights(prt_denoiser* dn, float* w, int64_t n)
{
  if (!dn || !w || n != int64_t(dn->nParams)) return dfail(PRT_ERR_INVALID, "weights need param_count floats");
  DN_TRY(hipSetDevice(dn->device));
  DN_TRY(hipMemcpy(w, dn->w, sizeof(float) * dn->nParams, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_denoiser_set_weights(prt_denoiser* dn, const float* w, int64_t n)
{
  if (!dn || !w || n != int64_t(dn->nParams)) return dfail(PRT_ERR_INVALID, "weights need param_count floats");
  DN_TRY(hipSetDevice(dn->device));
  DN_TRY(hipMemcpy(dn->w, w, sizeof(float) * dn->nParams, hipMemcpyHostToDevice));
  DN_TRY(hipMemset(dn->m, 0, sizeof(float) * dn->nParams));
  DN_TRY(hipMemset(dn->v, 0, sizeof(float) * dn->nParams));
  dn->step = 0;
  int rc = packWeights(dn);
  if (rc != PRT_OK) return rc;
  DN_TRY(hipStreamSynchronize(dn->stream));
  return PRT_OK;
}

int prt_denoiser_gradient(prt_denoiser* dn, const float* input, const float* target, int32_t n, int32_t h, int32_t w, double* loss,
                          float* grad, int64_t grad_count)
{
  if (!dn || !input || !target) return dfail(PRT_ERR_INVALID, "null argument");
  if (grad && grad_count != int64_t(dn->nParams)) return dfail(PRT_ERR_INVALID, "grad needs param_count floats");
  DN_TRY(hipSetDevice(dn->device));
  int rc = uploadBatch(dn, input, target, n, h, w);
  if (rc == PRT_OK) rc = lossAndGrad(dn, n, h, w, loss);
  if (rc != PRT_OK) return rc;
  if (grad) DN_TRY(hipMemcpy(grad, dn->g, sizeof(float) * dn->nParams, hipMemcpyDeviceToHost));
  return PRT_OK;
}

int prt_denoiser_train_step(prt_denoiser* dn, const float* input, const float* target, int32_t n, int32_t h, int32_t w, float lr,
                            double* loss, double* ms)
{
  if (!dn || !input || !target) return dfail(PRT_ERR_INVALID, "null argument");
  DN_TRY(hipSetDevice(dn->device));
  int rc = uploadBatch(dn, input, target, n, h, w);
  if (rc != PRT_OK) return rc;
  DN_TRY(hipEventRecord(dn->ev[0], dn->stream));
  rc = lossAndGrad(dn, n, h, w, loss);
  if (rc != PRT_OK) return rc;
  ++dn->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - std::pow(b1, float(dn->step))), c2 = 1.0f / (1.0f - std::pow(b2, float(dn->step)));
  kAdam<<<blocksFor(dn->nParams), kBlock, 0, dn->stream>>>(dn->w, dn->m, dn->v, dn->g, dn->nParams, lr, b1, b2, c1, c2, 1e-8f);
  rc = packWeights(dn);
  if (rc != PRT_OK) return rc;
  DN_TRY(hipEventRecord(dn->ev[1], dn->stream));
  DN_TRY(hipStreamSynchronize(dn->stream));
  float t = 0.f;
  DN_TRY(hipEventElapsedTime(&t, dn->ev[0], dn->ev[1]));
  if (ms) *ms = t;
  return PRT_OK;
}

int prt_denoiser_init(prt_denoiser* dn, const prt_denoiser_params* p)
{
  if (!dn || !p) return dfail(PRT_ERR_INVALID, "null argument");
  if (!p->color || !p->output || p->width < 1 || p->height < 1) return dfail(PRT_ERR_INVALID, "colour, output and size are required");
  if (!p->albedo || !p->normal) return dfail(PRT_ERR_INVALID, "this filter needs the albedo and normal images");
  dn->params = *p;
  dn->inited = true;
  return PRT_OK;
}

int prt_denoiser_execute(prt_denoiser* dn, double* seconds)
{
  if (!dn || !dn->inited) return dfail(PRT_ERR_STATE, "init the denoiser first");
  DN_TRY(hipSetDevice(dn->device));
  const auto t0 = std::chrono::high_resolution_clock::now();
  const prt_denoiser_params& p = dn->params;
  const int W = p.width, H = p.height;
  const int Wp = (W + 15) / 16 * 16, Hp = (H + 15) / 16 * 16;  // padded by edge replication
  std::vector<float> in9(size_t(Wp) * Hp * 9);
  for (int y = 0; y < Hp; ++y)
    for (int x = 0; x < Wp; ++x)
    {
      const size_t s = size_t(std::min(y, H - 1)) * W + std::min(x, W - 1), d = size_t(y) * Wp + x;
      for (int c = 0; c < 3; ++c)
      {
        in9[d * 9 + c] = p.color[s * 3 + c];
        in9[d * 9 + 3 + c] = p.albedo[s * 3 + c];
        in9[d * 9 + 6 + c] = p.normal[s * 3 + c];
      }
    }
  int rc = ensureActs(dn, 1, Hp, Wp);
  if (rc != PRT_OK) return rc;
  const size_t npix = size_t(Hp) * Wp;
  const float scale = std::isnan(p.hdrScale) || p.hdrScale <= 0.f ? 1.f : p.hdrScale;
  DN_TRY(hipMemcpyAsync(dn->in9, in9.data(), sizeof(float) * npix * 9, hipMemcpyHostToDevice, dn->stream));
  DN_TRY(hipEventRecord(dn->ev[0], dn->stream));
  kInput<<<blocksFor(npix), kBlock, 0, dn->stream>>>(dn->in9, npix, scale, dn->x);
  forward(dn, 1, Hp, Wp);
  kOutput<<<blocksFor(npix * 3), kBlock, 0, dn->stream>>>(dn->pred, npix, 1.0f / scale, dn->outF);
  DN_TRY(hipEventRecord(dn->ev[1], dn->stream));
  std::vector<float> out(npix * 3);
  DN_TRY(hipMemcpyAsync(out.data(), dn->outF, sizeof(float) * npix * 3, hipMemcpyDeviceToHost, dn->stream));
  DN_TRY(hipStreamSynchronize(dn->stream));
  DN_TRY(hipGetLastError());
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c) p.output[(size_t(y) * W + x) * 3 + c] = out[(size_t(y) * Wp + x) * 3 + c];
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return PRT_OK;
}

int prt_denoiser_kernel_ms(prt_denoiser* dn, double* ms)
{
  if (!dn || !ms) return dfail(PRT_ERR_INVALID, "null argument");
  float t = 0.f;
  DN_TRY(hipEventElapsedTime(&t, dn->ev[0], dn->ev[1]));
  *ms = t;
  return PRT_OK;
}

}  // extern "C"
