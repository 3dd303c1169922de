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
// Convolutions are implicit GEMMs on LDS-staged tiles.  Forward and data gradient (kConvTile):
// a block computes a 16 x 32 pixel tile for up to 64 output channels, M = output channels,
// N = pixels, K = 9 taps x input channels; the data gradient is the same kernel on the output
// gradient with the weights transposed and flipped, and its output is masked by the ReLU of the
// layer input.  Weight gradient (kWgradTile): M = output channels, N = input channels, K =
// pixels, fragments transposed on the matrix cores, reduced with FP32 atomics; the bias
// gradient comes out of the same kernel.
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

// Denoiser input images on the device: colour, albedo and normal as 3-channel FP32 images of
// W x H pixels, element (channel k, pixel i) at k * kStride + i * iStride (a frame's planar sums:
// kStride = W H, iStride = 1; interleaved means: 1 and 3), each optionally divided by a per-pixel
// sample weight (sum * (weight == 0 ? FLT_MAX : 1 / weight), as a blit forms the mean,
// image/accum_buffer.h:607-649).  A missing auxiliary image is a zero channel.
struct FrameIn
{
  const float *color, *colorW, *albedo, *albedoW, *normal, *normalW;
  int W, H, kStride, iStride;
};
__device__ inline float frameMean(const float* s, const float* w, const FrameIn& f, size_t i, int k)
{
  const float v = s[size_t(k) * f.kStride + i * f.iStride];
  if (!w) return v;
  const float wt = w[i];
  return v * (wt == 0.f ? 3.402823466e38f : 1.0f / wt);
}

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
  const uint16_t* lo;    // optional: input channels [0, loC) from this half-resolution tensor
  int loC;               // (loC channels, nearest-upsampled); channels [loC, cinPad) from x
  uint16_t* pool;        // optional: 2x2 max pooling of the BF16 output (coutPad channels)
  // output layer of an inference: expm1(value) * invScale written to out (outW x outH x 3, the
  // padded image cropped)
  float* out;
  int outW, outH;
  float invScale;
};

// 8 BF16 values <-> one 16-byte word
struct Bf8
{
  uint16_t h[8];
};
__device__ inline Bf8 loadBf8(const uint16_t* p)
{
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  Bf8 r;
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; ++k)
  {
    r.h[2 * k] = uint16_t(w[k] & 0xffffu);
    r.h[2 * k + 1] = uint16_t(w[k] >> 16);
  }
  return r;
}
__device__ inline void storeBf8(uint16_t* p, const Bf8& r)
{
  uint4 v;
  v.x = uint32_t(r.h[0]) | uint32_t(r.h[1]) << 16;
  v.y = uint32_t(r.h[2]) | uint32_t(r.h[3]) << 16;
  v.z = uint32_t(r.h[4]) | uint32_t(r.h[5]) << 16;
  v.w = uint32_t(r.h[6]) | uint32_t(r.h[7]) << 16;
  *reinterpret_cast<uint4*>(p) = v;
}

// Tiled convolution (forward and data gradient) on v_mfma_f32_16x16x32_bf16.  A block of ROWS/2
// waves computes ROWS x 32 pixels of one image for 16*NB output channels; each wave owns two
// image rows (4 pixel blocks of 16) for all output blocks.  K runs over 16-channel input slices:
// per slice the haloed (ROWS+2) x 34 input tile and the slice's weights (9 taps x 16 NB outputs)
// sit in LDS, double-buffered, so one barrier per slice separates staging from the MFMAs, and
// the next slice's global loads are in flight while the current one is multiplied.  One MFMA
// covers two taps: lanes 0-31 carry tap t (channels 0-7 and 8-15), lanes 32-63 tap t+1; the
// ninth tap pairs with zeros.  Both LDS arrays are stored half-major (the 8-channel half of a
// pixel or weight row is the outer index), so the 16 lanes of an operand read 16 consecutive
// 16-byte words.  The MFMA takes the weights as the row operand and the pixels as the column
// operand, so each lane ends with 4 consecutive output channels of one pixel and stores them as
// one 8-byte word.
// Optional fusions: the input's first loC channels come from a half-resolution tensor `lo`,
// upsampled by nearest neighbour (the decoder's upsample + concatenation, cat = [up(lo), x]);
// and with `pool`, the 2x2 max pooling of the (BF16) output is written too (a wave's two rows
// hold whole pooling windows; H and W are even).
constexpr int kCvX = 32, kCvIX = kCvX + 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int NB, int ROWS>
constexpr int convStageWords()  // 16-byte words of one LDS stage (input tile + weight slice)
{
  return (ROWS + 2) * kCvIX * 2 + 9 * 16 * NB * 2;
}

template <int NB, int ROWS>
__global__ void __launch_bounds__(ROWS * 32) kConv(ConvArgs a)
{
  constexpr int kThreads = ROWS * 32;
  constexpr int kInPix = (ROWS + 2) * kCvIX;
  constexpr int kInChunks = kInPix * 2;      // 16-byte words of the input tile
  constexpr int kWChunks = 9 * 16 * NB * 2;  // 16-byte words of the weight slice
  constexpr int kXItems = (kInChunks + kThreads - 1) / kThreads;
  constexpr int kWItems = (kWChunks + kThreads - 1) / kThreads;
  // one stage per buffer: the input tile [half * kInPix + pixel], then the weight slice
  // [((tap * NB + b) * 2 + half) * 16 + row]; a layer of one 16-channel slice gets one buffer
  extern __shared__ uint4 smem[];
  constexpr int kStage = convStageWords<NB, ROWS>();
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tilesX = (a.W + kCvX - 1) / kCvX, tilesY = (a.H + ROWS - 1) / ROWS;
  int t = blockIdx.x;
  const int tx = t % tilesX; t /= tilesX;
  const int ty = t % tilesY;
  const int n = t / tilesY;
  const int co0 = blockIdx.y * 16 * NB;
  const int nb = min(NB, (a.coutPad - co0) / 16);
  const int x0 = tx * kCvX, y0 = ty * ROWS;
  const int li = lane & 15, g = lane >> 4, tapSel = g >> 1, half = g & 1, lk = 4 * g;
  f32x4 acc[4][NB];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[p][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int Hl = a.H >> 1, Wl = a.W >> 1;
  uint4 xr[kXItems], wr[kWItems];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int k = 0; k < kXItems; ++k)
    {
      const int i = tid + k * kThreads, hf = i & 1, pix = i >> 1;
      const int yy = y0 + pix / kCvIX - 1, xx = x0 + pix % kCvIX - 1;
      xr[k] = make_uint4(0u, 0u, 0u, 0u);
      if (i < kInChunks && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
      {
        const uint16_t* src = c0 < a.loC ? a.lo + ((size_t(n) * Hl + (yy >> 1)) * Wl + (xx >> 1)) * a.loC + c0 + 8 * hf
                                         : a.x + ((size_t(n) * a.H + yy) * a.W + xx) * a.xC + (c0 - a.loC) + 8 * hf;
        xr[k] = *reinterpret_cast<const uint4*>(src);
      }
    }
#pragma unroll
    for (int k = 0; k < kWItems; ++k)
    {
      const int i = tid + k * kThreads, hf = i & 1, r = i >> 1;
      const int co = co0 + r % (16 * NB), tap = r / (16 * NB);
      wr[k] = make_uint4(0u, 0u, 0u, 0u);
      if (i < kWChunks && co < a.coutPad)
        wr[k] = *reinterpret_cast<const uint4*>(a.w + (size_t(tap) * a.coutPad + co) * a.cinPad + c0 + 8 * hf);
    }
  };
  auto stage = [&](int buf) {
    uint4* const xsb = smem + buf * kStage;
    uint4* const wsb = xsb + kInChunks;
#pragma unroll
    for (int k = 0; k < kXItems; ++k)
    {
      const int i = tid + k * kThreads;
      if (i < kInChunks) xsb[(i & 1) * kInPix + (i >> 1)] = xr[k];
    }
#pragma unroll
    for (int k = 0; k < kWItems; ++k)
    {
      const int i = tid + k * kThreads, r = i >> 1;
      if (i < kWChunks)
      {
        const int rr = r % (16 * NB), tap = r / (16 * NB);
        wsb[((tap * NB + (rr >> 4)) * 2 + (i & 1)) * 16 + (rr & 15)] = wr[k];
      }
    }
  };
  fetch(0);
  stage(0);
  __syncthreads();
  const int slices = a.cinPad / 16;
  for (int sl = 0; sl < slices; ++sl)
  {
    const int buf = sl & 1;
    const uint4* const xsb = smem + buf * kStage;
    const uint4* const wsb = xsb + kInChunks;
    if (sl + 1 < slices) fetch((sl + 1) * 16);
#pragma unroll
    for (int q = 0; q < 5; ++q)
    {
      const int tapRaw = 2 * q + tapSel;
      const bool valid = tapRaw < 9;
      const int tap = valid ? tapRaw : 8;
      const int oy = tap / 3, ox = tap % 3;
      bf16x8 B[4];
#pragma unroll
      for (int p = 0; p < 4; ++p)
      {
        const int ly = 2 * wave + (p >> 1) + oy, lx = 16 * (p & 1) + li + ox;
        uint4 v = xsb[half * kInPix + ly * kCvIX + lx];
        if (!valid) v = make_uint4(0u, 0u, 0u, 0u);
        B[p] = __builtin_bit_cast(bf16x8, v);
      }
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (b < nb)
        {
          uint4 v = wsb[((tap * NB + b) * 2 + half) * 16 + li];
          if (!valid) v = make_uint4(0u, 0u, 0u, 0u);
          const bf16x8 A = __builtin_bit_cast(bf16x8, v);
#pragma unroll
          for (int p = 0; p < 4; ++p) acc[p][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[p], acc[p][b], 0, 0, 0);
        }
    }
    if (sl + 1 < slices) stage(buf ^ 1);
    __syncthreads();
  }
  // D of (p, b): lane holds pixel li of pixel block p, output channels co0 + 16 b + lk + j.
  // Epilogue: bias, ReLU, zero padding channels, the mask; the BF16 values are kept for pooling.
  float bias[NB][4];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
      const int co = co0 + 16 * b + lk + j;
      bias[b][j] = a.bias && co < a.coutReal ? a.bias[co] : 0.f;
    }
  uint16_t hv[4][NB][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
  {
    const int y = y0 + 2 * wave + (p >> 1), x = x0 + 16 * (p & 1) + li;
    const bool inside = y < a.H && x < a.W;
    const size_t pix = (size_t(n) * a.H + y) * a.W + x;
#pragma unroll
    for (int b = 0; b < NB; ++b)
    {
      const int co = co0 + 16 * b + lk;
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        float u = acc[p][b][j] + bias[b][j];
        if (a.relu) u = fmaxf(u, 0.f);
        v[j] = co + j < a.coutReal ? u : 0.f;
        hv[p][b][j] = toBf16(v[j]);
      }
      if (!inside || b >= nb) continue;
      const size_t o = pix * a.yC + a.yOff + co;
      if (a.mask)
      {
        const uint2 m = *reinterpret_cast<const uint2*>(a.mask + o);
        const uint32_t mw[2] = {m.x, m.y};
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (!(fromBf16(uint16_t(mw[j >> 1] >> (16 * (j & 1)))) > 0.f)) { v[j] = 0.f; hv[p][b][j] = 0; }
      }
      if (a.y)
        *reinterpret_cast<uint2*>(a.y + o) = make_uint2(uint32_t(hv[p][b][0]) | uint32_t(hv[p][b][1]) << 16,
                                                         uint32_t(hv[p][b][2]) | uint32_t(hv[p][b][3]) << 16);
      if (a.yf)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (co + j < a.coutReal) a.yf[pix * a.coutReal + co + j] = v[j];
      if (a.out && y < a.outH && x < a.outW)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (co + j < a.coutReal)
            a.out[((size_t(n) * a.outH + y) * a.outW + x) * a.coutReal + co + j] = expm1f(v[j]) * a.invScale;
    }
  }
  if (a.pool)
  {
    // windows (2 wave, 2 i) .. (2 wave + 1, 2 i + 1): rows from pixel blocks p and p + 2, columns
    // from lanes li and li ^ 1; the maximum is taken in kPool's order on the stored BF16 values
    const int y = y0 + 2 * wave;
#pragma unroll
    for (int p = 0; p < 2; ++p)
    {
      const int x = x0 + 16 * p + li;
#pragma unroll
      for (int b = 0; b < NB; ++b)
      {
        if (b >= nb) continue;
        uint16_t m[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
          const uint16_t v00 = hv[p][b][j], v10 = hv[p + 2][b][j];
          const uint16_t v01 = uint16_t(__shfl_xor(int(v00), 1)), v11 = uint16_t(__shfl_xor(int(v10), 1));
          uint16_t r = v00;
          if (fromBf16(v01) > fromBf16(r)) r = v01;
          if (fromBf16(v10) > fromBf16(r)) r = v10;
          if (fromBf16(v11) > fromBf16(r)) r = v11;
          m[j] = r;
        }
        if (!(li & 1) && y < a.H && x < a.W)
        {
          const size_t pp = (size_t(n) * Hl + (y >> 1)) * Wl + (x >> 1);
          *reinterpret_cast<uint2*>(a.pool + pp * a.coutPad + co0 + 16 * b + lk) =
              make_uint2(uint32_t(m[0]) | uint32_t(m[1]) << 16, uint32_t(m[2]) | uint32_t(m[3]) << 16);
        }
      }
    }
  }
}

struct WgradArgs
{
  const uint16_t* dy;  // [pixels][coutPad] (masked output gradient)
  const uint16_t* x;   // [pixels][cinPad] layer input
  int N, H, W, cinPad, coutPad, rows;
  float* part;   // per (image, row chunk) partial sums: [N * chunks][9][coutPad][cinPad]
  float* partB;  // per (image, row chunk) partial bias sums: [N * chunks][coutPad] (null: no bias)
};

// 16 x 16 BF16 fragment transpose on the matrix cores: with the identity as the column operand,
// D = A, and D's lane layout (column = lane % 16, rows 4 (lane / 16) + j) is the transpose of the
// row operand's (row = lane % 16, k = 4 (lane / 16) + j).  Exact: one product x * 1 per output.
__device__ inline bf16x4 transposeFrag(const bf16x4 a, const bf16x4 ident)
{
  const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ident, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  bf16x4 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) r[j] = short(__float_as_uint(d[j]) >> 16);
  return r;
}

// Weight gradient dW[tap][co][ci] = sum over pixels p of dy[p][co] * x[p + tap offset][ci].
// A block (6 waves) owns 64 output x 32 input channels of all 9 taps and `rows` image rows of
// one image.  Per 64-pixel row segment it copies dy (64 pixels x 64 channels) and the three
// input rows the taps read (66 pixels x 32 channels each) into LDS in their NHWC layout (padded
// pixel strides, so the 16 pixels of a fragment hit distinct banks); the next segment's global
// loads are issued before the current segment's MFMAs.  The MFMA wants K = pixels in both
// operands: each 16 x 16 fragment is read with channels contiguous and transposed on the matrix
// cores (transposeFrag).  Wave w computes the three taps of kernel row w % 3 for the input
// channels 16 (w / 3) .. +15 of the group; the blocks of the first input-channel group also sum
// dy into the bias gradient (a product with ones).  Pixels beyond the image width have dy = 0.
// Each block stores its sums in its own slice of a partial buffer (no atomics: thousands of
// blocks would contend for the same few thousand weights); kWgradReduce adds the slices.
constexpr int kWgThreads = 384, kWgSeg = 64;
constexpr int kWgDyStride = 68, kWgXStride = 36;  // LDS pixel strides in BF16 elements
constexpr int kWgDyItems = (kWgSeg * 8 + kWgThreads - 1) / kWgThreads;
constexpr int kWgXItems = (3 * (kWgSeg + 2) * 4 + kWgThreads - 1) / kWgThreads;

__global__ void __launch_bounds__(kWgThreads) __attribute__((amdgpu_waves_per_eu(3))) kWgradTile(WgradArgs a)
{
  __shared__ __align__(16) uint16_t dyS[kWgSeg * kWgDyStride];
  __shared__ __align__(16) uint16_t xS[3 * (kWgSeg + 2) * kWgXStride];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int krow = wave % 3, half = wave / 3;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  const int coGroups = (a.coutPad + 63) / 64, ciGroups = (a.cinPad + 31) / 32;
  const int chunks = (a.H + a.rows - 1) / a.rows;
  int t = blockIdx.x;
  const int cg = t % coGroups; t /= coGroups;
  const int ig = t % ciGroups; t /= ciGroups;
  const int ch = t % chunks;
  const int n = t / chunks;
  const int co0 = cg * 64, ci0 = ig * 32;
  const int ncb = min(4, (a.coutPad - co0) / 16);
  const bool active = ci0 + 16 * half < a.cinPad;  // this wave's 16 input channels exist
  const bool doBias = a.partB && ig == 0 && wave == 0;
  bf16x4 ident, ones;
#pragma unroll
  for (int j = 0; j < 4; ++j)
  {
    ident[j] = short(lk + j == li ? 0x3F80 : 0);
    ones[j] = short(0x3F80);
  }
  f32x4 acc[3][4], accB[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
  {
    accB[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 3; ++s) acc[s][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const int yBeg = ch * a.rows, yEnd = min(a.H, yBeg + a.rows);
  const int segsPerRow = (a.W + kWgSeg - 1) / kWgSeg;
  const int nSeg = (yEnd - yBeg) * segsPerRow;
  uint4 dr[kWgDyItems], xr[kWgXItems];
  auto fetch = [&](int sIdx) {
    const int y = yBeg + sIdx / segsPerRow, x0 = (sIdx % segsPerRow) * kWgSeg;
#pragma unroll
    for (int k = 0; k < kWgDyItems; ++k)
    {
      const int i = tid + k * kWgThreads, px = i >> 3, oct = i & 7, x = x0 + px;
      dr[k] = make_uint4(0u, 0u, 0u, 0u);
      if (i < kWgSeg * 8 && x < a.W && co0 + 8 * oct < a.coutPad)
        dr[k] = *reinterpret_cast<const uint4*>(a.dy + ((size_t(n) * a.H + y) * a.W + x) * a.coutPad + co0 + 8 * oct);
    }
#pragma unroll
    for (int k = 0; k < kWgXItems; ++k)
    {
      const int i = tid + k * kWgThreads, oct = i & 3, r = i >> 2;
      const int row = r / (kWgSeg + 2), q = r % (kWgSeg + 2);
      const int yy = y + row - 1, xx = x0 - 1 + q;
      xr[k] = make_uint4(0u, 0u, 0u, 0u);
      if (i < 3 * (kWgSeg + 2) * 4 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W && ci0 + 8 * oct < a.cinPad)
        xr[k] = *reinterpret_cast<const uint4*>(a.x + ((size_t(n) * a.H + yy) * a.W + xx) * a.cinPad + ci0 + 8 * oct);
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int k = 0; k < kWgDyItems; ++k)
    {
      const int i = tid + k * kWgThreads;
      if (i < kWgSeg * 8)
      {
        uint2* d = reinterpret_cast<uint2*>(dyS + (i >> 3) * kWgDyStride + 8 * (i & 7));
        d[0] = make_uint2(dr[k].x, dr[k].y);
        d[1] = make_uint2(dr[k].z, dr[k].w);
      }
    }
#pragma unroll
    for (int k = 0; k < kWgXItems; ++k)
    {
      const int i = tid + k * kWgThreads;
      if (i < 3 * (kWgSeg + 2) * 4)
      {
        uint2* d = reinterpret_cast<uint2*>(xS + (i >> 2) * kWgXStride + 8 * (i & 3));
        d[0] = make_uint2(xr[k].x, xr[k].y);
        d[1] = make_uint2(xr[k].z, xr[k].w);
      }
    }
  };
  if (nSeg > 0) fetch(0);
  for (int sIdx = 0; sIdx < nSeg; ++sIdx)
  {
    __syncthreads();
    stage();
    __syncthreads();
    if (sIdx + 1 < nSeg) fetch(sIdx + 1);
    if (!active) continue;
    for (int k0 = 0; k0 < kWgSeg; k0 += 16)
    {
      bf16x4 A[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < ncb)
        {
          A[c] = transposeFrag(*reinterpret_cast<const bf16x4*>(dyS + (k0 + li) * kWgDyStride + 16 * c + lk), ident);
          if (doBias) accB[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c], ones, accB[c], 0, 0, 0);
        }
#pragma unroll
      for (int s = 0; s < 3; ++s)
      {
        const bf16x4 B = transposeFrag(
            *reinterpret_cast<const bf16x4*>(xS + (krow * (kWgSeg + 2) + k0 + li + s) * kWgXStride + 16 * half + lk), ident);
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c < ncb) acc[s][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[c], B, acc[s][c], 0, 0, 0);
      }
    }
  }
  // D of (s, c): lane holds input channel ci0 + 16 half + li for output channels co0 + 16 c + lk + r
  const size_t pIdx = size_t(n) * chunks + ch;
  if (active)
  {
    float* dst = a.part + pIdx * (size_t(9) * a.coutPad * a.cinPad);
    const int ci = ci0 + 16 * half + li;
#pragma unroll
    for (int s = 0; s < 3; ++s)
    {
      const int tap = krow * 3 + s;
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        if (c >= ncb) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          const int co = co0 + 16 * c + lk + r;
          dst[(size_t(tap) * a.coutPad + co) * a.cinPad + ci] = acc[s][c][r];
        }
      }
    }
  }
  // bias: every column of accB holds the sum over pixels; column 0 (lanes 0, 16, 32, 48) stores it
  if (doBias && li == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ncb)
#pragma unroll
        for (int r = 0; r < 4; ++r) a.partB[pIdx * a.coutPad + co0 + 16 * c + lk + r] = accB[c][r];
}

// Sum of the per-block partial weight (and bias) gradients.  A block takes 32 consecutive
// entries (weights, then biases); its 8 thread rows each sum every 8th partial, and the rows are
// added in LDS in a fixed order.
constexpr int kRedEl = 32, kRedRows = kBlock / kRedEl;

__global__ void __launch_bounds__(kBlock) kWgradReduce(const float* __restrict__ part, const float* __restrict__ partB, int nPart,
                                                       size_t size, int coutPad, int coutReal, float* __restrict__ dw,
                                                       float* __restrict__ db)
{
  __shared__ float red[kRedRows][kRedEl];
  const int el = threadIdx.x % kRedEl, row = threadIdx.x / kRedEl;
  const size_t e = size_t(blockIdx.x) * kRedEl + el;
  const bool isW = e < size, isB = !isW && partB && e < size + size_t(coutReal);
  float t = 0.f;
  if (isW)
  {
#pragma unroll 4
    for (int p = row; p < nPart; p += kRedRows) t += part[size_t(p) * size + e];
  }
  else if (isB)
  {
    const int co = int(e - size);
#pragma unroll 4
    for (int p = row; p < nPart; p += kRedRows) t += partB[size_t(p) * coutPad + co];
  }
  red[row][el] = t;
  __syncthreads();
  if (row == 0 && (isW || isB))
  {
    float s = 0.f;
    for (int r = 0; r < kRedRows; ++r) s += red[r][el];
    if (isW) dw[e] = s;
    else db[e - size] = s;
  }
}

// Encoder level backward: de = (pool gradient routed to the first maximum of each 2x2 window
// + the skip slice of the decoder's concat gradient), masked by ReLU(e) > 0.  A thread takes
// 8 channels of one 2x2 window.
__global__ void __launch_bounds__(kBlock) kPoolBwd(const uint16_t* __restrict__ e, const uint16_t* __restrict__ dp,
                                                   const uint16_t* __restrict__ dcat, int catC, int catOff, int N, int H, int W, int C,
                                                   uint16_t* __restrict__ de)
{
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2, octs = C / 8;
  if (i >= size_t(N) * Ho * Wo * octs) return;
  const int c = int(i % octs) * 8;
  size_t p = i / octs;
  const int xo = int(p % Wo); p /= Wo;
  const int yo = int(p % Ho);
  const int n = int(p / Ho);
  Bf8 v[4];
  size_t pix[4];
  for (int k = 0; k < 4; ++k)
  {
    pix[k] = (size_t(n) * H + 2 * yo + k / 2) * W + 2 * xo + k % 2;
    v[k] = loadBf8(e + pix[k] * C + c);
  }
  Bf8 g{};
  if (dp) g = loadBf8(dp + ((size_t(n) * Ho + yo) * Wo + xo) * C + c);
  int first[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
  {
    int best = 0;
    float bv = fromBf16(v[0].h[j]);
    for (int k = 1; k < 4; ++k)
      if (fromBf16(v[k].h[j]) > bv) { bv = fromBf16(v[k].h[j]); best = k; }
    first[j] = best;
  }
  for (int k = 0; k < 4; ++k)
  {
    Bf8 sk{};
    if (dcat) sk = loadBf8(dcat + pix[k] * catC + catOff + c);
    Bf8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
      float t = 0.f;
      if (fromBf16(v[k].h[j]) > 0.f)
      {
        if (dp && first[j] == k) t += fromBf16(g.h[j]);
        if (dcat) t += fromBf16(sk.h[j]);
      }
      o.h[j] = toBf16(t);
    }
    storeBf8(de + pix[k] * C + c, o);
  }
}

// Decoder input: cat[.., 0:Cl] = nearest 2x upsampling of the lower level, cat[.., Cl:] = skip;
// a thread copies 8 channels.
__global__ void __launch_bounds__(kBlock) kUpcat(const uint16_t* __restrict__ low, int Cl, const uint16_t* __restrict__ skip, int Cs,
                                                 int N, int H, int W, uint16_t* __restrict__ cat)
{
  const int C = Cl + Cs, octs = C / 8;
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= size_t(N) * H * W * octs) return;
  const int c = int(i % octs) * 8;
  size_t p = i / octs;
  const int x = int(p % W); p /= W;
  const int y = int(p % H);
  const int n = int(p / H);
  const uint16_t* src = c < Cl ? low + ((size_t(n) * (H / 2) + y / 2) * (W / 2) + x / 2) * Cl + c
                               : skip + ((size_t(n) * H + y) * W + x) * Cs + (c - Cl);
  *reinterpret_cast<uint4*>(cat + i * 8) = *reinterpret_cast<const uint4*>(src);
}

// Gradient of the upsampled part of a concat: the sum over each 2x2 block (already masked).
__global__ void __launch_bounds__(kBlock) kUpBwd(const uint16_t* __restrict__ dcat, int catC, int Cl, int N, int H, int W,
                                                 uint16_t* __restrict__ dlow)
{
  const int Ho = H / 2, Wo = W / 2, octs = Cl / 8;
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= size_t(N) * Ho * Wo * octs) return;
  const int c = int(i % octs) * 8;
  size_t p = i / octs;
  const int xo = int(p % Wo); p /= Wo;
  const int yo = int(p % Ho);
  const int n = int(p / Ho);
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < 4; ++k)
  {
    const Bf8 v = loadBf8(dcat + ((size_t(n) * H + 2 * yo + k / 2) * W + 2 * xo + k % 2) * catC + c);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += fromBf16(v.h[j]);
  }
  Bf8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o.h[j] = toBf16(s[j]);
  storeBf8(dlow + i * 8, o);
}

// Network input: log(1 + colour * hdrScale), albedo, normal -> 16 BF16 channels (9 used) of
// the Wp x Hp input, pixels beyond the images' W x H repeating their last row / column.
__global__ void __launch_bounds__(kBlock) kInput(FrameIn f, int Wp, int Hp, float hdrScale, uint16_t* __restrict__ x)
{
  const size_t p = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= size_t(Wp) * Hp) return;
  const int py = int(p / Wp), px = int(p % Wp);
  const size_t i = size_t(min(py, f.H - 1)) * f.W + min(px, f.W - 1);
  Bf8 lo{}, hi{};
  float v[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    v[c] = log1pf(fmaxf(frameMean(f.color, f.colorW, f, i, c) * hdrScale, 0.f));
    v[3 + c] = f.albedo ? frameMean(f.albedo, f.albedoW, f, i, c) : 0.f;
    v[6 + c] = f.normal ? frameMean(f.normal, f.normalW, f, i, c) : 0.f;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) lo.h[c] = toBf16(v[c]);
  hi.h[0] = toBf16(v[8]);
  storeBf8(x + p * 16, lo);
  storeBf8(x + p * 16 + 8, hi);
}

// L1 loss in log(1 + colour) space: loss sum (one atomic per block) and the BF16 output gradient.
__global__ void __launch_bounds__(kBlock) kLoss(const float* __restrict__ pred, const float* __restrict__ target, size_t npix,
                                                float scale, uint16_t* __restrict__ dout, float* __restrict__ loss)
{
  __shared__ float part[kBlock / 64];
  const size_t p = size_t(blockIdx.x) * kBlock + threadIdx.x;
  float l = 0.f;
  if (p < npix)
  {
    Bf8 g{}, z{};
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      const float d = pred[p * 3 + c] - log1pf(fmaxf(target[p * 3 + c], 0.f));
      l += fabsf(d);
      g.h[c] = toBf16(d > 0.f ? scale : (d < 0.f ? -scale : 0.f));
    }
    storeBf8(dout + p * 16, g);
    storeBf8(dout + p * 16 + 8, z);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    float t = 0.f;
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
    if (t != 0.f) atomicAdd(loss, t);
  }
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
  float* wgPart = nullptr;  // partial weight / bias gradients of the weight-gradient blocks
  int step = 0;
  // activation buffers for the current (N, H, W)
  int N = 0, H = 0, W = 0;
  std::vector<void*> bufs;
  uint16_t *x = nullptr, *e[5] = {}, *p[4] = {}, *cat[4] = {}, *d[4] = {};
  uint16_t *dout = nullptr, *dd[4] = {}, *dcat[4] = {}, *de[5] = {}, *dp[4] = {};
  float *pred = nullptr, *in9 = nullptr, *tgt = nullptr, *loss = nullptr, *outF = nullptr;
  // result of the last execute on a rendered frame: W x H x 3 on the device
  float* outDev = nullptr;
  size_t outDevCap = 0;
  int outW = 0, outH = 0;
  hipEvent_t evIn = nullptr;  // the frame's producer stream -> the denoiser stream
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

// Image rows per weight-gradient block: about 1024 blocks per launch.
int wgradRows(int N, int H, int coutPad, int cinPad)
{
  const size_t groups = size_t((coutPad + 63) / 64) * ((cinPad + 31) / 32);
  return int(std::min<size_t>(size_t(H), std::max<size_t>(1, (size_t(N) * H * groups + 1023) / 1024)));
}

// Floats of partial sums the weight-gradient blocks of a layer at one level need.
size_t wgradPartFloats(int N, int H, int coutPad, int cinPad)
{
  const int rows = wgradRows(N, H, coutPad, cinPad);
  const size_t nPart = size_t(N) * ((H + rows - 1) / rows);
  return nPart * (size_t(9) * coutPad * cinPad + coutPad);
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
  size_t part = 0;
  for (int i = 0; i < 10; ++i)
  {
    const int lvl = i < 5 ? i : (i < 9 ? 8 - i : 0);  // enc0..enc4, dec3..dec0, out
    part = std::max(part, wgradPartFloats(N, H >> lvl, dn->L[i].coutPad, dn->L[i].cinPad));
  }
  DN_TRY(actAlloc(dn, dn->wgPart, part));
  dn->N = N; dn->H = H; dn->W = W;
  return PRT_OK;
}

template <int NB, int ROWS>
void launchConvT(prt_denoiser* dn, const ConvArgs& a, int groups)
{
  const unsigned tiles = unsigned(a.N * ((a.H + ROWS - 1) / ROWS) * ((a.W + kCvX - 1) / kCvX));
  const size_t lds = size_t(a.cinPad > 16 ? 2 : 1) * convStageWords<NB, ROWS>() * 16;
  static const bool ldsAttr = hipFuncSetAttribute(reinterpret_cast<const void*>(&kConv<NB, ROWS>),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  2 * convStageWords<NB, ROWS>() * 16) == hipSuccess;
  (void)ldsAttr;  // a refusal surfaces as the launch's error
  kConv<NB, ROWS><<<dim3(tiles, unsigned(groups)), ROWS * 32, lds, dn->stream>>>(a);
}

template <int NB>
void launchConvNB(prt_denoiser* dn, const ConvArgs& a, int rows, int groups)
{
  switch (rows)
  {
    case 16: launchConvT<NB, 16>(dn, a, groups); break;
    case 8: launchConvT<NB, 8>(dn, a, groups); break;
    case 4: launchConvT<NB, 4>(dn, a, groups); break;
    default: launchConvT<NB, 2>(dn, a, groups); break;
  }
}

// Output channels per block (16 NB, NB <= 4, split evenly) and tile height: 16 rows of 32
// pixels, halved (down to 2) while the launch would leave CUs without a block; every block
// stages the whole weight slice, so smaller tiles only pay where the level is too small to
// fill the GPU.
void launchConv(prt_denoiser* dn, const ConvArgs& a)
{
  const int blocks16 = a.coutPad / 16;
  const int groups = (blocks16 + 3) / 4;
  const int NB = (blocks16 + groups - 1) / groups;
  int rows = 16;
  while (rows > 2 && size_t(a.N) * ((a.H + rows - 1) / rows) * ((a.W + kCvX - 1) / kCvX) * groups < 256) rows /= 2;
  switch (NB)
  {
    case 1: launchConvNB<1>(dn, a, rows, groups); break;
    case 2: launchConvNB<2>(dn, a, rows, groups); break;
    case 3: launchConvNB<3>(dn, a, rows, groups); break;
    default: launchConvNB<4>(dn, a, rows, groups); break;
  }
}

// Forward convolution of layer `layer`.  With lo, the input is [up(lo) (loC channels), x]; with
// pool, the 2x2 max pooling of y is written too.
void conv(prt_denoiser* dn, int layer, const uint16_t* x, int N, int H, int W, uint16_t* y, int yC, int yOff, bool relu, float* yf,
          const uint16_t* lo = nullptr, int loC = 0, uint16_t* pool = nullptr)
{
  const DenoiseLayer& l = dn->L[layer];
  ConvArgs a{};
  a.x = x; a.xC = l.cinPad - loC; a.w = dn->wf[layer]; a.bias = dn->w + l.wOff + size_t(l.cout) * l.cin * 9;
  a.N = N; a.H = H; a.W = W; a.cinPad = l.cinPad; a.coutPad = l.coutPad; a.coutReal = l.cout;
  a.relu = relu ? 1 : 0; a.y = y; a.yC = yC; a.yOff = yOff; a.yf = yf; a.mask = nullptr;
  a.lo = lo; a.loC = loC; a.pool = pool;
  launchConv(dn, a);
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
  launchConv(dn, a);
}

void wgrad(prt_denoiser* dn, int layer, const uint16_t* dy, const uint16_t* x, int N, int H, int W)
{
  const DenoiseLayer& l = dn->L[layer];
  const int rows = wgradRows(N, H, l.coutPad, l.cinPad);
  const int chunks = (H + rows - 1) / rows;
  const int nPart = N * chunks;
  const size_t size = size_t(9) * l.coutPad * l.cinPad;
  WgradArgs a{dy, x, N, H, W, l.cinPad, l.coutPad, rows, dn->wgPart, dn->wgPart + size_t(nPart) * size};
  const size_t groups = size_t((l.coutPad + 63) / 64) * ((l.cinPad + 31) / 32);
  kWgradTile<<<unsigned(groups * nPart), kWgThreads, 0, dn->stream>>>(a);
  kWgradReduce<<<unsigned((size + l.cout + kRedEl - 1) / kRedEl), kBlock, 0, dn->stream>>>(a.part, a.partB, nPart, size, l.coutPad, l.cout,
                                                                     dn->dwp[layer], dn->g + l.wOff + size_t(l.cout) * l.cin * 9);
}

// Source and destination of an inference: the input is formed from `fin` (padded to the
// network's size by edge replication), the output layer writes the cropped image to `out`.
struct NetIO
{
  FrameIn fin;
  float hdrScale;
  float* out;
  int outW, outH;
};

// Forward pass.  The encoder convolutions write their pooled outputs too; the decoder
// convolutions read the upsampled lower level and the skip directly.  Training (io null) reads
// the BF16 input dn->x its caller formed and writes the FP32 prediction dn->pred, and also
// materialises each concatenation, which the weight gradient and the ReLU mask read.
void forward(prt_denoiser* dn, int N, int H, int W, const NetIO* io)
{
  const bool keepCat = io == nullptr;
  const int encC[5] = {32, 48, 64, 96, 128};
  const uint16_t* in = dn->x;
  for (int l = 0; l < 5; ++l)
  {
    if (l == 0 && io)
      kInput<<<blocksFor(size_t(N) * H * W), kBlock, 0, dn->stream>>>(io->fin, W, N * H, io->hdrScale, dn->x);
    conv(dn, l, in, N, H >> l, W >> l, dn->e[l], encC[l], 0, true, nullptr, nullptr, 0, l < 4 ? dn->p[l] : nullptr);
    if (l < 4) in = dn->p[l];
  }
  const int decC[4] = {32, 48, 64, 96};
  const uint16_t* low = dn->e[4];
  int lowC = 128;
  for (int l = 3; l >= 0; --l)
  {
    const int h = H >> l, w = W >> l;
    if (keepCat)
    {
      const size_t n = size_t(N) * h * w * ((lowC + encC[l]) / 8);
      kUpcat<<<blocksFor(n), kBlock, 0, dn->stream>>>(low, lowC, dn->e[l], encC[l], N, h, w, dn->cat[l]);
    }
    conv(dn, 8 - l, dn->e[l], N, h, w, dn->d[l], decC[l], 0, true, nullptr, low, lowC);  // dec3 = layer 5 ... dec0 = layer 8
    low = dn->d[l];
    lowC = decC[l];
  }
  if (io)
  {
    const DenoiseLayer& L9 = dn->L[9];
    ConvArgs a{};
    a.x = dn->d[0]; a.xC = L9.cinPad; a.w = dn->wf[9]; a.bias = dn->w + L9.wOff + size_t(L9.cout) * L9.cin * 9;
    a.N = N; a.H = H; a.W = W; a.cinPad = L9.cinPad; a.coutPad = L9.coutPad; a.coutReal = L9.cout;
    a.out = io->out; a.outW = io->outW; a.outH = io->outH; a.invScale = 1.0f / io->hdrScale;
    launchConv(dn, a);
  }
  else
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
    const size_t n = size_t(N) * (h / 2) * (w / 2) * (lowC / 8);
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
      const size_t n = size_t(N) * h * w * (encC[l - 1] / 8);  // 2x2 windows of level l - 1
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
  // the batch as one image of W x (N H) pixels, 9 interleaved channels (colour, albedo, normal)
  const FrameIn f{dn->in9, nullptr, dn->in9 + 3, nullptr, dn->in9 + 6, nullptr, W, N * H, 1, 9};
  kInput<<<blocksFor(npix), kBlock, 0, dn->stream>>>(f, W, N * H, 1.f, dn->x);
  forward(dn, N, H, W, nullptr);
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
  if (e == hipSuccess) e = hipEventCreateWithFlags(&dn->evIn, hipEventDisableTiming);
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
  if (dn->outDev) hipFree(dn->outDev);
  if (dn->evIn) (void)hipEventDestroy(dn->evIn);
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
  // Denoiser::Params makes the auxiliary images optional (image/denoiser.h:14-29); as with OIDN a
  // normal image needs an albedo image (image/oidn_denoiser.cpp).  Missing ones enter the network
  // as zero channels, the input the weights for that input set were trained on.
  if (p->normal && !p->albedo) return dfail(PRT_ERR_INVALID, "a normal image needs an albedo image");
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
  const int Wp = (W + 15) / 16 * 16, Hp = (H + 15) / 16 * 16;  // the network pads by edge replication
  int rc = ensureActs(dn, 1, Hp, Wp);
  if (rc != PRT_OK) return rc;
  const size_t n = size_t(W) * H;
  const float scale = std::isnan(p.hdrScale) || p.hdrScale <= 0.f ? 1.f : p.hdrScale;
  // the images as given (interleaved H x W x 3) into dn->in9: colour, then albedo, then normal
  DN_TRY(hipMemcpyAsync(dn->in9, p.color, sizeof(float) * n * 3, hipMemcpyHostToDevice, dn->stream));
  if (p.albedo) DN_TRY(hipMemcpyAsync(dn->in9 + 3 * n, p.albedo, sizeof(float) * n * 3, hipMemcpyHostToDevice, dn->stream));
  if (p.normal) DN_TRY(hipMemcpyAsync(dn->in9 + 6 * n, p.normal, sizeof(float) * n * 3, hipMemcpyHostToDevice, dn->stream));
  NetIO io{};
  io.fin = FrameIn{dn->in9, nullptr, p.albedo ? dn->in9 + 3 * n : nullptr, nullptr, p.normal ? dn->in9 + 6 * n : nullptr, nullptr,
                   W, H, 1, 3};
  io.hdrScale = scale;
  io.out = dn->outF;
  io.outW = W;
  io.outH = H;
  DN_TRY(hipEventRecord(dn->ev[0], dn->stream));
  forward(dn, 1, Hp, Wp, &io);
  DN_TRY(hipEventRecord(dn->ev[1], dn->stream));
  DN_TRY(hipMemcpyAsync(p.output, dn->outF, sizeof(float) * n * 3, hipMemcpyDeviceToHost, dn->stream));
  DN_TRY(hipStreamSynchronize(dn->stream));
  DN_TRY(hipGetLastError());
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return PRT_OK;
}

}  // extern "C"

namespace {

// Denoises planar images already on the device into dn->outDev (W x H x 3).
int executeImages(prt_denoiser* dn, const FrameImages& f, int32_t inputs, float hdrScale, double* seconds)
{
  if (f.device != dn->device) return dfail(PRT_ERR_INVALID, "the frame is on another device than the denoiser");
  if (!f.color || f.W < 1 || f.H < 1) return dfail(PRT_ERR_STATE, "the frame has no colour image");
  if (inputs & ~(PRT_DENOISE_ALBEDO | PRT_DENOISE_NORMAL)) return dfail(PRT_ERR_INVALID, "unknown input flags");
  if ((inputs & PRT_DENOISE_NORMAL) && !(inputs & PRT_DENOISE_ALBEDO))
    return dfail(PRT_ERR_INVALID, "a normal image needs an albedo image");
  if ((inputs & PRT_DENOISE_ALBEDO) && !f.albedo)
    return dfail(PRT_ERR_INVALID, "the frame has no first-hit albedo (enable PRT_AOV_ALBEDO1)");
  if ((inputs & PRT_DENOISE_NORMAL) && !f.normal)
    return dfail(PRT_ERR_INVALID, "the frame has no first-hit normal (enable PRT_AOV_NORMAL1)");
  DN_TRY(hipSetDevice(dn->device));
  const auto t0 = std::chrono::high_resolution_clock::now();
  const int W = f.W, H = f.H;
  const int Wp = (W + 15) / 16 * 16, Hp = (H + 15) / 16 * 16;
  int rc = ensureActs(dn, 1, Hp, Wp);
  if (rc != PRT_OK) return rc;
  const size_t nOut = size_t(W) * H * 3;
  if (dn->outDevCap < nOut)
  {
    if (dn->outDev) DN_TRY(hipFree(dn->outDev));
    dn->outDev = nullptr;
    dn->outDevCap = 0;
    DN_TRY(hipMalloc(&dn->outDev, sizeof(float) * nOut));
    dn->outDevCap = nOut;
  }
  NetIO io{};
  io.fin = FrameIn{f.color, f.colorW, (inputs & PRT_DENOISE_ALBEDO) ? f.albedo : nullptr,
                   (inputs & PRT_DENOISE_ALBEDO) ? f.albedoW : nullptr, (inputs & PRT_DENOISE_NORMAL) ? f.normal : nullptr,
                   (inputs & PRT_DENOISE_NORMAL) ? f.normalW : nullptr, W, H, W * H, 1};
  io.hdrScale = std::isnan(hdrScale) || hdrScale <= 0.f ? 1.f : hdrScale;
  io.out = dn->outDev;
  io.outW = W;
  io.outH = H;
  DN_TRY(hipEventRecord(dn->evIn, f.stream));
  DN_TRY(hipStreamWaitEvent(dn->stream, dn->evIn, 0));
  DN_TRY(hipEventRecord(dn->ev[0], dn->stream));
  forward(dn, 1, Hp, Wp, &io);
  DN_TRY(hipEventRecord(dn->ev[1], dn->stream));
  DN_TRY(hipStreamSynchronize(dn->stream));
  DN_TRY(hipGetLastError());
  dn->outW = W;
  dn->outH = H;
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return PRT_OK;
}

}  // namespace

extern "C" {

int prt_denoiser_execute_frame(prt_denoiser* dn, prt_context* ctx, int32_t inputs, float hdr_scale, double* seconds)
{
  if (!dn || !ctx) return dfail(PRT_ERR_INVALID, "null argument");
  FrameImages f;
  int rc = contextFrameImages(ctx, &f);
  if (rc != PRT_OK) return dfail(rc, prt_last_error());
  return executeImages(dn, f, inputs, hdr_scale, seconds);
}

int prt_denoiser_execute_batch(prt_denoiser* dn, prt_batch* batch, int32_t frame, int32_t buffer, int32_t inputs, float hdr_scale,
                               double* seconds)
{
  if (!dn || !batch) return dfail(PRT_ERR_INVALID, "null argument");
  FrameImages f;
  int rc = batchFrameImages(batch, frame, buffer, &f);
  if (rc != PRT_OK) return dfail(rc, prt_last_error());
  return executeImages(dn, f, inputs, hdr_scale, seconds);
}

int prt_denoiser_get_output(prt_denoiser* dn, float* dst, int32_t width, int32_t height)
{
  if (!dn || !dst) return dfail(PRT_ERR_INVALID, "null argument");
  if (!dn->outW) return dfail(PRT_ERR_STATE, "no frame has been denoised");
  if (width != dn->outW || height != dn->outH) return dfail(PRT_ERR_INVALID, "size differs from the denoised frame's");
  DN_TRY(hipSetDevice(dn->device));
  DN_TRY(hipMemcpy(dst, dn->outDev, sizeof(float) * size_t(width) * height * 3, hipMemcpyDeviceToHost));
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
