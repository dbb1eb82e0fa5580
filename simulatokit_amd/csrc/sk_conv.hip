// sk_conv.hip — the layers of the HAADF surrogate (build-only; the reference has no learned part): 5 x 5
// convolutions with circular padding (the cells are periodic), bfloat16 weights and float32 activations and
// accumulation, their two backward passes, the fixed separable resampling of the output to the scan grid, the squared
// error's gradient and an Adam step that keeps float32 master weights and publishes their bfloat16 rounding.
// Every reduction runs in a fixed order, so a training run is reproducible bit for bit.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sk_internal.h"
#include "sk_mfma.h"
#include "../../include/simulato_surrogate.h"

namespace sk {
namespace {

constexpr int kTile = 16;  // output tile edge
constexpr int kCo = 8;     // output channels per block
constexpr int kCi = 8;     // input channels per LDS chunk
constexpr int kHalo = kTile + 4;

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// y[b][co][p] = bias[co] + sum_{ci, ky, kx} w[co][ci][ky][kx] x[b][ci][p + (ky - 2, kx - 2) mod (h, w)], then ReLU if
// asked. Block: a 16 x 16 tile of one image and 8 output channels; the input tile with its 2-pixel halo and the
// weights (as [ci][k][co], so one read serves eight output channels) go through LDS in chunks of 8 input channels.
__global__ void __launch_bounds__(256) conv5_forward_kernel(const float* __restrict__ x, const uint16_t* __restrict__ wt,
                                                            const float* __restrict__ bias, int cin, int cout, int h,
                                                            int w, int relu, float* __restrict__ y) {
  __shared__ float tile[kCi][kHalo][kHalo + 1];
  __shared__ float4 wsh[kCi][25][kCo / 4];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int tiles_x = w / kTile;
  const int x0 = (blockIdx.x % tiles_x) * kTile, y0 = (blockIdx.x / tiles_x) * kTile;
  const int co0 = blockIdx.y * kCo;
  const long long b = blockIdx.z;
  float acc[kCo];
#pragma unroll
  for (int c = 0; c < kCo; ++c) acc[c] = (bias && co0 + c < cout) ? bias[co0 + c] : 0.f;
  for (int ci0 = 0; ci0 < cin; ci0 += kCi) {
    for (int idx = threadIdx.x; idx < kCi * kHalo * kHalo; idx += 256) {
      const int c = idx / (kHalo * kHalo), r = idx % (kHalo * kHalo), yy = r / kHalo, xx = r % kHalo;
      float v = 0.f;
      if (ci0 + c < cin)
        v = x[((b * cin + ci0 + c) * h + wrap(y0 + yy - 2, h)) * (long long)w + wrap(x0 + xx - 2, w)];
      tile[c][yy][xx] = v;
    }
    for (int idx = threadIdx.x; idx < kCi * 25 * kCo; idx += 256) {
      const int c = idx / (25 * kCo), r = idx % (25 * kCo), k = r / kCo, o = r % kCo;
      float v = 0.f;
      if (ci0 + c < cin && co0 + o < cout) v = bf16_to_f32(wt[((size_t)(co0 + o) * cin + ci0 + c) * 25 + k]);
      reinterpret_cast<float*>(&wsh[c][k][0])[o] = v;
    }
    __syncthreads();
    for (int c = 0; c < kCi; ++c)
#pragma unroll
      for (int ky = 0; ky < 5; ++ky)
#pragma unroll
        for (int kx = 0; kx < 5; ++kx) {
          const float v = tile[c][ty + ky][tx + kx];
          const float4 wa = wsh[c][ky * 5 + kx][0], wb = wsh[c][ky * 5 + kx][1];
          acc[0] += wa.x * v;
          acc[1] += wa.y * v;
          acc[2] += wa.z * v;
          acc[3] += wa.w * v;
          acc[4] += wb.x * v;
          acc[5] += wb.y * v;
          acc[6] += wb.z * v;
          acc[7] += wb.w * v;
        }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < kCo; ++c) {
    if (co0 + c >= cout) break;
    float v = acc[c];
    if (relu) v = v > 0.f ? v : 0.f;
    y[((b * cout + co0 + c) * h + y0 + ty) * (long long)w + x0 + tx] = v;
  }
}

// Matrix-core forward convolution for up to 32 input and 32 output channels (every layer of the surrogate, and the
// backward pass over the data of each): an implicit GEMM, y[co][p] = sum over k = (tap, ci) of W[co][k] x[ci][p + tap],
// on v_mfma_f32_32x32x16_bf16 (A = the weights, 32 output channels x 16 k; B = 16 k x 32 pixels of one output row).
// The weights are bfloat16 already, so A is exact; each activation enters as a bfloat16 pair x = hi + lo (lo =
// bf16(x - hi), leaving a relative 2^-18), two products accumulated in float32. Block: 4 wavefronts, each two output
// rows of 32 pixels (two 32 x 32 tiles); the input rows with their circular halo (12 rows x 36 columns x CINP
// channels, the channels past cin zero) sit in LDS as the hi and lo halves, [row][column][channel] with the channels
// padded to CINP + 8 (a 16-byte read per lane; the padding spreads a row's 32 pixels over all 64 banks). The weights
// come as wtm[tap][32][CINP] (sk_conv5_weights_tap_major: the rows past cout and the columns past cin zero), 8
// consecutive channels per lane and k-step; only the cout real output channels are stored.
constexpr int kMfWaves = 4, kMfRowsPerWave = 2, kMfRows = kMfWaves * kMfRowsPerWave, kMfCols = 32;
constexpr int kMfHR = kMfRows + 4, kMfHC = kMfCols + 4;

template <int CINP>
__global__ void __launch_bounds__(64 * kMfWaves) conv5_forward_mfma_kernel(const float* __restrict__ x,
                                                                           const __bf16* __restrict__ wtm,
                                                                           const float* __restrict__ bias, int cin,
                                                                           int cout, int h, int w, int relu,
                                                                           float* __restrict__ y) {
  constexpr int CP = CINP + 8;
  __shared__ __align__(16) __bf16 xhi[kMfHR * kMfHC * CP];
  __shared__ __align__(16) __bf16 xlo[kMfHR * kMfHC * CP];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tiles_x = w / kMfCols;
  const int x0 = (blockIdx.x % tiles_x) * kMfCols, y0 = (blockIdx.x / tiles_x) * kMfRows;
  const long long b = blockIdx.y;
  // the input rows y0 - 2 .. y0 + kMfRows + 1, columns x0 - 2 .. x0 + 33 (circular), every channel: a task is 8
  // channels of one pixel (consecutive threads take consecutive columns), written to LDS as one 16-byte hi and lo
  for (int e = tid; e < (CINP / 8) * kMfHR * kMfHC; e += 64 * kMfWaves) {
    const int cc = e % kMfHC, r = e / kMfHC, rr = r % kMfHR, c8 = 8 * (r / kMfHR);
    const long long pix = (long long)wrap(y0 + rr - 2, h) * w + wrap(x0 + cc - 2, w);
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = c8 + j < cin ? x[(b * cin + c8 + j) * (long long)h * w + pix] : 0.f;
    bf16x8 hi, lo;
    split_bf16(v, hi, lo);
    *reinterpret_cast<bf16x8*>(xhi + (rr * kMfHC + cc) * CP + c8) = hi;
    *reinterpret_cast<bf16x8*>(xlo + (rr * kMfHC + cc) * CP + c8) = lo;
  }
  __syncthreads();
  const int n = lane & 31, half = lane >> 5;
  f32x16 acc[kMfRowsPerWave][2];
#pragma unroll
  for (int t = 0; t < kMfRowsPerWave; ++t)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[t][0][q] = acc[t][1][q] = 0.f;
  for (int tap = 0; tap < 25; ++tap) {
    const int ky = tap / 5, kx = tap % 5;
#pragma unroll
    for (int c16 = 0; c16 < CINP; c16 += 16) {
      const int ci = c16 + 8 * half;
      // A[m = output channel n][k = 8 half + j]: channels ci .. ci + 7 of tap
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(wtm + ((size_t)tap * 32 + n) * CINP + ci);
#pragma unroll
      for (int t = 0; t < kMfRowsPerWave; ++t) {
        const int rr = wv * kMfRowsPerWave + t + ky;  // halo row of output row y0 + wv * 2 + t, tap row ky
        const int slot = (rr * kMfHC + n + kx) * CP + ci;
        const bf16x8 bh = *reinterpret_cast<const bf16x8*>(xhi + slot);
        const bf16x8 bl = *reinterpret_cast<const bf16x8*>(xlo + slot);
        acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bl, acc[t][1], 0, 0, 0);
        acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bh, acc[t][0], 0, 0, 0);
      }
    }
  }
  // D register q of lane l: D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32]: output channel, pixel n
#pragma unroll
  for (int t = 0; t < kMfRowsPerWave; ++t) {
    const int oy = y0 + wv * kMfRowsPerWave + t;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int co = 8 * (q / 4) + 4 * half + (q % 4);
      if (co >= cout) continue;
      float v = acc[t][1][q] + acc[t][0][q] + (bias ? bias[co] : 0.f);
      if (relu) v = v > 0.f ? v : 0.f;
      y[((b * cout + co) * h + oy) * (long long)w + x0 + n] = v;
    }
  }
}

// The tap-major bfloat16 layout the matrix-core kernel reads, wtm[tap][32][kp] with kp = 16 for up to 16 input
// channels of the convolution it serves and 32 above: wtm[tap][co][ci] = w[co][ci][tap] (flip = 0; rows = cout,
// columns = cin), or the flipped, transposed kernel of the backward pass over the data, wtm[tap][ci][co] =
// w[co][ci][24 - tap] (flip = 1; rows = cin, columns = cout); the rows and columns past those zero.
__global__ void weights_tap_major_kernel(const uint16_t* __restrict__ wt, int cin, int cout, int flip, int kp,
                                         uint16_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 25 * 32 * kp) return;
  const int col = i % kp, row = (i / kp) % 32, tap = i / (32 * kp);
  const int co = flip ? col : row, ci = flip ? row : col, src_tap = flip ? 24 - tap : tap;
  out[i] = (co < cout && ci < cin) ? wt[((size_t)co * cin + ci) * 25 + src_tap] : uint16_t(0);
}

// Matrix-core gradient of the weights for up to 32 input and 32 output channels: dw[co][ci][tap] = sum over the
// pixels p of every image of dy[co][p] x[ci][p + tap], a GEMM of M = 32 output channels by N = 32 input channels per
// tap with the pixels as K, on v_mfma_f32_32x32x16_bf16 with both operands as bfloat16 hi + lo pairs and three
// products (lo hi, hi lo, hi hi; the dropped lo lo is a relative 2^-18). Work unit: a strip of 4 rows x 32 columns
// of one image; the strips go round-robin to a fixed number of blocks (kWgBlocks), each of which keeps its sums in
// registers and writes them to its own slice of `partial`; a second kernel adds the slices in block order, so the
// result does not depend on the device or on timing. Block: 5 wavefronts, wavefront ky the five taps (ky, 0..4).
// LDS: x rows y0 - 2 .. y0 + 5, columns x0 - 8 .. x0 + 39 (circular; groups of 8 columns keep every 16-byte read
// aligned), [row][channel][column] with a 56-column stride, and dy rows y0 .. y0 + 3, columns x0 .. x0 + 31 with a
// 40-column stride (112- and 80-byte strides put 16 lanes' 16-byte reads on distinct banks); both as hi and lo.
// For output row r and k-step t (16 pixels from column 16 t), lane half h reads the 24 staged x columns 16 t + 8 h ..
// + 23 of its channel once; tap kx takes elements 6 + kx .. 13 + kx of them (odd kx by a 16-bit funnel shift).
constexpr int kWgWaves = 5, kWgRows = 4, kWgCols = 32, kWgXRows = kWgRows + 4;
constexpr int kWgXStride = 56, kWgDStride = 40, kWgXGroups = (kWgCols + 16) / 8, kWgDGroups = kWgCols / 8;
constexpr int kWgBlocks = 512, kWgOut = 25 * 32 * 32;
static_assert(kWgBlocks * kWgOut + 2 * 32 * 64 == SK_CONV5_MFMA_WORKSPACE_FLOATS, "workspace size");

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split8(const float* __restrict__ src, bool valid, u32x4& hi, u32x4& lo) {
  float v[8];
  if (valid) {
    const float4 a = reinterpret_cast<const float4*>(src)[0], c = reinterpret_cast<const float4*>(src)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = c.x; v[5] = c.y; v[6] = c.z; v[7] = c.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = 0.f;
  }
  bf16x8 h, l;
  split_bf16(v, h, l);
  hi = __builtin_bit_cast(u32x4, h);
  lo = __builtin_bit_cast(u32x4, l);
}

// elements o .. o + 7 of the 24-element window held as 12 packed pairs
template <int O>
__device__ __forceinline__ bf16x8 window8(const uint32_t (&wv)[12]) {
  u32x4 r;
  if constexpr (O % 2 == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = wv[O / 2 + i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (wv[O / 2 + i] >> 16) | (wv[O / 2 + i + 1] << 16);
  }
  return __builtin_bit_cast(bf16x8, r);
}

__global__ void __launch_bounds__(64 * kWgWaves, 2) conv5_wgrad_mfma_kernel(const float* __restrict__ x,
                                                                            const float* __restrict__ dy, int cin,
                                                                            int cout, int h, int w, long long strips,
                                                                            float* __restrict__ partial) {
  __shared__ __align__(16) uint16_t xs[2][kWgXRows * 32 * kWgXStride];
  __shared__ __align__(16) uint16_t ds[2][kWgRows * 32 * kWgDStride];
  const int tid = threadIdx.x, lane = tid & 63, ky = tid >> 6;
  const int n = lane & 31, half = lane >> 5;
  f32x16 acc[5];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[k][q] = 0.f;
  const int sx = w / kWgCols, per_image = sx * (h / kWgRows);
  for (long long s = blockIdx.x; s < strips; s += gridDim.x) {
    const long long b = s / per_image;
    const int r0 = int(s % per_image), y0 = (r0 / sx) * kWgRows, x0 = (r0 % sx) * kWgCols;
    __syncthreads();  // the previous strip's reads are done
    for (int e = tid; e < kWgXRows * 32 * kWgXGroups; e += 64 * kWgWaves) {
      const int g = e % kWgXGroups, rc = e / kWgXGroups, ci = rc % 32, rr = rc / 32;
      const float* src =
          x + ((b * cin + (ci < cin ? ci : 0)) * h + wrap(y0 + rr - 2, h)) * (long long)w + wrap(x0 - 8 + 8 * g, w);
      u32x4 hi, lo;
      split8(src, ci < cin, hi, lo);
      const int slot = (rr * 32 + ci) * kWgXStride + 8 * g;
      *reinterpret_cast<u32x4*>(xs[0] + slot) = hi;
      *reinterpret_cast<u32x4*>(xs[1] + slot) = lo;
    }
    for (int e = tid; e < kWgRows * 32 * kWgDGroups; e += 64 * kWgWaves) {
      const int g = e % kWgDGroups, rc = e / kWgDGroups, co = rc % 32, rr = rc / 32;
      const float* src = dy + ((b * cout + (co < cout ? co : 0)) * h + y0 + rr) * (long long)w + x0 + 8 * g;
      u32x4 hi, lo;
      split8(src, co < cout, hi, lo);
      const int slot = (rr * 32 + co) * kWgDStride + 8 * g;
      *reinterpret_cast<u32x4*>(ds[0] + slot) = hi;
      *reinterpret_cast<u32x4*>(ds[1] + slot) = lo;
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < kWgRows; ++r) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int da = (r * 32 + n) * kWgDStride + 16 * t + 8 * half;
        const bf16x8 ah = *reinterpret_cast<const bf16x8*>(ds[0] + da);
        const bf16x8 al = *reinterpret_cast<const bf16x8*>(ds[1] + da);
        const int xa = ((r + ky) * 32 + n) * kWgXStride + 16 * t + 8 * half;
        uint32_t wh[12], wl[12];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const u32x4 vh = *reinterpret_cast<const u32x4*>(xs[0] + xa + 8 * i);
          const u32x4 vl = *reinterpret_cast<const u32x4*>(xs[1] + xa + 8 * i);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            wh[4 * i + j] = vh[j];
            wl[4 * i + j] = vl[j];
          }
        }
        acc[0] = mfma3(ah, al, window8<6>(wh), window8<6>(wl), acc[0]);
        acc[1] = mfma3(ah, al, window8<7>(wh), window8<7>(wl), acc[1]);
        acc[2] = mfma3(ah, al, window8<8>(wh), window8<8>(wl), acc[2]);
        acc[3] = mfma3(ah, al, window8<9>(wh), window8<9>(wl), acc[3]);
        acc[4] = mfma3(ah, al, window8<10>(wh), window8<10>(wl), acc[4]);
      }
    }
  }
  // D register q of lane l: D[co = 8 (q / 4) + 4 (l / 32) + q % 4][ci = l % 32]
  float* out = partial + (size_t)blockIdx.x * kWgOut;
#pragma unroll
  for (int kx = 0; kx < 5; ++kx)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int co = 8 * (q / 4) + 4 * half + (q % 4);
      out[((ky * 5 + kx) * 32 + co) * 32 + n] = acc[kx][q];
    }
}

// dw[co][ci][tap] = the sum of the blocks' slices in a fixed order, for the real channels: a block takes 64
// consecutive outputs, its four wavefronts a quarter of the slices each (in block order, float64), and the quarters
// are added in order
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ partial, int blocks, int cin,
                                                           int cout, float* __restrict__ dw) {
  __shared__ double part[4][64];
  const int o = threadIdx.x & 63, quarter = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + o;  // [tap][co][ci]
  const int per = (blocks + 3) / 4, g0 = quarter * per, g1 = g0 + per < blocks ? g0 + per : blocks;
  double s = 0.0;
#pragma unroll 8
  for (int g = g0; g < g1; ++g) s += partial[(size_t)g * kWgOut + i];
  part[quarter][o] = s;
  __syncthreads();
  if (quarter) return;
  const int ci = i % 32, co = (i / 32) % 32, tap = i / 1024;
  if (ci < cin && co < cout)
    dw[((size_t)co * cin + ci) * 25 + tap] = float(((part[0][o] + part[1][o]) + part[2][o]) + part[3][o]);
}

// db[co] = sum over images and pixels of dy[b][co][p], in a fixed order: block (co, chunk) sums one of
// kBgChunks contiguous ranges of the batch's pixels into partial[co][chunk] (float64), then bias_finish_kernel adds
// the chunks in order
constexpr int kBgChunks = 64;
__global__ void __launch_bounds__(256) bias_partial_kernel(const float* __restrict__ dy, int batch, int cout,
                                                           long long hw, double* __restrict__ partial) {
  __shared__ double red[256];
  const int co = blockIdx.x, chunk = blockIdx.y;
  const long long total = (long long)batch * hw, len = (total + kBgChunks - 1) / kBgChunks;
  const long long i0 = chunk * len, i1 = i0 + len < total ? i0 + len : total;
  double s = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) s += dy[((i / hw) * cout + co) * hw + i % hw];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[co * kBgChunks + chunk] = red[0];
}

__global__ void bias_finish_kernel(const double* __restrict__ partial, int cout, float* __restrict__ db) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  double s = 0.0;
  for (int c = 0; c < kBgChunks; ++c) s += partial[co * kBgChunks + c];
  db[co] = float(s);
}

// w'[ci][co][ky][kx] = w[co][ci][4 - ky][4 - kx]: the backward pass over the data is the forward convolution of dy
// with the flipped, transposed kernel.
__global__ void flip_transpose_kernel(const uint16_t* __restrict__ wt, int cin, int cout, uint16_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cin * cout * 25) return;
  const int k = i % 25, r = i / 25, co = r % cout, ci = r / cout;
  out[i] = wt[((size_t)co * cin + ci) * 25 + (24 - k)];
}

// dw[co][ci][ky][kx] = sum_{b, p} dy[b][co][p] x[b][ci][p + (ky - 2, kx - 2)]: one block per (co, ci), a fixed-order
// reduction (each thread's strided partials, then a tree over the block); db[co] the same over dy alone.
__global__ void __launch_bounds__(256) conv5_backward_weights_kernel(const float* __restrict__ x,
                                                                     const float* __restrict__ dy, int batch, int cin,
                                                                     int cout, int h, int w, float* __restrict__ dw,
                                                                     float* __restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x / cin, ci = blockIdx.x % cin;
  float acc[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) acc[k] = 0.f;
  float accb = 0.f;
  const long long hw = (long long)h * w;
  for (long long idx = threadIdx.x; idx < batch * hw; idx += 256) {
    const long long b = idx / hw, p = idx % hw;
    const int py = int(p / w), px = int(p % w);
    const float g = dy[(b * cout + co) * hw + p];
    accb += g;
    const float* xs = x + (b * cin + ci) * hw;
#pragma unroll
    for (int ky = 0; ky < 5; ++ky) {
      const float* row = xs + (long long)wrap(py + ky - 2, h) * w;
#pragma unroll
      for (int kx = 0; kx < 5; ++kx) acc[ky * 5 + kx] += g * row[wrap(px + kx - 2, w)];
    }
  }
  for (int k = 0; k <= 25; ++k) {
    red[threadIdx.x] = k < 25 ? acc[k] : accb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (k < 25) dw[(size_t)blockIdx.x * 25 + k] = red[0];
      else if (ci == 0 && db) db[co] = red[0];
    }
    __syncthreads();
  }
}

// dy *= (y > 0): the ReLU's backward pass, y the layer's output after the ReLU
__global__ void relu_mask_kernel(float* dy, const float* y, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n && !(y[i] > 0.f)) dy[i] = 0.f;
}

// out[b][u][v] = sum_{i, j} A[u][i] A[v][j] in[b][i][j], A m x n, in n x n, out m x m (two passes through tmp)
__global__ void sep_rows_kernel(const float* in, const float* A, int m, int n, int batch, float* tmp) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;  // tmp[b][i][v], i < n, v < m
  if (i >= (long long)batch * n * m) return;
  const int v = int(i % m);
  const long long r = i / m;  // b * n + row
  const float* src = in + r * n;
  float s = 0.f;
  for (int j = 0; j < n; ++j) s += A[(size_t)v * n + j] * src[j];
  tmp[i] = s;
}
__global__ void sep_cols_kernel(const float* tmp, const float* A, int m, int n, int batch, float* out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;  // out[b][u][v]
  if (i >= (long long)batch * m * m) return;
  const int v = int(i % m), u = int((i / m) % m);
  const long long b = i / ((long long)m * m);
  const float* src = tmp + b * n * m + v;
  float s = 0.f;
  for (int r = 0; r < n; ++r) s += A[(size_t)u * n + r] * src[(long long)r * m];
  out[i] = s;
}

// g = 2 (pred - label) / count; loss partials per block (fixed-order sums)
// g = 2 w (pred - label) / n and the block's sum of w (pred - label)^2, w = weights[i / per] (1 when weights is null)
__global__ void mse_grad_kernel(const float* pred, const float* label, long long n, const float* weights,
                                long long per, float* g, double* partial) {
  __shared__ double red[256];
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double e = 0;
  if (i < n) {
    const float d = pred[i] - label[i];
    const float wt = weights ? weights[i / per] : 1.f;
    g[i] = 2.f * wt * d / float(n);
    e = (double)wt * d * d;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, uint16_t* p_bf16, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (p_bf16) p_bf16[i] = f32_to_bf16(pi);
}

__global__ void to_bf16_kernel(const float* in, uint16_t* out, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = f32_to_bf16(in[i]);
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

hipStream_t pick(void* stream) { return stream ? static_cast<hipStream_t>(stream) : nullptr; }

void check_conv(int batch, int cin, int cout, int h, int w) {
  if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, std::string("conv5: ") + sk_hip_last_error());
  if (batch < 1 || cin < 1 || cout < 1) fail(SK_ERR_INVALID, "conv5: empty shape");
  if (h % kTile || w % kTile || h < kTile || w < kTile) fail(SK_ERR_INVALID, "conv5: h and w must be multiples of 16");
  if (batch > 65535) fail(SK_ERR_INVALID, "conv5: batch above 65535");
}

}  // namespace
}  // namespace sk

extern "C" {

int sk_conv5_forward(const float* x, int batch, int cin, int h, int w, const uint16_t* wt, const float* bias, int cout,
                     int relu, float* y, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!x || !wt || !y) fail(SK_ERR_INVALID, "conv5_forward: null pointer");
    check_conv(batch, cin, cout, h, w);
    dim3 grid((h / kTile) * (w / kTile), (cout + kCo - 1) / kCo, batch);
    hipLaunchKernelGGL(conv5_forward_kernel, grid, dim3(256), 0, pick(stream), x, wt, bias, cin, cout, h, w, relu, y);
    SK_HIP(hipGetLastError());
  });
}

int sk_conv5_weights_tap_major(const uint16_t* wt, int cin, int cout, int flip, uint16_t* out, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!wt || !out || cin < 1 || cout < 1 || cin > 32 || cout > 32)
      fail(SK_ERR_INVALID, "conv5_weights_tap_major: null pointer, or channels outside 1 .. 32");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    const int kp = (flip ? cout : cin) <= 16 ? 16 : 32;
    hipLaunchKernelGGL(weights_tap_major_kernel, dim3(blocks_for(25LL * 32 * kp)), dim3(256), 0, pick(stream), wt, cin,
                       cout, flip ? 1 : 0, kp, out);
    SK_HIP(hipGetLastError());
  });
}

int sk_conv5_forward_tap_major(const float* x, int batch, int cin, int h, int w, const uint16_t* wtm,
                               const float* bias, int cout, int relu, float* y, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!x || !wtm || !y) fail(SK_ERR_INVALID, "conv5_forward_tap_major: null pointer");
    check_conv(batch, cin, cout, h, w);
    if (cin > 32 || cout > 32 || w % kMfCols || h % kMfRows)
      fail(SK_ERR_INVALID, "conv5_forward_tap_major: at most 32 input and 32 output channels, w a multiple of 32 "
                           "and h of 8");
    dim3 grid((h / kMfRows) * (w / kMfCols), batch);
    const __bf16* wm = reinterpret_cast<const __bf16*>(wtm);
    if (cin <= 16)
      hipLaunchKernelGGL(conv5_forward_mfma_kernel<16>, grid, dim3(64 * kMfWaves), 0, pick(stream), x, wm, bias, cin,
                         cout, h, w, relu, y);
    else
      hipLaunchKernelGGL(conv5_forward_mfma_kernel<32>, grid, dim3(64 * kMfWaves), 0, pick(stream), x, wm, bias, cin,
                         cout, h, w, relu, y);
    SK_HIP(hipGetLastError());
  });
}

int sk_conv5_backward_weights_mfma(const float* x, const float* dy, int batch, int cin, int h, int w, int cout,
                                   float* dw, float* db, float* workspace, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!x || !dy || !dw || !workspace) fail(SK_ERR_INVALID, "conv5_backward_weights_mfma: null pointer");
    check_conv(batch, cin, cout, h, w);
    if (cin > 32 || cout > 32 || w % kWgCols || h % kWgRows)
      fail(SK_ERR_INVALID, "conv5_backward_weights_mfma: at most 32 input and 32 output channels, w a multiple of 32");
    hipStream_t s = pick(stream);
    const long long strips = (long long)batch * (h / kWgRows) * (w / kWgCols);
    const int blocks = (int)(strips < kWgBlocks ? strips : kWgBlocks);
    hipLaunchKernelGGL(conv5_wgrad_mfma_kernel, dim3(blocks), dim3(64 * kWgWaves), 0, s, x, dy, cin, cout, h, w,
                       strips, workspace);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(kWgOut / 64), dim3(256), 0, s, workspace, blocks, cin, cout, dw);
    if (db) {
      double* bias_partial = reinterpret_cast<double*>(workspace + (size_t)kWgBlocks * kWgOut);
      hipLaunchKernelGGL(bias_partial_kernel, dim3(cout, kBgChunks), dim3(256), 0, s, dy, batch, cout,
                         (long long)h * w, bias_partial);
      hipLaunchKernelGGL(bias_finish_kernel, dim3(1), dim3(32), 0, s, bias_partial, cout, db);
    }
    SK_HIP(hipGetLastError());
  });
}

int sk_conv5_backward_data(const float* dy, int batch, int cin, int h, int w, const uint16_t* wt, int cout, float* dx,
                           uint16_t* wt_scratch, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!dy || !wt || !dx || !wt_scratch) fail(SK_ERR_INVALID, "conv5_backward_data: null pointer");
    check_conv(batch, cin, cout, h, w);
    hipStream_t s = pick(stream);
    hipLaunchKernelGGL(flip_transpose_kernel, dim3(blocks_for((long long)cin * cout * 25)), dim3(256), 0, s, wt, cin,
                       cout, wt_scratch);
    dim3 grid((h / kTile) * (w / kTile), (cin + kCo - 1) / kCo, batch);
    hipLaunchKernelGGL(conv5_forward_kernel, grid, dim3(256), 0, s, dy, wt_scratch, nullptr, cout, cin, h, w, 0, dx);
    SK_HIP(hipGetLastError());
  });
}

int sk_conv5_backward_weights(const float* x, const float* dy, int batch, int cin, int h, int w, int cout, float* dw,
                              float* db, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!x || !dy || !dw) fail(SK_ERR_INVALID, "conv5_backward_weights: null pointer");
    check_conv(batch, cin, cout, h, w);
    hipLaunchKernelGGL(conv5_backward_weights_kernel, dim3(cin * cout), dim3(256), 0, pick(stream), x, dy, batch, cin,
                       cout, h, w, dw, db);
    SK_HIP(hipGetLastError());
  });
}

int sk_relu_backward(float* dy, const float* y, long long n, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!dy || !y || n < 0) fail(SK_ERR_INVALID, "relu_backward: invalid arguments");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    if (n == 0) return;
    hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(n)), dim3(256), 0, pick(stream), dy, y, n);
    SK_HIP(hipGetLastError());
  });
}

int sk_separable_map(const float* in, int batch, int n, const float* A, int m, float* tmp, float* out, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!in || !A || !tmp || !out || batch < 1 || n < 1 || m < 1) fail(SK_ERR_INVALID, "separable_map: invalid");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    hipStream_t s = pick(stream);
    hipLaunchKernelGGL(sep_rows_kernel, dim3(blocks_for((long long)batch * n * m)), dim3(256), 0, s, in, A, m, n,
                       batch, tmp);
    hipLaunchKernelGGL(sep_cols_kernel, dim3(blocks_for((long long)batch * m * m)), dim3(256), 0, s, tmp, A, m, n,
                       batch, out);
    SK_HIP(hipGetLastError());
  });
}

int sk_mse_grad(const float* pred, const float* label, long long n, float* grad, double* partials, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!pred || !label || !grad || !partials || n < 1) fail(SK_ERR_INVALID, "mse_grad: invalid");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(n)), dim3(256), 0, pick(stream), pred, label, n,
                       (const float*)nullptr, 1LL, grad, partials);
    SK_HIP(hipGetLastError());
  });
}

int sk_weighted_mse_grad(const float* pred, const float* label, const float* weights, long long per, long long n,
                         float* grad, double* partials, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!pred || !label || !weights || !grad || !partials || n < 1 || per < 1 || n % per)
      fail(SK_ERR_INVALID, "weighted_mse_grad: invalid");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(n)), dim3(256), 0, pick(stream), pred, label, n, weights,
                       per, grad, partials);
    SK_HIP(hipGetLastError());
  });
}

int sk_adam_step(float* p, const float* g, float* m, float* v, uint16_t* p_bf16, long long n, float lr, float b1,
                 float b2, float eps, int step, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!p || !g || !m || !v || n < 1 || step < 1) fail(SK_ERR_INVALID, "adam_step: invalid");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, pick(stream), p, g, m, v, p_bf16, n, lr, b1,
                       b2, eps, c1, c2);
    SK_HIP(hipGetLastError());
  });
}

int sk_to_bf16(const float* in, uint16_t* out, long long n, void* stream) {
  using namespace sk;
  return guarded_call([&] {
    if (!in || !out || n < 1) fail(SK_ERR_INVALID, "to_bf16: invalid");
    if (sk_hip_available() != 1) fail(SK_ERR_UNAVAILABLE, sk_hip_last_error());
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, pick(stream), in, out, n);
    SK_HIP(hipGetLastError());
  });
}

}  // extern "C"
