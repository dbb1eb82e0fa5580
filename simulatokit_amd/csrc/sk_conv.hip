// sk_conv.hip — the layers of the HAADF surrogate (build-only; the reference has no learned part): 5 x 5
// convolutions with circular padding (the cells are periodic), bfloat16 weights and float32 activations and
// accumulation, their two backward passes, the fixed separable resampling of the output to the scan grid, the squared
// error's gradient and an Adam step that keeps float32 master weights and publishes their bfloat16 rounding.
// Every reduction runs in a fixed order, so a training run is reproducible bit for bit.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sk_internal.h"
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
