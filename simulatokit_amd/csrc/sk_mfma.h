// sk_mfma.h — split-BF16 MFMA line transforms (device), for the generic passes.
//
// A line of n = 16 * n2 points (n2 = 16, 24, 32 or 64) is transformed in two dense stages on the matrix cores
// (four-step): j = n2 j1 + j2, k = k1 + 16 k2,
//   Y[k1][j2] = sum_{j1} W16^{j1 k1} x[n2 j1 + j2]              (stage A: DFT-16 of the n2 strided sub-lines)
//   X[k1 + 16 k2] = sum_{j2} Wn2^{j2 k2} (Wn^{j2 k1} Y[k1][j2])  (twiddle, stage C: DFT-n2 of the 16 rows)
// Each stage is a real GEMM of the data (vectors x [re, im interleaved]) with the DFT's 2L x 2L real block matrix
// [[Wr, Wi], [-Wi, Wr]] (rows: input re/im, columns: output re/im). Data and matrix enter as bfloat16 pairs, x =
// hi + lo with hi = bf16(x), lo = bf16(x - hi); three products (lo*hi, hi*lo, hi*hi) accumulate in float32, dropping
// lo*lo (a relative 2^-16). v_mfma_f32_32x32x16_bf16 operand layout (lane l, element j): A[m = l % 32][k = 8 (l / 32)
// + j], B[k = 8 (l / 32) + j][n = l % 32], D register j = D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
// A block's vectors fill 32-row tiles; the rows past the block's last vector load zeros and are not stored. At
// n2 = 24 the DFT-24 matrix has 48 real columns, padded with zero columns to two 32-column tiles.
#pragma once

#include <hip/hip_runtime.h>

namespace sk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 h = (__bf16)x[j];
    hi[j] = h;
    lo[j] = (__bf16)(x[j] - (float)h);
  }
}

__device__ __forceinline__ f32x16 mfma3(const bf16x8& ahi, const bf16x8& alo, const bf16x8& bhi, const bf16x8& blo,
                                        f32x16 acc) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
  return acc;
}

// The 8 matrix values lane `lane` holds for k-step s and output tile t of the DFT of length L, sign index si
// (0: e^{-2 pi i}, 1: e^{+2 pi i}), part h (0 hi, 1 lo): tables laid out [si][s][t][h][lane].
__device__ __forceinline__ bf16x8 mfma_matrix(const bf16x8* tab, int L, int si, int s, int t, int h, int lane) {
  const int ksteps = L / 8, ntiles = (L + 15) / 16;
  return tab[((((size_t)si * ksteps + s) * ntiles + t) * 2 + h) * 64 + lane];
}

// Transforms `lines` lines (stride ls, in `a`) of length n = 16 n2 in place of `a`, using `b` as the intermediate;
// 256 threads, lines * n <= 2048. Returns `a`. The caller has synchronised `a`.
template <int N2>
__device__ float2* mfma_lines(float2* a, float2* b, int lines, int ls, const bf16x8* tab16, const bf16x8* tabn2,
                              const float2* tw, int sign, int tid) {
  static_assert(N2 % 8 == 0 && N2 <= 64, "n2 must be a multiple of 8, at most 64");
  const int lane = tid & 63, w = tid >> 6, half = lane >> 5, i = lane & 31;
  const int si = sign > 0 ? 1 : 0;
  // stage A: vectors (line, j2), DFT-16 over j1; K = 32 (2 k-steps), N = 32 (1 tile); lines * N2 <= 128 vectors,
  // one 32-row tile per wave
  {
    const int nva = lines * N2;
    const int v = 32 * w + i;
    const bool ok = v < nva;
    const int line = ok ? v / N2 : 0, j2 = v % N2;
    const float* src = reinterpret_cast<const float*>(a + (size_t)line * ls);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 16 * s + 8 * half + j;
        x[j] = ok ? src[2 * (N2 * (k >> 1) + j2) + (k & 1)] : 0.f;
      }
      bf16x8 ahi, alo;
      split_bf16(x, ahi, alo);
      acc = mfma3(ahi, alo, mfma_matrix(tab16, 16, si, s, 0, 0, lane), mfma_matrix(tab16, 16, si, s, 0, 1, lane), acc);
    }
    const int k1 = i >> 1, q = i & 1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int vv = 32 * w + 8 * (r >> 2) + 4 * half + (r & 3);
      if (vv < nva) {
        float* dst = reinterpret_cast<float*>(b + (size_t)(vv / N2) * ls);
        dst[2 * (k1 * N2 + vv % N2) + q] = acc[r];
      }
    }
  }
  __syncthreads();
  // stage C: vectors (line, k1), twiddled, DFT-N2 over j2; K = 2 N2 (N2 / 8 k-steps), N = 2 N2 (ceil(N2 / 16)
  // tiles, the last one zero-padded when N2 = 24); lines * 16 vectors
  constexpr int ksteps = N2 / 8, ntiles = (N2 + 15) / 16;
  const int nvc = lines * 16;
  const int mtiles = (nvc + 31) / 32;
  for (int pr = w; pr < mtiles * ntiles; pr += 4) {
    const int mt = pr / ntiles, t = pr % ntiles;
    const int v = 32 * mt + i;
    const bool ok = v < nvc;
    const int line = ok ? v / 16 : 0, k1 = v % 16;
    const float2* Y = b + (size_t)line * ls + k1 * N2;
    const float2* twr = tw + k1 * N2;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int s = 0; s < ksteps; ++s) {
      float x[8];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int j2 = 8 * s + 4 * half + jj;
        float2 c = twr[j2];
        if (sign > 0) c.y = -c.y;
        const float2 y = ok ? Y[j2] : make_float2(0.f, 0.f);
        x[2 * jj] = y.x * c.x - y.y * c.y;
        x[2 * jj + 1] = y.x * c.y + y.y * c.x;
      }
      bf16x8 ahi, alo;
      split_bf16(x, ahi, alo);
      acc = mfma3(ahi, alo, mfma_matrix(tabn2, N2, si, s, t, 0, lane), mfma_matrix(tabn2, N2, si, s, t, 1, lane), acc);
    }
    const int nn = 32 * t + i, k2 = nn >> 1, q = nn & 1;
    if (k2 < N2) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int vv = 32 * mt + 8 * (r >> 2) + 4 * half + (r & 3);
        if (vv < nvc) {
          float* dst = reinterpret_cast<float*>(a + (size_t)(vv / 16) * ls);
          dst[2 * ((vv % 16) + 16 * k2) + q] = acc[r];
        }
      }
    }
  }
  __syncthreads();
  return a;
}

}  // namespace sk
