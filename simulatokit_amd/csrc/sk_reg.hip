// sk_reg.hip — register-resident row and column passes of the slice loop for lines of 384 and 1024 points, and for
// the grid lengths 257..512 that Bluestein's algorithm transforms on 1024 points (e.g. 509).
//
// A line of M points is held by T = M / V lanes, V values per lane: lane i holds elements i + T r, r < V (384: V = 24,
// T = 16, four lines per wavefront; 1024: V = 16, T = 64, one line per wavefront). The transform is a three-stage
// Stockham FFT (384 = 8 * 8 * 6, 1024 = 16 * 8 * 8): the first stage works from the registers, the last ends in the
// registers in natural order, and the two exchanges between them go through the line's own LDS, which only its
// wavefront touches (no block barrier). The line's LDS slots are XOR-swizzled, e -> e ^ (((e >> 3) ^ (e >> 7)) & 31),
// which a bank model of every exchange (64 banks of 4 bytes, groups of 32 lanes) finds free of conflicts at both
// lengths (with a line stride of 400 slots at 384). Stage twiddles come from per-block LDS tables of the line's
// float64-derived roots.
//
// Bluestein (1970) for a logical length n <= 512 on M = 1024: X_k = c_k sum_j (x_j c_j) conj(c_{k - j}), c_k =
// e^{-i pi k^2 / n} (conjugated for the inverse), the convolution done in the registers as the forward 1024-point
// transform of x c (zero past n), the product with the kernel's transform, the inverse 1024-point transform, and the
// closing chirp product.
//
//   row pass:    blocks of 2 wavefronts; each line is one row x of one position, IFFT_y -> * t(x, y) -> FFT_y (and
//                the other row modes of the fast passes), the block's rows for PB positions in turn (RegTuning: t in
//                registers or read per position)
//   column pass: FFT_x -> * P -> IFFT_x (middle slices), the probe-generating first pass and the detecting last one;
//                a block moves TC compact columns of one position between HBM and LDS in row segments of TC values,
//                each column's lanes transform it inside the block's LDS, P in registers for the block's PB positions.
//                At 384 and 1024 points the block phases also run the first stage of the first transform (from the
//                values they load) and the last stage of the inverse (into the values they store).

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <utility>

#include "sk_dft.h"
#include "sk_internal.h"

namespace sk {
namespace {

template <int M>
struct RegPlan;
template <>
struct RegPlan<1024> {
  static constexpr int V = 16, R1 = 16, R2 = 8, R3 = 8;
  static constexpr int LS = 1024;  // line stride (slots) in the row pass
};
template <>
struct RegPlan<384> {
  static constexpr int V = 24, R1 = 8, R2 = 8, R3 = 6;
  static constexpr int LS = 400;   // 384 + 16: the four lines of a wavefront on different bank halves
};

// Register use per length, measured on an MI355X (HAADF geometry, mean row / column milliseconds per launch of 4096
// positions, 1024 at 1024^2, over every combination):
// - T_REGS: the row pass keeps t(x, .) in registers for the block's positions (else reads it per position, from the
//   L1 cache). 384: 1.51 ms reading against 2.22 ms holding (24 values a lane); 1024: 2.87 against 4.52 ms holding;
//   Bluestein 509: 10.29 against 11.42 ms holding.
// - OPAQUE: the lane index is made opaque to the compiler per position, so the lane's LDS slot addresses are
//   recomputed in each position instead of held in registers across the loop. It costs the row passes (384: 2.53
//   against 1.51 ms). In the column passes it helped the Bluestein pass (10.30 against 13.30 ms); since the block
//   phases run the first and last transform stages (whose lane index is opaque in every pass), it keeps the 384-point
//   middle pass at 255 VGPRs (two wavefronts per SIMD, against one with 256 VGPRs and 52 AGPRs): 1.57 ms against
//   2.09-2.11 ms before the fused stages. At 1024 points it would save registers that buy no occupancy (176 against
//   202 VGPRs, two wavefronts per SIMD either way) for 30 % more vector instructions: 13.03-13.07 ms without it,
//   13.72-13.78 with it.
// - The column tile (TC columns per block), with the next position's loads in flight during the transforms: 384,
//   8 columns (2.28 against 2.32 ms for 16); 1024, 8 columns (3.95 against 5.36 ms for 4); Bluestein, 4 columns
//   (9.27 against 10.61 ms for 8). With the blocks in XCD order (xcd_unit), per 4096 positions at every length:
//   384, 2.02-2.07 against 2.22-2.25 ms for 16; 1024, 13.69-13.71 against 14.91 ms for 4; Bluestein 509,
//   8.21-8.23 against 8.85-8.86 ms for 8; with the fused first and last stages, 1024 at 8 columns 13.67-13.77 ms
//   against 15.24 for 4, and 17.64 for 4 held to three wavefronts per SIMD (168 VGPRs, 48 bytes spilled). Blocks of
//   1, 2 or 4 row wavefronts differ by under 2 %.
// - PB: positions per block of both passes (t or P loaded, and the block's tables built, once per PB positions):
//   16 at 1024 points (1024^2 174,400 -> 184,100, Bluestein 509 244,800 -> 251,300 position-slices/s against 8),
//   8 at 384 (1.287 M against 1.252 M for 16).
template <int M, bool BLUE>
struct RegTuning {
  static constexpr bool T_REGS = M == 1024, ROW_OPAQUE = false, COL_OPAQUE = M == 384 || BLUE;
  static constexpr int COL_TILE = BLUE ? 4 : 8;
  static constexpr int PB = M == 1024 ? 16 : 8;
};

__device__ __forceinline__ int rslot(int e) { return e ^ (((e >> 3) ^ (e >> 7)) & 31); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int R>
__device__ __forceinline__ void dft_r(float2 (&u)[R], int sign) {
  if constexpr (R == 6) dft6(u, sign);
  else dft_pow2<R>(u, sign);
}

// The twiddles of stages 2 and 3 in LDS: tab[(j - 1) L2 + k] = w_{L2 R2}^{j k}, then tab[S2 + (j - 1) L3 + k] =
// w_{L3 R3}^{j k}, from roots[m] = e^{-2 pi i m / M}.
template <int M>
struct RegTw {
  using P = RegPlan<M>;
  static constexpr int L2 = P::R1, L3 = P::R1 * P::R2;
  static constexpr int S2 = (P::R2 - 1) * L2, kSize = S2 + (P::R3 - 1) * L3;
  __device__ static void build(float2* tab, const float2* __restrict__ roots, int tid, int nthreads) {
    for (int e = tid; e < kSize; e += nthreads) {
      if (e < S2) {
        const int j = e / L2 + 1, k = e % L2;
        tab[e] = roots[j * k * (M / (L2 * P::R2))];
      } else {
        const int f = e - S2, j = f / L3 + 1, k = f % L3;
        tab[e] = roots[j * k * (M / (L3 * P::R3))];
      }
    }
  }
};

// Whether the elements i + T r of an M-point line (n = M, not Bluestein's) lie, for every lane i, in the band that
// the antialias aperture zeroes on every grid: |k| > floor(M / 3), since the kept columns and the propagator's rows
// satisfy |k| <= 0.3333333 n. Their values are known zeros on the way in (compact inputs, products with P) and are
// never stored on the way out, so the code for them folds away (1024: r = 6..9; 384: r = 9..15).
template <int M>
__host__ __device__ constexpr bool in_band(int r) {
  constexpr int T = M / RegPlan<M>::V, K = M / 3;
  return r * T > K && r * T + T - 1 < M - K;
}
static_assert(!in_band<1024>(5) && in_band<1024>(6) && in_band<1024>(9) && !in_band<1024>(10), "1024-point band");
static_assert(!in_band<384>(8) && in_band<384>(9) && in_band<384>(15) && !in_band<384>(16), "384-point band");

// The first stage's known zeros in that band: bit j of band_mask<M, R>(q) when butterfly q's input j (element
// i + T (q + NB j)) lies in it.
template <int M, int R>
__host__ __device__ constexpr int band_mask(int q) {
  constexpr int NB = RegPlan<M>::V / R;
  int z = 0;
  for (int j = 0; j < R; ++j)
    if (in_band<M>(q + NB * j)) z |= 1 << j;
  return z;
}
static_assert(band_mask<1024, 16>(0) == 0x3c0, "1024: inputs 6..9");
static_assert(band_mask<384, 8>(0) == 0x38 && band_mask<384, 8>(1) == 0x18 && band_mask<384, 8>(2) == 0x18,
              "384: inputs 3..5 of the first butterfly, 3..4 of the others");

template <int M, int R, int Q>
__device__ __forceinline__ void band_dft(float2 (&u)[R], int sign) {
  constexpr int Z = band_mask<M, R>(Q);
  if constexpr (R == 16) dft16_z<Z>(u, sign);
  else if constexpr (R == 8) dft8_z<Z>(u, sign);
  else static_assert(R == 8 || R == 16, "a radix-8 or radix-16 first stage");
}
template <int M, int R, int NB, int... Q>
__device__ __forceinline__ void band_dfts(float2 (&u)[NB][R], int sign, std::integer_sequence<int, Q...>) {
  (band_dft<M, R, Q>(u[Q], sign), ...);
}

// One Stockham stage of radix R at sub-length L: butterfly bi = i + T q (q < V / R) reads x[bi + (M / R) j],
// multiplies by w_{L R}^{j k} (k = bi mod L; tw = the stage's table), takes the DFT and writes y[(bi - k) R + k + L j].
// FROM_REG: the inputs are the registers (element i + T r in v[r]); TO_REG: the outputs go to the registers.
// HALF_IN (a radix-16 first stage): the inputs j >= 8 are zero and are not read. BAND (a first stage): the inputs in
// the antialias band (in_band) are zero and are not read.
template <int M, int R, int L, int SIGN, bool FROM_REG, bool TO_REG, class Slot, bool HALF_IN = false,
          bool BAND = false>
__device__ __forceinline__ void vstage(float2 (&v)[RegPlan<M>::V], float2* __restrict__ buf, int i,
                                      const float2* __restrict__ tw, Slot slot) {
  constexpr int V = RegPlan<M>::V, T = M / V, NB = V / R, STRIDE = M / R;
  static_assert(V % R == 0, "the lane's butterflies");
  float2 u[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int bi = i + T * q;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if constexpr (FROM_REG)
        u[q][j] = ((HALF_IN && j >= R / 2) || (BAND && in_band<M>(q + NB * j))) ? make_float2(0.f, 0.f)
                                                                                : v[q + NB * j];
      else u[q][j] = buf[slot(bi + STRIDE * j)];
    }
  }
  if constexpr (!FROM_REG) wave_sync();  // every lane has read the line before any lane overwrites it
  if constexpr (BAND) {
    static_assert(FROM_REG && L == 1 && !HALF_IN, "the band's zeros: a first stage from the registers");
    band_dfts<M, R, NB>(u, SIGN, std::make_integer_sequence<int, NB>{});
  }
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int bi = i + T * q;
    const int k = (L == 1) ? 0 : (bi % L);
    if constexpr (L > 1) {
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const float2 w = tw[(j - 1) * L + k];
        u[q][j] = SIGN < 0 ? cmul(u[q][j], w) : cmulc(u[q][j], w);
      }
    }
    if constexpr (HALF_IN) {
      static_assert(R == 16 && L == 1 && FROM_REG, "the zero upper half: a radix-16 first stage");
      dft16_half(u[q], SIGN);
    } else if constexpr (!BAND) {  // BAND: transformed above
      dft_r<R>(u[q], SIGN);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if constexpr (TO_REG) v[q + NB * j] = u[q][j];
      else buf[slot((bi - k) * R + k + j * L)] = u[q][j];
    }
  }
  if constexpr (!TO_REG) wave_sync();
}

// The last stage of the inverse transform (radix R3 at sub-length M / R3, whose outputs land where its inputs were)
// for lane i of a line in LDS, one butterfly at a time, each butterfly's outputs (elements bi + (M / R3) j) stored
// straight to out[element * stride]: the values of one butterfly are live at a time, not the lane's V.
template <int M, class Slot>
__device__ __forceinline__ void last_stage_store(const float2* __restrict__ buf, int i, const float2* __restrict__ tw,
                                                 Slot slot, float2* out, long long stride) {
  using P = RegPlan<M>;
  constexpr int V = P::V, T = M / V, R = P::R3, L = P::R1 * P::R2, NB = V / R, STRIDE = M / R;
  static_assert(L * R == M, "the last stage");
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int bi = i + T * q;  // k = bi mod L = bi
    float2 u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) u[j] = buf[slot(bi + STRIDE * j)];
#pragma unroll
    for (int j = 1; j < R; ++j) u[j] = cmulc(u[j], tw[(j - 1) * L + bi]);
    dft_r<R>(u, +1);
#pragma unroll
    for (int j = 0; j < R; ++j) out[(long long)(bi + STRIDE * j) * stride] = u[j];
  }
}

// Unscaled transform of the line in the registers (SIGN -1 forward, +1 inverse). HALF_IN: the elements of the upper
// half of the line (r >= V / 2 in every lane) are zero. BAND: the elements in the antialias band (in_band) are zero.
template <int M, int SIGN, class Slot, bool HALF_IN = false, bool BAND = false>
__device__ __forceinline__ void vfft(float2 (&v)[RegPlan<M>::V], float2* buf, int i, const float2* tab, Slot slot) {
  using P = RegPlan<M>;
  vstage<M, P::R1, 1, SIGN, true, false, Slot, HALF_IN, BAND>(v, buf, i, nullptr, slot);
  vstage<M, P::R2, P::R1, SIGN, false, false>(v, buf, i, tab, slot);
  vstage<M, P::R3, P::R1 * P::R2, SIGN, false, true>(v, buf, i, tab + RegTw<M>::S2, slot);
}

// The line transform of logical length n: the M-point transform itself, or (BLUE, M = 1024) Bluestein's for n <= 512
// with the chirp cs[n] and the kernel transform ks[1024] (both in LDS; conjugated for the inverse). Values at
// elements e >= n are ignored on entry and undefined on return.
// BAND (not Bluestein's): the input is zero in the antialias band.
// CHIRP_IN / CHIRP_OUT (Bluestein's): false leaves out the opening / closing chirp product. An inverse transform
// without its closing product followed, after a pointwise product, by a forward one without its opening product
// (or a forward one then an inverse one) computes the same: the two chirps are conj(c_e) and c_e (or c_e and
// conj(c_e)), whose product is 1; an intensity |X_k|^2 does not see the closing product either. Without CHIRP_IN
// the elements e >= n are set to zero here as the opening product does.
template <int M, int SIGN, bool BLUE, class Slot, bool BAND = false, bool CHIRP_IN = true, bool CHIRP_OUT = true>
__device__ __forceinline__ void line_fft(float2 (&v)[RegPlan<M>::V], float2* buf, int i, const float2* tab,
                                         const float2* cs, const float2* ks, int n, Slot slot) {
  static_assert(!(BAND && BLUE), "the band's zeros are known at compile time for M-point lines only");
  if constexpr (!BLUE) {
    vfft<M, SIGN, Slot, false, BAND>(v, buf, i, tab, slot);
  } else {
    constexpr int V = RegPlan<M>::V, T = M / V;
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const int e = i + T * r;
      if (e < n) {
        if constexpr (CHIRP_IN) {
          const float2 c = cs[e];
          v[r] = SIGN < 0 ? cmul(v[r], c) : cmulc(v[r], c);
        }
      } else {
        v[r] = make_float2(0.f, 0.f);
      }
    }
    // the zero-padded input: elements i + T r with r >= V / 2 lie at or past 512 >= n
    vfft<M, -1, Slot, true>(v, buf, i, tab, slot);
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const float2 kh = ks[i + T * r];
      v[r] = SIGN < 0 ? cmul(v[r], kh) : cmulc(v[r], kh);
    }
    vfft<M, +1>(v, buf, i, tab, slot);
    // only the first n <= 512 outputs are kept: the upper half is dropped here, so that the last stage's work for it
    // is dead code
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const int e = i + T * r;
      if (r >= V / 2) {
        v[r] = make_float2(0.f, 0.f);
      } else if (CHIRP_OUT && e < n) {
        const float2 c = cs[e];
        v[r] = SIGN < 0 ? cmul(v[r], c) : cmulc(v[r], c);
      }
    }
  }
}

constexpr int kRowWaves = 2;   // wavefronts per row-pass block

__device__ __forceinline__ int compact_col(const RowArgs& a, int iy) {
  return iy <= a.kept_hi ? iy : (iy >= a.kept_lo ? iy - a.kept_shift : -1);
}

// Row pass: LINES = (64 / T) x kRowWaves rows x per block, positions [b0, b0 + PB). Rows past nx (a grid length
// that is not a multiple of LINES) compute on the last row and store nothing.
template <int M, int FLAGS, bool BLUE>
__global__ void __launch_bounds__(64 * kRowWaves) row_reg_kernel(RowArgs a, const float2* __restrict__ roots,
                                                                 const float2* __restrict__ chirp,
                                                                 const float2* __restrict__ khat, int batch) {
  using P = RegPlan<M>;
  constexpr int V = P::V, T = M / V, LINES = (64 / T) * kRowWaves, LS = P::LS, NT = 64 * kRowWaves;
  __shared__ float2 lds[LINES * LS];
  __shared__ float2 tab[RegTw<M>::kSize];
  __shared__ float2 cs[BLUE ? 512 : 1];
  __shared__ float2 ks[BLUE ? 1024 : 1];
  const int tid = threadIdx.x, line = tid / T, i = tid - line * T;
  const int n = BLUE ? a.ny : M;
  RegTw<M>::build(tab, roots, tid, NT);
  if constexpr (BLUE) {
    for (int e = tid; e < n; e += NT) cs[e] = chirp[e];
    for (int e = tid; e < 1024; e += NT) ks[e] = khat[e];
  }
  __syncthreads();
  const int x = blockIdx.x * LINES + line;
  const bool live = x < a.nx;
  const long long xs = live ? x : a.nx - 1;
  constexpr int PB = RegTuning<M, BLUE>::PB;
  const int b0 = blockIdx.y * PB;
  float2* buf = lds + line * LS;
  auto slot = [](int e) { return rslot(e); };
  const long long gw = (long long)a.nx * a.nc;
  // the compact column of element e of the lane's line (or -1: pruned, or past n)
  auto cidx = [&](int li, int r) {
    const int e = li + T * r;
    return e < n ? compact_col(a, e) : -1;
  };
  const float2* tr = a.t + xs * n;  // t(x, .): in registers (T_REGS) or read per position from the L1 cache
  constexpr bool TREG = RegTuning<M, BLUE>::T_REGS && (FLAGS & ROW_MUL_T) != 0;
  float2 treg[TREG ? V : 1];
  if constexpr (TREG) {
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const int e = i + T * r;
      treg[r] = e < n ? tr[e] : make_float2(0.f, 0.f);
    }
  }
  // Bluestein's inverse, the product with t and Bluestein's forward in a row (no exit plane between them): the
  // inverse's closing chirp and the forward's opening one cancel, and neither is applied (t is zero past n)
  constexpr bool CHIRP_PAIR = BLUE && (FLAGS & ROW_INV) != 0 && (FLAGS & ROW_MUL_T) != 0 && (FLAGS & ROW_FWD) != 0 &&
                              (FLAGS & ROW_WRITE_PLANE) == 0;
  for (int p = 0; p < PB; ++p) {
    const long long b = b0 + p;
    if (b >= batch) break;  // uniform over the block
    // ROW_OPAQUE: the lane index as a value the compiler cannot follow across positions, so the lane's LDS slot
    // addresses are recomputed in each position rather than kept live over the loop (one register each)
    int il = i;
    if constexpr (RegTuning<M, BLUE>::ROW_OPAQUE) asm volatile("" : "+v"(il));
    float2 v[V];
    if constexpr ((FLAGS & ROW_IN_COMPACT) != 0) {
      const float2* g = a.G + b * gw + xs * a.nc;
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int c = cidx(il, r);
        if (!BLUE && in_band<M>(r)) v[r] = make_float2(0.f, 0.f);
        else v[r] = c >= 0 ? g[c] : make_float2(0.f, 0.f);
      }
    } else {
      const float2* g = a.in_full + b * a.in_full_stride + xs * n;
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int e = il + T * r;
        v[r] = e < n ? g[e] : make_float2(0.f, 0.f);
      }
    }
    if constexpr ((FLAGS & ROW_INV) != 0) {
      // from the compact wave, the band's elements are zero
      constexpr bool band = !BLUE && (FLAGS & ROW_IN_COMPACT) != 0;
      line_fft<M, +1, BLUE, decltype(slot), band, true, !CHIRP_PAIR>(v, buf, il, tab, cs, ks, n, slot);
    }
    if constexpr ((FLAGS & ROW_WRITE_PLANE) != 0) {
      if (live) {
        float2* pl = a.plane + b * a.plane_stride + xs * n;
#pragma unroll
        for (int r = 0; r < V; ++r) {
          const int e = il + T * r;
          if (e < n) pl[e] = cscale(v[r], a.out_scale);
        }
      }
    }
    if constexpr ((FLAGS & ROW_MUL_T) != 0) {
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int e = il + T * r;
        if constexpr (TREG) v[r] = cmul(v[r], treg[r]);
        else v[r] = e < n ? cmul(v[r], tr[e]) : make_float2(0.f, 0.f);
      }
    }
    if constexpr ((FLAGS & ROW_FWD) != 0)
      line_fft<M, -1, BLUE, decltype(slot), false, !CHIRP_PAIR>(v, buf, il, tab, cs, ks, n, slot);
    if constexpr ((FLAGS & ROW_OUT_COMPACT) != 0) {
      if (live) {
        float2* g = a.G + b * gw + xs * a.nc;
#pragma unroll
        for (int r = 0; r < V; ++r) {
          if (!BLUE && in_band<M>(r)) continue;  // never kept: the forward transform's outputs there are dead
          const int c = cidx(il, r);
          if (c >= 0) g[c] = v[r];
        }
      }
    }
  }
}

// The column pass's tile: TC columns of one position, column c at slots [c CS, c CS + M), element e at slot
// rslot(e) (Bluestein: the 1024-point convolution buffer, which holds the column's n values before the transform).
// CS, from a bank model of 64 banks of 4 bytes in groups of 32 lanes: 1024 points, CS = 1028 (TC = 8) or 1032
// (TC = 4) keeps the block phases' row segments free of conflicts; 384 points, CS = 400 keeps the exchanges free
// of conflicts (four columns per wavefront).
template <int M, int TC>
constexpr int col_stride() {
  return M == 384 ? 400 : (TC == 8 ? 1028 : 1032);
}

// The tile slot of element e of column c. At 384 points the element slots of column c are XORed with c (16-column
// tiles) or 4 c (8-column tiles): the block phases' row segments then fall on distinct banks, which CS = 400 alone
// does not give (a model of 64 banks finds the 16-column row segments 8-way conflicted without it, the 8-column ones
// 4-way, and both free with it, the exchanges free either way; measured 4.7e8 bank-conflict cycles per 4096-position
// launch of 16-column tiles without it).
template <int M, int TC>
__device__ __forceinline__ int tile_elem(int e, int c) {
  if constexpr (M == 384) return rslot(e) ^ ((TC == 8 ? 4 * c : c) & 31);
  else return rslot(e);
}

// Column pass: block = TC compact columns [TC tile, + TC) of positions [b0, b0 + PB) of its unit (xcd_unit);
// TC x T threads, the T lanes of column c = tid / T. MODE 0: a middle slice (COL_IN_G | COL_FWD | COL_MUL_P | COL_INV |
// COL_OUT); 1: the first pass, whose input is the probe spectrum times the position's ramps, generated in the block
// phase (COL_GENERATE | COL_INV | COL_OUT); 2: the last pass, which ends in the detectors (COL_IN_G | COL_FWD |
// COL_MUL_P | COL_DETECT): each thread sums its elements' contributions to one detector slot at a time, pixels in a
// fixed order, then a tree over the block writes the block's partial sum (partial[position][slot][tile]).
template <int M, bool BLUE, int TC, int MODE>
__global__ void __launch_bounds__(TC * (M / RegPlan<M>::V)) col_reg_kernel(ColArgs a, const float2* __restrict__ roots,
                                                                       const float2* __restrict__ chirp,
                                                                       const float2* __restrict__ khat, int batch) {
  using P = RegPlan<M>;
  constexpr int V = P::V, T = M / V, CS = col_stride<M, TC>(), NT = TC * T;
  constexpr int RSTEP = NT / TC;  // rows per block-phase round
  __shared__ float2 tile[TC * CS];
  __shared__ float2 tab[RegTw<M>::kSize];
  __shared__ float2 cs[BLUE ? 512 : 1];
  __shared__ float2 ks[BLUE ? 1024 : 1];
  __shared__ double red[MODE == 2 ? (NT / 64) * kMaxSlots : 1];
  const int tid = threadIdx.x;
  const int n = BLUE ? a.nx : M;
  RegTw<M>::build(tab, roots, tid, NT);
  if constexpr (BLUE) {
    for (int e = tid; e < n; e += NT) cs[e] = chirp[e];
    for (int e = tid; e < 1024; e += NT) ks[e] = khat[e];
  }
  const int2 unit = xcd_unit();
  const int c0 = unit.x * TC;
  // block phases: thread -> column cb of the tile, rows xi + RSTEP r
  const int cb = tid % TC, xi = tid / TC;
  // column lanes: column wc, lane i
  const int wc = tid / T, i = tid - wc * T;
  float2* col = tile + wc * CS;
  auto slot = [wc](int e) { return tile_elem<M, TC>(e, wc); };
  const long long gw = (long long)a.nx * a.nc;
  // P of the lane's elements of column c0 + wc, for every position of the block
  float2 pv[MODE == 1 ? 1 : V];
  if constexpr (MODE != 1) {
    const float2* Pc = a.P + c0 + wc;
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const int e = i + T * r;
      if (!BLUE && in_band<M>(r)) pv[r] = make_float2(0.f, 0.f);
      else pv[r] = e < n ? Pc[(long long)e * a.nc] : make_float2(0.f, 0.f);
    }
  }
  const int iyb = a.keep[c0 + cb];
  const bool kept = iyb >= 0;  // padding columns: their values are never used (the block phases store zeros there)
  constexpr int PB = RegTuning<M, BLUE>::PB;
  const int b0 = unit.y * PB;
  // The block phase's values of the next position are loaded into registers while this position transforms: the
  // loads' latency then overlaps the arithmetic (measured: 1024 points, 8 columns, 6.76 -> 3.95 ms per launch of
  // 1024 positions; 384 points 3.03 -> 2.28 ms per 4096).
  // the rows a block-phase thread moves: every row of the M-point line, or (Bluestein's) the rows below 512 >= n
  constexpr int NL = (BLUE ? 512 : M) / RSTEP;
  static_assert((BLUE ? 512 : M) % RSTEP == 0, "whole rounds of rows");
  float2 pre[MODE == 1 ? 1 : NL];
  // The first pass generates each position's probe where it writes the tile (prefetching the three factors of
  // every element measured slower at 384 points and the same at 1024).
  auto generate = [&](long long bb, int x) {
    return kept ? cmul(cmul(a.spec[(long long)x * a.ny + iyb], a.rx[bb * a.nx + x]), a.ry[bb * a.ny + iyb])
                : make_float2(0.f, 0.f);
  };
  auto masked = [&](int r) {  // Bluestein's prefetched value of row xi + RSTEP r, or zero
    return (kept && xi + RSTEP * r < n) ? pre[r] : make_float2(0.f, 0.f);
  };
  auto prefetch = [&](long long bb) {
    if constexpr (MODE != 1) {
      const float2* gg = a.G + bb * gw + c0 + cb;
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        // branch-free loads of every column (padding columns are read too, and their values replaced by zeros)
        const int x = xi + RSTEP * r;
        if constexpr (BLUE) {
          // rows past n (Bluestein's rows n..511) read row n - 1; the zeros are put in where the values are used
          // (masked), since a select here made the compiler wait for these loads at once
          pre[r] = gg[(long long)(x < n ? x : n - 1) * a.nc];
        } else {
          const float2 t = gg[(long long)x * a.nc];
          pre[r] = kept ? t : make_float2(0.f, 0.f);
        }
      }
    }
  };
  // FUSE (not Bluestein's): a block-phase thread (column cb, rows xi + T r) holds what lane xi of column cb holds
  // for the line's first transform stage (RSTEP = T), so the block phase runs that stage from its loaded values into
  // the tile, and the last stage of the inverse from the tile into the values it stores: the columns' own lanes run
  // the stages between, and two of the tile's passes through LDS per position (the raw values in and out) are gone.
  constexpr bool FUSE = !BLUE;
  static_assert(!FUSE || (RSTEP == T && NL == V), "the block phase's rows are a lane's elements");
  float2* bcol = tile + cb * CS;
  auto bslot = [cb](int e) { return tile_elem<M, TC>(e, cb); };
  if (b0 < batch) prefetch(b0);
  // The first position's loads complete before the loop. At the loop's head the compiler waits for the loads of
  // the position about to run; the paths into the head are the first entry (those loads outstanding, nothing after
  // them) and the back edge (those loads, then the previous position's stores): merged, they would make it wait for
  // every outstanding access, the stores just issued included, once per position. With the first entry's loads
  // done, the head waits only for the loads (s_waitcnt vmcnt(NL), the NL stores still in flight); the stores and
  // the loads take no branch for the same reason (below).
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  for (int p = 0; p < PB; ++p) {
    const long long b = b0 + p;
    if (b >= batch) break;  // uniform over the block
    float2* g = a.G + b * gw + c0 + cb;
    __syncthreads();  // the tile's previous position is stored (and, first, the tables are built)
    // the block phase's lane index, opaque to the compiler per position: its LDS slot addresses are recomputed here
    // rather than held in registers across the loop
    int xl = xi;
    asm volatile("" : "+v"(xl));
    if constexpr (FUSE) {
      if constexpr (MODE == 1) {
        float2 gv[V];
#pragma unroll
        for (int r = 0; r < V; ++r) gv[r] = generate(b, xl + T * r);
        vstage<M, P::R1, 1, +1, true, false>(gv, bcol, xl, nullptr, bslot);
      } else {
        vstage<M, P::R1, 1, -1, true, false>(pre, bcol, xl, nullptr, bslot);
      }
    } else {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int x = xi + RSTEP * r;  // below 512: a slot of the tile; the transform ignores the elements past n
        if constexpr (MODE == 1) {
          if (x < n) tile[cb * CS + tile_elem<M, TC>(x, cb)] = generate(b, x);
        } else {
          tile[cb * CS + tile_elem<M, TC>(x, cb)] = masked(r);
        }
      }
    }
    if (p + 1 < PB && b + 1 < batch) prefetch(b + 1);
    __syncthreads();
    int il = i;  // COL_OPAQUE: as ROW_OPAQUE in the row pass
    if constexpr (RegTuning<M, BLUE>::COL_OPAQUE) asm volatile("" : "+v"(il));
    float2 v[V];
    if constexpr (FUSE) {
      if constexpr (MODE != 1) {
        // the forward transform's second and last stages
        vstage<M, P::R2, P::R1, -1, false, false>(v, col, il, tab, slot);
        vstage<M, P::R3, P::R1 * P::R2, -1, false, true>(v, col, il, tab + RegTw<M>::S2, slot);
      }
    } else {
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int e = il + T * r;
        v[r] = e < n ? col[slot(e)] : make_float2(0.f, 0.f);
      }
      wave_sync();  // every lane has read its column before the exchanges overwrite it
      // Bluestein's forward without its closing chirp: the inverse after the product with P (zero past n) leaves
      // out its opening one (mode 0), and the detectors sum intensities (mode 2)
      if constexpr (MODE != 1)
        line_fft<M, -1, BLUE, decltype(slot), false, true, false>(v, col, il, tab, cs, ks, n, slot);
    }
    if constexpr (MODE != 1) {
#pragma unroll
      for (int r = 0; r < V; ++r) {
        if (!BLUE && in_band<M>(r)) v[r] = make_float2(0.f, 0.f);  // P is zero there
        else v[r] = cmul(v[r], pv[r]);
      }
    }
    if constexpr (MODE == 2) {
      // the detectors: the lane's elements x = il + T r of column c0 + wc (ky = keep); the pixelated accumulators,
      // then each slot summed over the lane's elements, over the wavefront (shuffles, fixed order) and over the
      // block's wavefronts in order
      const DetPlan& d = a.det;
      const int iy = a.keep[c0 + wc];
      const long long pos = first_position(a.pos0, a.pos_dev) + b;
      // The lane's values go back to its own slots of the column, and the sums read them from there in a loop that
      // is not unrolled: held in registers through the slot loop, float64 intensities of all V values spilled.
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int e = il + T * r;
        if (e < n) col[slot(e)] = v[r];
      }
      auto intensity = [&](int e) {
        const float2 u = col[slot(e)];
        return ((double)u.x * u.x + (double)u.y * u.y) * d.scale;
      };
      if (iy >= 0 && d.npix > 0) {
#pragma unroll 1
        for (int e = il; e < n; e += T) detect_pixelated(d, e, iy, intensity(e), pos);
      }
      const int lane = tid & 63, wave = tid >> 6;
      const double sy = (iy >= 0 && d.sqy) ? d.sqy[iy] : 0.0;
      for (int q = 0; q < d.nslots; ++q) {
        const int kind = d.slot_kind[q];
        double val = 0.0;
        if (iy >= 0) {
          if (kind == 0) {
            const double i2 = d.slot_i2[q], o2 = d.slot_o2[q];
#pragma unroll 1
            for (int e = il; e < n; e += T) {
              const double a2 = __dadd_rn(d.sqx[e], sy);  // float64, as the host rounds it
              if (a2 >= i2 && a2 < o2) val += intensity(e);
            }
          } else {
#pragma unroll 1
            for (int e = il; e < n; e += T) val += slot_value(d, q, e, iy, 0.0, intensity(e));
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) val += __shfl_xor(val, off);
        if (lane == 0) red[wave * kMaxSlots + q] = val;
      }
      __syncthreads();
      if (tid < d.nslots) {
        double t = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) t += red[wv * kMaxSlots + tid];
        d.partial[(b * d.nslots + tid) * d.blocks_per_wave + unit.x] = t;
      }
      continue;  // nothing is stored; the next position's block phase waits at the loop's first barrier
    } else if constexpr (FUSE) {
      // the inverse transform's stages before the last: after the product with P the band's elements are zero (not
      // in the first pass, whose generated probe spectrum need not be, and whose first stage ran in the block phase)
      if constexpr (MODE == 0)
        vstage<M, P::R1, 1, +1, true, false, decltype(slot), false, true>(v, col, il, nullptr, slot);
      vstage<M, P::R2, P::R1, +1, false, false>(v, col, il, tab, slot);
      __syncthreads();
      // the last stage in the block phase: thread (cb, xi) runs lane xi's butterflies of column cb and stores their
      // outputs
      int xs = xi;
      asm volatile("" : "+v"(xs));
      // every column is stored, padding columns too (nothing reads them, and their values are zeros from zero
      // inputs): a store that only kept columns take would put a branch in the loop and bring back the head's wait
      // for all outstanding accesses
      last_stage_store<M>(bcol, xs, tab + RegTw<M>::S2, bslot, g, a.nc);
    } else {
      line_fft<M, +1, BLUE, decltype(slot), !BLUE && MODE == 0, MODE != 0>(v, col, il, tab, cs, ks, n, slot);
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const int e = il + T * r;
        if (e < n) col[slot(e)] = v[r];
      }
      __syncthreads();
      // every column and NL rows a thread, as the FUSE stores (padding columns get zeros): a row past n stores the
      // lane's last row below n again, the same value to the same address (n >= 257 > RSTEP: every lane has one)
      const int xlast = xi + RSTEP * ((n - 1 - xi) / RSTEP);
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int x0 = xi + RSTEP * r, x = x0 < n ? x0 : xlast;
        g[(long long)x * a.nc] = tile[cb * CS + tile_elem<M, TC>(x, cb)];
      }
    }
  }
}

template <class K, class... A>
void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, A... args) {
  if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

constexpr int ROW_MID = ROW_IN_COMPACT | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_MID_PLANE = ROW_MID | ROW_WRITE_PLANE;
constexpr int ROW_ENTRY = ROW_IN_FULL | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_FROM_FULL = ROW_IN_FULL | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_EXIT = ROW_IN_COMPACT | ROW_INV | ROW_WRITE_PLANE;
constexpr int COL_MID = COL_IN_G | COL_FWD | COL_MUL_P | COL_INV | COL_OUT;

// the transform length the register passes use for a grid length n: n itself (384, 1024), 1024 for Bluestein's
// lengths 257..512 that are not 256 or 384 and have no fast pass, 0 for none
int reg_length(const LineFFT& f) {
  if (f.n == 384 || f.n == 1024) return f.blue ? 0 : f.n;
  if (f.blue && f.m == 1024 && f.n >= 257 && f.n <= 512) return 1024;
  return 0;
}

template <int M, int F, bool BLUE>
void row_go(const RowArgs& a, int batch, hipStream_t s) {
  constexpr int LINES = (64 / (M / RegPlan<M>::V)) * kRowWaves;
  constexpr int PB = RegTuning<M, BLUE>::PB;
  dim3 grid((a.nx + LINES - 1) / LINES, (batch + PB - 1) / PB);
  launch_timed(row_reg_kernel<M, F, BLUE>, grid, dim3(64 * kRowWaves), s, a.ev_start, a.ev_stop, a, a.fy.roots,
               a.fy.chirp, a.fy.khat, batch);
}

template <int M, bool BLUE>
void row_launch(const RowArgs& a, int batch, hipStream_t s) {
  switch (a.flags) {
    case ROW_MID: row_go<M, ROW_MID, BLUE>(a, batch, s); break;
    case ROW_MID_PLANE: row_go<M, ROW_MID_PLANE, BLUE>(a, batch, s); break;
    case ROW_ENTRY: row_go<M, ROW_ENTRY, BLUE>(a, batch, s); break;
    case ROW_FROM_FULL: row_go<M, ROW_FROM_FULL, BLUE>(a, batch, s); break;
    default: row_go<M, ROW_EXIT, BLUE>(a, batch, s); break;
  }
  SK_HIP(hipGetLastError());
}

// columns per column-pass block: the caller's sk_options.col_tile where this length has that width, else the
// default (RegTuning::COL_TILE)
int reg_col_tile(int m, bool blue, int requested) {
  if (requested == 0)
    return m == 384 ? RegTuning<384, false>::COL_TILE
                    : (blue ? RegTuning<1024, true>::COL_TILE : RegTuning<1024, false>::COL_TILE);
  if (m == 384 ? (requested == 8 || requested == 16) : (requested == 4 || requested == 8)) return requested;
  fail(SK_ERR_INVALID, "col_tile " + std::to_string(requested) + " is not a column-pass width for " +
                           std::to_string(m) + "-point lines (384: 8 or 16; 1024: 4 or 8)");
}

constexpr int COL_FIRST = COL_GENERATE | COL_INV | COL_OUT;
constexpr int COL_LAST = COL_IN_G | COL_FWD | COL_MUL_P | COL_DETECT;

template <int M, bool BLUE, int TC, int MODE>
void col_go_mode(const ColArgs& a, int batch, hipStream_t s) {
  constexpr int NT = TC * (M / RegPlan<M>::V);
  constexpr int PB = RegTuning<M, BLUE>::PB;
  dim3 grid(a.nc / TC, (batch + PB - 1) / PB);
  launch_timed(col_reg_kernel<M, BLUE, TC, MODE>, grid, dim3(NT), s, a.ev_start, a.ev_stop, a, a.fx.roots,
               a.fx.chirp, a.fx.khat, batch);
  SK_HIP(hipGetLastError());
}
template <int M, bool BLUE, int TC>
void col_go(const ColArgs& a, int batch, hipStream_t s) {
  if (a.flags == COL_FIRST) col_go_mode<M, BLUE, TC, 1>(a, batch, s);
  else if (a.flags == COL_LAST) col_go_mode<M, BLUE, TC, 2>(a, batch, s);
  else col_go_mode<M, BLUE, TC, 0>(a, batch, s);
}

}  // namespace

bool reg_row_supported(const RowArgs& a) {
  if (a.fy.use_mfma || reg_length(a.fy) == 0 || a.fy.n != a.ny) return false;
  const int f = a.flags;
  return f == ROW_MID || f == ROW_MID_PLANE || f == ROW_ENTRY || f == ROW_FROM_FULL || f == ROW_EXIT;
}

bool reg_col_supported(const ColArgs& a) {
  if (a.fx.use_mfma || a.fx.n != a.nx) return false;
  if (a.flags != COL_MID && a.flags != COL_FIRST && a.flags != COL_LAST) return false;
  const int m = reg_length(a.fx);
  return m != 0 && a.nc % reg_col_tile(m, a.fx.blue, a.col_tile) == 0;
}

int reg_col_blocks_per_wave(const ColArgs& a) {
  return a.nc / reg_col_tile(reg_length(a.fx), a.fx.blue, a.col_tile);
}

void launch_row_reg(const RowArgs& a, int batch, hipStream_t s) {
  if (a.fy.blue) row_launch<1024, true>(a, batch, s);
  else if (a.fy.n == 384) row_launch<384, false>(a, batch, s);
  else row_launch<1024, false>(a, batch, s);
}

void launch_col_reg(const ColArgs& a, int batch, hipStream_t s) {
  const int m = reg_length(a.fx), tc = reg_col_tile(m, a.fx.blue, a.col_tile);
  if (a.fx.blue) {
    if (tc == 4) col_go<1024, true, 4>(a, batch, s);
    else col_go<1024, true, 8>(a, batch, s);
  } else if (m == 384) {
    if (tc == 16) col_go<384, false, 16>(a, batch, s);
    else col_go<384, false, 8>(a, batch, s);
  } else {
    if (tc == 4) col_go<1024, false, 4>(a, batch, s);
    else col_go<1024, false, 8>(a, batch, s);
  }
}

}  // namespace sk
