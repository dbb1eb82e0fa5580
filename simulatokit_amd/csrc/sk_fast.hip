// sk_fast.hip — register-resident row and column passes for lines of 256 and 512 points.
//
// Each line (a row along y, or a column along x) is held by T = N / 8 lanes, eight values per lane: lane i holds
// the elements i + T r, r = 0..7. The line transform is a Stockham autosort FFT (radix 8, 8, 8 for 512; 8, 8, 4 for
// 256): the first stage works straight from registers, the last stage leaves the result in registers in natural
// order, and the two exchanges between stages go through LDS. The pointwise work of the slice (the transmission
// multiply in the row pass, the propagator multiply in the column pass) and the detector reduction of the final
// pass happen on the registers between the transforms, so every pass reads and writes each stored value once.
//
//   row pass:    one wavefront per block and row x; IFFT_y -> * t(x, y) -> FFT_y for 8 positions in turn, with the
//                zero band of the 2/3 bandlimit pruned (row_fast_kernel)
//   column pass: at 512, blocks of 8 compact columns, one column per wavefront, running balanced spans of
//                (tile, position) units; the block phases move 64-byte row segments between HBM and an LDS tile and
//                run the first forward and last inverse stage, the wavefronts the middle stages and * P(kx, ky) or
//                the detectors (col512_span_kernel; col512_lanes_kernel for the lane pipeline; col512_kernel with
//                16-column blocks for batches that stream from HBM). At 256, 16 columns per block (col_fast_kernel).
//
// A block keeps the lane's twiddle factors (registers or LDS) and its share of t (row pass) or P (column pass) for
// all the positions it runs. The row lines and the 512-point column tiles use XOR-swizzled LDS layouts, each found by
// search to be free of bank conflicts for every access pattern of its kernel; the 256-point column pass keeps its
// tile as [x][16 columns]. Twiddles come from the line's float64-derived root table.

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "sk_dft.h"
#include "sk_internal.h"

namespace sk {
namespace {

// The stage plan of a fast length: radices (R1, R2, R3) with sub-lengths L = 1, R1, R1 R2.
template <int N>
struct FastPlan;
template <>
struct FastPlan<512> {
  static constexpr int R1 = 8, R2 = 8, R3 = 8;
};
template <>
struct FastPlan<256> {
  static constexpr int R1 = 8, R2 = 8, R3 = 4;
};

// The twiddle factors a lane applies in stages 2 and 3 (stage 1 has L = 1 and none): w[q][j] = w_{L R}^{j k},
// k = (i + T q) mod L, for butterfly q of the lane. Forward factors; the inverse uses their conjugates.
template <int N>
struct LaneTwiddles {
  static constexpr int T = N / 8;
  static constexpr int R2 = FastPlan<N>::R2, R3 = FastPlan<N>::R3;
  static constexpr int L2 = FastPlan<N>::R1, L3 = FastPlan<N>::R1 * FastPlan<N>::R2;
  float2 w2[8 / R2][R2];
  float2 w3[8 / R3][R3];
  __device__ __forceinline__ void load(const float2* __restrict__ tw, int i) {
#pragma unroll
    for (int q = 0; q < 8 / R2; ++q) {
      const int k = (i + T * q) % L2;
#pragma unroll
      for (int j = 0; j < R2; ++j) w2[q][j] = tw[j * k * (N / (L2 * R2))];
    }
#pragma unroll
    for (int q = 0; q < 8 / R3; ++q) {
      const int k = (i + T * q) % L3;
#pragma unroll
      for (int j = 0; j < R3; ++j) w3[q][j] = tw[j * k * (N / (L3 * R3))];
    }
  }
  __device__ __forceinline__ float2 s2(int q, int j) const { return w2[q][j]; }
  __device__ __forceinline__ float2 s3(int q, int j) const { return w3[q][j]; }
};

// The same factors read from per-block LDS tables t2[j][k] (k < L2) and t3[j][k] (k < L3): lanes read consecutive
// entries (stage 3) or share them (stage 2), so the reads are free of bank conflicts. Frees the lane's registers.
template <int N>
struct LdsTwiddles {
  static constexpr int T = N / 8;
  static constexpr int R2 = FastPlan<N>::R2, R3 = FastPlan<N>::R3;
  static constexpr int L2 = FastPlan<N>::R1, L3 = FastPlan<N>::R1 * FastPlan<N>::R2;
  static constexpr int kSize = R2 * L2 + R3 * L3;
  const float2* t2;
  const float2* t3;
  int i;
  // fill the tables (all threads of the block; the caller synchronises)
  __device__ static void build(float2* table, const float2* __restrict__ tw, int tid, int nthreads) {
    for (int e = tid; e < kSize; e += nthreads) {
      if (e < R2 * L2) {
        const int j = e / L2, k = e % L2;
        table[e] = tw[j * k * (N / (L2 * R2))];
      } else {
        const int f = e - R2 * L2, j = f / L3, k = f % L3;
        table[e] = tw[j * k * (N / (L3 * R3))];
      }
    }
  }
  __device__ __forceinline__ float2 s2(int q, int j) const { return t2[j * L2 + (i + T * q) % L2]; }
  __device__ __forceinline__ float2 s3(int q, int j) const { return t3[j * L3 + (i + T * q) % L3]; }
};

// Synchronisation of an exchange that stays inside one wavefront: its lanes execute in lockstep and the LDS serves
// one wavefront's instructions in order, so only the compiler must not move LDS accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool WAVE>
__device__ __forceinline__ void exchange_sync() {
  if constexpr (WAVE) wave_sync();
  else __syncthreads();
}

// LDS slot of element j of a row line: an XOR swizzle under which every access pattern of the 256- and 512-point
// stages (stores 8i + j, loads i + T j, stage-2 stores (i / 8) 64 + i mod 8 + 8 j) hits distinct bank pairs in every
// group of 16 and of 32 lanes.
__device__ __forceinline__ int row_slot(int j) { return j ^ ((j >> 3) & 31); }

// LDS slot of (x, column) in a 512 x TC column tile. Found by search among shift-XOR swizzles to be free of bank
// conflicts, in groups of 16 and of 32 lanes, for every access of the column pass: the block phases' rows
// (lanes that share an x take consecutive columns, x = xi + 64 r), the block phases' first-stage writes
// (x = 8 xi + j), and each column wavefront's exchanges (reads x = l + 64 j, writes 8 l + j and
// 64 (l / 8) + l mod 8 + 8 j).
template <int TC>
__device__ __forceinline__ int tile_slot(int x, int c) {
  if constexpr (TC == 16) {
    const int h = ((x >> 3) ^ (x >> 7)) & 1;
    const int g = (x ^ (x >> 2) ^ (x >> 3)) & 15;
    return (x ^ h) * 16 + (c ^ g);
  } else {
    static_assert(TC == 8, "column tiles of 8 or 16");
    const int h = ((x >> 3) ^ (x >> 6)) & 3;
    const int g = (x ^ (x >> 4)) & 7;
    return (x ^ h) * 8 + (c ^ g);
  }
}

// One Stockham stage of radix R at sub-length L (see the file comment); W: the lane's factors for this stage
// (ignored when L = 1). FROM_REG: inputs are the registers; TO_REG: outputs go to the registers. WAVE: the line
// lives in one wavefront, so the exchanges need no block barrier. Z34: the first radix-8 stage of a line whose
// registers 3 and 4 are zero in every lane (they are not read).
template <int N, int R, int L, int SIGN, bool FROM_REG, bool TO_REG, bool WAVE = false, bool Z34 = false, class Idx,
          class Tw>
__device__ __forceinline__ void stage(float2 (&v)[8], float2* __restrict__ buf, int i, Tw W, Idx idx) {
  static_assert(!Z34 || (R == 8 && L == 1 && FROM_REG), "Z34: the first stage of radix 8, from the registers");
  constexpr int T = N / 8;
  constexpr int NB = 8 / R;
  constexpr int STRIDE = N / R;
  float2 u[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int bi = i + T * q;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if constexpr (FROM_REG) u[q][j] = v[q + NB * j];
      else u[q][j] = buf[idx(bi + STRIDE * j)];
    }
  }
  if constexpr (!FROM_REG) exchange_sync<WAVE>();  // every lane has read the line before any lane overwrites it
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int bi = i + T * q;
    const int k = (L == 1) ? 0 : (bi % L);
    if constexpr (L > 1) {
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const float2 wf = W(q, j);
        u[q][j] = SIGN < 0 ? cmul(u[q][j], wf) : cmulc(u[q][j], wf);
      }
    }
    if constexpr (Z34) dft8_z34(u[q], SIGN);
    else dft_pow2<R>(u[q], SIGN);
    if constexpr (TO_REG) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[q + NB * j] = u[q][j];
    } else {
      const int base = (bi - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) buf[idx(base + j * L)] = u[q][j];
    }
  }
  if constexpr (!TO_REG) exchange_sync<WAVE>();
}

// Unscaled transform of the line in registers (SIGN -1 forward, +1 inverse). Z34: registers 3 and 4 hold zeros in
// every lane (elements 3T..5T-1 of the line).
template <int N, int SIGN, bool WAVE = false, bool Z34 = false, class Idx, class TW>
__device__ __forceinline__ void fft_reg(float2 (&v)[8], float2* buf, int i, const TW& w, Idx idx) {
  using P = FastPlan<N>;
  auto none = [](int, int) { return make_float2(1.f, 0.f); };
  auto f2 = [&w](int q, int j) { return w.s2(q, j); };
  auto f3 = [&w](int q, int j) { return w.s3(q, j); };
  stage<N, P::R1, 1, SIGN, true, false, WAVE, Z34>(v, buf, i, none, idx);
  stage<N, P::R2, P::R1, SIGN, false, false, WAVE>(v, buf, i, f2, idx);
  stage<N, P::R3, P::R1 * P::R2, SIGN, false, true, WAVE>(v, buf, i, f3, idx);
}

constexpr int kColTile = 16;  // compact columns per column-pass block

constexpr int ROW_MID = ROW_IN_COMPACT | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_MID_PLANE = ROW_MID | ROW_WRITE_PLANE;
constexpr int ROW_ENTRY = ROW_IN_FULL | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_FROM_FULL = ROW_IN_FULL | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
constexpr int ROW_EXIT = ROW_IN_COMPACT | ROW_INV | ROW_WRITE_PLANE;

constexpr int COL_MID = COL_IN_G | COL_FWD | COL_MUL_P | COL_INV | COL_OUT;
constexpr int COL_INIT = COL_GENERATE | COL_INV | COL_OUT;
constexpr int COL_FINAL = COL_IN_G | COL_FWD | COL_MUL_P | COL_DETECT;

// compact column of iy, or -1 for a pruned ky
__device__ __forceinline__ int compact_of(const RowArgs& a, int iy) {
  return iy <= a.kept_hi ? iy : (iy >= a.kept_lo ? iy - a.kept_shift : -1);
}

// Row pass: block = one wavefront = 64 / T rows x of positions [b0, b0 + PB). PIPE: the positions belong to one lane
// of the pipeline, whose descriptor for the current step gives the transmission function and the position count.
// BAND (compact input and output): the kept set has the shape of a 2/3 bandlimit, 2T-1 <= kept_hi < 3T and
// 5T <= kept_lo <= 6T (the host checks), so of the lane's elements i + T r those of r = 0, 1, 6, 7 are always kept,
// r = 3, 4 never (the inverse transform starts from zeros there, the forward one's values there are dropped), and
// only r = 2 and 5 depend on the lane.
template <int N, int FLAGS, int PB, bool PIPE = false, bool BAND = false>
__global__ void __launch_bounds__(64) row_fast_kernel(RowArgs a, const float2* __restrict__ tw, int batch) {
  static_assert(!BAND || (FLAGS & (ROW_IN_COMPACT | ROW_OUT_COMPACT | ROW_INV)) ==
                             (ROW_IN_COMPACT | ROW_OUT_COMPACT | ROW_INV), "BAND: compact in and out");
  constexpr int T = N / 8, ROWS = 64 / T, LD = N;
  __shared__ float2 lds[ROWS * LD];
  const int tid = threadIdx.x, row = tid / T, i = tid - row * T;
  const int x = blockIdx.x * ROWS + row;  // the host launches nx / ROWS blocks along x: every x is in range
  const int b0 = blockIdx.y * PB;
  const float2* tsrc = a.t;
  long long limit = batch;
  if constexpr (PIPE) {
    const int lane = b0 / a.pipe.lane_waves;  // lane_waves is a multiple of PB: the block stays in one lane
    const LaneDesc d = a.pipe.descs[(long long)pipe_step(a.pipe) * a.pipe.lanes + lane];
    if (d.t == nullptr) return;  // idle lane (uniform over the block)
    tsrc = d.t;
    limit = (long long)lane * a.pipe.lane_waves + d.count;
  }
  float2* buf = lds + row * LD;
  auto idx = [](int j) { return row_slot(j); };
  const long long gw = (long long)a.nx * a.nc;
  LaneTwiddles<N> w;
  w.load(tw, i);
  int cidx[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) cidx[r] = compact_of(a, i + T * r);
  float2 t[8];
  if constexpr ((FLAGS & ROW_MUL_T) != 0) {
    const float2* tr = tsrc + (long long)x * N;
#pragma unroll
    for (int r = 0; r < 8; ++r) t[r] = tr[i + T * r];
  }
  for (int p = 0; p < PB; ++p) {
    const long long b = b0 + p;
    if (b >= limit) break;  // uniform over the block
    float2 v[8];
    if constexpr (BAND) {
      const float2* g = a.G + b * gw + (long long)x * a.nc;
      v[0] = g[i];
      v[1] = g[i + T];
      v[2] = cidx[2] >= 0 ? g[cidx[2]] : make_float2(0.f, 0.f);
      v[3] = v[4] = make_float2(0.f, 0.f);
      v[5] = cidx[5] >= 0 ? g[cidx[5]] : make_float2(0.f, 0.f);
      v[6] = g[cidx[6]];
      v[7] = g[cidx[7]];
    } else if constexpr ((FLAGS & ROW_IN_COMPACT) != 0) {
      const float2* g = a.G + b * gw + (long long)x * a.nc;
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = cidx[r] >= 0 ? g[cidx[r]] : make_float2(0.f, 0.f);
    } else {
      const float2* g = a.in_full + b * a.in_full_stride + (long long)x * N;
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = g[i + T * r];
    }
    if constexpr ((FLAGS & ROW_INV) != 0) fft_reg<N, +1, true, BAND>(v, buf, i, w, idx);
    if constexpr ((FLAGS & ROW_WRITE_PLANE) != 0) {
      float2* pl = a.plane + b * a.plane_stride + (long long)x * N;
#pragma unroll
      for (int r = 0; r < 8; ++r) pl[i + T * r] = cscale(v[r], a.out_scale);
    }
    if constexpr ((FLAGS & ROW_MUL_T) != 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = cmul(v[r], t[r]);
    }
    if constexpr ((FLAGS & ROW_FWD) != 0) fft_reg<N, -1, true>(v, buf, i, w, idx);
    if constexpr (BAND) {
      float2* g = a.G + b * gw + (long long)x * a.nc;
      g[i] = v[0];
      g[i + T] = v[1];
      // A lane whose element 2 or 5 is not kept stores a zero into one of the row's padding columns instead (lanes
      // sharing one store the same zero), so that the row is written in whole lines, padding included; without
      // padding columns it stores its element 0 or 1 again (the same value to the same address).
      const int nkept = a.ny - a.kept_shift, npad = a.nc - nkept;
      const int cpad = npad > 0 ? nkept + i % npad : -1;
      g[cidx[2] >= 0 ? cidx[2] : (npad > 0 ? cpad : i)] =
          cidx[2] >= 0 ? v[2] : (npad > 0 ? make_float2(0.f, 0.f) : v[0]);
      g[cidx[5] >= 0 ? cidx[5] : (npad > 0 ? cpad : i + T)] =
          cidx[5] >= 0 ? v[5] : (npad > 0 ? make_float2(0.f, 0.f) : v[1]);
      g[cidx[6]] = v[6];
      g[cidx[7]] = v[7];
    } else if constexpr ((FLAGS & ROW_OUT_COMPACT) != 0) {
      float2* g = a.G + b * gw + (long long)x * a.nc;
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (cidx[r] >= 0) g[cidx[r]] = v[r];
    }
  }
}

// The detectors of one wave's column tile (final pass): |Psi(k)|^2 of the lane's eight pixels (zero on a padding
// column) into the pixelated accumulators of position `pos`, and one scalar sum per slot, reduced over the block
// deterministically (butterflies within each wavefront, then the wavefronts in order) into the partial sums of wave
// `b`. Ends with a barrier, so the LDS is free again.
template <int N, int NT>
__device__ __forceinline__ void detect_block(const float2 (&v)[8], const DetPlan& d, long long pos, long long b, int iy,
                                             int i, int tid, float2* lds, int tile) {
  constexpr int T = N / 8;
  double I[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) I[r] = iy >= 0 ? ((double)v[r].x * v[r].x + (double)v[r].y * v[r].y) * d.scale : 0.0;
  if (iy >= 0) {
    for (int q = 0; q < d.npix; ++q) {
      const int oy = d.pix_mapy[q][iy];
      if (oy < 0) continue;
      double* acc = d.pix_acc[q] + pos * d.pix_stride[q] + oy;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int ox = d.pix_mapx[q][i + T * r];
        if (ox >= 0) acc[(long long)ox * d.pix_my[q]] += I[r];
      }
    }
  }
  double* red = reinterpret_cast<double*>(lds);
  const int lane = tid & 63, wave = tid >> 6;
  const double sy = (iy >= 0 && d.sqy) ? d.sqy[iy] : 0.0;
  for (int q = 0; q < d.nslots; ++q) {
    const int kind = d.slot_kind[q], arg = d.slot_arg[q];
    double val = 0.0;
    if (iy >= 0) {
      if (kind == 0) {
        const double i2 = d.slot_i2[q], o2 = d.slot_o2[q];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const double a2 = __dadd_rn(d.sqx[i + T * r], sy);  // float64, as the host rounds it
          if (a2 >= i2 && a2 < o2) val += I[r];
        }
      } else if (kind == 3) {
        const int* m = d.seg_map[arg] + iy;
        const int bin = d.slot_bin[q];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (m[(long long)(i + T * r) * d.ny] == bin) val += I[r];
      } else if (d.com_mapy[arg][iy] >= 0) {
        const double wy = d.com_wy[arg][iy];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int x = i + T * r;
          if (d.com_mapx[arg][x] >= 0) val += I[r] * (kind == 1 ? d.com_wx[arg][x] : wy);
        }
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
    d.partial[(b * d.nslots + tid) * d.blocks_per_wave + tile] = t;
  }
  __syncthreads();  // `red` is read before the next transform reuses the LDS
}

// Column pass: block = 16 compact columns (lanes c = tid mod 16, i = tid / 16) of positions [b0, b0 + PB).
template <int N, int FLAGS, int PB>
__global__ void __launch_bounds__(kColTile * (N / 8)) col_fast_kernel(ColArgs a, const float2* __restrict__ tw,
                                                                        int batch) {
  constexpr int T = N / 8, NT = kColTile * T;
  __shared__ float2 lds[N * kColTile];  // [x][column]
  const int tid = threadIdx.x, cl = tid % kColTile, i = tid / kColTile;
  const int c = blockIdx.x * kColTile + cl;  // the host launches nc / 16 blocks along x: every c is in range
  const int b0 = blockIdx.y * PB;
  const int iy = a.keep[c];
  auto idx = [cl](int j) { return j * kColTile + cl; };
  const long long gw = (long long)a.nx * a.nc;
  LaneTwiddles<N> w;
  w.load(tw, i);
  float2 pv[8];
  if constexpr ((FLAGS & COL_MUL_P) != 0) {
    const float2* P = a.P + c;
#pragma unroll
    for (int r = 0; r < 8; ++r) pv[r] = P[(long long)(i + T * r) * a.nc];
  }
  for (int p = 0; p < PB; ++p) {
    const long long b = b0 + p;
    if (b >= batch) break;  // uniform over the block
    float2 v[8];
    if constexpr ((FLAGS & COL_IN_G) != 0) {
      const float2* g = a.G + b * gw + c;
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = iy >= 0 ? g[(long long)(i + T * r) * a.nc] : make_float2(0.f, 0.f);
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int x = i + T * r;
        v[r] = iy >= 0 ? cmul(cmul(a.spec[(long long)x * a.ny + iy], a.rx[b * a.nx + x]), a.ry[b * a.ny + iy])
                       : make_float2(0.f, 0.f);
      }
    }
    if constexpr ((FLAGS & COL_FWD) != 0) fft_reg<N, -1>(v, lds, i, w, idx);
    if constexpr ((FLAGS & COL_MUL_P) != 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = cmul(v[r], pv[r]);
    }
    if constexpr ((FLAGS & COL_DETECT) != 0) {
      detect_block<N, NT>(v, a.det, first_position(a.pos0, a.pos_dev) + b, b, iy, i, tid, lds, blockIdx.x);
      continue;
    }
    if constexpr ((FLAGS & COL_INV) != 0) fft_reg<N, +1>(v, lds, i, w, idx);
    if constexpr ((FLAGS & COL_OUT) != 0) {
      float2* g = a.G + b * gw + c;
#pragma unroll
      for (int r = 0; r < 8; ++r) g[(long long)(i + T * r) * a.nc] = v[r];
    }
  }
}

// Column pass of 512-point columns, one column per wavefront. The block's TC columns move between HBM and a
// swizzled LDS tile in coalesced row segments of TC values (block phases); each wavefront then transforms its own
// column with wavefront-local exchanges inside the tile (no block barrier), and the next position's loads are in
// flight meanwhile. Three block barriers per position. With TC = 8 two blocks share a CU, so one block's barriers
// overlap the other's work.
//
// One work unit of the 512-point column pass: wave `b` of the batch on the tile of columns [TC tx, TC tx + TC).
struct ColUnit {
  int tx, b;
  bool on;            // false: nothing to do for this unit (uniform over the block)
  const float2* P;    // the wave's propagator [nx][nc]
  long long pos;      // detectors: where the wave accumulates
  int ramp;           // generated probes: the row of rx and ry
};

// col512_run: work units [u0, u1) as where(u) describes them; consecutive units of one tile and propagator share
// its P, which is (re)loaded when either changes. The twiddles are in LDS already, and every wavefront is past its
// last LDS access on entry and again on exit. A unit's description is recomputed where it is used rather than kept
// live across the transforms. ONE_TILE: every unit is on one tile with one P (loaded once, before the loop). SKIPS:
// some units may be off (otherwise every unit is taken to be on, and the next unit is always prefetched). FUSE: the
// block phases run the first and last stages of the transforms (below); otherwise they only move the tile between
// HBM and LDS, and each column wavefront runs whole transforms (fewer registers, for 1024-thread blocks).
template <int FLAGS, int TC, bool ONE_TILE, bool SKIPS, bool FUSE, class Where>
__device__ __forceinline__ void col512_run(const ColArgs& a, int u0, int u1, Where where, float2* tile, float2* ptile,
                                           const LdsTwiddles<512>& w, double* red) {
  constexpr int N = 512, T = 64, NT = TC * 64;
  constexpr bool HAS_P = (FLAGS & COL_MUL_P) != 0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cb = tid % TC, xi = tid / TC;  // block phases: column cb of the tile, rows xi + 64 r
  const long long gw = (long long)a.nx * a.nc;
  auto tile_slot = [](int x, int c) { return sk::tile_slot<TC>(x, c); };
  auto widx = [wave, tile_slot](int j) { return tile_slot(j, wave); };
  // Block-phase element r of the lane in the [x][nc] rows of the tile that starts at t0: row xi + T r, column cb.
  // Tile starts are uniform over the block, so the accesses take a scalar base plus the lane's 32-bit byte offset.
  const unsigned loff = (unsigned)(xi * a.nc + cb) * (unsigned)sizeof(float2);
  const long long rstep = (long long)T * a.nc * (long long)sizeof(float2);
  auto elem = [loff, rstep](const float2* t0, int r) {
    return reinterpret_cast<const float2*>(reinterpret_cast<const char*>(t0) + r * rstep + loff);
  };
  // the block-phase load of unit u: column cb of its tile, rows xi + 64 r. From G every column is loaded, padding
  // columns too (gk false: their values are replaced by zeros where they are used), without a branch that would
  // make the compiler's waits for these loads cover every access before them
  bool gk = false;
  auto gload = [&](int u, float2 (&g)[8]) {
    const ColUnit U = where(u);
    const int c = U.tx * TC + cb, iy = a.keep[c];
    if constexpr ((FLAGS & COL_IN_G) != 0) {
      const float2* t0 = a.G + U.b * gw + U.tx * TC;
      gk = iy >= 0;
#pragma unroll
      for (int r = 0; r < 8; ++r) g[r] = *elem(t0, r);
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int x = xi + T * r;
        g[r] = iy >= 0 ? cmul(cmul(a.spec[(long long)x * a.ny + iy], a.rx[(long long)U.ramp * a.nx + x]),
                              a.ry[(long long)U.ramp * a.ny + iy])
                       : make_float2(0.f, 0.f);
      }
    }
  };
  // P is zero on the rows x = 3T .. 5T-1 (r = 3, 4 of every lane): there |kx| >= 0.375 / sx exceeds the antialias
  // cutoff 0.3333333 / max(sx, sy), and the propagator carries that aperture. Those rows are neither loaded nor
  // multiplied, and the inverse transform starts from zeros there.
  auto pload = [&](int tx, const float2* P) {
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (r != 3 && r != 4) ptile[tile_slot(xi + T * r, cb)] = *elem(P + tx * TC, r);
  };
  float2 g[8];
  int gu = -1;  // the unit whose block-phase values g holds
  int tcur = -1;
  const float2* pcur = nullptr;
  if (u0 < u1) {
    const ColUnit U = where(u0);
    if (!SKIPS || U.on) {
      gload(u0, g);
      gu = u0;
    }
    if constexpr (HAS_P && ONE_TILE) {
      tcur = U.tx;
      pload(tcur, U.P);
    }
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // the first unit's loads, before the loop
  for (int u = u0; u < u1; ++u) {
    if constexpr (SKIPS) {
      if (!where(u).on) continue;  // uniform over the block
    }
    if constexpr (HAS_P && !ONE_TILE) {
      const ColUnit U = where(u);
      // uniform; the previous unit's LDS reads ended before its last barrier. P changes between units only where
      // units belong to different lanes (SKIPS); elsewhere it is the launch's one propagator.
      if (U.tx != tcur || (SKIPS && U.P != pcur)) {
        tcur = U.tx;
        pcur = U.P;
        pload(U.tx, U.P);
      }
    }
    if constexpr (SKIPS) {
      if (gu != u) gload(u, g);
    }
    // The first stage of the line's transform combines the eight elements i + 64 r of a column, which a block-phase
    // thread holds for i = xi, column cb: the block phases run it from their loaded values and write its outputs to
    // the tile (then a block barrier); the column wavefronts run the middle stages with wavefront-local exchanges;
    // the block phases run the last stage of the inverse from the tile into the values they store.
    auto none = [](int, int) { return make_float2(1.f, 0.f); };
    auto f2 = [&w](int q, int j) { return w.s2(q, j); };
    auto f3 = [&w](int q, int j) { return w.s3(q, j); };
    auto bidx = [cb, tile_slot](int j) { return tile_slot(j, cb); };
    if constexpr ((FLAGS & COL_IN_G) != 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) g[r] = gk ? g[r] : make_float2(0.f, 0.f);
    }
    if constexpr (!FUSE) {
#pragma unroll
      for (int r = 0; r < 8; ++r) tile[tile_slot(xi + T * r, cb)] = g[r];
      __syncthreads();
    } else if constexpr ((FLAGS & COL_FWD) != 0) {
      stage<N, 8, 1, -1, true, false, false>(g, tile, xi, none, bidx);
    } else {
      stage<N, 8, 1, +1, true, false, false>(g, tile, xi, none, bidx);
    }
    if (u + 1 < u1 && (!SKIPS || where(u + 1).on)) {  // in flight during the transforms
      gload(u + 1, g);
      gu = u + 1;
    }
    float2 v[8];
    if constexpr (!FUSE) {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = tile[tile_slot(lane + T * r, wave)];
      if constexpr ((FLAGS & COL_FWD) != 0) fft_reg<N, -1, true>(v, tile, lane, w, widx);
    } else if constexpr ((FLAGS & COL_FWD) != 0) {
      stage<N, 8, 8, -1, false, false, true>(v, tile, lane, f2, widx);
      stage<N, 8, 64, -1, false, true, true>(v, tile, lane, f3, widx);
    }
    if constexpr (HAS_P) {
#pragma unroll
      for (int r = 0; r < 8; ++r)
        v[r] = (r == 3 || r == 4) ? make_float2(0.f, 0.f) : cmul(v[r], ptile[tile_slot(lane + T * r, wave)]);
    }
    if constexpr ((FLAGS & COL_DETECT) != 0) {
      const ColUnit U = where(u);
      detect_block<N, NT>(v, a.det, U.pos, U.b, a.keep[U.tx * TC + wave], lane, tid, reinterpret_cast<float2*>(red),
                          U.tx);
      continue;  // detect_block ends with a barrier: the tile is free again
    }
    if constexpr (!FUSE) {
      if constexpr ((FLAGS & COL_INV) != 0) fft_reg<N, +1, true, HAS_P>(v, tile, lane, w, widx);
#pragma unroll
      for (int r = 0; r < 8; ++r) tile[tile_slot(lane + T * r, wave)] = v[r];
      __syncthreads();
      const ColUnit U = where(u);
      const float2* t0 = a.G + U.b * gw + U.tx * TC;
      // every column is stored, as in the FUSE variant below (padding columns get zeros)
#pragma unroll
      for (int r = 0; r < 8; ++r) *const_cast<float2*>(elem(t0, r)) = tile[tile_slot(xi + T * r, cb)];
      __syncthreads();
      continue;
    }
    // the inverse: its first stage from the registers (after a forward transform), its second in the wavefront
    if constexpr ((FLAGS & COL_FWD) != 0) stage<N, 8, 1, +1, true, false, true, HAS_P>(v, tile, lane, none, widx);
    stage<N, 8, 8, +1, false, false, true>(v, tile, lane, f2, widx);
    __syncthreads();
    {
      // last stage in the block phases; its barrier between the reads and the next unit's writes frees the tile
      const LdsTwiddles<N> wb{w.t2, w.t3, xi};
      auto f3b = [&wb](int q, int j) { return wb.s3(q, j); };
      float2 o[8];
      stage<N, 8, 64, +1, false, true, false>(o, tile, xi, f3b, bidx);
      const ColUnit U = where(u);
      const float2* t0 = a.G + U.b * gw + U.tx * TC;
      // every column is stored: padding columns get zeros (zero inputs, and P is zero there), which no pass reads;
      // a store only kept columns take would put a branch here and bring back a wait for these stores at the head
#pragma unroll
      for (int r = 0; r < 8; ++r) *const_cast<float2*>(elem(t0, r)) = o[r];
    }
  }
}

// 2-D grid (16-column tiles): block (bx, by) runs tile bx over positions [PB by, PB by + PB).
template <int FLAGS, int PB, int TC>
__global__ void __launch_bounds__(TC * 64) col512_kernel(ColArgs a, const float2* __restrict__ tw, int batch) {
  constexpr int N = 512, NT = TC * 64;
  constexpr bool HAS_P = (FLAGS & COL_MUL_P) != 0;
  __shared__ float2 tile[N * TC];
  __shared__ float2 ptile[HAS_P ? N * TC : 1];  // the block's columns of P (swizzled as the tile)
  __shared__ float2 twt[LdsTwiddles<N>::kSize];
  __shared__ double red[(NT / 64) * kMaxSlots];
  const int bx = blockIdx.x, by = blockIdx.y;  // the host launches nc / TC blocks along x: every column is in range
  const int b0 = by * PB;
  const int b1 = b0 + PB < batch ? b0 + PB : batch;
  LdsTwiddles<N>::build(twt, tw, threadIdx.x, NT);
  const LdsTwiddles<N> w{twt, twt + LdsTwiddles<N>::R2 * LdsTwiddles<N>::L2, (int)(threadIdx.x & 63)};
  const long long first = first_position(a.pos0, a.pos_dev);
  col512_run<FLAGS, TC, true, false, false>(
      a, b0, b1, [&a, bx, first](int u) { return ColUnit{bx, u, true, a.P, first + u, u}; }, tile,
      ptile, w, red);
}

// Balanced 1-D grid of 8-column tiles, sized by the host to one round of resident blocks. The work is the list of
// (tile pair m, position b) units in the order m * batch + b, cut into `ranges` contiguous spans of equal length (to
// one unit); the blocks 2 r and 2 r + 1 run span r on tiles 2 m and 2 m + 1, whose 64-byte halves make up the same
// 128-byte lines. Renumbered XCD-major (work-groups are dealt round-robin over the 8 XCDs), the two blocks of a span
// run side by side on one XCD and share its L2. A span that crosses a pair boundary reloads P for its next tile.
template <int FLAGS>
__global__ void __launch_bounds__(8 * 64, 4)
    col512_span_kernel(ColArgs a, const float2* __restrict__ tw, int batch, int ranges) {
  constexpr int N = 512, TC = 8, NT = TC * 64;
  constexpr bool HAS_P = (FLAGS & COL_MUL_P) != 0;
  __shared__ float2 tile[N * TC];
  __shared__ float2 ptile[HAS_P ? N * TC : 1];
  __shared__ float2 twt[LdsTwiddles<N>::kSize];
  __shared__ double red[(NT / 64) * kMaxSlots];
  const int G = gridDim.x;  // 2 * ranges
  int j = blockIdx.x;
  if (G % 8 == 0) j = (j % 8) * (G / 8) + j / 8;
  const int h = j & 1, r = j >> 1;
  const long long U = (long long)(a.nc / (2 * TC)) * batch;  // below 2^31: the host checks
  const int u0 = (int)(U * r / ranges), u1 = (int)(U * (r + 1) / ranges);
  LdsTwiddles<N>::build(twt, tw, threadIdx.x, NT);
  const LdsTwiddles<N> w{twt, twt + LdsTwiddles<N>::R2 * LdsTwiddles<N>::L2, (int)(threadIdx.x & 63)};
  const long long first = first_position(a.pos0, a.pos_dev);
  col512_run<FLAGS, TC, false, false, true>(
      a, u0, u1,
      [&a, batch, h, first](int u) {
        const int m = u / batch, b = u - m * batch;
        return ColUnit{2 * m + h, b, true, a.P, first + b, b};
      },
      tile, ptile, w, red);
}

// Lane pipeline, 512-point columns: the balanced spans of col512_span_kernel over the pipeline's whole batch, each
// unit taking the descriptor of its lane for the current step. FLAGS selects the lanes served: COL_MID the lanes in
// a mid slice, COL_FINAL those finishing an item (FINAL, FINAL_INIT), COL_INIT those starting one (INIT, FINAL_INIT:
// launched after COL_FINAL, into the same storage). Units of other lanes, and waves past a lane's count, are skipped.
template <int FLAGS>
__global__ void __launch_bounds__(8 * 64, (FLAGS & COL_DETECT) != 0 ? 1 : 4)
    col512_lanes_kernel(ColArgs a, const float2* __restrict__ tw, int batch, int ranges) {
  constexpr int N = 512, TC = 8, NT = TC * 64;
  constexpr bool HAS_P = (FLAGS & COL_MUL_P) != 0;
  __shared__ float2 tile[N * TC];
  __shared__ float2 ptile[HAS_P ? N * TC : 1];
  __shared__ float2 twt[LdsTwiddles<N>::kSize];
  __shared__ double red[(NT / 64) * kMaxSlots];
  const int G = gridDim.x;  // 2 * ranges
  int j = blockIdx.x;
  if (G % 8 == 0) j = (j % 8) * (G / 8) + j / 8;
  const int h = j & 1, r = j >> 1;
  const long long U = (long long)(a.nc / (2 * TC)) * batch;  // below 2^31: the host checks
  const int u0 = (int)(U * r / ranges), u1 = (int)(U * (r + 1) / ranges);
  LdsTwiddles<N>::build(twt, tw, threadIdx.x, NT);
  const LdsTwiddles<N> w{twt, twt + LdsTwiddles<N>::R2 * LdsTwiddles<N>::L2, (int)(threadIdx.x & 63)};
  const LaneDesc* descs = a.pipe.descs + (long long)pipe_step(a.pipe) * a.pipe.lanes;
  const int W = a.pipe.lane_waves;
  col512_run<FLAGS, TC, false, true, true>(
      a, u0, u1,
      [descs, batch, h, W](int u) {
        const int m = u / batch, b = u - m * batch, ln = b / W, local = b - ln * W;
        const LaneDesc& d = descs[ln];
        ColUnit c{2 * m + h, b, false, d.P, (long long)d.pos + d.acc_off + local, d.next_pos + local};
        if constexpr ((FLAGS & COL_DETECT) != 0)
          c.on = (d.mode == LANE_FINAL || d.mode == LANE_FINAL_INIT) && local < d.count;
        else if constexpr ((FLAGS & COL_GENERATE) != 0)
          c.on = (d.mode == LANE_INIT || d.mode == LANE_FINAL_INIT) && local < d.next_count;
        else
          c.on = d.mode == LANE_MID && local < d.count;
        return c;
      },
      tile, ptile, w, red);
}

constexpr int kPipePB = 8;  // positions per block in pipelined launches; lane_waves is a multiple of it

// Pipelined column pass of the 256-point grids (512 uses col512_lanes_kernel): the lane of the block's positions gives
// the mode of this step —
//   MID:        FFT_x, * P, IFFT_x                       (a slice inside the item)
//   FINAL:      FFT_x, * P, detectors                    (the item's last slice; the lane then idles)
//   FINAL_INIT: FFT_x, * P, detectors, then the probe of the lane's next item generated into the same storage
//   INIT:       the probe of the lane's first item
template <int N, int TC>
__global__ void __launch_bounds__(TC * (N / 8)) col_pipe_kernel(ColArgs a, const float2* __restrict__ tw) {
  constexpr int T = N / 8, NT = TC * T;
  __shared__ float2 lds[N * TC];  // [x][column]
  const int tid = threadIdx.x, cl = tid % TC, i = tid / TC;
  const int c = blockIdx.x * TC + cl;
  const int b0 = blockIdx.y * kPipePB;
  const int lane = b0 / a.pipe.lane_waves, local0 = b0 - lane * a.pipe.lane_waves;
  const LaneDesc d = a.pipe.descs[(long long)pipe_step(a.pipe) * a.pipe.lanes + lane];
  if (d.mode == LANE_IDLE) return;  // uniform over the block
  const int iy = a.keep[c];
  auto idx = [cl](int j) { return j * TC + cl; };
  const long long gw = (long long)a.nx * a.nc;
  LaneTwiddles<N> w;
  w.load(tw, i);
  float2 pv[8];
  if (d.mode != LANE_INIT) {
    const float2* P = d.P + c;
#pragma unroll
    for (int r = 0; r < 8; ++r) pv[r] = P[(long long)(i + T * r) * a.nc];
  }
  for (int p = 0; p < kPipePB; ++p) {
    const int local = local0 + p;
    const long long b = b0 + p;
    float2* g = a.G + b * gw + c;
    float2 v[8];
    if (d.mode != LANE_INIT && local < d.count) {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = iy >= 0 ? g[(long long)(i + T * r) * a.nc] : make_float2(0.f, 0.f);
      fft_reg<N, -1>(v, lds, i, w, idx);
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = cmul(v[r], pv[r]);
      if (d.mode == LANE_MID) {
        fft_reg<N, +1>(v, lds, i, w, idx);
#pragma unroll
        for (int r = 0; r < 8; ++r) g[(long long)(i + T * r) * a.nc] = v[r];
        continue;
      }
      detect_block<N, NT>(v, a.det, (long long)d.pos + d.acc_off + local, b, iy, i, tid, lds, blockIdx.x);
    }
    if ((d.mode == LANE_FINAL_INIT || d.mode == LANE_INIT) && local < d.next_count) {
      const long long q = (long long)d.next_pos + local;
      const float2* rx = a.pipe.rx_all + q * a.nx;
      const float2 ryv = iy >= 0 ? a.pipe.ry_all[q * a.ny + iy] : make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int x = i + T * r;
        v[r] = iy >= 0 ? cmul(cmul(a.spec[(long long)x * a.ny + iy], rx[x]), ryv) : make_float2(0.f, 0.f);
      }
      fft_reg<N, +1>(v, lds, i, w, idx);
#pragma unroll
      for (int r = 0; r < 8; ++r) g[(long long)(i + T * r) * a.nc] = v[r];
    }
  }
}

// Sum the partials of every wave whose lane finished an item this step into its position's accumulators.
__global__ void reduce_pipe_kernel(DetPlan d, PipeArgs pa, double* acc) {
  const long long b = blockIdx.x;
  const int lane = int(b / pa.lane_waves), local = int(b - (long long)lane * pa.lane_waves);
  const LaneDesc ld = pa.descs[(long long)pipe_step(pa) * pa.lanes + lane];
  if (!(ld.mode == LANE_FINAL || ld.mode == LANE_FINAL_INIT) || local >= ld.count) return;
  const long long pos = (long long)ld.pos + ld.acc_off + local;
  for (int q = threadIdx.x; q < d.nslots; q += blockDim.x) {
    const double* p = d.partial + (b * d.nslots + q) * d.blocks_per_wave;
    double s = 0.0;
    for (int k = 0; k < d.blocks_per_wave; ++k) s += p[k];
    acc[pos * d.nslots + q] += s;
  }
}

// The lanes' accumulator copies into the scan's accumulators, in the fixed order of the copies.
__global__ void sum_copies_kernel(const double* __restrict__ in, long long n, int copies, double* __restrict__ out) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = out[e];
  for (int c = 0; c < copies; ++c) s += in[c * n + e];
  out[e] = s;
}

__global__ void step_increment_kernel(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

bool is_fast_length(int n) { return n == 256 || n == 512; }

// A launch timed by the events of its own dispatch when the caller asks for timing (e0 not null): the events then
// bracket the kernel alone, not the stream around it.
template <class K, class... A>
void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, A... args) {
  if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

constexpr int kRowPB = 8;  // positions per row-pass block
constexpr int kColPB = 8;  // positions per column-pass block (P and the twiddles loaded once per 8 positions)

// The kept set has the shape the BAND row kernels assume (a 2/3 bandlimit of an N-point line, T = N / 8).
bool band_shape(const RowArgs& a, int N) {
  const int T = N / 8;
  return a.kept_hi >= 2 * T - 1 && a.kept_hi < 3 * T && a.kept_lo >= 5 * T && a.kept_lo <= 6 * T;
}

template <int N, int F>
void row_go(const RowArgs& a, int batch, hipStream_t s, const float2* tw) {
  constexpr int ROWS = 64 / (N / 8);
  dim3 grid(a.nx / ROWS, (batch + kRowPB - 1) / kRowPB);
  if constexpr (F == ROW_MID) {
    if (band_shape(a, N)) {
      launch_timed(row_fast_kernel<N, F, kRowPB, false, true>, grid, dim3(64), s, a.ev_start, a.ev_stop, a, tw,
                   batch);
      return;
    }
  }
  launch_timed(row_fast_kernel<N, F, kRowPB>, grid, dim3(64), s, a.ev_start, a.ev_stop, a, tw, batch);
}

template <int N>
void row_launch(const RowArgs& a, int batch, hipStream_t s, const float2* tw) {
  switch (a.flags) {
    case ROW_MID: row_go<N, ROW_MID>(a, batch, s, tw); break;
    case ROW_MID_PLANE: row_go<N, ROW_MID_PLANE>(a, batch, s, tw); break;
    case ROW_ENTRY: row_go<N, ROW_ENTRY>(a, batch, s, tw); break;
    case ROW_FROM_FULL: row_go<N, ROW_FROM_FULL>(a, batch, s, tw); break;
    default: row_go<N, ROW_EXIT>(a, batch, s, tw); break;
  }
  SK_HIP(hipGetLastError());
}

// Spans of the balanced 8-column pass of mode F over `units` (tile pair, position) units: one round of resident
// blocks, two per span, and a multiple of 8 spans where there are enough units, so that the XCD-major renumbering
// keeps both blocks of every span on one XCD.
template <int F, bool LANES = false>
int span_count(long long units) {
  // blocks of this kernel the device holds at once; a static initialised once, under the language's guarantee, since
  // a graph-replayed scan launches its first batch on one host thread while a second records the same launches
  static const int resident = [] {
    int dev = 0, cus = 0, occ = 0;
    SK_HIP(hipGetDevice(&dev));
    SK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if constexpr (LANES)
      SK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, col512_lanes_kernel<F>, 8 * 64, 0));
    else
      SK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, col512_span_kernel<F>, 8 * 64, 0));
    return std::max(2, cus * std::max(1, occ));
  }();
  long long r = std::min<long long>(units, resident / 2);
  if (r >= 8) r -= r % 8;
  return (int)std::max<long long>(1, r);
}

template <int N, int F>
void col_go(const ColArgs& a, int batch, hipStream_t s, const float2* tw) {
  const int by = (batch + kColPB - 1) / kColPB;
  if constexpr (N == 512) {
    // the detecting pass always runs 8-column tiles: its reductions need more than the 128 VGPRs of a 1024-thread
    // block, and 512-thread blocks give it 256 without spilling
    const int tc = F == COL_FINAL ? 8 : fast_col_tile(512, a.col_tile, double(batch) * a.nx * a.nc * sizeof(float2));
    if (tc == 8) {
      const long long units = (long long)(a.nc / 16) * batch;
      if (units >= (1LL << 31)) fail(SK_ERR_INVALID, "column pass: too many positions in one launch");
      const int ranges = span_count<F>(units);
      launch_timed(col512_span_kernel<F>, dim3(2 * ranges), dim3(8 * 64), s, a.ev_start, a.ev_stop, a, tw, batch,
                   ranges);
    } else if constexpr (F != COL_FINAL) {
      launch_timed(col512_kernel<F, kColPB, 16>, dim3(a.nc / 16, by), dim3(16 * 64), s, a.ev_start, a.ev_stop, a, tw,
                   batch);
    }
  } else {
    launch_timed(col_fast_kernel<N, F, kColPB>, dim3(a.nc / kColTile, by), dim3(kColTile * (N / 8)), s, a.ev_start,
                 a.ev_stop, a, tw, batch);
  }
}

template <int N>
void col_launch(const ColArgs& a, int batch, hipStream_t s, const float2* tw) {
  switch (a.flags) {
    case COL_MID: col_go<N, COL_MID>(a, batch, s, tw); break;
    case COL_INIT: col_go<N, COL_INIT>(a, batch, s, tw); break;
    default: col_go<N, COL_FINAL>(a, batch, s, tw); break;
  }
  SK_HIP(hipGetLastError());
}

}  // namespace

bool fast_row_supported(int nx, int ny, int flags) {
  if (!is_fast_length(ny)) return false;
  const int rows = 64 / (ny / 8);
  if (nx % rows != 0) return false;
  return flags == ROW_MID || flags == ROW_MID_PLANE || flags == ROW_ENTRY || flags == ROW_FROM_FULL ||
         flags == ROW_EXIT;
}

bool fast_col_supported(int nx, int nc, int flags) {
  if (!is_fast_length(nx) || nc % kColTile != 0) return false;
  return flags == COL_MID || flags == COL_INIT || flags == COL_FINAL;
}

int fast_col_tile(int nx, int requested, double wave_bytes) {
  if (nx != 512) return kColTile;
  if (requested == 8 || requested == 16) return requested;
  if (requested != 0) fail(SK_ERR_INVALID, "col_tile 4 is not a column-pass width for 512-point lines (8 or 16)");
  // 8-column spans when the batch stays in the last-level cache (they fill every CU for any batch); 16-column
  // blocks when it streams from HBM, where their full 128-byte row segments measured faster
  return wave_bytes <= (double)kCachedWaveBytes ? 8 : 16;  // the 8-column spans also fill every CU at any batch
}

int fast_col_blocks_per_wave(int nx, int nc, int pipeline) {
  (void)pipeline;  // at 512 both the detecting pass and the lane pipeline's column pass use 8-column tiles
  return nx == 512 ? nc / 8 : nc / kColTile;
}

void launch_row_fast(const RowArgs& a, int batch, hipStream_t s) {
  const float2* tw = line_plan(a.ny, s).dev.roots;
  if (a.ny == 512) row_launch<512>(a, batch, s, tw);
  else row_launch<256>(a, batch, s, tw);
}

void launch_col_fast(const ColArgs& a, int batch, hipStream_t s) {
  const float2* tw = line_plan(a.nx, s).dev.roots;
  if (a.nx == 512) col_launch<512>(a, batch, s, tw);
  else col_launch<256>(a, batch, s, tw);
}

// ---- lane pipeline launches --------------------------------------------------------------------------------------
bool pipeline_supported(int nx, int ny, int nc, int lane_waves) {
  return is_fast_length(nx) && is_fast_length(ny) && nc % kColTile == 0 && nx % (64 / (ny / 8)) == 0 &&
         lane_waves > 0 && lane_waves % kPipePB == 0;
}

int pipe_positions_per_block() { return kPipePB; }

void launch_row_pipe(const RowArgs& a, int batch, hipStream_t s) {
  const float2* tw = line_plan(a.ny, s).dev.roots;
  constexpr int F = ROW_MID;
  const bool band = band_shape(a, a.ny);
  if (a.ny == 512) {
    dim3 grid(a.nx / 1, batch / kPipePB);
    if (band) hipLaunchKernelGGL((row_fast_kernel<512, F, kPipePB, true, true>), grid, dim3(64), 0, s, a, tw, batch);
    else hipLaunchKernelGGL((row_fast_kernel<512, F, kPipePB, true>), grid, dim3(64), 0, s, a, tw, batch);
  } else {
    dim3 grid(a.nx / 2, batch / kPipePB);
    if (band) hipLaunchKernelGGL((row_fast_kernel<256, F, kPipePB, true, true>), grid, dim3(64), 0, s, a, tw, batch);
    else hipLaunchKernelGGL((row_fast_kernel<256, F, kPipePB, true>), grid, dim3(64), 0, s, a, tw, batch);
  }
  SK_HIP(hipGetLastError());
}

void launch_col_pipe(const ColArgs& a, int batch, hipStream_t s, int modes) {
  if (modes == 0) return;
  const float2* tw = line_plan(a.nx, s).dev.roots;
  if (a.nx == 512) {
    // balanced spans, one launch per mode: mid slices, items finishing (detectors), items starting (probes); the
    // probes of FINAL_INIT lanes overwrite the waves their detectors have just read
    const long long units = (long long)(a.nc / 16) * batch;
    if (units >= (1LL << 31)) fail(SK_ERR_INVALID, "column pass: too many positions in one launch");
    if (modes & 1) {
      const int ranges = span_count<COL_MID, true>(units);
      hipLaunchKernelGGL((col512_lanes_kernel<COL_MID>), dim3(2 * ranges), dim3(8 * 64), 0, s, a, tw, batch, ranges);
    }
    if (modes & 2) {
      const int ranges = span_count<COL_FINAL, true>(units);
      hipLaunchKernelGGL((col512_lanes_kernel<COL_FINAL>), dim3(2 * ranges), dim3(8 * 64), 0, s, a, tw, batch,
                         ranges);
    }
    if (modes & 4) {
      ColArgs g = a;  // generated probes: the ramps of every position, indexed by position
      g.rx = a.pipe.rx_all;
      g.ry = a.pipe.ry_all;
      const int ranges = span_count<COL_INIT, true>(units);
      hipLaunchKernelGGL((col512_lanes_kernel<COL_INIT>), dim3(2 * ranges), dim3(8 * 64), 0, s, g, tw, batch, ranges);
    }
  } else {
    hipLaunchKernelGGL((col_pipe_kernel<256, kColTile>), dim3(a.nc / kColTile, batch / kPipePB), dim3(kColTile * 32),
                       0, s, a, tw);
  }
  SK_HIP(hipGetLastError());
}

void launch_reduce_pipe(const DetPlan& d, const PipeArgs& p, int batch, double* acc, hipStream_t s) {
  if (d.nslots == 0) return;
  hipLaunchKernelGGL(reduce_pipe_kernel, dim3(batch), dim3(64), 0, s, d, p, acc);
  SK_HIP(hipGetLastError());
}

void launch_sum_copies(const double* in, long long n, int copies, double* out, hipStream_t s) {
  if (n <= 0 || copies <= 0) return;
  hipLaunchKernelGGL(sum_copies_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n, copies, out);
  SK_HIP(hipGetLastError());
}

void launch_step_increment(int* step, hipStream_t s) {
  hipLaunchKernelGGL(step_increment_kernel, dim3(1), dim3(64), 0, s, step);
  SK_HIP(hipGetLastError());
}

}  // namespace sk
