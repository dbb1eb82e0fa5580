// sk_passes.hip — the generic row and column passes of the batched multislice, for any grid length.
//
// One slice of the reference's loop is psi <- IFFT2(P * FFT2(t * psi)) (Multislice.swift:141-155). With the 2-D
// transform split into a transform along y (rows, contiguous) and one along x (columns), a slice becomes two
// passes over the batch of waves, each reading and writing every wave once:
//   row pass:    inverse along y (previous slice) -> * t_j -> forward along y
//   column pass: forward along x -> * P -> inverse along x
// The waves live between passes in "compact" form G[b][x][c]: only the columns c whose ky lies inside the antialias
// disk (Antialias.swift:18-39) are stored, since P is zero on every other column and they carry nothing into the
// next slice. The generic passes below stage each line in LDS and transform it there (Stockham autosort, radix
// 2/3/4/5/7/8/11/13, Bluestein's algorithm for other lengths); the fast passes (sk_fast.hip) keep lines in
// registers for the lengths that have them.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <complex>
#include <map>

#include "sk_dft.h"
#include "sk_internal.h"
#include "sk_mfma.h"

namespace sk {

// ---------------------------------------------------------------------------------------------------------------
// Quotient and remainder by a divisor the compiler does not know: n / d and n mod d for 0 <= n < 65536 and
// 0 < d <= 16384, from the float reciprocal of d. The truncated product n * (1 / d) lies within one of the quotient
// (relative error below 2^-22), and one correction step either way makes it exact; an exhaustive check over that
// range, also with 1 / d off by one ulp either way, finds no error. About 6 VALU instructions, against some 30 for
// a 32-bit integer division.
// ---------------------------------------------------------------------------------------------------------------
struct Div {
  int d;
  float inv;
  __device__ explicit Div(int dd) : d(dd), inv(1.0f / float(dd)) {}
  __device__ __forceinline__ int quot(int n, int& r) const {
    int q = int(float(n) * inv);
    r = n - q * d;
    if (r < 0) {
      --q;
      r += d;
    } else if (r >= d) {
      ++q;
      r -= d;
    }
    return q;
  }
};

// ---------------------------------------------------------------------------------------------------------------
// Small DFTs (sign -1 forward, +1 inverse, unscaled)
// ---------------------------------------------------------------------------------------------------------------
template <int R>
__device__ inline void dft_small(float2 (&u)[R], int sign, const float2* roots, int m) {
  if constexpr (R == 2 || R == 4 || R == 8 || R == 16) {
    dft_pow2<R>(u, sign);
  } else if constexpr (R == 6) {
    dft6(u, sign);
  } else {
    // direct DFT of an odd prime radix with the line's roots: e^{-2 pi i j q / R} = roots[(j q mod R) m / R]
    float2 v[R];
    const int step = m / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = u[0];
#pragma unroll
      for (int j = 1; j < R; ++j) {
        float2 w = roots[((j * q) % R) * step];
        if (sign > 0) w.y = -w.y;
        acc = cadd(acc, cmul(u[j], w));
      }
      v[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) u[q] = v[q];
  }
}

constexpr int kGenericThreads = 256;
constexpr int kGenericLdsComplex = 2048;  // unpadded complex values per LDS buffer (16 KiB); two buffers

// LDS layout of the float32 line transforms: entry q = line * ls + k of the unpadded lines lives at slot
// q + (q / 16), one pad slot per 16 entries. The Stockham stages read and write at power-of-two strides (the first
// stage of radix 8 writes at stride 8, the column loads step by the line stride); the pad spreads those strides
// over the LDS banks (a bank model of the 384-point stages: 2.3 -> 1.3 times the conflict-free time). The
// split-BF16 stages keep their lines unpadded.
__device__ __forceinline__ int pad16(int q) { return q + (q >> 4); }
template <int MF>
__device__ __forceinline__ int at(int q) {
  if constexpr (MF == 0) return pad16(q);
  else return q;
}
// entries of one LDS buffer holding `q` unpadded entries
__host__ __device__ inline int lds_entries(int q, bool padded) { return padded ? q + (q >> 4) : q; }

// One Stockham pass of radix R over `lines` lines of length m (line stride ls): for butterfly i in [0, m/R),
// k = i mod L, u_j = x[i + j m/R] * w_{L R}^{j k}, DFT_R(u), y[(i - k) R + k + j L] = u_j.
template <int R>
__device__ void stockham_stage(const float2* __restrict__ src, float2* __restrict__ dst, int lines, int m, int ls,
                               int L, const float2* __restrict__ roots, int sign, int tid, int T) {
  const int stride = m / R;
  const int step = m / (L * R);
  const int total = lines * stride;
  const Div dstride(stride), dL(L);
  for (int w = tid; w < total; w += T) {
    int i, k;
    const int line = dstride.quot(w, i);
    const int lb = line * ls;
    (void)dL.quot(i, k);
    float2 u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      float2 v = src[pad16(lb + i + j * stride)];
      if (j > 0 && k > 0) {
        float2 tw = roots[j * k * step];  // j k step < R L step = m
        if (sign > 0) tw.y = -tw.y;
        v = cmul(v, tw);
      }
      u[j] = v;
    }
    dft_small<R>(u, sign, roots, m);
    const int base = lb + (i - k) * R + k;
#pragma unroll
    for (int j = 0; j < R; ++j) dst[pad16(base + j * L)] = u[j];
  }
}

// Unscaled Stockham transform of length p.m of every line; the data starts in `a` (synchronised); returns the
// buffer that holds the result.
__device__ float2* lds_stockham(float2* a, float2* b, int lines, int ls, const LineFFT& p, int sign, int tid, int T) {
  int L = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int r = p.radix[s];
    switch (r) {
      case 2: stockham_stage<2>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 3: stockham_stage<3>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 4: stockham_stage<4>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 5: stockham_stage<5>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 7: stockham_stage<7>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 8: stockham_stage<8>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      case 11: stockham_stage<11>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
      default: stockham_stage<13>(a, b, lines, p.m, ls, L, p.roots, sign, tid, T); break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    L *= r;
  }
  return a;
}

// The same pass with the radix, sub-length, length and line count fixed at compile time (line stride M), so every
// index folds to shifts and constant multiplies.
// Pointwise work fused into a transform's first reads or last writes: hook(e, v) for element e of the line.
struct NoHook {
  __device__ float2 operator()(int, float2 v) const { return v; }
};

template <int R, int L, int M, int LINES, class Pre = NoHook, class Post = NoHook>
__device__ __forceinline__ void stockham_stage_fixed(const float2* __restrict__ src, float2* __restrict__ dst,
                                                     const float2* __restrict__ roots, int sign, int tid,
                                                     Pre pre = {}, Post post = {}) {
  constexpr int stride = M / R, step = M / (L * R), total = LINES * stride;
  for (int w = tid; w < total; w += kGenericThreads) {
    const int line = w / stride, i = w - line * stride;
    const int lb = line * M;
    const int k = i % L;
    float2 u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      float2 v = pre(i + j * stride, src[pad16(lb + i + j * stride)]);
      if constexpr (L > 1) {
        if (j > 0) {
          float2 tw = roots[j * k * step];  // roots[0] = 1 where k = 0
          if (sign > 0) tw.y = -tw.y;
          v = cmul(v, tw);
        }
      }
      u[j] = v;
    }
    dft_small<R>(u, sign, roots, M);
    const int e0 = (i - k) * R + k;
#pragma unroll
    for (int j = 0; j < R; ++j) dst[pad16(lb + e0 + j * L)] = post(e0 + j * L, u[j]);
  }
}

// Compile-time stage lists of the fixed lengths: 384 = 8 8 6 (a grid length of the precision study; the radix-6
// stage is the prime-factor 2 x 3 DFT, one LDS pass where radix 2 then 3 take two); the Bluestein lengths of the
// grid lengths 65..1024, 256 = 8 8 4, 512 = 8 8 8, 1024 = 8 8 16 (509 among them; dft16 as 4 x 4) and
// 2048 = 8 8 8 4, which are also the grid lengths 256, 512 and 1024 where the generic passes run them. Any order of
// radices uses the same root table of the length.
// `pre` applies to the first stage's reads, `post` to the last stage's writes.
template <int M, int LINES, class Pre = NoHook, class Post = NoHook>
__device__ float2* lds_stockham_fixed(float2* a, float2* b, const float2* roots, int sign, int tid, Pre pre = {},
                                      Post post = {}) {
  if constexpr (M == 384 || M == 256 || M == 512) {
    constexpr int R3 = M / 64;  // 6, 4 or 8
    stockham_stage_fixed<8, 1, M, LINES>(a, b, roots, sign, tid, pre);
    __syncthreads();
    stockham_stage_fixed<8, 8, M, LINES>(b, a, roots, sign, tid);
    __syncthreads();
    stockham_stage_fixed<R3, 64, M, LINES>(a, b, roots, sign, tid, NoHook{}, post);
    __syncthreads();
    return b;
  } else if constexpr (M == 2048) {
    stockham_stage_fixed<8, 1, M, LINES>(a, b, roots, sign, tid, pre);
    __syncthreads();
    stockham_stage_fixed<8, 8, M, LINES>(b, a, roots, sign, tid);
    __syncthreads();
    stockham_stage_fixed<8, 64, M, LINES>(a, b, roots, sign, tid);
    __syncthreads();
    stockham_stage_fixed<4, 512, M, LINES>(b, a, roots, sign, tid, NoHook{}, post);
    __syncthreads();
    return a;
  } else {
    static_assert(M == 1024, "fixed lengths 256, 384, 512, 1024 and 2048");
    stockham_stage_fixed<8, 1, M, LINES>(a, b, roots, sign, tid, pre);
    __syncthreads();
    stockham_stage_fixed<8, 8, M, LINES>(b, a, roots, sign, tid);
    __syncthreads();
    stockham_stage_fixed<16, 64, M, LINES>(a, b, roots, sign, tid, NoHook{}, post);
    __syncthreads();
    return b;
  }
}

// Unscaled transform of length p.m of every line: MF = 0, the float32 Stockham stages (FM > 0: of the length FM and
// the FL lines fixed at compile time); MF = n2 > 0, the split-BF16 matrix-core stages of length p.m = 16 n2 (the
// plan carries their tables). Returns the buffer holding the result.
template <int MF, int FM, int FL>
__device__ __forceinline__ float2* lds_length_m(float2* a, float2* b, int lines, int ls, const LineFFT& p, int sign,
                                                int tid, int T) {
  if constexpr (MF == 0 && FM > 0) {
    return lds_stockham_fixed<FM, FL>(a, b, p.roots, sign, tid);
  } else if constexpr (MF == 0) {
    return lds_stockham(a, b, lines, ls, p, sign, tid, T);
  } else {
    return mfma_lines<MF>(a, b, lines, ls, static_cast<const bf16x8*>(p.mtab16), static_cast<const bf16x8*>(p.mtabn2),
                          p.mtw, sign, tid);
  }
}

// Unscaled transform of logical length p.n of every line (entries [0, n) of each line of stride ls).
// Bluestein (1970): X_k = c_k sum_j (x_j c_j) conj(c_{k-j}), c_k = e^{sign pi i k^2 / n}, the convolution done with
// two power-of-two transforms of length m.
template <int MF, int FM, int FL>
__device__ __forceinline__ float2* lds_line_transform(float2* a, float2* b, int lines, int ls, const LineFFT& p,
                                                      int sign, int tid, int T) {
  if (!p.blue) return lds_length_m<MF, FM, FL>(a, b, lines, ls, p, sign, tid, T);
  if constexpr (MF == 0 && FM > 0) {
    // fixed-length Bluestein: the chirp product fused into the forward transform's first reads (zero past n), the
    // kernel product into its last writes, the closing chirp product into the inverse transform's last writes
    const int n = p.n;
    const float2* chirp = p.chirp;
    const float2* khat = p.khat;
    auto first = [=](int e, float2 v) {
      if (e >= n) return make_float2(0.f, 0.f);
      float2 c = chirp[e];
      if (sign > 0) c.y = -c.y;
      return cmul(v, c);
    };
    auto kernel = [=](int e, float2 v) {
      float2 kh = khat[e];
      if (sign > 0) kh.y = -kh.y;
      return cmul(v, kh);
    };
    auto last = [=](int e, float2 v) {
      if (e >= n) return v;
      float2 c = chirp[e];
      if (sign > 0) c.y = -c.y;
      return cmul(v, c);
    };
    float2* r = lds_stockham_fixed<FM, FL>(a, b, p.roots, -1, tid, first, kernel);
    float2* o = (r == a) ? b : a;
    return lds_stockham_fixed<FM, FL>(r, o, p.roots, +1, tid, NoHook{}, last);
  }
  const int n = p.n, m = p.m;
  const Div dm(m), dn(n);
  for (int idx = tid; idx < lines * m; idx += T) {
    int k;
    const int l = dm.quot(idx, k);
    float2 v = make_float2(0.f, 0.f);
    if (k < n) {
      float2 c = p.chirp[k];
      if (sign > 0) c.y = -c.y;
      v = cmul(a[at<MF>(l * ls + k)], c);
    }
    a[at<MF>(l * ls + k)] = v;
  }
  __syncthreads();
  float2* r = lds_length_m<MF, FM, FL>(a, b, lines, ls, p, -1, tid, T);
  for (int idx = tid; idx < lines * m; idx += T) {
    int k;
    const int l = dm.quot(idx, k);
    float2 kh = p.khat[k];
    if (sign > 0) kh.y = -kh.y;
    r[at<MF>(l * ls + k)] = cmul(r[at<MF>(l * ls + k)], kh);
  }
  __syncthreads();
  float2* o = (r == a) ? b : a;
  r = lds_length_m<MF, FM, FL>(r, o, lines, ls, p, +1, tid, T);
  for (int idx = tid; idx < lines * n; idx += T) {
    int k;
    const int l = dn.quot(idx, k);
    float2 c = p.chirp[k];
    if (sign > 0) c.y = -c.y;
    r[at<MF>(l * ls + k)] = cmul(r[at<MF>(l * ls + k)], c);
  }
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------------------------------------------
// Generic row pass: block = `rows` lines along y of one wave (blockIdx.y). MF and FM select the line transform (MF = 0:
// float32 Stockham, of the fixed length FM when FM > 0; else the split-BF16 stages of n2 = MF), one kernel each so
// the float32 kernels keep their registers.
// ---------------------------------------------------------------------------------------------------------------
inline int row_lines(int m) { return m >= kGenericLdsComplex ? 1 : kGenericLdsComplex / m; }
// Columns per block of the generic column pass. The float32 384-point kernel takes 8 (64-byte row segments of the
// compact wave, 52 KiB of LDS) where the 2048-point rule gives 5 (40-byte segments), the 1024-point one 3 (52 KiB)
// where the rule gives 2.
constexpr int kCols384 = 8;
constexpr int kCols1024 = 3;
constexpr int fixed_col_lines(int fm) {
  return fm == 384 ? kCols384 : fm == 1024 ? kCols1024 : fm > 0 ? kGenericLdsComplex / fm : 0;
}
int col_generic_lines(const LineFFT& p) {
  if (!p.use_mfma && p.m == 384) return kCols384;
  if (!p.use_mfma && p.m == 1024) return kCols1024;
  return p.m >= kGenericLdsComplex ? 1 : kGenericLdsComplex / p.m;
}

template <int MF, int FM>
__global__ void __launch_bounds__(kGenericThreads) row_generic_kernel(RowArgs a, int rows) {
  extern __shared__ float2 lds[];
  const int m = FM > 0 ? FM : a.fy.m, ls = m;
  float2* A = lds;
  float2* B = lds + lds_entries(rows * ls, MF == 0);
  const int tid = threadIdx.x, T = blockDim.x;
  const int x0 = blockIdx.x * rows;
  const long long b = blockIdx.y;
  const int f = a.flags;
  const long long gw = (long long)a.nx * a.nc;
  const Div dm(m), dny(a.ny), dnc(a.nc);

  for (int idx = tid; idx < rows * m; idx += T) {
    int k;
    const int l = dm.quot(idx, k), x = x0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (x < a.nx && k < a.ny) {
      if (f & ROW_IN_COMPACT) {
        const int c = a.keepinv[k];
        if (c >= 0) v = a.G[b * gw + (long long)x * a.nc + c];
      } else {
        v = a.in_full[b * a.in_full_stride + (long long)x * a.ny + k];
      }
    }
    A[at<MF>(l * ls + k)] = v;
  }
  __syncthreads();
  float2* cur = A;
  float2* oth = B;
  if (f & ROW_INV) {
    cur = lds_line_transform<MF, FM, (FM > 0 ? kGenericLdsComplex / FM : 0)>(cur, oth, rows, ls, a.fy, +1, tid, T);
    oth = (cur == A) ? B : A;
  }
  if (f & (ROW_WRITE_PLANE | ROW_MUL_T)) {
    for (int idx = tid; idx < rows * a.ny; idx += T) {
      int k;
      const int l = dny.quot(idx, k), x = x0 + l;
      if (x >= a.nx) continue;
      float2 v = cur[at<MF>(l * ls + k)];
      if (f & ROW_WRITE_PLANE) a.plane[b * a.plane_stride + (long long)x * a.ny + k] = cscale(v, a.out_scale);
      if (f & ROW_MUL_T) cur[at<MF>(l * ls + k)] = cmul(v, a.t[(long long)x * a.ny + k]);
    }
    __syncthreads();
  }
  if (f & ROW_FWD) cur = lds_line_transform<MF, FM, (FM > 0 ? kGenericLdsComplex / FM : 0)>(cur, oth, rows, ls, a.fy, -1, tid, T);
  if (f & ROW_OUT_COMPACT) {
    for (int idx = tid; idx < rows * a.nc; idx += T) {
      int c;
      const int l = dnc.quot(idx, c), x = x0 + l;
      if (x >= a.nx) continue;
      const int iy = a.keep[c];
      a.G[b * gw + (long long)x * a.nc + c] = iy >= 0 ? cur[at<MF>(l * ls + iy)] : make_float2(0.f, 0.f);
    }
  }
  if (f & ROW_OUT_FULL) {
    for (int idx = tid; idx < rows * a.ny; idx += T) {
      int k;
      const int l = dny.quot(idx, k), x = x0 + l;
      if (x >= a.nx) continue;
      a.out_full[b * a.out_full_stride + (long long)x * a.ny + k] = cscale(cur[at<MF>(l * ls + k)], a.out_scale);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Detector reduction of one block's pixels (final column pass). `pix(idx)` yields (ix, iy, intensity) or iy < 0.
// ---------------------------------------------------------------------------------------------------------------
__device__ inline void block_reduce(double s, double* scratch, int tid, int T, double* out) {
  // deterministic tree reduction; `scratch` holds T doubles
  scratch[tid] = s;
  __syncthreads();
  for (int w = T / 2; w > 0; w >>= 1) {
    if (tid < w) scratch[tid] += scratch[tid + w];
    __syncthreads();
  }
  if (tid == 0) *out = scratch[0];
  __syncthreads();
}

// ---------------------------------------------------------------------------------------------------------------
// Generic column pass: block = `cols` compact columns (lines along x) of one wave; MF and FM as in the row pass.
// ---------------------------------------------------------------------------------------------------------------
template <int MF, int FM>
__global__ void __launch_bounds__(kGenericThreads) col_generic_kernel(ColArgs a, int cols) {
  extern __shared__ float2 lds[];
  const int m = FM > 0 ? FM : a.fx.m, ls = m;
  float2* A = lds;
  float2* B = lds + lds_entries(cols * ls, MF == 0);
  const int tid = threadIdx.x, T = blockDim.x;
  const int2 unit = xcd_unit();  // tiles of one position on one XCD: their row segments share 128-byte lines
  const int c0 = unit.x * cols;
  const long long b = unit.y;
  const int f = a.flags;
  const long long gw = (long long)a.nx * a.nc;
  const Div dcols(cols);

  for (int idx = tid; idx < cols * m; idx += T) {
    int l;
    const int k = dcols.quot(idx, l), c = c0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (k < a.nx && c < a.nc) {
      if (f & COL_IN_G) {
        v = a.G[b * gw + (long long)k * a.nc + c];
      } else if (f & COL_GENERATE) {
        const int iy = a.keep[c];
        if (iy >= 0)
          v = cmul(cmul(a.spec[(long long)k * a.ny + iy], a.rx[b * a.nx + k]), a.ry[b * a.ny + iy]);
      }
    }
    A[at<MF>(l * ls + k)] = v;
  }
  __syncthreads();
  float2* cur = A;
  float2* oth = B;
  if (f & COL_FWD) {
    cur = lds_line_transform<MF, FM, fixed_col_lines(FM)>(cur, oth, cols, ls, a.fx, -1, tid, T);
    oth = (cur == A) ? B : A;
  }
  if (f & COL_MUL_P) {
    for (int idx = tid; idx < cols * a.nx; idx += T) {
      int l;
      const int k = dcols.quot(idx, l), c = c0 + l;
      if (c < a.nc) cur[at<MF>(l * ls + k)] = cmul(cur[at<MF>(l * ls + k)], a.P[(long long)k * a.nc + c]);
    }
    __syncthreads();
  }
  if (f & COL_DETECT) {
    // one slot at a time: each thread sums its pixels in index order, then the block's tree; the pixelated
    // accumulators after the slots
    const long long pos = first_position(a.pos0, a.pos_dev) + b;
    double* scratch = reinterpret_cast<double*>(oth);
    double* out = a.det.partial + (b * a.det.nslots) * a.det.blocks_per_wave + unit.x;
    for (int q = 0; q < a.det.nslots; ++q) {
      double s = 0.0;
      for (int idx = tid; idx < cols * a.nx; idx += T) {
        int l;
        const int k = dcols.quot(idx, l), c = c0 + l;
        if (c >= a.nc) continue;
        const int iy = a.keep[c];
        if (iy < 0) continue;
        const float2 v = cur[at<MF>(l * ls + k)];
        const double I = ((double)v.x * v.x + (double)v.y * v.y) * a.det.scale;
        s += slot_value(a.det, q, k, iy, annular_a2(a.det, k, iy), I);
      }
      block_reduce(s, scratch, tid, T, out + (long long)q * a.det.blocks_per_wave);
    }
    if (a.det.npix > 0) {
      for (int idx = tid; idx < cols * a.nx; idx += T) {
        int l;
        const int k = dcols.quot(idx, l), c = c0 + l;
        if (c >= a.nc) continue;
        const int iy = a.keep[c];
        if (iy < 0) continue;
        const float2 v = cur[at<MF>(l * ls + k)];
        detect_pixelated(a.det, k, iy, ((double)v.x * v.x + (double)v.y * v.y) * a.det.scale, pos);
      }
    }
    return;
  }
  if (f & COL_INV) cur = lds_line_transform<MF, FM, fixed_col_lines(FM)>(cur, oth, cols, ls, a.fx, +1, tid, T);
  if (f & COL_OUT) {
    for (int idx = tid; idx < cols * a.nx; idx += T) {
      int l;
      const int k = dcols.quot(idx, l), c = c0 + l;
      if (c < a.nc) a.G[b * gw + (long long)k * a.nc + c] = cur[at<MF>(l * ls + k)];
    }
  }
}

// Sum the per-block partials of every wave in block order and add them to the position's accumulators.
__global__ void reduce_partials_kernel(DetPlan d, int pos0, const int* pos_dev, double* acc) {
  const long long b = blockIdx.x;
  const long long pos = first_position(pos0, pos_dev) + b;
  for (int q = threadIdx.x; q < d.nslots; q += blockDim.x) {
    const double* p = d.partial + (b * d.nslots + q) * d.blocks_per_wave;
    double s = 0.0;
    for (int i = 0; i < d.blocks_per_wave; ++i) s += p[i];
    acc[pos * d.nslots + q] += s;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------------------------
// The kernel variant of a plan's line transform: MF = 0 float32, else n2 = m / 16 of the split-BF16 stages; FM = the
// float32 lengths with compile-time stages (256, 384, 512, 1024, 2048), else 0.
int transform_kind(const LineFFT& p) {
  if (!p.use_mfma) return 0;
  if (p.m != 256 && p.m != 384 && p.m != 512 && p.m != 1024)
    fail(SK_ERR_INVALID, "split-BF16 plan of unsupported length " + std::to_string(p.m));
  return p.m / 16;
}
int fixed_length(const LineFFT& p) {
  if (p.use_mfma) return 0;
  switch (p.m) {
    case 256: case 384: case 512: case 1024: case 2048: return p.m;
    default: return 0;
  }
}

// The kernel of the plan's variant: Pick<MF, FM>::kernel() for the plan's MF and FM.
template <template <int, int> class Pick>
auto generic_variant(const LineFFT& p) {
  switch (transform_kind(p)) {
    case 0:
      switch (fixed_length(p)) {
        case 256: return Pick<0, 256>::kernel();
        case 384: return Pick<0, 384>::kernel();
        case 512: return Pick<0, 512>::kernel();
        case 1024: return Pick<0, 1024>::kernel();
        case 2048: return Pick<0, 2048>::kernel();
        default: return Pick<0, 0>::kernel();
      }
    case 16: return Pick<16, 0>::kernel();
    case 24: return Pick<24, 0>::kernel();
    case 32: return Pick<32, 0>::kernel();
    default: return Pick<64, 0>::kernel();
  }
}
template <int MF, int FM>
struct RowKernel {
  static auto kernel() { return &row_generic_kernel<MF, FM>; }
};
template <int MF, int FM>
struct ColKernel {
  static auto kernel() { return &col_generic_kernel<MF, FM>; }
};

void launch_row_generic(const RowArgs& a, int batch, hipStream_t s) {
  const int rows = row_lines(a.fy.m);
  dim3 grid((a.nx + rows - 1) / rows, batch);
  const size_t lds = 2 * sizeof(float2) * lds_entries(rows * a.fy.m, !a.fy.use_mfma);
  const auto kernel = generic_variant<RowKernel>(a.fy);
  hipLaunchKernelGGL(kernel, grid, dim3(kGenericThreads), lds, s, a, rows);
  SK_HIP(hipGetLastError());
}

void launch_col_generic(const ColArgs& a, int batch, hipStream_t s) {
  const int cols = col_generic_lines(a.fx);
  dim3 grid((a.nc + cols - 1) / cols, batch);
  const size_t buf = sizeof(float2) * lds_entries(cols * a.fx.m, !a.fx.use_mfma);
  size_t lds = 2 * buf;
  if (a.flags & COL_DETECT) lds = std::max(lds, buf + sizeof(double) * kGenericThreads);
  const auto kernel = generic_variant<ColKernel>(a.fx);
  hipLaunchKernelGGL(kernel, grid, dim3(kGenericThreads), lds, s, a, cols);
  SK_HIP(hipGetLastError());
}

void launch_reduce_partials(const DetPlan& d, int batch, int pos0, const int* pos_dev, double* acc, hipStream_t s) {
  if (d.nslots == 0) return;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(batch), dim3(64), 0, s, d, pos0, pos_dev, acc);
  SK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------------------------
// Plans (host): radices, roots and Bluestein tables, computed in float64 and rounded to complex64.
// ---------------------------------------------------------------------------------------------------------------
namespace {

using cd = std::complex<double>;

std::vector<int> smooth_radices(int n) {
  std::vector<int> r;
  while (n % 8 == 0) { r.push_back(8); n /= 8; }
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0) { r.push_back(p); n /= p; }
  if (n != 1) r.clear();
  return r;
}

cd root(long long k, long long n) {  // e^{-2 pi i k / n}
  k %= n;
  double a = 2.0 * M_PI * double(k) / double(n);
  return cd(std::cos(a), -std::sin(a));
}

void host_fft_pow2(std::vector<cd>& a) {  // forward, unscaled, iterative radix-2 (tables only)
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1)
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < len / 2; ++j) {
        cd w = root((long long)(j * (n / len)), (long long)n);
        cd u = a[i + j], v = a[i + j + len / 2] * w;
        a[i + j] = u + v;
        a[i + j + len / 2] = u - v;
      }
}

std::vector<float2> to_f2(const std::vector<cd>& v) {
  std::vector<float2> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = make_float2(float(v[i].real()), float(v[i].imag()));
  return o;
}

// The length of the device transform of a line of n points: n when n factors into the Stockham radices, else the
// Bluestein length, the smallest power of two (at least 16) holding the 2 n - 1 point convolution.
int device_length(int n) {
  if (n <= 1 || !smooth_radices(n).empty()) return n;
  int m = 16;
  while (m < 2 * n - 1) m <<= 1;
  return m;
}

std::map<int, std::unique_ptr<LinePlan>>& plans() {
  static auto* p = new std::map<int, std::unique_ptr<LinePlan>>;  // process lifetime: never freed at exit
  return *p;
}

}  // namespace

const LinePlan& line_plan(int n, hipStream_t s) {
  auto& ps = plans();
  auto it = ps.find(n);
  if (it != ps.end()) return *it->second;
  if (n < 1) fail(SK_ERR_INVALID, "grid length must be at least 1");
  auto lp = std::make_unique<LinePlan>();
  LineFFT& d = lp->dev;
  d.n = n;
  const int m = device_length(n);
  d.blue = m != n;
  auto r = smooth_radices(m);
  if (m > 8192) fail(SK_ERR_INVALID, "grid length " + std::to_string(n) + " exceeds the device FFT's LDS tile");
  if ((int)r.size() > kMaxStages) fail(SK_ERR_INVALID, "too many FFT stages");
  d.m = m;
  d.nst = (int)r.size();
  for (size_t i = 0; i < r.size(); ++i) d.radix[i] = r[i];
  std::vector<cd> roots(m);
  for (int k = 0; k < m; ++k) roots[k] = root(k, m);
  lp->roots.upload(to_f2(roots), s);
  d.roots = lp->roots.p;
  if (d.blue) {
    std::vector<cd> chirp(n);
    for (int k = 0; k < n; ++k) {
      long long k2 = (long long)k * k % (2LL * n);
      double a = M_PI * double(k2) / double(n);
      chirp[k] = cd(std::cos(a), -std::sin(a));
    }
    std::vector<cd> bk(m, cd(0, 0));
    bk[0] = std::conj(chirp[0]);
    for (int k = 1; k < n; ++k) bk[k] = bk[m - k] = std::conj(chirp[k]);
    host_fft_pow2(bk);
    for (auto& v : bk) v /= double(m);
    lp->chirp.upload(to_f2(chirp), s);
    lp->khat.upload(to_f2(bk), s);
    d.chirp = lp->chirp.p;
    d.khat = lp->khat.p;
  } else {
    d.chirp = nullptr;
    d.khat = nullptr;
  }
  SK_HIP(hipStreamSynchronize(s));
  auto& ref = *lp;
  ps[n] = std::move(lp);
  return ref;
}

bool mfma_supported(int n) {
  const int m = device_length(n);
  return m == 256 || m == 384 || m == 512 || m == 1024;
}

namespace {
// float -> bfloat16 bits, round to nearest even (finite inputs)
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
float bf16_value(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// The real block matrix of the length-L DFT in MFMA B-operand order [sign][k-step][tile][hi/lo][lane][8]:
// row k = 16 s + 8 (lane / 32) + j (input j1 = k / 2, re/im k % 2), column 32 t + lane % 32 (output k1, re/im);
// the columns past 2 L (the padding of the last tile when L is not a multiple of 16) are zero.
std::vector<uint16_t> dft_block_table(int L) {
  const int ksteps = L / 8, ntiles = (L + 15) / 16;
  std::vector<uint16_t> out((size_t)2 * ksteps * ntiles * 2 * 64 * 8);
  for (int si = 0; si < 2; ++si)
    for (int s = 0; s < ksteps; ++s)
      for (int t = 0; t < ntiles; ++t)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int k = 16 * s + 8 * (lane / 32) + j, nn = 32 * t + lane % 32;
            const int j1 = k >> 1, p = k & 1, k1 = nn >> 1, q = nn & 1;
            const double ang = (si ? 2.0 : -2.0) * M_PI * double((long long)j1 * k1 % L) / double(L);
            const double wr = std::cos(ang), wi = std::sin(ang);
            const double R = k1 >= L ? 0.0 : p == 0 ? (q == 0 ? wr : wi) : (q == 0 ? -wi : wr);
            const uint16_t hi = bf16_bits((float)R);
            const uint16_t lo = bf16_bits((float)(R - (double)bf16_value(hi)));
            const size_t base = ((((size_t)si * ksteps + s) * ntiles + t) * 2) * 64 * 8;
            out[base + (size_t)lane * 8 + j] = hi;
            out[base + 64 * 8 + (size_t)lane * 8 + j] = lo;
          }
  return out;
}
struct MfmaPlan {
  LineFFT dev{};
  DevBuf<uint16_t> t16, tn2;
  DevBuf<float2> tw;
};
}  // namespace

const LineFFT& mfma_line_plan(int n, hipStream_t s) {
  static std::map<int, std::unique_ptr<MfmaPlan>>* plans = new std::map<int, std::unique_ptr<MfmaPlan>>;
  auto it = plans->find(n);
  if (it != plans->end()) return it->second->dev;
  if (!mfma_supported(n))
    fail(SK_ERR_INVALID, "the split-BF16 MFMA transforms support grid lengths 256, 384, 512 and 1024, and lengths "
                         "done by Bluestein's algorithm from 65 to 512 (grid length " + std::to_string(n) + ")");
  auto mp = std::make_unique<MfmaPlan>();
  mp->dev = line_plan(n, s).dev;
  const int m = mp->dev.m, n2 = m / 16;  // the matrix-core stages transform the device length m (Bluestein's, or n)
  mp->t16.upload(dft_block_table(16), s);
  mp->tn2.upload(dft_block_table(n2), s);
  std::vector<cd> tw((size_t)16 * n2);
  for (int k1 = 0; k1 < 16; ++k1)
    for (int j2 = 0; j2 < n2; ++j2) tw[(size_t)k1 * n2 + j2] = root((long long)k1 * j2 % m, m);
  mp->tw.upload(to_f2(tw), s);
  mp->dev.use_mfma = 1;
  mp->dev.mtab16 = mp->t16.p;
  mp->dev.mtabn2 = mp->tn2.p;
  mp->dev.mtw = mp->tw.p;
  const LineFFT& d = mp->dev;
  (*plans)[n] = std::move(mp);
  return d;
}

}  // namespace sk
