// sk_dft.h — small unscaled DFTs (sign -1 forward, +1 inverse) and the column passes' block order, shared by the
// generic, fast and register passes.
#pragma once

#include <hip/hip_runtime.h>

#include "sk_internal.h"

namespace sk {

// The (tile, position) unit of a column-pass block in a 2-D grid of tiles x positions (or position groups). The
// device deals work-groups out round-robin over its 8 XCDs, each with its own L2, in dispatch order L = blockIdx.x +
// gridDim.x blockIdx.y. Taken in that order, block L runs unit (L mod 8) U / 8 + L / 8 of the U = gridDim.x gridDim.y
// units (tile fastest) when 8 divides U: each XCD walks consecutive tiles of its own positions, so tiles whose row
// segments (a few columns, 16 to 64 bytes) make up the same 128-byte lines are read through one L2 rather than
// fetched by two or more XCDs.
__device__ __forceinline__ int2 xcd_unit() {
  const int gx = gridDim.x, units = gx * gridDim.y;
  const int L = blockIdx.x + gx * blockIdx.y;
  const int u = units % 8 == 0 ? (L % 8) * (units / 8) + L / 8 : L;
  return make_int2(u % gx, u / gx);
}

// The immediate of s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15) (gfx9 fields): a wait for vector memory accesses only,
// for __builtin_amdgcn_s_waitcnt. A column pass completes its first unit's prefetch with it before its loop, so that
// the compiler's wait at the loop's head covers the prefetched loads and not the stores issued after them.
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ inline float2 mul_si(float2 a, int sign) {  // a * (sign * i)
  return sign < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

__device__ inline void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

__device__ inline void dft4(float2& a0, float2& a1, float2& a2, float2& a3, int sign) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_si(csub(a1, a3), sign);
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// DFT of length 8 as two DFT4s (even and odd samples) and the w8^k combination, w8 = e^{sign i pi / 4}
__device__ inline void dft8(float2 (&u)[8], int sign) {
  float2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
  float2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
  dft4(e0, e1, e2, e3, sign);
  dft4(o0, o1, o2, o3, sign);
  const float h = 0.70710678118654752f;
  const float s = float(sign);
  float2 w1 = make_float2((o1.x - s * o1.y) * h, (o1.y + s * o1.x) * h);
  float2 w2 = mul_si(o2, sign);
  float2 w3 = make_float2((-o3.x - s * o3.y) * h, (-o3.y + s * o3.x) * h);
  u[0] = cadd(e0, o0);
  u[4] = csub(e0, o0);
  u[1] = cadd(e1, w1);
  u[5] = csub(e1, w1);
  u[2] = cadd(e2, w2);
  u[6] = csub(e2, w2);
  u[3] = cadd(e3, w3);
  u[7] = csub(e3, w3);
}

// dft8 of a vector whose samples 3 and 4 are zero (not read): the two DFT4s drop the four additions with a zero
// operand. Equal to dft8 on such input up to the sign of exact zeros.
__device__ inline void dft8_z34(float2 (&u)[8], int sign) {
  // even samples (u0, u2, 0, u6)
  const float2 et2 = cadd(u[2], u[6]), et3 = mul_si(csub(u[2], u[6]), sign);
  const float2 e0 = cadd(u[0], et2), e1 = cadd(u[0], et3), e2 = csub(u[0], et2), e3 = csub(u[0], et3);
  // odd samples (u1, 0, u5, u7)
  const float2 ot0 = cadd(u[1], u[5]), ot1 = csub(u[1], u[5]);
  const float2 ot3 = mul_si(make_float2(-u[7].x, -u[7].y), sign);
  const float2 o0 = cadd(ot0, u[7]), o1 = cadd(ot1, ot3), o2 = csub(ot0, u[7]), o3 = csub(ot1, ot3);
  const float h = 0.70710678118654752f;
  const float s = float(sign);
  const float2 w1 = make_float2((o1.x - s * o1.y) * h, (o1.y + s * o1.x) * h);
  const float2 w2 = mul_si(o2, sign);
  const float2 w3 = make_float2((-o3.x - s * o3.y) * h, (-o3.y + s * o3.x) * h);
  u[0] = cadd(e0, o0);
  u[4] = csub(e0, o0);
  u[1] = cadd(e1, w1);
  u[5] = csub(e1, w1);
  u[2] = cadd(e2, w2);
  u[6] = csub(e2, w2);
  u[3] = cadd(e3, w3);
  u[7] = csub(e3, w3);
}

// The final pass's detection (AnnularDetector, CenterOfMassDetector, SegmentedDetector, PixelatedDetector;
// Detectors.swift:119-205), one scalar slot at a time so that a thread's running sum is one register: the
// contribution of intensity I at pixel (ix, iy) to slot q, with a2 = ax^2 + ay^2 from annular_a2.
__device__ inline double annular_a2(const DetPlan& d, int ix, int iy) {
  // float64 exactly as the host rounds it (no contraction)
  return d.sqx ? __dadd_rn(d.sqx[ix], d.sqy[iy]) : 0.0;
}
__device__ inline double slot_value(const DetPlan& d, int q, int ix, int iy, double a2, double I) {
  const int kind = d.slot_kind[q], arg = d.slot_arg[q];
  if (kind == 0) return (a2 >= d.slot_i2[q] && a2 < d.slot_o2[q]) ? I : 0.0;
  if (kind == 3) return d.seg_map[arg][(long long)ix * d.ny + iy] == d.slot_bin[q] ? I : 0.0;
  const int mx = d.com_mapx[arg][ix], my = d.com_mapy[arg][iy];
  return (mx >= 0 && my >= 0) ? I * (kind == 1 ? d.com_wx[arg][ix] : d.com_wy[arg][iy]) : 0.0;
}
// The pixelated accumulators of pixel (ix, iy) of position pos
__device__ inline void detect_pixelated(const DetPlan& d, int ix, int iy, double I, long long pos) {
  for (int p = 0; p < d.npix; ++p) {
    const int ox = d.pix_mapx[p][ix], oy = d.pix_mapy[p][iy];
    if (ox >= 0 && oy >= 0) d.pix_acc[p][pos * d.pix_stride[p] + (long long)ox * d.pix_my[p] + oy] += I;
  }
}

// DFT of length 3: with w = e^{sign 2 pi i / 3} = -1/2 + sign i sqrt(3)/2, Y1,2 = a - (b + c)/2 +- sign i sqrt(3)/2 (b - c)
__device__ inline void dft3(float2& a, float2& b, float2& c, int sign) {
  const float h = 0.866025403784438647f;  // sqrt(3) / 2
  const float2 s = cadd(b, c);
  const float2 t = mul_si(csub(b, c), sign);
  const float2 m = make_float2(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
  a = cadd(a, s);
  b = make_float2(m.x + h * t.x, m.y + h * t.y);
  c = make_float2(m.x - h * t.x, m.y - h * t.y);
}

// DFT of length 6 by the prime-factor map of 6 = 2 x 3 (Good-Thomas, no twiddles): input n = (3 n1 + 2 n2) mod 6,
// output k = (3 k1 + 4 k2) mod 6, so that w6^{n k} = w2^{n1 k1} w3^{n2 k2}: two DFT3s over n2, then three DFT2s over n1.
__device__ inline void dft6(float2 (&u)[6], int sign) {
  float2 a0 = u[0], a1 = u[2], a2 = u[4];  // n1 = 0: n = 0, 2, 4
  float2 b0 = u[3], b1 = u[5], b2 = u[1];  // n1 = 1: n = 3, 5, 1
  dft3(a0, a1, a2, sign);
  dft3(b0, b1, b2, sign);
  u[0] = cadd(a0, b0);  // k2 = 0: k = 0, 3
  u[3] = csub(a0, b0);
  u[4] = cadd(a1, b1);  // k2 = 1: k = 4, 1
  u[1] = csub(a1, b1);
  u[2] = cadd(a2, b2);  // k2 = 2: k = 2, 5
  u[5] = csub(a2, b2);
}

// DFT of length 16 as 4 x 4 (Cooley-Tukey): n = 4 n1 + n2, k = k1 + 4 k2; DFT4 over n1 for each n2, the twiddle
// w16^{n2 k1} (w16 = e^{sign 2 pi i / 16}), then DFT4 over n2 for each k1.
__device__ inline void dft16(float2 (&u)[16], int sign) {
  const float c1 = 0.923879532511286756f, s1 = 0.382683432365089772f, r2 = 0.707106781186547524f;
  float2 y[4][4];  // y[n2][k1]
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 a0 = u[n2], a1 = u[4 + n2], a2 = u[8 + n2], a3 = u[12 + n2];
    dft4(a0, a1, a2, a3, sign);
    y[n2][0] = a0;
    y[n2][1] = a1;
    y[n2][2] = a2;
    y[n2][3] = a3;
  }
  const float sg = sign < 0 ? -1.f : 1.f;
  // w16^m for m = n2 k1 in {1, 2, 3, 4, 6, 9}: (cos(2 pi m / 16), sign sin(2 pi m / 16))
  y[1][1] = cmul(y[1][1], make_float2(c1, sg * s1));
  y[1][2] = cmul(y[1][2], make_float2(r2, sg * r2));
  y[1][3] = cmul(y[1][3], make_float2(s1, sg * c1));
  y[2][1] = cmul(y[2][1], make_float2(r2, sg * r2));
  y[2][2] = mul_si(y[2][2], sign);
  y[2][3] = cmul(y[2][3], make_float2(-r2, sg * r2));
  y[3][1] = cmul(y[3][1], make_float2(s1, sg * c1));
  y[3][2] = cmul(y[3][2], make_float2(-r2, sg * r2));
  y[3][3] = cmul(y[3][3], make_float2(-c1, -sg * s1));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
    dft4(b0, b1, b2, b3, sign);
    u[k1] = b0;
    u[k1 + 4] = b1;
    u[k1 + 8] = b2;
    u[k1 + 12] = b3;
  }
}

// dft16 of inputs whose upper half u[8..15] is zero (Bluestein's zero-padded forward transform): the first
// layer's DFT4s take two inputs, a0 +- a1 and a0 +- (sign i) a1; the twiddles and the second layer as in dft16.
__device__ inline void dft16_half(float2 (&u)[16], int sign) {
  const float c1 = 0.923879532511286756f, s1 = 0.382683432365089772f, r2 = 0.707106781186547524f;
  float2 y[4][4];  // y[n2][k1]
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    const float2 a0 = u[n2], a1 = u[4 + n2], t3 = mul_si(a1, sign);
    y[n2][0] = cadd(a0, a1);
    y[n2][1] = cadd(a0, t3);
    y[n2][2] = csub(a0, a1);
    y[n2][3] = csub(a0, t3);
  }
  const float sg = sign < 0 ? -1.f : 1.f;
  y[1][1] = cmul(y[1][1], make_float2(c1, sg * s1));
  y[1][2] = cmul(y[1][2], make_float2(r2, sg * r2));
  y[1][3] = cmul(y[1][3], make_float2(s1, sg * c1));
  y[2][1] = cmul(y[2][1], make_float2(r2, sg * r2));
  y[2][2] = mul_si(y[2][2], sign);
  y[2][3] = cmul(y[2][3], make_float2(-r2, sg * r2));
  y[3][1] = cmul(y[3][1], make_float2(s1, sg * c1));
  y[3][2] = cmul(y[3][2], make_float2(-r2, sg * r2));
  y[3][3] = cmul(y[3][3], make_float2(-c1, -sg * s1));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
    dft4(b0, b1, b2, b3, sign);
    u[k1] = b0;
    u[k1 + 4] = b1;
    u[k1 + 8] = b2;
    u[k1 + 12] = b3;
  }
}

// Complex sums whose operands may be known zeros (compile-time flags ZA, ZB): a known zero costs nothing. A result
// is a known zero when both operands are.
template <bool ZA, bool ZB>
__device__ __forceinline__ float2 zadd(float2 a, float2 b) {
  if constexpr (ZA && ZB) return make_float2(0.f, 0.f);
  else if constexpr (ZA) return b;
  else if constexpr (ZB) return a;
  else return cadd(a, b);
}
template <bool ZA, bool ZB>
__device__ __forceinline__ float2 zsub(float2 a, float2 b) {
  if constexpr (ZA && ZB) return make_float2(0.f, 0.f);
  else if constexpr (ZA) return make_float2(-b.x, -b.y);
  else if constexpr (ZB) return a;
  else return csub(a, b);
}

// dft4 whose inputs j with bit j of Z set are known zeros (not read); with Z = 0 the operations of dft4
template <int Z>
__device__ __forceinline__ void dft4_z(float2& a0, float2& a1, float2& a2, float2& a3, int sign) {
  constexpr bool z0 = Z & 1, z1 = (Z >> 1) & 1, z2 = (Z >> 2) & 1, z3 = (Z >> 3) & 1;
  constexpr bool zt0 = z0 && z2, zt2 = z1 && z3;  // t1 is zero with t0, t3 with t2
  const float2 t0 = zadd<z0, z2>(a0, a2), t1 = zsub<z0, z2>(a0, a2), t2 = zadd<z1, z3>(a1, a3);
  const float2 t3 = zt2 ? make_float2(0.f, 0.f) : mul_si(zsub<z1, z3>(a1, a3), sign);
  a0 = zadd<zt0, zt2>(t0, t2);
  a2 = zsub<zt0, zt2>(t0, t2);
  a1 = zadd<zt0, zt2>(t1, t3);
  a3 = zsub<zt0, zt2>(t1, t3);
}

// dft8 whose inputs with bit j of Z set are known zeros; with Z = 0 the operations of dft8
template <int Z>
__device__ __forceinline__ void dft8_z(float2 (&u)[8], int sign) {
  constexpr int ZE = (Z & 1) | ((Z >> 1) & 2) | ((Z >> 2) & 4) | ((Z >> 3) & 8);         // samples 0, 2, 4, 6
  constexpr int ZO = ((Z >> 1) & 1) | ((Z >> 2) & 2) | ((Z >> 3) & 4) | ((Z >> 4) & 8);  // samples 1, 3, 5, 7
  constexpr bool ze = ZE == 15, zo = ZO == 15;
  float2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
  float2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
  dft4_z<ZE>(e0, e1, e2, e3, sign);
  dft4_z<ZO>(o0, o1, o2, o3, sign);
  const float h = 0.70710678118654752f;
  const float s = float(sign);
  const float2 w1 = make_float2((o1.x - s * o1.y) * h, (o1.y + s * o1.x) * h);
  const float2 w2 = mul_si(o2, sign);
  const float2 w3 = make_float2((-o3.x - s * o3.y) * h, (-o3.y + s * o3.x) * h);
  u[0] = zadd<ze, zo>(e0, o0);
  u[4] = zsub<ze, zo>(e0, o0);
  u[1] = zadd<ze, zo>(e1, w1);
  u[5] = zsub<ze, zo>(e1, w1);
  u[2] = zadd<ze, zo>(e2, w2);
  u[6] = zsub<ze, zo>(e2, w2);
  u[3] = zadd<ze, zo>(e3, w3);
  u[7] = zsub<ze, zo>(e3, w3);
}

// dft16 (4 x 4, as dft16) whose inputs with bit j of Z set are known zeros; with Z = 0 the operations of dft16
template <int Z>
__device__ __forceinline__ void dft16_z(float2 (&u)[16], int sign) {
  const float c1 = 0.923879532511286756f, s1 = 0.382683432365089772f, r2 = 0.707106781186547524f;
  // first layer over n1 for each n2: inputs n2, 4 + n2, 8 + n2, 12 + n2
  constexpr int Z0 = (Z & 1) | ((Z >> 3) & 2) | ((Z >> 6) & 4) | ((Z >> 9) & 8);
  constexpr int Z1 = ((Z >> 1) & 1) | ((Z >> 4) & 2) | ((Z >> 7) & 4) | ((Z >> 10) & 8);
  constexpr int Z2 = ((Z >> 2) & 1) | ((Z >> 5) & 2) | ((Z >> 8) & 4) | ((Z >> 11) & 8);
  constexpr int Z3 = ((Z >> 3) & 1) | ((Z >> 6) & 2) | ((Z >> 9) & 4) | ((Z >> 12) & 8);
  float2 y[4][4];  // y[n2][k1]
  y[0][0] = u[0], y[0][1] = u[4], y[0][2] = u[8], y[0][3] = u[12];
  y[1][0] = u[1], y[1][1] = u[5], y[1][2] = u[9], y[1][3] = u[13];
  y[2][0] = u[2], y[2][1] = u[6], y[2][2] = u[10], y[2][3] = u[14];
  y[3][0] = u[3], y[3][1] = u[7], y[3][2] = u[11], y[3][3] = u[15];
  dft4_z<Z0>(y[0][0], y[0][1], y[0][2], y[0][3], sign);
  dft4_z<Z1>(y[1][0], y[1][1], y[1][2], y[1][3], sign);
  dft4_z<Z2>(y[2][0], y[2][1], y[2][2], y[2][3], sign);
  dft4_z<Z3>(y[3][0], y[3][1], y[3][2], y[3][3], sign);
  constexpr bool zy1 = Z1 == 15, zy2 = Z2 == 15, zy3 = Z3 == 15, zy0 = Z0 == 15;
  const float sg = sign < 0 ? -1.f : 1.f;
  if constexpr (!zy1) {
    y[1][1] = cmul(y[1][1], make_float2(c1, sg * s1));
    y[1][2] = cmul(y[1][2], make_float2(r2, sg * r2));
    y[1][3] = cmul(y[1][3], make_float2(s1, sg * c1));
  }
  if constexpr (!zy2) {
    y[2][1] = cmul(y[2][1], make_float2(r2, sg * r2));
    y[2][2] = mul_si(y[2][2], sign);
    y[2][3] = cmul(y[2][3], make_float2(-r2, sg * r2));
  }
  if constexpr (!zy3) {
    y[3][1] = cmul(y[3][1], make_float2(s1, sg * c1));
    y[3][2] = cmul(y[3][2], make_float2(-r2, sg * r2));
    y[3][3] = cmul(y[3][3], make_float2(-c1, -sg * s1));
  }
  constexpr int ZK = (zy0 ? 1 : 0) | (zy1 ? 2 : 0) | (zy2 ? 4 : 0) | (zy3 ? 8 : 0);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
    dft4_z<ZK>(b0, b1, b2, b3, sign);
    u[k1] = b0;
    u[k1 + 4] = b1;
    u[k1 + 8] = b2;
    u[k1 + 12] = b3;
  }
}

template <int R>
__device__ inline void dft_pow2(float2 (&u)[R], int sign) {
  if constexpr (R == 2) {
    dft2(u[0], u[1]);
  } else if constexpr (R == 4) {
    dft4(u[0], u[1], u[2], u[3], sign);
  } else if constexpr (R == 16) {
    dft16(u, sign);
  } else {
    static_assert(R == 8, "radix 2, 4, 8 or 16");
    dft8(u, sign);
  }
}

}  // namespace sk
