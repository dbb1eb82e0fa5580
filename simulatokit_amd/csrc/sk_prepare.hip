// sk_prepare.hip — per-potential and per-scan set-up kernels: the quantities the reference computes once in float64
// on the host before its GPU loop (MetalEngine.swift:277-332, 539-553) are computed here on the device, in float64
// where the phase must stay exact, and rounded to complex64 once.

#include <hip/hip_runtime.h>

#include <cmath>

#include "sk_internal.h"

namespace sk {

namespace {
constexpr int kThreads = 256;
inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

__device__ inline double fftfreq_d(int i, int n, double d) {  // Grid.swift:82-86
  const double val = 1.0 / (double(n) * d);
  const int positive = (n - 1) / 2 + 1;
  return double(i < positive ? i : i - n) * val;
}
}  // namespace

// t0 = e^{i sigma V} with the argument sigma V in float64 (Multislice.swift:90-94)
__global__ void phase_to_complex_kernel(const double* V, double sigma, float2* out, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  sincos(sigma * V[i], &s, &c);
  out[i] = make_float2((float)c, (float)s);
}
void launch_phase_to_complex(const double* V, double sigma, float2* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(phase_to_complex_kernel, grid_for(n), dim3(kThreads), 0, s, V, sigma, out, n);
  SK_HIP(hipGetLastError());
}

__global__ void real_to_complex_kernel(const double* re, const double* im, float2* out, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = make_float2((float)re[i], im ? (float)im[i] : 0.f);
}
void launch_real_to_complex(const double* re, const double* im, float2* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(real_to_complex_kernel, grid_for(n), dim3(kThreads), 0, s, re, im, out, n);
  SK_HIP(hipGetLastError());
}

__global__ void complex_to_split_kernel(const float2* in, double* re, double* im, size_t n, double scale) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  re[i] = (double)in[i].x * scale;
  im[i] = (double)in[i].y * scale;
}
void launch_complex_to_split(const float2* in, double* re, double* im, size_t n, double scale, hipStream_t s) {
  hipLaunchKernelGGL(complex_to_split_kernel, grid_for(n), dim3(kThreads), 0, s, in, re, im, n, scale);
  SK_HIP(hipGetLastError());
}

// |FFT2(psi_exit)|^2 of `batch` waves held as compact spectra G [batch][nx][nc] (the final column pass's output,
// spectrum / N): every grid point in FFT order [ix * ny + iy], float64, zero at the columns the bandlimit prunes
// (the propagator is zero there), times `scale` (N^2) (DiffractionIntensity, Detectors.swift:28-36).
__global__ void compact_intensity_kernel(const float2* G, int nx, int ny, int nc, const int* keepinv, size_t total,
                                         double scale, double* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const size_t N = (size_t)nx * ny;
  const size_t b = i / N, r = i - b * N;
  const int ix = int(r / ny), iy = int(r - (size_t)ix * ny);
  const int c = keepinv[iy];
  double v = 0.0;
  if (c >= 0) {
    const float2 z = G[(b * nx + ix) * nc + c];
    v = ((double)z.x * z.x + (double)z.y * z.y) * scale;
  }
  out[i] = v;
}
void launch_compact_intensity(const float2* G, int nx, int ny, int nc, const int* keepinv, size_t batch, double scale,
                              double* out, hipStream_t s) {
  const size_t total = batch * (size_t)nx * ny;
  hipLaunchKernelGGL(compact_intensity_kernel, grid_for(total), dim3(kThreads), 0, s, G, nx, ny, nc, keepinv, total,
                     scale, out);
  SK_HIP(hipGetLastError());
}

// Antialias.aperture (Antialias.swift:18-39), float64
__global__ void antialias_kernel(int nx, int ny, double ex, double ey, double* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int ix = int(i / ny), iy = int(i % ny);
  const double sx = ex / nx, sy = ey / ny;
  const double dmax = fmax(sx, sy);
  const double kc = 0.6666666 / dmax / 2;
  const double w = 0.01 / dmax;
  const double kx = fftfreq_d(ix, nx, sx), ky = fftfreq_d(iy, ny, sy);
  const double r = sqrt(__dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky)));
  double v;
  if (r > kc) v = 0;
  else if (r > kc - w) v = 0.5 * (1 + cos(M_PI * (r - kc + w) / w));
  else v = 1;
  out[i] = v;
}
void launch_antialias(int nx, int ny, double ex, double ey, double* out, hipStream_t s) {
  hipLaunchKernelGGL(antialias_kernel, grid_for((size_t)nx * ny), dim3(kThreads), 0, s, nx, ny, ex, ey, out);
  SK_HIP(hipGetLastError());
}

// Multislice.propagator (Multislice.swift:108-130) on the compact columns, times `scale`; phase in float64.
__global__ void propagator_kernel(int nx, int ny, double ex, double ey, double lambda, double dz, double tx,
                                  double ty, const double* aperture, const int* keep, int nc, double scale,
                                  float2* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * nc) return;
  const int ix = int(i / nc), c = int(i % nc);
  const int iy = keep[c];
  if (iy < 0) {
    out[i] = make_float2(0.f, 0.f);
    return;
  }
  const double sx = ex / nx, sy = ey / ny;
  const double kx = fftfreq_d(ix, nx, sx), ky = fftfreq_d(iy, ny, sy);
  const double k2 = __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
  const double x = -k2 * M_PI * dz * lambda;
  const double ap = aperture[(size_t)ix * ny + iy];
  double sn, cs;
  sincos(x, &sn, &cs);
  double re = cs * ap, im = sn * ap;
  if (tx != 0 || ty != 0) {
    const double t = -2 * M_PI * dz * (kx * tan(tx / 1e3) + ky * tan(ty / 1e3));
    double st, ct;
    sincos(t, &st, &ct);
    const double nr = re * ct - im * st, ni = re * st + im * ct;
    re = nr;
    im = ni;
  }
  out[i] = make_float2((float)(re * scale), (float)(im * scale));
}
void launch_propagator(int nx, int ny, double ex, double ey, double lambda, double dz, double tx, double ty,
                       const double* aperture, const int* keep, int nc, double scale, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(propagator_kernel, grid_for((size_t)nx * nc), dim3(kThreads), 0, s, nx, ny, ex, ey, lambda, dz,
                     tx, ty, aperture, keep, nc, scale, out);
  SK_HIP(hipGetLastError());
}

__global__ void aperture_as_complex_kernel(size_t n, const double* aperture, double scale, float2* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = make_float2((float)(aperture[i] * scale), 0.f);
}
void launch_aperture_as_complex(int nx, int ny, const double* aperture, double scale, float2* out, hipStream_t s) {
  const size_t n = (size_t)nx * ny;
  hipLaunchKernelGGL(aperture_as_complex_kernel, grid_for(n), dim3(kThreads), 0, s, n, aperture, scale, out);
  SK_HIP(hipGetLastError());
}

// chi(alpha, phi) (Aberrations.swift:39-59), coefficients in the sk_probe order
__device__ double chi_d(const double* c, double alpha, double phi, double lambda) {
  double x = 0.0;
  if (c[0] != 0 || c[1] != 0 || c[2] != 0) x = x + 1.0 / 2 * pow(alpha, 2) * (c[0] + c[1] * cos(2 * (phi - c[2])));
  if (c[3] != 0 || c[4] != 0 || c[5] != 0 || c[6] != 0)
    x = x + 1.0 / 3 * pow(alpha, 3) * (c[3] * cos(phi - c[4]) + c[5] * cos(3 * (phi - c[6])));
  if (c[7] != 0 || c[8] != 0 || c[9] != 0 || c[10] != 0 || c[11] != 0)
    x = x + 1.0 / 4 * pow(alpha, 4) * (c[7] + c[8] * cos(2 * (phi - c[9])) + c[10] * cos(4 * (phi - c[11])));
  if (c[12] != 0 || c[13] != 0 || c[14] != 0 || c[15] != 0 || c[16] != 0 || c[17] != 0)
    x = x + 1.0 / 5 * pow(alpha, 5) *
                (c[12] * cos(phi - c[13]) + c[14] * cos(3 * (phi - c[15])) + c[16] * cos(5 * (phi - c[17])));
  if (c[18] != 0 || c[19] != 0 || c[20] != 0 || c[21] != 0 || c[22] != 0 || c[23] != 0 || c[24] != 0)
    x = x + 1.0 / 6 * pow(alpha, 6) *
                (c[18] + c[19] * cos(2 * (phi - c[20])) + c[21] * cos(4 * (phi - c[22])) +
                 c[23] * cos(6 * (phi - c[24])));
  return x * (2 * M_PI / lambda);
}

struct Aberr {
  double c[25];
};

// Probe.centredSpectrum before normalisation (Probe.swift:113-129; Aperture.values, Probe.swift:25-48)
// Aperture.values (Aperture.swift:25-48) at frequency (ix, iy): 1 inside alpha_c; the soft edge
// clip((alpha_c - alpha) / d + 1/2, 0, 1) with d the angular step along the direction phi, A(0) = 1; +inf = no aperture
__device__ double aperture_d(int ix, int iy, int nx, int ny, double sx, double sy, double alpha, double phi,
                             double lambda, double cutoff_mrad, int soft) {
  if (isinf(cutoff_mrad)) return 1.0;
  const double cutoff = cutoff_mrad * 1e-3;
  if (!soft) return alpha <= cutoff ? 1.0 : 0.0;
  if (ix == 0 && iy == 0) return 1.0;
  const double rsx = 1 / (double(nx) * sx), rsy = 1 / (double(ny) * sy);
  const double dax = rsx * 1e3 * lambda * 1e-3, day = rsy * 1e3 * lambda * 1e-3;
  const double a = cos(phi) * dax, b = sin(phi) * day;
  const double d = sqrt(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)));
  return fmin(fmax((cutoff - alpha) / d + 0.5, 0.0), 1.0);
}

__global__ void probe_spectrum_kernel(int nx, int ny, double ex, double ey, double lambda, double cutoff_mrad,
                                      int soft, Aberr ab, double* re, double* im) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int ix = int(i / ny), iy = int(i % ny);
  const double sx = ex / nx, sy = ey / ny;
  const double kx = fftfreq_d(ix, nx, sx), ky = fftfreq_d(iy, ny, sy);
  const double alpha = sqrt(__dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky))) * lambda;
  const double phi = atan2(ky, kx);
  const double ap = aperture_d(ix, iy, nx, ny, sx, sy, alpha, phi, lambda, cutoff_mrad, soft);
  const double ch = chi_d(ab.c, alpha, phi, lambda);
  double sn, cs;
  sincos(-ch, &sn, &cs);
  re[i] = ap * cs;
  im[i] = ap * sn;
}
void launch_probe_spectrum(int nx, int ny, double ex, double ey, double lambda, double cutoff, int soft,
                           const double* abv, double* re, double* im, hipStream_t s) {
  Aberr ab;
  for (int k = 0; k < 25; ++k) ab.c[k] = abv[k];
  hipLaunchKernelGGL(probe_spectrum_kernel, grid_for((size_t)nx * ny), dim3(kThreads), 0, s, nx, ny, ex, ey, lambda,
                     cutoff, soft, ab, re, im);
  SK_HIP(hipGetLastError());
}

// abTEM's dchi/dalpha and azimuthal term dchi/dphi (ContrastTransferFunction.swift:12-33), coefficients in the
// sk_probe order C10 C12 phi12 C21 phi21 C23 phi23 C30 C32 phi32 C34 phi34 C41 phi41 C43 phi43 C45 phi45 C50 C52
// phi52 C54 phi54 C56 phi56
__device__ void chi_derivatives_d(const double* c, double a, double p, double lambda, double& dk, double& dp) {
  const double a2 = a * a, a3 = a2 * a, a4 = a3 * a, a5 = a4 * a;
  const double t = (c[1] * cos(2 * (p - c[2])) + c[0]) * a +
                   (c[5] * cos(3 * (p - c[6])) + c[3] * cos(p - c[4])) * a2 +
                   (c[10] * cos(4 * (p - c[11])) + c[8] * cos(2 * (p - c[9])) + c[7]) * a3 +
                   (c[16] * cos(5 * (p - c[17])) + c[14] * cos(3 * (p - c[15])) + c[12] * cos(p - c[13])) * a4 +
                   (c[23] * cos(6 * (p - c[24])) + c[21] * cos(4 * (p - c[22])) + c[19] * cos(2 * (p - c[20])) +
                    c[18]) * a5;
  dk = 2 * M_PI / lambda * t;
  const double u = 1 / 2.0 * (2.0 * c[1] * sin(2.0 * (p - c[2]))) * a +
                   1 / 3.0 * (3.0 * c[5] * sin(3.0 * (p - c[6])) + 1.0 * c[3] * sin(1.0 * (p - c[4]))) * a2 +
                   1 / 4.0 * (4.0 * c[10] * sin(4.0 * (p - c[11])) + 2.0 * c[8] * sin(2.0 * (p - c[9]))) * a3 +
                   1 / 5.0 * (5.0 * c[16] * sin(5.0 * (p - c[17])) + 3.0 * c[14] * sin(3.0 * (p - c[15])) +
                              1.0 * c[12] * sin(1.0 * (p - c[13]))) * a4 +
                   (1 / 6.0) * (6.0 * c[23] * sin(6.0 * (p - c[24])) + 4.0 * c[21] * sin(4.0 * (p - c[22])) +
                                2.0 * c[19] * sin(2.0 * (p - c[20]))) * a5;
  dp = -2 * M_PI / lambda * u;
}

// ContrastTransferFunction.kernel (ContrastTransferFunction.swift:84-111) in float64: e^{-i chi} times the spatial
// envelope exp(-sgn(beta) (beta/2)^2 ((dchi/dalpha)^2 + (dchi/dphi)^2)) (beta != 0, mrad / 1e3), the temporal
// envelope exp(-((pi/2/lambda) Delta alpha^2)^2) (Delta != 0) and the aperture; times `scale` into out (complex64)
// and, when re is given, unscaled into re / im (float64).
__global__ void ctf_kernel_kernel(int nx, int ny, double ex, double ey, double lambda, CtfParams c, float scale,
                                  float2* out, double* re, double* im) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int ix = int(i / ny), iy = int(i % ny);
  const double sx = ex / nx, sy = ey / ny;
  const double kx = fftfreq_d(ix, nx, sx), ky = fftfreq_d(iy, ny, sy);
  const double alpha = sqrt(__dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky))) * lambda;
  const double phi = atan2(ky, kx);
  const double x = chi_d(c.ab, alpha, phi, lambda);
  double sn, cs;
  sincos(-x, &sn, &cs);
  double factor = 1.0;
  if (c.angular_spread != 0) {
    const double b = c.angular_spread / 1e3;
    double dk, dp;
    chi_derivatives_d(c.ab, alpha, phi, lambda, dk, dp);
    const double sgn = b > 0 ? 1.0 : (b < 0 ? -1.0 : 0.0);
    factor *= exp(-sgn * (b / 2) * (b / 2) * (dk * dk + dp * dp));
  }
  if (c.focal_spread != 0) {
    const double t = 0.5 * M_PI / lambda * c.focal_spread * alpha * alpha;
    factor *= exp(-(t * t));
  }
  factor *= aperture_d(ix, iy, nx, ny, sx, sy, alpha, phi, lambda, c.cutoff_mrad, c.soft);
  if (out) out[i] = make_float2((float)(cs * factor * scale), (float)(sn * factor * scale));
  if (re) {
    re[i] = cs * factor;
    im[i] = sn * factor;
  }
}
void launch_ctf_kernel(int nx, int ny, double ex, double ey, double lambda, const CtfParams& c, float scale,
                       float2* out, double* re, double* im, hipStream_t s) {
  hipLaunchKernelGGL(ctf_kernel_kernel, grid_for((size_t)nx * ny), dim3(kThreads), 0, s, nx, ny, ex, ey, lambda, c,
                     scale, out, re, im);
  SK_HIP(hipGetLastError());
}

// sum of re^2 + im^2 in float64, deterministic (fixed block partials, then one block)
__global__ void sum_squares_partial_kernel(const double* re, const double* im, size_t n, double* partial) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc += re[i] * re[i] + im[i] * im[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
__global__ void sum_partials_kernel(const double* partial, int blocks, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double acc = 0.0;
    for (int i = 0; i < blocks; ++i) acc += partial[i];
    *out = acc;
  }
}
void launch_sum_squares(const double* re, const double* im, size_t n, double* partial, int blocks, double* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(sum_squares_partial_kernel, dim3(blocks), dim3(kThreads), 0, s, re, im, n, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(64), 0, s, partial, blocks, out);
  SK_HIP(hipGetLastError());
}

// Psi0 / ||Psi0|| * extra, rounded to complex64 (Probe.swift:130-135)
__global__ void normalise_spectrum_kernel(const double* re, const double* im, const double* norm2, double extra,
                                          size_t n, float2* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = sqrt(*norm2);
  out[i] = make_float2((float)(re[i] / f * extra), (float)(im[i] / f * extra));
}
void launch_normalise_spectrum(const double* re, const double* im, const double* norm2, double extra, size_t n,
                               float2* out, hipStream_t s) {
  hipLaunchKernelGGL(normalise_spectrum_kernel, grid_for(n), dim3(kThreads), 0, s, re, im, norm2, extra, n, out);
  SK_HIP(hipGetLastError());
}

// Probe.shiftRamps (Probe.swift:140-156): r_x[ix] = e^{-2 pi i p_x[ix] u_x}, p = fftfreq(n, 1), u = position / sampling
__global__ void ramps_kernel(int nx, int ny, double sx, double sy, const double* positions, int pos0,
                             const int* pos_dev, int count, float2* rx, float2* ry) {
  const int b = blockIdx.y;
  if (b >= count) return;
  const long long q = first_position(pos0, pos_dev) + b;
  const double ux = positions[2 * q] / sx, uy = positions[2 * q + 1] / sy;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx + ny; i += gridDim.x * blockDim.x) {
    double sn, cs;
    if (i < nx) {
      const double a = -2 * M_PI * fftfreq_d(i, nx, 1.0) * ux;
      sincos(a, &sn, &cs);
      rx[(size_t)b * nx + i] = make_float2((float)cs, (float)sn);
    } else {
      const int iy = i - nx;
      const double a = -2 * M_PI * fftfreq_d(iy, ny, 1.0) * uy;
      sincos(a, &sn, &cs);
      ry[(size_t)b * ny + iy] = make_float2((float)cs, (float)sn);
    }
  }
}
void launch_ramps(int nx, int ny, double sx, double sy, const double* positions, int pos0, const int* pos_dev,
                  int count, float2* rx, float2* ry, hipStream_t s) {
  dim3 grid((nx + ny + kThreads - 1) / kThreads, count);
  hipLaunchKernelGGL(ramps_kernel, grid, dim3(kThreads), 0, s, nx, ny, sx, sy, positions, pos0, pos_dev, count, rx,
                     ry);
  SK_HIP(hipGetLastError());
}

// deltas: out[idx[i]] = w[i] (unique indices, merged on the host in float64; out cleared first)
__global__ void scatter_deltas_kernel(const int* idx, const double* w, size_t n, float2* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[idx[i]] = make_float2((float)w[i], 0.f);
}
void launch_scatter_deltas(const int* idx, const double* w, size_t n, float2* out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(scatter_deltas_kernel, grid_for(n), dim3(kThreads), 0, s, idx, w, n, out);
  SK_HIP(hipGetLastError());
}

// f_Z(k^2) / S(k) per element (LobatoParameterization.swift:45-63, ProjectedPotential.swift:130-153), float64
// table: 103 rows of (a1..a5, b1..b5)
__global__ void lobato_kernels_kernel(int nx, int ny, double ex, double ey, const int* zs, int nz, int pi_mode,
                                      const double* table, float* out) {
  const size_t N = (size_t)nx * ny;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (i >= N || e >= nz) return;
  const int ix = int(i / ny), iy = int(i % ny);
  const double dx = ex / nx, dy = ey / ny;
  const double kx = fftfreq_d(ix, nx, dx), ky = fftfreq_d(iy, ny, dy);
  const double k2 = __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
  const double* p = table + 10 * (zs[e] - 1);
  const double kappa = 0.0208865737082965;
  const double pi = pi_mode ? (double)(float)M_PI : M_PI;
  const double pi2 = pi_mode ? (double)(float)(M_PI * M_PI) : M_PI * M_PI;
  const double fourPi2 = 4 * pi2, eightPi = 8 * pi;
  const double q = fourPi2 * k2;
  double sum = 0.0;
  for (int j = 0; j < 5; ++j) {
    const double A = (M_PI * M_PI) * p[j] / pow(p[5 + j], 1.5) / kappa;
    const double B = 2 * M_PI / sqrt(p[5 + j]);
    const double d = q + B * B;
    sum = sum + (A / B / d + A * B / (d * d));
  }
  const double f = eightPi * sum;
  const double a = kx * dx, b = ky * dy;
  const double k = sqrt(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)));
  const double dk2 = dx * dy;
  const double S = (ix == 0 && iy == 0) ? dk2 : sin(k) / k * dk2;
  out[(size_t)e * N + i] = (float)(f / S);
}
void launch_lobato_kernels(int nx, int ny, double ex, double ey, const int* zs, int nz, int pi_mode,
                           const double* table, float* out, hipStream_t s) {
  dim3 grid((unsigned)(((size_t)nx * ny + kThreads - 1) / kThreads), nz);
  hipLaunchKernelGGL(lobato_kernels_kernel, grid, dim3(kThreads), 0, s, nx, ny, ex, ey, zs, nz, pi_mode, table, out);
  SK_HIP(hipGetLastError());
}

// Finite projection (abTEM's Potential(projection="finite"); no counterpart in the reference): the spectrum of slice
// s is sum_a FFT(bilinear deltas of atom a)(k) * F_{a,s}(k) / S(k), with F the in-plane transform of the atom's 3-D
// Lobato potential V3(r) = sum_i A_i (2/(B_i r) + 1) e^{-B_i r} integrated over the slice's z range [z0, z1]:
//   F = G(z1 - z_a) - G(z0 - z_a),  G(T) = sgn(T) [f(k)/2 - c(k, |T|)],
//   c(k, T) = 2 pi sum_i A_i e^{-T kappa_i} [2 / (B_i kappa_i^2) + B_i (2 + T kappa_i) / kappa_i^4],
//   kappa_i = sqrt(B_i^2 + 4 pi^2 k^2)  (f(k) = 2 G(+inf), the infinite projection).
// One thread per frequency and slice, float64 throughout; the atoms of a slice are those within the host's z cutoff
// (atoms sorted by z, [lo, hi) per slice). The delta transform is the atom's corner phase times
// w0 + w1 e^{-2 pi i ix/nx} + w2 e^{-2 pi i iy/ny} + w3 e^{-2 pi i (ix/nx + iy/ny)}, the corner phase taken from integer
// products reduced modulo the grid, as the deltas land modulo the grid.
__device__ double lobato_slab(const double* AB, double pi, double q, double t0, double t1) {
  double f = 0, c0 = 0, c1 = 0;
  const double T0 = fabs(t0), T1 = fabs(t1);
  for (int i = 0; i < 5; ++i) {
    const double A = AB[i], B = AB[5 + i];
    const double k2 = B * B + q, kap = sqrt(k2), k4 = k2 * k2;
    f += A / B / k2 + A * B / k4;
    c0 += A * exp(-T0 * kap) * (2 / (B * k2) + B * (2 + T0 * kap) / k4);
    c1 += A * exp(-T1 * kap) * (2 / (B * k2) + B * (2 + T1 * kap) / k4);
  }
  f *= 8 * pi;
  c0 *= 2 * pi;
  c1 *= 2 * pi;
  const double s1 = double((t1 > 0) - (t1 < 0)), s0 = double((t0 > 0) - (t0 < 0));
  return s1 * (f / 2 - c1) - s0 * (f / 2 - c0);
}

__global__ void finite_slices_kernel(int nx, int ny, double ex, double ey, int pi_mode, FiniteAtoms at,
                                     const double* z0, const double* z1, const int* lo, const int* hi, int cnt,
                                     float2* out) {
  const size_t N = (size_t)nx * ny;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int sl = blockIdx.y;
  if (i >= N || sl >= cnt) return;
  const int ix = int(i / ny), iy = int(i % ny);
  const double dx = ex / nx, dy = ey / ny;
  const double kx = fftfreq_d(ix, nx, dx), ky = fftfreq_d(iy, ny, dy);
  const double pi = pi_mode ? (double)(float)M_PI : M_PI;
  const double pi2 = pi_mode ? (double)(float)(M_PI * M_PI) : M_PI * M_PI;
  const double q = 4 * pi2 * __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
  const double a = kx * dx, b = ky * dy;
  const double kk = sqrt(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)));
  const double S = (ix == 0 && iy == 0) ? dx * dy : sin(kk) / kk * (dx * dy);
  double sxr, sxi, syr, syi;  // e^{-2 pi i ix/nx}, e^{-2 pi i iy/ny}
  sincospi(-2.0 * ix / nx, &sxi, &sxr);
  sincospi(-2.0 * iy / ny, &syi, &syr);
  const double sxyr = sxr * syr - sxi * syi, sxyi = sxr * syi + sxi * syr;
  double accr = 0, acci = 0;
  const double zl = z0[sl], zh = z1[sl];
  for (int k = lo[sl]; k < hi[sl]; ++k) {
    const int rx = int(((long long)at.r[k] * ix) % nx), cy = int(((long long)at.c[k] * iy) % ny);
    double pr, pim;
    sincospi(-2.0 * ((double)rx / nx + (double)cy / ny), &pim, &pr);
    const double* w = at.w + 4 * (size_t)k;
    const double dr = w[0] + w[1] * sxr + w[2] * syr + w[3] * sxyr;
    const double di = w[1] * sxi + w[2] * syi + w[3] * sxyi;
    const double F = lobato_slab(at.AB + 10 * at.e[k], pi, q, zl - at.z[k], zh - at.z[k]);
    accr += F * (pr * dr - pim * di);
    acci += F * (pr * di + pim * dr);
  }
  out[(size_t)sl * N + i] = make_float2((float)(accr / S), (float)(acci / S));
}
void launch_finite_slices(int nx, int ny, double ex, double ey, int pi_mode, const FiniteAtoms& at, const double* z0,
                          const double* z1, const int* lo, const int* hi, int cnt, float2* out, hipStream_t s) {
  dim3 grid((unsigned)(((size_t)nx * ny + kThreads - 1) / kThreads), cnt);
  hipLaunchKernelGGL(finite_slices_kernel, grid, dim3(kThreads), 0, s, nx, ny, ex, ey, pi_mode, at, z0, z1, lo, hi,
                     cnt, out);
  SK_HIP(hipGetLastError());
}

// Spec[s] = sum_e K_e * D[s, e]  (ProjectedPotential.swift:164-171, linearity of the inverse transform)
__global__ void combine_elements_kernel(const float2* D, const float* K, int nslices, int nz, size_t N, float2* out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= N || s >= nslices) return;
  float2 acc = make_float2(0.f, 0.f);
  for (int e = 0; e < nz; ++e) {
    const float k = K[(size_t)e * N + i];
    const float2 d = D[((size_t)s * nz + e) * N + i];
    acc.x += k * d.x;
    acc.y += k * d.y;
  }
  out[(size_t)s * N + i] = acc;
}
void launch_combine_elements(const float2* D, const float* K, int nslices, int nz, size_t N, float2* out,
                             hipStream_t s) {
  dim3 grid((unsigned)((N + kThreads - 1) / kThreads), nslices);
  hipLaunchKernelGGL(combine_elements_kernel, grid, dim3(kThreads), 0, s, D, K, nslices, nz, N, out);
  SK_HIP(hipGetLastError());
}

__global__ void real_part_kernel(const float2* in, double* out, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (double)in[i].x;
}
void launch_real_part(const float2* in, double* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(real_part_kernel, grid_for(n), dim3(kThreads), 0, s, in, out, n);
  SK_HIP(hipGetLastError());
}

// out[b][dx][dy] = in[b][srcx[dx]][srcy[dy]], zero where a source is -1 (a Fourier crop or zero-pad)
__global__ void resample_gather_kernel(const float2* in, int nx, int ny, const int* srcx, const int* srcy, int rows,
                                       int cols, size_t batch, float2* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t M = (size_t)rows * cols;
  if (i >= batch * M) return;
  const size_t b = i / M, r = i % M;
  const int dx = int(r / cols), dy = int(r % cols);
  const int sx = srcx[dx], sy = srcy[dy];
  out[i] = (sx >= 0 && sy >= 0) ? in[b * (size_t)nx * ny + (size_t)sx * ny + sy] : make_float2(0.f, 0.f);
}
void launch_resample_gather(const float2* in, int nx, int ny, const int* srcx, const int* srcy, int rows, int cols,
                            size_t batch, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(resample_gather_kernel, grid_for(batch * rows * cols), dim3(kThreads), 0, s, in, nx, ny, srcx,
                     srcy, rows, cols, batch, out);
  SK_HIP(hipGetLastError());
}

// out = acc * inv (float64 or float32)
__global__ void finalise_kernel(const double* acc, size_t n, double inv, void* out, int f32) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = acc[i] * inv;
  if (f32) static_cast<float*>(out)[i] = (float)v;
  else static_cast<double*>(out)[i] = v;
}
void launch_finalise_outputs(const double* acc, size_t n, double inv, void* out, int f32, hipStream_t s) {
  hipLaunchKernelGGL(finalise_kernel, grid_for(n), dim3(kThreads), 0, s, acc, n, inv, out, f32);
  SK_HIP(hipGetLastError());
}

// out[pos * stride + j] = acc[pos * nslots + slot] * inv  (one scalar slot of every position)
__global__ void gather_scalar_kernel(const double* acc, int nslots, int slot, int npos, double inv, void* out,
                                     int stride, int f32) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npos) return;
  const double v = acc[(size_t)p * nslots + slot] * inv;
  if (f32) static_cast<float*>(out)[(size_t)p * stride] = (float)v;
  else static_cast<double*>(out)[(size_t)p * stride] = v;
}
void launch_gather_scalar(const double* acc, int nslots, int slot, int npos, double inv, void* out, int stride,
                          int f32, hipStream_t s) {
  hipLaunchKernelGGL(gather_scalar_kernel, dim3((npos + kThreads - 1) / kThreads), dim3(kThreads), 0, s, acc, nslots,
                     slot, npos, inv, out, stride, f32);
  SK_HIP(hipGetLastError());
}

// words[0 .. n) = 0: clears device memory with a kernel (no memset node, so a captured stream replays it correctly)
__global__ void fill_zero_kernel(uint32_t* words, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) words[i] = 0u;
}
void launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / sizeof(uint32_t);
  if (bytes % sizeof(uint32_t)) fail(SK_ERR_INVALID, "fill_zero: size is not a multiple of 4 bytes");
  if (!n) return;
  hipLaunchKernelGGL(fill_zero_kernel, grid_for(n), dim3(kThreads), 0, s, static_cast<uint32_t*>(p), n);
  SK_HIP(hipGetLastError());
}

// *flag = 1 when the float64 spectrum (re, im) [nx][ny] is non-zero on a column the bandlimit prunes
// (keepinv[iy] < 0); the caller clears *flag first. Every writer stores the same value.
__global__ void pruned_support_kernel(const double* re, const double* im, const int* keepinv, int nx, int ny,
                                      int* flag) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  if (keepinv[i % ny] < 0 && (re[i] != 0.0 || im[i] != 0.0)) *flag = 1;
}
void launch_pruned_support(const double* re, const double* im, const int* keepinv, int nx, int ny, int* flag,
                           hipStream_t s) {
  hipLaunchKernelGGL(pruned_support_kernel, grid_for((size_t)nx * ny), dim3(kThreads), 0, s, re, im, keepinv, nx,
                     ny, flag);
  SK_HIP(hipGetLastError());
}

}  // namespace sk
