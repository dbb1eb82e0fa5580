// simulato_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// A float64 CPU restatement of SimulatoKit's multislice STEM path (the reference's `.cpu` backend), against which the
// tests hold the HIP library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
// product library never links, loads or calls it.
//
// Every function cites the reference file:line it follows (paths relative to the SimulatoKit checkout). The
// reference's CPU FFT is Apple vDSP (closed, macOS only); this file has its own float64 FFT (Stockham mixed radix
// with radix-2/3/4/5 butterflies, generic odd radix, Bluestein for lengths with a prime factor above 13), which any
// exact DFT may replace (the fixtures pin values to 1e-10, not bits).
//
// Pinning (what checks each part; see tests/test_oracle_*.py):
//   constants, lambda, sigma, f(k2)          slice-1 fixtures `beam`, `f` (1e-14)
//   slicing, V, t, multislice, probe          slice-1 arrays A1/A2/B1/B2/C1/D1 (1e-10 with abTEM's float32 pi)
//   odd grids, tilt                           slice-4 arrays (1e-7)
//   STEM scan, annular/pixelated detectors    slice-2 arrays S1/S2/S3, S4 HAADF (1e-10)
//   frozen phonons (given displacements)      slice-2 F1 (1e-10)
//   pixelated cutoff crop, centre of mass     slice-4-scan dp, com_mrad, com_inv (1e-6 / 1e-5)
//   FFT                                       direct O(n^2) DFT, numpy.fft (1e-12)
//   GaussianGenerator                         published SplitMix64 / xoshiro256** known answers + statistics

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include <chrono>

using cd = std::complex<double>;

namespace {

// ---------------------------------------------------------------------------------------------------------------
// Constants (PhysicalConstants.swift:8-31)
// ---------------------------------------------------------------------------------------------------------------
constexpr double kPlanck = 6.62607004e-34;
constexpr double kElectronMass = 9.10938356e-31;
constexpr double kSpeedOfLight = 299792458.0;
constexpr double kElementaryCharge = 1.6021766208e-19;
constexpr double kKappa = 0.0208865737082965;
constexpr double kAseKg = 6.0221408585491615e+26;
constexpr double kAseCoulomb = 6.241509125883258e+18;
constexpr double kAseSecond = 98226947884640.62;
constexpr double kAseJoule = 6.241509125883258e+18;
constexpr double kPi = 3.141592653589793;

struct LobatoRow {
  int present;
  double a[5];
  double b[5];
};
const LobatoRow kLobato[103] = {
#include "../simulatokit_amd/csrc/lobato_table.inc"
};

// ElectronBeam.swift:129-134
double wavelength(double energy) {
  double c2 = kSpeedOfLight * kSpeedOfLight;
  return kPlanck * kSpeedOfLight / std::sqrt(energy * (2 * kElectronMass * c2 / kElementaryCharge + energy)) /
         kElementaryCharge * 1.0e10;
}

// ElectronBeam.swift:115-121
double sigma(double energy) {
  double mass =
      (1 + kElementaryCharge * energy / (kElectronMass * (kSpeedOfLight * kSpeedOfLight))) * kElectronMass;
  double hsj = kPlanck * kAseSecond * kAseJoule;
  return 2 * kPi * mass * kAseKg * kElementaryCharge * kAseCoulomb * wavelength(energy) / (hsj * hsj);
}

// Grid.swift:82-86 (NumPy fftfreq, bit for bit)
std::vector<double> fftfreq(int n, double d) {
  double val = 1.0 / (double(n) * d);
  int positive = (n - 1) / 2 + 1;
  std::vector<double> out(n);
  for (int i = 0; i < n; ++i) out[i] = double(i < positive ? i : i - n) * val;
  return out;
}

struct Grid {
  int nx, ny;
  double ex, ey;
  double sx() const { return ex / nx; }  // Grid.swift:45-46
  double sy() const { return ey / ny; }
  double rx() const { return 1 / (double(nx) * sx()); }  // Grid.swift:52-53
  double ry() const { return 1 / (double(ny) * sy()); }
  int count() const { return nx * ny; }
  std::vector<double> kx() const { return fftfreq(nx, sx()); }
  std::vector<double> ky() const { return fftfreq(ny, sy()); }
};

// ---------------------------------------------------------------------------------------------------------------
// FFT (semantics of FFT2D.swift:198-215: forward unscaled e^{-2 pi i}, inverse scaled 1/(nx ny); rows (y) then
// columns (x)). Algorithm: Stockham autosort, mixed radix; Bluestein (1970) for lengths with a prime factor > 13.
// ---------------------------------------------------------------------------------------------------------------
cd root_of_unity(long long k, long long n, int sign) {
  // e^{sign 2 pi i k / n} with k reduced mod n and the angle taken from the nearest octant, for accuracy.
  k %= n;
  if (k < 0) k += n;
  double a = 2.0 * kPi * double(k) / double(n);
  return cd(std::cos(a), sign * std::sin(a));
}

struct Plan1D {
  int n = 0;
  std::vector<int> radices;
  std::vector<cd> roots;  // e^{-2 pi i k / n}
  bool bluestein = false;
  int m = 0;  // Bluestein length
  std::vector<cd> chirp;  // e^{-pi i k^2 / n}
  std::vector<cd> kernel_hat;  // FFT_m of conj chirp, wrapped
  std::unique_ptr<Plan1D> inner;
};

std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p = 3; p * p <= n; p += 2)
    while (n % p == 0) { f.push_back(p); n /= p; }
  if (n > 1) f.push_back(n);
  return f;
}

void stockham(const Plan1D& p, cd* x, cd* work, int sign);

std::unique_ptr<Plan1D> make_plan(int n) {
  auto p = std::make_unique<Plan1D>();
  p->n = n;
  auto f = factorize(n);
  int largest = f.empty() ? 1 : *std::max_element(f.begin(), f.end());
  if (largest > 13) {
    p->bluestein = true;
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    if (m < 16) m = 16;
    p->m = m;
    p->inner = make_plan(m);
    p->chirp.resize(n);
    for (int k = 0; k < n; ++k) {
      long long k2 = (long long)k * k % (2LL * n);
      double a = kPi * double(k2) / double(n);
      p->chirp[k] = cd(std::cos(a), -std::sin(a));
    }
    std::vector<cd> b(m, cd(0, 0)), w(m);
    b[0] = std::conj(p->chirp[0]);
    for (int k = 1; k < n; ++k) b[k] = b[m - k] = std::conj(p->chirp[k]);
    stockham(*p->inner, b.data(), w.data(), -1);
    p->kernel_hat = b;
  } else {
    p->radices = f;
    p->roots.resize(n);
    for (int k = 0; k < n; ++k) p->roots[k] = root_of_unity(k, n, -1);
  }
  return p;
}

// One Stockham pass per radix r: for i in [0, n/r): k = i mod L, u_j = x[i + j n/r] * w_{L r}^{j k}, DFT_r(u),
// y[(i - k) r + k + j L] = u_j; L *= r.
void stockham(const Plan1D& p, cd* x, cd* work, int sign) {
  const int n = p.n;
  if (n == 1) return;
  cd* src = x;
  cd* dst = work;
  int L = 1;
  std::vector<cd> u(16), v(16);
  for (int r : p.radices) {
    const int stride = n / r;
    const long long step = n / (L * r);  // root index step for w_{L r}
    if (int(u.size()) < r) { u.resize(r); v.resize(r); }
    for (int i = 0; i < stride; ++i) {
      const int k = i % L;
      for (int j = 0; j < r; ++j) {
        cd val = src[i + j * stride];
        if (j && k) {
          cd w = p.roots[(long long)j * k * step % n];
          if (sign > 0) w = std::conj(w);
          val *= w;
        }
        u[j] = val;
      }
      if (r == 2) {
        v[0] = u[0] + u[1];
        v[1] = u[0] - u[1];
      } else if (r == 4) {
        cd a0 = u[0] + u[2], a1 = u[0] - u[2], a2 = u[1] + u[3], a3 = u[1] - u[3];
        cd a3r = sign < 0 ? cd(a3.imag(), -a3.real()) : cd(-a3.imag(), a3.real());  // a3 * (sign i)
        v[0] = a0 + a2;
        v[2] = a0 - a2;
        v[1] = a1 + a3r;
        v[3] = a1 - a3r;
      } else {
        const long long rstep = n / r;
        for (int q = 0; q < r; ++q) {
          cd s = u[0];
          for (int j = 1; j < r; ++j) {
            cd w = p.roots[(long long)j * q % r * rstep];
            if (sign > 0) w = std::conj(w);
            s += u[j] * w;
          }
          v[q] = s;
        }
      }
      const int base = (i - k) * r + k;
      for (int j = 0; j < r; ++j) dst[base + j * L] = v[j];
    }
    std::swap(src, dst);
    L *= r;
  }
  if (src != x) std::memcpy(x, src, sizeof(cd) * n);
}

// Unscaled transform of one line (sign -1 forward, +1 inverse without the 1/n).
void transform_line(const Plan1D& p, cd* x, std::vector<cd>& scratch, int sign) {
  if (!p.bluestein) {
    if ((int)scratch.size() < p.n) scratch.resize(p.n);
    stockham(p, x, scratch.data(), sign);
    return;
  }
  const int n = p.n, m = p.m;
  if ((int)scratch.size() < 2 * m) scratch.resize(2 * m);
  cd* a = scratch.data();
  cd* w = scratch.data() + m;
  for (int k = 0; k < n; ++k) {
    cd c = sign < 0 ? p.chirp[k] : std::conj(p.chirp[k]);
    a[k] = x[k] * c;
  }
  for (int k = n; k < m; ++k) a[k] = 0;
  stockham(*p.inner, a, w, -1);
  for (int k = 0; k < m; ++k) {
    cd kh = p.kernel_hat[k];
    if (sign > 0) kh = std::conj(kh);  // FFT of the conjugated chirp kernel for the inverse direction
    a[k] *= kh;
  }
  // inverse of length m (unscaled) then 1/m
  stockham(*p.inner, a, w, +1);
  const double invm = 1.0 / m;
  for (int k = 0; k < n; ++k) {
    cd c = sign < 0 ? p.chirp[k] : std::conj(p.chirp[k]);
    x[k] = a[k] * invm * c;
  }
}

struct FFT2 {
  int nx, ny;
  std::unique_ptr<Plan1D> px, py;
  FFT2(int nx_, int ny_) : nx(nx_), ny(ny_), px(make_plan(nx_)), py(make_plan(ny_)) {}
  // FFT2D.swift:217-264 (rows along y, then columns along x)
  void run(cd* a, int sign) const {
    std::vector<cd> scratch, col(nx);
    for (int ix = 0; ix < nx; ++ix) transform_line(*py, a + (size_t)ix * ny, scratch, sign);
    for (int iy = 0; iy < ny; ++iy) {
      for (int ix = 0; ix < nx; ++ix) col[ix] = a[(size_t)ix * ny + iy];
      transform_line(*px, col.data(), scratch, sign);
      for (int ix = 0; ix < nx; ++ix) a[(size_t)ix * ny + iy] = col[ix];
    }
    if (sign > 0) {
      const double s = 1.0 / (double(nx) * double(ny));
      for (size_t i = 0; i < (size_t)nx * ny; ++i) a[i] *= s;
    }
  }
  void forward(cd* a) const { run(a, -1); }
  void inverse(cd* a) const { run(a, +1); }
};

// ---------------------------------------------------------------------------------------------------------------
// Antialias (Antialias.swift:18-52) and detector geometry (Detectors.swift:39-67, Waves.swift:90-96)
// ---------------------------------------------------------------------------------------------------------------
std::vector<double> antialias(const Grid& g) {
  double dmax = std::max(g.sx(), g.sy());
  double kc = 0.6666666 / dmax / 2;
  double w = 0.01 / dmax;
  auto kx = g.kx(), ky = g.ky();
  std::vector<double> out(g.count());
  for (int ix = 0; ix < g.nx; ++ix)
    for (int iy = 0; iy < g.ny; ++iy) {
      double r = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy]);
      double v;
      if (r > kc) v = 0;
      else if (r > kc - w) v = 0.5 * (1 + std::cos(kPi * (r - kc + w) / w));
      else v = 1;
      out[(size_t)ix * g.ny + iy] = v;
    }
  return out;
}

double round_decimals(double x, int d) {
  double f = std::pow(10.0, double(d));
  return std::nearbyint(x * f) / f;  // default rounding: to nearest, ties to even
}

// Detectors.swift:54-62 (grid parity) and Antialias.swift:209-218 (odd parity)
void cutoff_gpts(const Grid& g, bool odd, int& mx, int& my) {
  double kcut = 2.0 / 3.0 / std::max(g.sx(), g.sy());
  auto size = [&](int n, double extent) {
    int m = (int)std::floor(round_decimals(kcut * extent, 7));
    if (m % 2 != n % 2) m += 1;
    if (odd && m % 2 == 0) m += 1;
    return m;
  };
  mx = size(g.nx, double(g.nx) * g.sx());
  my = size(g.ny, double(g.ny) * g.sy());
}

std::vector<int> crop_indices(int n, int m) {
  std::vector<int> out;
  if (m >= n) {
    for (int i = 0; i < n; ++i) out.push_back(i);
    return out;
  }
  int low = m % 2 == 1 ? m / 2 + 1 : m / 2;
  int high = m / 2;
  for (int i = 0; i < low; ++i) out.push_back(i);
  for (int i = n - high; i < n; ++i) out.push_back(i);
  return out;
}

// Detectors.swift:70-80 — crop in FFT order, then fftshift
std::vector<double> cropped(const Grid& g, const double* values, int mx, int my) {
  auto sx = crop_indices(g.nx, mx), sy = crop_indices(g.ny, my);
  int cx_n = (int)sx.size(), cy_n = (int)sy.size();
  std::vector<double> out((size_t)cx_n * cy_n);
  for (int cx = 0; cx < cx_n; ++cx) {
    int ox = (cx + cx_n / 2) % cx_n;
    for (int cy = 0; cy < cy_n; ++cy)
      out[(size_t)ox * cy_n + (cy + cy_n / 2) % cy_n] = values[(size_t)sx[cx] * g.ny + sy[cy]];
  }
  return out;
}

// ---------------------------------------------------------------------------------------------------------------
// Lobato projected scattering factor (LobatoParameterization.swift:38-69)
// ---------------------------------------------------------------------------------------------------------------
bool lobato_supports(int z) { return z >= 1 && z <= 103 && kLobato[z - 1].present; }

void lobato_f(int z, int pi_mode, const double* k2, size_t n, double* out) {
  const LobatoRow& p = kLobato[z - 1];
  double A[5], B[5];
  double pi2e = kPi * kPi;
  for (int i = 0; i < 5; ++i) {
    A[i] = pi2e * p.a[i] / std::pow(p.b[i], 1.5) / kKappa;
    B[i] = 2 * kPi / std::sqrt(p.b[i]);
  }
  double pi, pi2;
  if (pi_mode == 1) {  // abTEM's float32 pi (LobatoParameterization.swift:46-47)
    pi = double(float(kPi));
    pi2 = double(float(kPi * kPi));
  } else {
    pi = kPi;
    pi2 = kPi * kPi;
  }
  double fourPi2 = 4 * pi2, eightPi = 8 * pi;
  for (size_t j = 0; j < n; ++j) {
    double q = fourPi2 * k2[j];
    double sum = 0.0;
    for (int i = 0; i < 5; ++i) {
      double d = q + B[i] * B[i];
      sum = sum + (A[i] / B[i] / d + A[i] * B[i] / (d * d));
    }
    out[j] = eightPi * sum;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Slicing (Slicing.swift:17-104)
// ---------------------------------------------------------------------------------------------------------------
struct Atom {
  int z;
  double x, y, zc;
};

std::vector<Atom> wrapped(const std::vector<Atom>& atoms, const double box[3]) {
  std::vector<Atom> out = atoms;
  for (auto& a : out) {
    double* p[3] = {&a.x, &a.y, &a.zc};
    for (int d = 0; d < 3; ++d) {
      double f = std::fmod(*p[d] / box[d], 1.0);
      if (f < 0) f += 1.0;
      if (f == 1.0) f = 0;
      *p[d] = f * box[d];
    }
    if (a.zc > box[2] - 1e-10) a.zc = 0;
  }
  return out;
}

std::vector<int> slice_indices(const std::vector<Atom>& atoms, const std::vector<double>& thick) {
  std::vector<double> edges;
  double acc = 0.0;
  for (double s : thick) {
    acc += s;
    edges.push_back(acc);
  }
  for (auto& e : edges) e = e - 1e-12;
  std::vector<int> out;
  for (auto& a : atoms) {
    int lo = 0, hi = (int)edges.size();
    while (lo < hi) {
      int mid = (lo + hi) / 2;
      if (edges[mid] <= a.zc) lo = mid + 1;
      else hi = mid;
    }
    out.push_back(std::min(lo, (int)thick.size() - 1));
  }
  return out;
}

// ---------------------------------------------------------------------------------------------------------------
// Projected potential (ProjectedPotential.swift:117-201)
// ---------------------------------------------------------------------------------------------------------------
struct SliceBuilder {
  Grid g;
  int pi_mode;
  FFT2 fft;
  std::vector<double> k2, comp;
  std::map<int, std::vector<double>> kernels;
  SliceBuilder(const Grid& g_, int pi_mode_) : g(g_), pi_mode(pi_mode_), fft(g_.nx, g_.ny) {
    auto kx = g.kx(), ky = g.ky();
    k2.resize(g.count());
    for (int ix = 0; ix < g.nx; ++ix)
      for (int iy = 0; iy < g.ny; ++iy) k2[(size_t)ix * g.ny + iy] = kx[ix] * kx[ix] + ky[iy] * ky[iy];
    // abTEM's sinc compensation S(k) (ProjectedPotential.swift:130-143)
    double dx = g.sx(), dy = g.sy(), dk2 = dx * dy;
    comp.resize(g.count());
    for (int ix = 0; ix < g.nx; ++ix)
      for (int iy = 0; iy < g.ny; ++iy) {
        double a = kx[ix] * dx, b = ky[iy] * dy;
        double k = std::sqrt(a * a + b * b);
        comp[(size_t)ix * g.ny + iy] = (ix == 0 && iy == 0) ? dk2 : std::sin(k) / k * dk2;
      }
  }
  const std::vector<double>& kernel(int z) {
    auto it = kernels.find(z);
    if (it != kernels.end()) return it->second;
    std::vector<double> f(g.count());
    lobato_f(z, pi_mode, k2.data(), k2.size(), f.data());
    for (size_t i = 0; i < f.size(); ++i) f[i] = f[i] / comp[i];
    return kernels[z] = f;
  }
  // superposeDeltas (ProjectedPotential.swift:178-200): NumPy add.at order, weight-major then atom
  void superpose(const std::vector<Atom>& atoms, std::vector<double>& arr) const {
    size_t n = atoms.size();
    std::vector<int> rows(n), cols(n);
    std::vector<double> w[4];
    for (auto& v : w) v.resize(n);
    for (size_t a = 0; a < n; ++a) {
      double px = atoms[a].x / g.sx(), py = atoms[a].y / g.sy();
      int r = (int)std::floor(px), c = (int)std::floor(py);
      double x = px - r, y = py - c, xy = x * y;
      rows[a] = r;
      cols[a] = c;
      w[0][a] = 1 + xy - y - x;
      w[1][a] = x - xy;
      w[2][a] = y - xy;
      w[3][a] = xy;
    }
    auto mod = [](int a, int m) { int r = a % m; return r < 0 ? r + m : r; };
    for (int k = 0; k < 4; ++k)
      for (size_t a = 0; a < n; ++a) {
        int i = mod(rows[a] + ((k == 1 || k == 3) ? 1 : 0), g.nx);
        int j = mod(cols[a] + (k >= 2 ? 1 : 0), g.ny);
        arr[(size_t)i * g.ny + j] += w[k][a];
      }
  }
  // potential(of:) (ProjectedPotential.swift:156-174)
  std::vector<double> potential(const std::vector<Atom>& atoms) {
    std::vector<double> v(g.count(), 0.0);
    std::vector<int> zs;
    for (auto& a : atoms) zs.push_back(a.z);
    std::sort(zs.begin(), zs.end());
    zs.erase(std::unique(zs.begin(), zs.end()), zs.end());
    std::vector<cd> field(g.count());
    std::vector<double> re(g.count());
    for (int z : zs) {
      std::vector<Atom> sel;
      for (auto& a : atoms)
        if (a.z == z) sel.push_back(a);
      std::fill(re.begin(), re.end(), 0.0);
      superpose(sel, re);
      for (size_t i = 0; i < re.size(); ++i) field[i] = cd(re[i], 0);
      fft.forward(field.data());
      const auto& k = kernel(z);
      for (size_t i = 0; i < k.size(); ++i) field[i] = cd(field[i].real() * k[i], field[i].imag() * k[i]);
      fft.inverse(field.data());
      for (size_t i = 0; i < v.size(); ++i) v[i] += field[i].real();
    }
    return v;
  }
};

// ---------------------------------------------------------------------------------------------------------------
// Transmission, propagator, multislice (Multislice.swift:83-166)
// ---------------------------------------------------------------------------------------------------------------
std::vector<cd> transmission(const Grid& g, const FFT2& fft, const std::vector<double>& aperture, const double* v,
                             double sig) {
  std::vector<cd> t(g.count());
  for (size_t i = 0; i < t.size(); ++i) {
    double phase = sig * v[i];
    t[i] = cd(std::cos(phase), std::sin(phase));
  }
  fft.forward(t.data());
  for (size_t i = 0; i < t.size(); ++i) t[i] = cd(t[i].real() * aperture[i], t[i].imag() * aperture[i]);
  fft.inverse(t.data());
  return t;
}

std::vector<cd> propagator(const Grid& g, double lambda, double dz, double tx, double ty) {
  auto kx = g.kx(), ky = g.ky();
  auto ap = antialias(g);
  std::vector<cd> p(g.count());
  for (int ix = 0; ix < g.nx; ++ix)
    for (int iy = 0; iy < g.ny; ++iy) {
      size_t i = (size_t)ix * g.ny + iy;
      double k2 = kx[ix] * kx[ix] + ky[iy] * ky[iy];
      double x = -k2 * kPi * dz * lambda;
      double re = std::cos(x) * ap[i], im = std::sin(x) * ap[i];
      if (tx != 0 || ty != 0) {
        double t = -2 * kPi * dz * (kx[ix] * std::tan(tx / 1e3) + ky[iy] * std::tan(ty / 1e3));
        double c = std::cos(t), s = std::sin(t);
        double nr = re * c - im * s, ni = re * s + im * c;
        re = nr;
        im = ni;
      }
      p[i] = cd(re, im);
    }
  return p;
}

inline void cmul_inplace(std::vector<cd>& psi, const std::vector<cd>& t) {
  // Multislice.swift:160-166: (a + ib)(c + id) = (ac - bd) + i(ad + bc)
  for (size_t i = 0; i < psi.size(); ++i) {
    double a = psi[i].real(), b = psi[i].imag(), c = t[i].real(), d = t[i].imag();
    psi[i] = cd(a * c - b * d, a * d + b * c);
  }
}

struct PreparedPotential {
  Grid g;
  double lambda;
  std::vector<std::vector<cd>> t;
  std::vector<double> dz;
  std::map<double, std::vector<cd>> props;  // per distinct dz (Multislice.swift:136,145-148)
};

PreparedPotential prepare(const Grid& g, double energy, int nslices, const double* V, const double* dz, double tx,
                          double ty) {
  PreparedPotential pp;
  pp.g = g;
  pp.lambda = wavelength(energy);
  FFT2 fft(g.nx, g.ny);
  auto ap = antialias(g);
  double sig = sigma(energy);
  for (int j = 0; j < nslices; ++j) {
    pp.t.push_back(transmission(g, fft, ap, V + (size_t)j * g.count(), sig));
    pp.dz.push_back(dz[j]);
    if (!pp.props.count(dz[j])) pp.props[dz[j]] = propagator(g, pp.lambda, dz[j], tx, ty);
  }
  return pp;
}

// runCPU (Multislice.swift:132-157); planes ascending, -1 = incident
void run_cpu(const PreparedPotential& pp, const FFT2& fft, std::vector<cd>& psi, const int* planes, int nplanes,
             cd* out) {
  int q = 0;
  size_t N = pp.g.count();
  if (nplanes > 0 && planes[0] == -1) {
    std::copy(psi.begin(), psi.end(), out);
    q = 1;
  }
  for (size_t j = 0; j < pp.t.size(); ++j) {
    cmul_inplace(psi, pp.t[j]);
    fft.forward(psi.data());
    cmul_inplace(psi, pp.props.at(pp.dz[j]));
    fft.inverse(psi.data());
    if (q < nplanes && planes[q] == (int)j) {
      std::copy(psi.begin(), psi.end(), out + (size_t)q * N);
      ++q;
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Probe (Probe.swift:25-156, Aberrations.swift:39-59)
// ---------------------------------------------------------------------------------------------------------------
struct ProbeParams {
  double energy, cutoff_mrad;
  int soft;
  double ab[25];  // C10 C12 phi12 C21 phi21 C23 phi23 C30 C32 phi32 C34 phi34 C41 phi41 C43 phi43 C45 phi45
                  // C50 C52 phi52 C54 phi54 C56 phi56
};

double chi(const double* c, double alpha, double phi, double lambda) {
  enum { C10, C12, phi12, C21, phi21, C23, phi23, C30, C32, phi32, C34, phi34, C41, phi41, C43, phi43, C45, phi45,
         C50, C52, phi52, C54, phi54, C56, phi56 };
  double x = 0.0;
  if (c[C10] != 0 || c[C12] != 0 || c[phi12] != 0)
    x = x + 1.0 / 2 * std::pow(alpha, 2) * (c[C10] + c[C12] * std::cos(2 * (phi - c[phi12])));
  if (c[C21] != 0 || c[phi21] != 0 || c[C23] != 0 || c[phi23] != 0)
    x = x + 1.0 / 3 * std::pow(alpha, 3) *
                (c[C21] * std::cos(phi - c[phi21]) + c[C23] * std::cos(3 * (phi - c[phi23])));
  if (c[C30] != 0 || c[C32] != 0 || c[phi32] != 0 || c[C34] != 0 || c[phi34] != 0)
    x = x + 1.0 / 4 * std::pow(alpha, 4) *
                (c[C30] + c[C32] * std::cos(2 * (phi - c[phi32])) + c[C34] * std::cos(4 * (phi - c[phi34])));
  if (c[C41] != 0 || c[phi41] != 0 || c[C43] != 0 || c[phi43] != 0 || c[C45] != 0 || c[phi45] != 0)
    x = x + 1.0 / 5 * std::pow(alpha, 5) *
                (c[C41] * std::cos(phi - c[phi41]) + c[C43] * std::cos(3 * (phi - c[phi43])) +
                 c[C45] * std::cos(5 * (phi - c[phi45])));
  if (c[C50] != 0 || c[C52] != 0 || c[phi52] != 0 || c[C54] != 0 || c[phi54] != 0 || c[C56] != 0 ||
      c[phi56] != 0)
    x = x + 1.0 / 6 * std::pow(alpha, 6) *
                (c[C50] + c[C52] * std::cos(2 * (phi - c[phi52])) + c[C54] * std::cos(4 * (phi - c[phi54])) +
                 c[C56] * std::cos(6 * (phi - c[phi56])));
  return x * (2 * kPi / lambda);
}

std::vector<double> aperture_values(const Grid& g, double cutoff_mrad, int soft, double lambda) {
  std::vector<double> out(g.count(), 1.0);
  if (std::isinf(cutoff_mrad)) return out;
  auto kx = g.kx(), ky = g.ky();
  double cutoff = cutoff_mrad * 1e-3;
  double dax = g.rx() * 1e3 * lambda * 1e-3;
  double day = g.ry() * 1e3 * lambda * 1e-3;
  for (int ix = 0; ix < g.nx; ++ix)
    for (int iy = 0; iy < g.ny; ++iy) {
      size_t i = (size_t)ix * g.ny + iy;
      double alpha = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy]) * lambda;
      if (soft) {
        if (ix == 0 && iy == 0) { out[i] = 1; continue; }
        double phi = std::atan2(ky[iy], kx[ix]);
        double a = std::cos(phi) * dax, b = std::sin(phi) * day;
        double d = std::sqrt(a * a + b * b);
        out[i] = std::min(std::max((cutoff - alpha) / d + 0.5, 0.0), 1.0);
      } else {
        out[i] = alpha <= cutoff ? 1 : 0;
      }
    }
  return out;
}

// abTEM's dchi/dalpha and the azimuthal term dchi/dphi (ContrastTransferFunction.swift:12-33)
void chi_derivatives(const double* c, double a, double p, double lambda, double& dk, double& dp) {
  enum { C10, C12, phi12, C21, phi21, C23, phi23, C30, C32, phi32, C34, phi34, C41, phi41, C43, phi43, C45, phi45,
         C50, C52, phi52, C54, phi54, C56, phi56 };
  const double t1 = (c[C12] * std::cos(2 * (p - c[phi12])) + c[C10]) * a;
  const double t2 = (c[C23] * std::cos(3 * (p - c[phi23])) + c[C21] * std::cos(1 * (p - c[phi21]))) * std::pow(a, 2);
  const double t3 = (c[C34] * std::cos(4 * (p - c[phi34])) + c[C32] * std::cos(2 * (p - c[phi32])) + c[C30]) *
                    std::pow(a, 3);
  const double t4 = (c[C45] * std::cos(5 * (p - c[phi45])) + c[C43] * std::cos(3 * (p - c[phi43])) +
                     c[C41] * std::cos(1 * (p - c[phi41]))) * std::pow(a, 4);
  const double t5 = (c[C56] * std::cos(6 * (p - c[phi56])) + c[C54] * std::cos(4 * (p - c[phi54])) +
                     c[C52] * std::cos(2 * (p - c[phi52])) + c[C50]) * std::pow(a, 5);
  dk = 2 * kPi / lambda * (t1 + t2 + t3 + t4 + t5);
  const double u1 = 1 / 2.0 * (2.0 * c[C12] * std::sin(2.0 * (p - c[phi12]))) * a;
  const double u2 = 1 / 3.0 * (3.0 * c[C23] * std::sin(3.0 * (p - c[phi23])) + 1.0 * c[C21] * std::sin(1.0 * (p - c[phi21]))) *
                    std::pow(a, 2);
  const double u3 = 1 / 4.0 * (4.0 * c[C34] * std::sin(4.0 * (p - c[phi34])) + 2.0 * c[C32] * std::sin(2.0 * (p - c[phi32]))) *
                    std::pow(a, 3);
  const double u4 = 1 / 5.0 * (5.0 * c[C45] * std::sin(5.0 * (p - c[phi45])) + 3.0 * c[C43] * std::sin(3.0 * (p - c[phi43])) +
                               1.0 * c[C41] * std::sin(1.0 * (p - c[phi41]))) * std::pow(a, 4);
  const double u5 = (1 / 6.0) * (6.0 * c[C56] * std::sin(6.0 * (p - c[phi56])) + 4.0 * c[C54] * std::sin(4.0 * (p - c[phi54])) +
                                 2.0 * c[C52] * std::sin(2.0 * (p - c[phi52]))) * std::pow(a, 5);
  dp = -2 * kPi / lambda * (u1 + u2 + u3 + u4 + u5);
}

// ContrastTransferFunction.kernel (ContrastTransferFunction.swift:84-111): e^{-i chi} x E_s (beta != 0) x E_t
// (Delta != 0) x aperture (finite cutoff); E_t = exp(-((pi/2/lambda) Delta alpha^2)^2),
// E_s = exp(-sgn(beta) (beta/2)^2 ((dchi/dalpha)^2 + (dchi/dphi)^2)), beta in mrad / 1e3
std::vector<cd> ctf_kernel(const Grid& g, double lambda, const double* ab, double cutoff_mrad, int soft,
                           double focal_spread, double angular_spread) {
  const bool has_ap = std::isfinite(cutoff_mrad);
  std::vector<double> ap;
  if (has_ap) ap = aperture_values(g, cutoff_mrad, soft, lambda);
  auto kx = g.kx(), ky = g.ky();
  std::vector<cd> h(g.count());
  for (int ix = 0; ix < g.nx; ++ix)
    for (int iy = 0; iy < g.ny; ++iy) {
      const size_t i = (size_t)ix * g.ny + iy;
      const double alpha = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy]) * lambda;
      const double phi = std::atan2(ky[iy], kx[ix]);
      const double x = chi(ab, alpha, phi, lambda);
      double re = std::cos(-x), im = std::sin(-x);
      double factor = 1.0;
      if (angular_spread != 0) {
        const double b = angular_spread / 1e3;
        double dk, dp;
        chi_derivatives(ab, alpha, phi, lambda, dk, dp);
        const double sgn = b > 0 ? 1.0 : (b < 0 ? -1.0 : 0.0);
        factor *= std::exp(-sgn * (b / 2) * (b / 2) * (dk * dk + dp * dp));
      }
      if (focal_spread != 0) {
        const double t = 0.5 * kPi / lambda * focal_spread * alpha * alpha;
        factor *= std::exp(-(t * t));
      }
      if (has_ap) factor *= ap[i];
      h[i] = cd(re * factor, im * factor);
    }
  return h;
}

// Probe.build(at:) (Probe.swift:72-109): Psi(k) = shift * A * e^{-i chi}, normalised, then IFFT
std::vector<cd> probe_build(const Grid& g, const FFT2& fft, const ProbeParams& pr, double posx, double posy) {
  double lambda = wavelength(pr.energy);
  auto kx = g.kx(), ky = g.ky();
  auto ap = aperture_values(g, pr.cutoff_mrad, pr.soft, lambda);
  double ux = posx / g.sx(), uy = posy / g.sy();
  auto px = fftfreq(g.nx, 1.0), py = fftfreq(g.ny, 1.0);
  std::vector<cd> psi(g.count());
  for (int ix = 0; ix < g.nx; ++ix) {
    double sx = -2 * kPi * px[ix] * ux;
    double cx = std::cos(sx), snx = std::sin(sx);
    for (int iy = 0; iy < g.ny; ++iy) {
      size_t i = (size_t)ix * g.ny + iy;
      double sy = -2 * kPi * py[iy] * uy;
      double cy = std::cos(sy), sny = std::sin(sy);
      double re = cx * cy - snx * sny;
      double im = cx * sny + snx * cy;
      re *= ap[i];
      im *= ap[i];
      double alpha = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy]) * lambda;
      double phi = std::atan2(ky[iy], kx[ix]);
      double ch = chi(pr.ab, alpha, phi, lambda);
      double c = std::cos(-ch), s = std::sin(-ch);
      psi[i] = cd(re * c - im * s, re * s + im * c);
    }
  }
  double norm = 0.0;
  for (auto& v : psi) norm += v.real() * v.real() + v.imag() * v.imag();
  double f = std::sqrt(norm);
  for (auto& v : psi) v = cd(v.real() / f, v.imag() / f);
  fft.inverse(psi.data());
  return psi;
}

// ---------------------------------------------------------------------------------------------------------------
// Detectors (Detectors.swift:28-206)
// ---------------------------------------------------------------------------------------------------------------
enum DetKind { kAnnular = 0, kPixelated = 1, kCom = 2, kSegmented = 3 };
enum MaxAngle { kCutoff = 0, kValid = 1, kFull = 2 };

struct DetSpec {
  int kind;
  double inner;
  double outer;  // < 0: default outer
  int max_angle;
  int units;  // COM: 0 mrad, 1 inverse angstrom
  int nr = 1, nphi = 1;  // segmented: radial and azimuthal bins
  double rotation = 0;   // segmented: azimuth origin [rad]
};
DetSpec det_spec(const double* s) {
  DetSpec d{(int)s[0], s[1], s[2], (int)s[3], (int)s[4]};
  d.nr = (int)s[5];
  d.nphi = (int)s[6];
  d.rotation = s[7];
  return d;
}

void pixelated_size(const Grid& g, int max_angle, int& mx, int& my) {
  if (max_angle == kFull) { mx = g.nx; my = g.ny; return; }
  int cx, cy;
  cutoff_gpts(g, false, cx, cy);
  if (max_angle == kCutoff) { mx = cx; my = cy; return; }
  auto v = [](int m, int n) {
    int k = (int)std::floor(round_decimals(double(m) / std::sqrt(2.0), 7));
    if (k % 2 != n % 2) k += 1;
    return k;
  };
  mx = v(cx, g.nx);
  my = v(cy, g.ny);
}

int det_count(const Grid& g, const DetSpec& d) {
  if (d.kind == kAnnular) return 1;
  if (d.kind == kCom) return 2;
  if (d.kind == kSegmented) return d.nr * d.nphi;
  int mx, my;
  pixelated_size(g, d.max_angle, mx, my);
  return mx * my;
}

struct DetError {
  std::string msg;
};

// AnnularDetector.outerAngle (Detectors.swift:108-117)
double annular_outer(const Grid& g, double lambda, const DetSpec& d) {
  double ax = g.rx() * lambda * 1e3, ay = g.ry() * lambda * 1e3;
  double maximum = std::min(double(g.nx / 2) * ax, double(g.ny / 2) * ay);
  if (d.outer < 0) {
    int cx, cy;
    cutoff_gpts(g, false, cx, cy);
    return std::floor(std::min(double(cx / 2) * ax, double(cy / 2) * ay));
  }
  double o = d.outer;
  if (o > maximum && std::fabs(o - maximum) > 1e-5 + 1e-5 * std::fabs(o))
    throw DetError{"outer angle exceeds the maximum simulated angle"};
  return o;
}

void detect(const Grid& g, double lambda, const DetSpec& d, const double* P, double* out) {
  double ax = g.rx() * lambda * 1e3, ay = g.ry() * lambda * 1e3;
  if (d.kind == kAnnular) {
    double o = annular_outer(g, lambda, d);
    if (d.inner > o) throw DetError{"inner angle exceeds the outer angle"};
    auto angx = fftfreq(g.nx, 1 / ax / double(g.nx)), angy = fftfreq(g.ny, 1 / ay / double(g.ny));
    double i2 = d.inner * d.inner, o2 = o * o, sum = 0.0;
    for (int ix = 0; ix < g.nx; ++ix)
      for (int iy = 0; iy < g.ny; ++iy) {
        double a2 = angx[ix] * angx[ix] + angy[iy] * angy[iy];
        if (a2 >= i2 && a2 < o2) sum += P[(size_t)ix * g.ny + iy];
      }
    out[0] = sum;
    return;
  }
  if (d.kind == kSegmented) {
    // SegmentedDetector.detect (Detectors.swift:232-262): bins on the crop that holds the outer angle, in FFT order
    auto m = [&](int n, double dd) {
      int k = int(2 * std::ceil(d.outer / dd)) + 1;
      if (k % 2 != n % 2) k += 1;
      return std::min(k, n);
    };
    const int mx = m(g.nx, ax), my = m(g.ny, ay);
    auto sx = crop_indices(g.nx, mx), sy = crop_indices(g.ny, my);
    auto angx = fftfreq(mx, 1 / ax / double(mx)), angy = fftfreq(my, 1 / ay / double(my));
    std::fill(out, out + d.nr * d.nphi, 0.0);
    const double two_pi = 2 * kPi;
    for (int cx = 0; cx < mx; ++cx)
      for (int cy = 0; cy < my; ++cy) {
        const double alpha = std::sqrt(angx[cx] * angx[cx] + angy[cy] * angy[cy]);
        if (!(alpha >= d.inner && alpha < d.outer)) continue;
        double phi = std::fmod(std::atan2(angy[cy], angx[cx]) - d.rotation, two_pi);
        if (phi < 0) phi += two_pi;
        const int r = (int)std::floor(double(d.nr) * (alpha - d.inner) / (d.outer - d.inner));
        if (r >= d.nr) continue;
        const double f = std::floor(double(d.nphi) * (phi / two_pi));
        const int az = (int)std::min(std::max(f, 0.0), double(d.nphi - 1));
        out[az + r * d.nphi] += P[(size_t)sx[cx] * g.ny + sy[cy]];
      }
    return;
  }
  int mx, my;
  pixelated_size(g, d.max_angle, mx, my);
  auto pat = cropped(g, P, mx, my);
  if (d.kind == kPixelated) {
    std::copy(pat.begin(), pat.end(), out);
    return;
  }
  double per = d.units == 0 ? 1.0 : 1.0 / (lambda * 1e3);
  double cx = 0.0, cy = 0.0;
  for (int ix = 0; ix < mx; ++ix) {
    double axv = double(ix - mx / 2) * ax * per;
    for (int iy = 0; iy < my; ++iy) {
      double value = pat[(size_t)ix * my + iy];
      cx += value * axv;
      cy += value * (double(iy - my / 2) * ay * per);
    }
  }
  out[0] = cx;
  out[1] = cy;
}

// DiffractionIntensity.init (Detectors.swift:28-36): |FFT(psi)|^2, psi / N first for a plane wave
void intensity(const FFT2& fft, std::vector<cd>& psi, bool values_norm, double* out) {
  if (values_norm) {
    double n = double(psi.size());
    for (auto& v : psi) v = cd(v.real() / n, v.imag() / n);
  }
  fft.forward(psi.data());
  for (size_t i = 0; i < psi.size(); ++i) out[i] = psi[i].real() * psi[i].real() + psi[i].imag() * psi[i].imag();
}

thread_local std::string g_err;
double g_scan_seconds = 0.0;  // wall time of the last scan's position loop (the CPU baseline's timed region)

}  // namespace

// =================================================================================================================
// C ABI for the tests (ctypes). Complex arrays are interleaved (re, im) float64, row-major [ix * ny + iy].
// Every function returns 0 on success, -1 on a refusal (message from orc_last_error()).
// =================================================================================================================
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

double orc_wavelength(double e) { return wavelength(e); }
double orc_sigma(double e) { return sigma(e); }

void orc_fftfreq(int n, double d, double* out) {
  auto f = fftfreq(n, d);
  std::copy(f.begin(), f.end(), out);
}

int orc_lobato_f(int z, int pi_mode, const double* k2, long long n, double* out) {
  if (!lobato_supports(z)) { g_err = "unsupported element"; return -1; }
  lobato_f(z, pi_mode, k2, (size_t)n, out);
  return 0;
}

// 1-D unscaled transform of `count` contiguous lines of length n (sign -1 forward, +1 inverse, unscaled)
void orc_fft1(int n, long long count, int sign, double* data) {
  auto p = make_plan(n);
  std::vector<cd> scratch;
  cd* c = reinterpret_cast<cd*>(data);
  for (long long i = 0; i < count; ++i) transform_line(*p, c + (size_t)i * n, scratch, sign);
}

// 2-D FFT (FFT2D.swift:202-215): sign -1 forward (unscaled), +1 inverse (1/(nx ny))
void orc_fft2(int nx, int ny, int sign, double* data) {
  FFT2 f(nx, ny);
  f.run(reinterpret_cast<cd*>(data), sign);
}

void orc_antialias(int nx, int ny, double ex, double ey, double* out) {
  auto a = antialias(Grid{nx, ny, ex, ey});
  std::copy(a.begin(), a.end(), out);
}

// Slicing: wrap atoms and assign slices. xyz: natoms x 3 (in/out: replaced by the wrapped positions).
void orc_slice_atoms(int natoms, double* xyz, const double* box, int nslices, const double* thick,
                     int* slice_of_atom) {
  std::vector<Atom> atoms(natoms);
  for (int a = 0; a < natoms; ++a) atoms[a] = Atom{0, xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
  auto w = wrapped(atoms, box);
  std::vector<double> t(thick, thick + nslices);
  auto s = slice_indices(w, t);
  for (int a = 0; a < natoms; ++a) {
    xyz[3 * a] = w[a].x;
    xyz[3 * a + 1] = w[a].y;
    xyz[3 * a + 2] = w[a].zc;
    slice_of_atom[a] = s[a];
  }
}

// Potential.build (Potential.swift:285-293): V of every slice from the structure. V: nslices x nx x ny.
int orc_potential(int nx, int ny, double ex, double ey, int natoms, const int* Z, const double* xyz,
                  const double* box, int nslices, const double* thick, int pi_mode, double* V,
                  int* slice_of_atom) {
  for (int a = 0; a < natoms; ++a)
    if (!lobato_supports(Z[a])) { g_err = "unsupported element"; return -1; }
  Grid g{nx, ny, ex, ey};
  std::vector<Atom> atoms(natoms);
  for (int a = 0; a < natoms; ++a) atoms[a] = Atom{Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
  auto w = wrapped(atoms, box);
  std::vector<double> t(thick, thick + nslices);
  auto s = slice_indices(w, t);
  if (slice_of_atom) std::copy(s.begin(), s.end(), slice_of_atom);
  SliceBuilder sb(g, pi_mode);
  for (int j = 0; j < nslices; ++j) {
    std::vector<Atom> in;
    for (int a = 0; a < natoms; ++a)
      if (s[a] == j) in.push_back(w[a]);
    auto v = sb.potential(in);
    std::copy(v.begin(), v.end(), V + (size_t)j * g.count());
  }
  return 0;
}

// Finite projection (abTEM's Potential(projection="finite"); SimulatoKit has none, docs/parity.md:58 lists it as
// missing). Slice j holds the Lobato 3-D potential V3(r) = sum_i A_i (2/(B_i r) + 1) e^{-B_i r} of every atom
// integrated over z in [z0_j, z1_j] only, instead of over all z. With t0 = z0_j - z_a, t1 = z1_j - z_a, the in-plane
// transform of that integral is F(k) = G(t1) - G(t0), G odd in T:
//   G(T) = sgn(T) [f(k)/2 - c(k, |T|)],
//   c(k, T) = 2 pi sum_i A_i e^{-T kappa_i} [2 / (B_i kappa_i^2) + B_i (2 + T kappa_i) / kappa_i^4],
//   kappa_i = sqrt(B_i^2 + 4 pi^2 k^2),
// from the 2-D transform at height t of e^{-B r}/r (2 pi e^{-|t| kappa}/kappa) and of e^{-B r} (its -d/dB), integrated
// over t; c(k, 0) = f(k)/2 and G(+inf) = f(k)/2, so an atom deep inside a slab gets its infinite projection f(k).
// In-plane, each atom enters as the infinite projection does (SliceBuilder): its four bilinear deltas, transformed,
// divided by the same S(k); so V_j = Re IFFT(sum_a FFT(delta_a) F_{a,j} / S), and with F = f for the atoms of slice
// j this is exactly the infinite slice. z is not periodic: tails beyond the stack are dropped.
double lobato_slab(const double* A, const double* B, double pi, double q, double t0, double t1) {
  double f = 0, c0 = 0, c1 = 0;
  auto c = [&](int i, double k2, double kap, double T) {
    return A[i] * std::exp(-T * kap) * (2 / (B[i] * k2) + B[i] * (2 + T * kap) / (k2 * k2));
  };
  for (int i = 0; i < 5; ++i) {
    const double k2 = B[i] * B[i] + q, kap = std::sqrt(k2);
    f += A[i] / B[i] / k2 + A[i] * B[i] / (k2 * k2);
    c0 += c(i, k2, kap, std::fabs(t0));
    c1 += c(i, k2, kap, std::fabs(t1));
  }
  f *= 8 * pi;
  c0 *= 2 * pi;
  c1 *= 2 * pi;
  auto sgn = [](double v) { return double((v > 0) - (v < 0)); };
  return sgn(t1) * (f / 2 - c1) - sgn(t0) * (f / 2 - c0);
}

int orc_potential_finite(int nx, int ny, double ex, double ey, int natoms, const int* Z, const double* xyz,
                         const double* box, int nslices, const double* thick, int pi_mode, double* V) {
  for (int a = 0; a < natoms; ++a)
    if (!lobato_supports(Z[a])) { g_err = "unsupported element"; return -1; }
  Grid g{nx, ny, ex, ey};
  std::vector<Atom> atoms(natoms);
  for (int a = 0; a < natoms; ++a) atoms[a] = Atom{Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
  auto w = wrapped(atoms, box);
  SliceBuilder sb(g, pi_mode);  // for S(k)
  FFT2 fft(nx, ny);
  const double pi = pi_mode == 1 ? double(float(kPi)) : kPi;
  const double pi2 = pi_mode == 1 ? double(float(kPi * kPi)) : kPi * kPi;
  const double sx = g.sx(), sy = g.sy();
  auto kx = g.kx(), ky = g.ky();
  // per atom: Lobato A, B and the transform of its bilinear deltas (the four corner phases, weighted)
  std::vector<std::array<double, 10>> AB(natoms);
  std::vector<cd> D((size_t)natoms * g.count());
  for (int a = 0; a < natoms; ++a) {
    const LobatoRow& p = kLobato[w[a].z - 1];
    for (int i = 0; i < 5; ++i) {
      AB[a][i] = kPi * kPi * p.a[i] / std::pow(p.b[i], 1.5) / kKappa;
      AB[a][5 + i] = 2 * kPi / std::sqrt(p.b[i]);
    }
    const double px = w[a].x / sx, py = w[a].y / sy;
    const int r = (int)std::floor(px), c = (int)std::floor(py);
    const double x = px - r, y = py - c, xy = x * y;
    const double wt[4] = {1 + xy - y - x, x - xy, y - xy, xy};
    const int cx[4] = {r, r + 1, r, r + 1}, cy[4] = {c, c, c + 1, c + 1};
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy) {
        cd acc(0, 0);
        for (int q = 0; q < 4; ++q) {
          // the grid index of the corner modulo n (as the delta lands) times the frequency index, modulo n
          const long long mx = ((long long)cx[q] % nx + nx) % nx, my = ((long long)cy[q] % ny + ny) % ny;
          acc += wt[q] * root_of_unity((mx * ix) % nx * ny + (my * iy) % ny * nx, (long long)nx * ny, -1);
        }
        D[(size_t)a * g.count() + (size_t)ix * ny + iy] = acc;
      }
  }
  double z0 = 0;
  std::vector<cd> field(g.count());
  for (int j = 0; j < nslices; ++j) {
    const double z1 = z0 + thick[j];
    std::fill(field.begin(), field.end(), cd(0, 0));
    for (int a = 0; a < natoms; ++a) {
      const double t0 = z0 - w[a].zc, t1 = z1 - w[a].zc;
      for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
          const size_t i = (size_t)ix * ny + iy;
          const double q = 4 * pi2 * (kx[ix] * kx[ix] + ky[iy] * ky[iy]);
          const double F = lobato_slab(AB[a].data(), AB[a].data() + 5, pi, q, t0, t1);
          field[i] += D[(size_t)a * g.count() + i] * (F / sb.comp[i]);
        }
    }
    fft.inverse(field.data());
    for (size_t i = 0; i < field.size(); ++i) V[(size_t)j * g.count() + i] = field[i].real();
    z0 = z1;
  }
  return 0;
}

void orc_transmission(int nx, int ny, double ex, double ey, int nslices, const double* V, double energy,
                      double* t_out) {
  Grid g{nx, ny, ex, ey};
  FFT2 fft(nx, ny);
  auto ap = antialias(g);
  double sig = sigma(energy);
  for (int j = 0; j < nslices; ++j) {
    auto t = transmission(g, fft, ap, V + (size_t)j * g.count(), sig);
    std::memcpy(t_out + (size_t)j * g.count() * 2, t.data(), sizeof(cd) * t.size());
  }
}

void orc_propagator(int nx, int ny, double ex, double ey, double energy, double dz, double tx, double ty,
                    double* out) {
  auto p = propagator(Grid{nx, ny, ex, ey}, wavelength(energy), dz, tx, ty);
  std::memcpy(out, p.data(), sizeof(cd) * p.size());
}

// Probe.build(at:); aberrations: 25 coefficients (C10 C12 phi12 C21 phi21 C23 phi23 C30 C32 phi32 C34 phi34 C41
// phi41 C43 phi43 C45 phi45 C50 C52 phi52 C54 phi54 C56 phi56) or null.
void orc_probe(int nx, int ny, double ex, double ey, double energy, double cutoff_mrad, int soft,
               const double* aberr, double px, double py, double* out) {
  Grid g{nx, ny, ex, ey};
  ProbeParams pr{energy, cutoff_mrad, soft, {0}};
  if (aberr) std::copy(aberr, aberr + 25, pr.ab);
  FFT2 fft(nx, ny);
  auto psi = probe_build(g, fft, pr, px, py);
  std::memcpy(out, psi.data(), sizeof(cd) * psi.size());
}

// Probe.centredSpectrum() (Probe.swift:113-136)
void orc_probe_spectrum(int nx, int ny, double ex, double ey, double energy, double cutoff_mrad, int soft,
                        const double* aberr, double* out) {
  Grid g{nx, ny, ex, ey};
  double ab[25] = {0};
  if (aberr) std::copy(aberr, aberr + 25, ab);
  double lambda = wavelength(energy);
  auto kx = g.kx(), ky = g.ky();
  auto ap = aperture_values(g, cutoff_mrad, soft, lambda);
  std::vector<cd> psi(g.count());
  double norm = 0.0;
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy) {
      size_t i = (size_t)ix * ny + iy;
      double alpha = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy]) * lambda;
      double phi = std::atan2(ky[iy], kx[ix]);
      double ch = chi(ab, alpha, phi, lambda);
      psi[i] = cd(ap[i] * std::cos(-ch), ap[i] * std::sin(-ch));
      norm += psi[i].real() * psi[i].real() + psi[i].imag() * psi[i].imag();
    }
  double f = std::sqrt(norm);
  for (auto& v : psi) v = cd(v.real() / f, v.imag() / f);
  std::memcpy(out, psi.data(), sizeof(cd) * psi.size());
}

// Multislice.run (Multislice.swift:52-59, 132-157) of one incident wave through V (nslices x N) with thicknesses
// dz; planes ascending (-1 = incident). out: nplanes x N complex.
void orc_multislice(int nx, int ny, double ex, double ey, double energy, double tx, double ty, int nslices,
                    const double* V, const double* dz, const double* incident, const int* planes, int nplanes,
                    double* out) {
  Grid g{nx, ny, ex, ey};
  auto pp = prepare(g, energy, nslices, V, dz, tx, ty);
  FFT2 fft(nx, ny);
  std::vector<cd> psi(g.count());
  std::memcpy(psi.data(), incident, sizeof(cd) * psi.size());
  run_cpu(pp, fft, psi, planes, nplanes, reinterpret_cast<cd*>(out));
}

// Multislice.runBatch (MetalMultislice.swift:28-60): exit waves of `count` incident waves, strided over `threads`
// workers (each runs whole waves, bit for bit as a sequential run).
void orc_multislice_batch(int nx, int ny, double ex, double ey, double energy, double tx, double ty, int nslices,
                          const double* V, const double* dz, int count, const double* incidents, int threads,
                          double* out) {
  Grid g{nx, ny, ex, ey};
  auto pp = prepare(g, energy, nslices, V, dz, tx, ty);
  size_t N = g.count();
  int last = nslices - 1;
  int workers = std::max(1, std::min({threads, (int)std::thread::hardware_concurrency(), count}));
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w)
    pool.emplace_back([&, w] {
      FFT2 fft(nx, ny);
      std::vector<cd> psi(N);
      for (int q = w; q < count; q += workers) {
        std::memcpy(psi.data(), incidents + q * N * 2, sizeof(cd) * N);
        run_cpu(pp, fft, psi, &last, 1, reinterpret_cast<cd*>(out) + (size_t)q * N);
      }
    });
  for (auto& t : pool) t.join();
}

// DiffractionIntensity of one wave (normalization: 0 = reciprocalSpace (probe), 1 = values (plane wave)).
void orc_diffraction_intensity(int nx, int ny, const double* wave, int values_norm, double* out) {
  FFT2 fft(nx, ny);
  std::vector<cd> psi(nx * ny);
  std::memcpy(psi.data(), wave, sizeof(cd) * psi.size());
  intensity(fft, psi, values_norm != 0, out);
}

// Waves.diffractionPattern(maxAngle:) (Waves.swift:62-86): max_angle 0 cutoff (odd), 2 full. Returns mx, my.
void orc_diffraction_pattern(int nx, int ny, double ex, double ey, const double* wave, int values_norm,
                             int max_angle, int* mxo, int* myo, double* out) {
  Grid g{nx, ny, ex, ey};
  std::vector<double> P(g.count());
  orc_diffraction_intensity(nx, ny, wave, values_norm, P.data());
  int mx = nx, my = ny;
  if (max_angle == kCutoff) cutoff_gpts(g, true, mx, my);
  *mxo = mx;
  *myo = my;
  if (out) {
    auto c = cropped(g, P.data(), mx, my);
    std::copy(c.begin(), c.end(), out);
  }
}

// The CTF kernel H [nx x ny] interleaved complex (ContrastTransferFunction.swift:84-111); cutoff_mrad may be +inf.
void orc_ctf_kernel(int nx, int ny, double ex, double ey, double energy, const double* ab, double cutoff_mrad,
                    int soft, double focal_spread, double angular_spread, double* out) {
  auto h = ctf_kernel(Grid{nx, ny, ex, ey}, wavelength(energy), ab, cutoff_mrad, soft, focal_spread, angular_spread);
  std::copy(reinterpret_cast<const double*>(h.data()), reinterpret_cast<const double*>(h.data()) + 2 * h.size(), out);
}

// Waves.applying(ctf) (ContrastTransferFunction.swift:140-153): IFFT(FFT(psi) H), in place on `wave`.
void orc_apply_ctf(int nx, int ny, double ex, double ey, double energy, const double* ab, double cutoff_mrad, int soft,
                   double focal_spread, double angular_spread, double* wave) {
  Grid g{nx, ny, ex, ey};
  auto h = ctf_kernel(g, wavelength(energy), ab, cutoff_mrad, soft, focal_spread, angular_spread);
  FFT2 fft(nx, ny);
  cd* w = reinterpret_cast<cd*>(wave);
  fft.forward(w);
  for (size_t i = 0; i < h.size(); ++i) {
    const double a = w[i].real(), b = w[i].imag(), c = h[i].real(), d = h[i].imag();
    w[i] = cd(a * c - b * d, a * d + b * c);
  }
  fft.inverse(w);
}

int orc_pixelated_size(int nx, int ny, double ex, double ey, int max_angle, int* mx, int* my) {
  pixelated_size(Grid{nx, ny, ex, ey}, max_angle, *mx, *my);
  return 0;
}

// One detector on one intensity (Detectors.swift:119-262). spec: kind, inner, outer (<0 default), max_angle, units,
// radial bins, azimuthal bins, rotation
int orc_detect(int nx, int ny, double ex, double ey, double energy, const double* spec, const double* P,
               double* out) {
  DetSpec d = det_spec(spec);
  try {
    detect(Grid{nx, ny, ex, ey}, wavelength(energy), d, P, out);
  } catch (const DetError& e) {
    g_err = e.msg;
    return -1;
  }
  return 0;
}

// STEM.scan(.cpu) (STEM.swift:49-56, 95-115, 137-150): probe over positions through every potential, detectors
// on each exit intensity, summed in potential order and divided by the potential count.
// V: npots blocks of (nslices[k] x N); dz likewise concatenated. dets: ndet x 8 doubles (as orc_detect's spec). out: concatenated per
// detector, npos x count_d each.
int orc_stem_scan(int nx, int ny, double ex, double ey, double energy, double cutoff_mrad, int soft,
                  const double* aberr, double tx, double ty, int npots, const int* nslices, const double* V,
                  const double* dz, int npos, const double* positions, int ndet, const double* dets, int threads,
                  double* out) {
  Grid g{nx, ny, ex, ey};
  size_t N = g.count();
  std::vector<DetSpec> specs;
  std::vector<int> counts;
  std::vector<size_t> offsets;
  size_t off = 0;
  double lambda = wavelength(energy);
  for (int d = 0; d < ndet; ++d) {
    specs.push_back(det_spec(dets + 8 * d));
    counts.push_back(det_count(g, specs.back()));
    offsets.push_back(off);
    off += (size_t)counts.back() * npos;
  }
  // refuse bad detectors before work, as the reference does on its first call
  try {
    for (auto& s : specs)
      if (s.kind == kAnnular) {
        double o = annular_outer(g, lambda, s);
        if (s.inner > o) throw DetError{"inner angle exceeds the outer angle"};
      }
  } catch (const DetError& e) {
    g_err = e.msg;
    return -1;
  }
  std::vector<PreparedPotential> pps;
  size_t voff = 0, doff = 0;
  for (int k = 0; k < npots; ++k) {
    pps.push_back(prepare(g, energy, nslices[k], V + voff, dz + doff, tx, ty));
    voff += (size_t)nslices[k] * N;
    doff += nslices[k];
  }
  ProbeParams pr{energy, cutoff_mrad, soft, {0}};
  if (aberr) std::copy(aberr, aberr + 25, pr.ab);
  std::fill(out, out + off, 0.0);
  int workers = std::max(1, std::min({threads, (int)std::thread::hardware_concurrency(), npos}));
  std::vector<std::thread> pool;
  auto t0 = std::chrono::steady_clock::now();
  for (int w = 0; w < workers; ++w)
    pool.emplace_back([&, w] {
      FFT2 fft(nx, ny);
      std::vector<cd> psi(N);
      std::vector<double> P(N), vals;
      for (int q = w; q < npos; q += workers) {
        auto incident = probe_build(g, fft, pr, positions[2 * q], positions[2 * q + 1]);
        for (int k = 0; k < npots; ++k) {
          psi = incident;
          int last = nslices[k] - 1;
          std::vector<cd> exitw(N);
          run_cpu(pps[k], fft, psi, &last, 1, exitw.data());
          intensity(fft, exitw, false, P.data());
          for (int d = 0; d < ndet; ++d) {
            vals.assign(counts[d], 0.0);
            detect(g, lambda, specs[d], P.data(), vals.data());
            double* o = out + offsets[d] + (size_t)q * counts[d];
            for (int i = 0; i < counts[d]; ++i) o[i] += vals[i];
          }
        }
      }
    });
  for (auto& t : pool) t.join();
  g_scan_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (size_t i = 0; i < off; ++i) out[i] = out[i] / double(npots);
  return 0;
}

double orc_last_scan_seconds(void) { return g_scan_seconds; }

// GaussianGenerator (FrozenPhonons.swift:14-58): `count` standard normals from `seed`.
void orc_gaussian_normals(unsigned long long seed, long long count, double* out) {
  uint64_t x = seed;
  auto splitmix = [&]() {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  uint64_t s[4] = {splitmix(), splitmix(), splitmix(), splitmix()};
  auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
  auto next = [&]() {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  };
  bool has_spare = false;
  double spare = 0;
  for (long long i = 0; i < count; ++i) {
    if (has_spare) {
      out[i] = spare;
      has_spare = false;
      continue;
    }
    double u1 = (double(next() >> 11) + 1) * 0x1p-53;
    double u2 = double(next() >> 11) * 0x1p-53;
    double r = std::sqrt(-2 * std::log(u1));
    double theta = 2 * kPi * u2;
    spare = r * std::sin(theta);
    has_spare = true;
    out[i] = r * std::cos(theta);
  }
}

// The raw xoshiro256** stream seeded by SplitMix64 (for known-answer checks).
void orc_xoshiro_stream(unsigned long long seed, long long count, unsigned long long* out) {
  uint64_t x = seed;
  auto splitmix = [&]() {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  uint64_t s[4] = {splitmix(), splitmix(), splitmix(), splitmix()};
  auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
  for (long long i = 0; i < count; ++i) {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    out[i] = result;
  }
}

// xoshiro256** from a given state (known-answer checks of the step function).
void orc_xoshiro_from_state(const unsigned long long* state, long long count, unsigned long long* out) {
  uint64_t s[4] = {state[0], state[1], state[2], state[3]};
  auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
  for (long long i = 0; i < count; ++i) {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    out[i] = result;
  }
}

void orc_splitmix_stream(unsigned long long seed, long long count, unsigned long long* out) {
  uint64_t x = seed;
  for (long long i = 0; i < count; ++i) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    out[i] = z ^ (z >> 31);
  }
}

// FrozenPhonons.configurationSeed (FrozenPhonons.swift:122-130)
unsigned long long orc_configuration_seed(unsigned long long seed, long long c) {
  auto mix = [](uint64_t v) {
    uint64_t z = v + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  return mix(mix(seed) + (uint64_t)c);
}

}  // extern "C"
