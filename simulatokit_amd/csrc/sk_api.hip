// sk_api.hip — the C ABI (include/simulato_hip.h): potentials, multislice runs and the STEM scan on one MI355X.
//
// The host side here replaces the reference's Metal engine (MetalEngine.swift, MetalMultislice.swift): it prepares
// a potential once (transmission functions and compact propagators resident in HBM), then streams batches of waves
// through the row/column passes of every slice, reducing the detectors inside the final pass.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "sk_internal.h"

// The device memory of sk_hip_scan_noalloc: one array per device (each device that loads this library holds its own
// copy until the process ends). Host code never uses this name as an address: the library's set-up (require()) asks
// the runtime for the array's device address and every kernel argument, copy and clear uses that address.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

struct sk_potential {
  sk_grid grid{};
  int nslices = 0;
  double energy = 0, lambda = 0, sigma = 0;
  double tilt[2] = {0, 0};
  std::vector<double> dz;
  int nkept = 0, nc = 0;  // kept columns, compact width
  std::vector<int> keep, keepinv;
  sk::DevBuf<int> keep_d, keepinv_d;
  sk::DevBuf<float2> t;                // nslices x N
  std::vector<int> prop_of_slice;      // index into the compact propagators
  std::vector<double> prop_dz;
  sk::DevBuf<float2> props;            // nprops x nx x nc, scaled by 1/N
  sk::DevBuf<double> aperture;         // N
  sk::DevBuf<double> V;                // nslices x N (atoms form only)
  bool has_values = false;
};

namespace sk {

std::recursive_mutex& api_lock() {
  static std::recursive_mutex* m = new std::recursive_mutex;  // process lifetime, as the context
  return *m;
}

std::string& last_error() {
  thread_local std::string e;
  return e;
}

namespace {

// ---- process context ------------------------------------------------------------------------------------------
struct Context {
  bool init = false;
  int status = SK_ERR_UNAVAILABLE;
  std::string why;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> side;  // further streams of split scans, created on first use (default priority)
  std::vector<hipStream_t> capture;  // streams a graph is recorded on while the scan's own streams run (ditto)
  bool big_llc = false;   // the device has a 256 MiB last-level cache (MI300 / MI355 class: gfx94x, gfx950)
  DevBuf<double> lobato;  // 103 x 10
  // timing
  bool timing = false;
  sk_timing last{};
  int col_tile = 0;   // the running call's sk_options.col_tile
  int precision = 0;  // the running call's sk_options.precision
  // identity keep maps per length (full-width transforms)
  std::map<int, std::unique_ptr<DevBuf<int>>> identity;
  // device address of the scratch array on the library's device (null when the runtime gave none; scratch_why says
  // why), taken once at set-up with that device current
  unsigned char* scratch_dev = nullptr;
  std::string scratch_why;
};

std::recursive_mutex& lock() { return api_lock(); }

Context& ctx() {
  // process lifetime on purpose: never destroyed, so no HIP call runs after the runtime has shut down
  static Context* c = new Context;
  return *c;
}


struct LobatoRow {
  int present;
  double a[5];
  double b[5];
};
// Lobato & Van Dyck (2014) parameters (a1..a5, b1..b5) per Z = 1 ... 103
const LobatoRow kLobato[103] = {
#include "lobato_table.inc"
};

Context& require() {
  Context& c = ctx();
  if (!c.init) {
    c.init = true;
    const char* hide = std::getenv("SK_HIP_HIDE_DEVICE");
    int n = 0;
    if (hide && hide[0] && hide[0] != '0') {
      c.status = SK_ERR_UNAVAILABLE;
      c.why = "the HIP device is hidden by SK_HIP_HIDE_DEVICE";
    } else if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      c.status = SK_ERR_UNAVAILABLE;
      c.why = "no HIP device";
    } else {
      hipError_t e = hipSetDevice(0);
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
      if (e != hipSuccess) {
        c.status = SK_ERR_UNAVAILABLE;
        c.why = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
      } else {
        c.status = SK_OK;
        hipDeviceProp_t prop{};
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
          const std::string arch = prop.gcnArchName;
          c.big_llc = arch.rfind("gfx94", 0) == 0 || arch.rfind("gfx950", 0) == 0;
        }
        std::vector<double> tab(103 * 10);
        for (int z = 0; z < 103; ++z)
          for (int k = 0; k < 5; ++k) {
            tab[z * 10 + k] = kLobato[z].a[k];
            tab[z * 10 + 5 + k] = kLobato[z].b[k];
          }
        c.lobato.upload(tab, c.stream);
        SK_HIP(hipStreamSynchronize(c.stream));
        // the library runs on device 0, current here: its copy of the scratch array
        void* sp = nullptr;
        const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (se == hipSuccess && sp) c.scratch_dev = static_cast<unsigned char*>(sp);
        else c.scratch_why = se == hipSuccess ? "null address" : hipGetErrorString(se);
      }
    }
  }
  if (c.status != SK_OK) fail(SK_ERR_UNAVAILABLE, "the HIP backend is unavailable: " + c.why);
  return c;
}

template <class F>
int guarded(F&& f) {
  return guarded_call(std::forward<F>(f));
}

// ---- host geometry (Grid.swift, Antialias.swift, Detectors.swift) -----------------------------------------------
// entry i of NumPy's fftfreq(n, d) (Grid.swift:82-86)
inline double fftfreq_at(int i, int n, double d) {
  const double val = 1.0 / (double(n) * d);
  const int positive = (n - 1) / 2 + 1;
  return double(i < positive ? i : i - n) * val;
}

std::vector<double> fftfreq(int n, double d) {
  std::vector<double> o(n);
  for (int i = 0; i < n; ++i) o[i] = fftfreq_at(i, n, d);
  return o;
}

double wavelength(double energy) {  // ElectronBeam.swift:129-134
  const double h = 6.62607004e-34, me = 9.10938356e-31, c = 299792458.0, e = 1.6021766208e-19;
  const double c2 = c * c;
  return h * c / std::sqrt(energy * (2 * me * c2 / e + energy)) / e * 1.0e10;
}

double sigma_of(double energy) {  // ElectronBeam.swift:115-121
  const double h = 6.62607004e-34, me = 9.10938356e-31, c = 299792458.0, e = 1.6021766208e-19;
  const double kg = 6.0221408585491615e+26, coulomb = 6.241509125883258e+18, second = 98226947884640.62,
               joule = 6.241509125883258e+18;
  const double mass = (1 + e * energy / (me * (c * c))) * me;
  const double hsj = h * second * joule;
  return 2 * M_PI * mass * kg * e * coulomb * wavelength(energy) / (hsj * hsj);
}

void check_grid(const sk_grid* g) {
  if (!g) fail(SK_ERR_INVALID, "no grid");
  if (g->nx < 2 || g->ny < 2) fail(SK_ERR_INVALID, "grid lengths must be at least 2");
  if (!(std::isfinite(g->extent_x) && std::isfinite(g->extent_y) && g->extent_x > 0 && g->extent_y > 0))
    fail(SK_ERR_INVALID, "grid extent must be finite and positive");
}

bool same_grid(const sk_grid& a, const sk_grid& b) {  // Grid.matches (Grid.swift:73-79)
  auto close = [](double x, double y) {
    return x == y || std::fabs(x - y) <= 1e-12 * std::max(std::fabs(x), std::fabs(y));
  };
  return a.nx == b.nx && a.ny == b.ny && close(a.extent_x, b.extent_x) && close(a.extent_y, b.extent_y);
}

// The columns whose ky lies inside the antialias disk (P is zero on every other column).
void kept_columns(const sk_grid& g, std::vector<int>& keep, std::vector<int>& keepinv, int& nkept, int& nc) {
  const double sx = g.extent_x / g.nx, sy = g.extent_y / g.ny;
  const double kc = 0.6666666 / std::max(sx, sy) / 2;
  auto ky = fftfreq(g.ny, sy);
  keep.clear();
  keepinv.assign(g.ny, -1);
  for (int iy = 0; iy < g.ny; ++iy)
    if (std::fabs(ky[iy]) <= kc) {
      keepinv[iy] = (int)keep.size();
      keep.push_back(iy);
    }
  nkept = (int)keep.size();
  nc = std::min(g.ny, (nkept + 15) / 16 * 16);
  if (nc < nkept) nc = nkept;
  while ((int)keep.size() < nc) keep.push_back(-1);
}

double round_decimals(double x, int d) {
  const double f = std::pow(10.0, double(d));
  return std::nearbyint(x * f) / f;
}

void cutoff_gpts(const sk_grid& g, int& mx, int& my) {  // Detectors.swift:54-62
  const double sx = g.extent_x / g.nx, sy = g.extent_y / g.ny;
  const double kcut = 2.0 / 3.0 / std::max(sx, sy);
  auto size = [&](int n, double extent) {
    int m = (int)std::floor(round_decimals(kcut * extent, 7));
    if (m % 2 != n % 2) m += 1;
    return m;
  };
  mx = size(g.nx, double(g.nx) * sx);
  my = size(g.ny, double(g.ny) * sy);
}

void pixelated_size(const sk_grid& g, const sk_detector& d, int& mx, int& my) {  // Detectors.swift:146-159
  if (d.crop_nx > 0 || d.crop_ny > 0) {
    mx = d.crop_nx;
    my = d.crop_ny;
    return;
  }
  if (d.max_angle == SK_ANGLE_FULL) {
    mx = g.nx;
    my = g.ny;
    return;
  }
  int cx, cy;
  cutoff_gpts(g, cx, cy);
  if (d.max_angle == SK_ANGLE_CUTOFF) {
    mx = cx;
    my = cy;
    return;
  }
  auto v = [](int m, int n) {
    int k = (int)std::floor(round_decimals(double(m) / std::sqrt(2.0), 7));
    if (k % 2 != n % 2) k += 1;
    return k;
  };
  mx = v(cx, g.nx);
  my = v(cy, g.ny);
}

std::vector<int> crop_indices(int n, int m) {  // Waves.swift:90-96
  std::vector<int> o;
  if (m >= n) {
    for (int i = 0; i < n; ++i) o.push_back(i);
    return o;
  }
  const int low = m % 2 == 1 ? m / 2 + 1 : m / 2, high = m / 2;
  for (int i = 0; i < low; ++i) o.push_back(i);
  for (int i = n - high; i < n; ++i) o.push_back(i);
  return o;
}

// source index -> fftshifted crop index, or -1 (Detectors.swift:70-80)
std::vector<int> crop_map(int n, int m) {
  auto s = crop_indices(n, m);
  std::vector<int> map(n, -1);
  const int cnt = (int)s.size();
  for (int c = 0; c < cnt; ++c) map[s[c]] = (c + cnt / 2) % cnt;
  return map;
}

double annular_outer(const sk_grid& g, double lambda, const sk_detector& d) {  // Detectors.swift:108-117
  const double rx = 1 / (double(g.nx) * (g.extent_x / g.nx)), ry = 1 / (double(g.ny) * (g.extent_y / g.ny));
  const double ax = rx * lambda * 1e3, ay = ry * lambda * 1e3;
  const double maximum = std::min(double(g.nx / 2) * ax, double(g.ny / 2) * ay);
  if (d.outer < 0) {
    int cx, cy;
    cutoff_gpts(g, cx, cy);
    return std::floor(std::min(double(cx / 2) * ax, double(cy / 2) * ay));
  }
  const double o = d.outer;
  if (o > maximum && std::fabs(o - maximum) > 1e-5 + 1e-5 * std::fabs(o))
    fail(SK_ERR_INVALID, "invalid detector: outer " + std::to_string(o) + " mrad exceeds the maximum simulated angle " +
                             std::to_string(maximum) + " mrad");
  return o;
}

int detector_count(const sk_grid& g, double lambda, const sk_detector& d, int* shape) {
  if (d.kind == SK_DET_ANNULAR) {
    if (!(std::isfinite(d.inner) && d.inner >= 0)) fail(SK_ERR_INVALID, "invalid detector: inner must be finite and >= 0");
    const double o = annular_outer(g, lambda, d);
    if (d.inner > o) fail(SK_ERR_INVALID, "invalid detector: inner exceeds the outer angle");
    return 1;
  }
  if (d.kind == SK_DET_COM) {
    if (shape) shape[0] = 2;
    return 2;
  }
  if (d.kind == SK_DET_SEGMENTED) {  // SegmentedDetector.init (Detectors.swift:218-227)
    if (d.bins_radial < 1 || d.bins_azimuthal < 1) fail(SK_ERR_INVALID, "invalid detector: bins must be at least 1");
    if (!(std::isfinite(d.inner) && std::isfinite(d.outer) && d.inner >= 0 && d.outer > d.inner &&
          std::isfinite(d.rotation)))
      fail(SK_ERR_INVALID, "invalid detector: need 0 <= inner < outer, finite, and a finite rotation");
    if ((long long)d.bins_radial * d.bins_azimuthal > kMaxSlots)
      fail(SK_ERR_INVALID, "invalid detector: at most " + std::to_string(kMaxSlots) + " segments");
    if (shape) {
      shape[0] = d.bins_radial;
      shape[1] = d.bins_azimuthal;
    }
    return d.bins_radial * d.bins_azimuthal;
  }
  if (d.kind == SK_DET_PIXELATED) {
    int mx, my;
    pixelated_size(g, d, mx, my);
    if (mx < 1 || my < 1 || mx > g.nx + 1 || my > g.ny + 1) fail(SK_ERR_INVALID, "invalid detector: crop size");
    if (shape) {
      shape[0] = mx;
      shape[1] = my;
    }
    return mx * my;
  }
  fail(SK_ERR_INVALID, "invalid detector kind");
}

// SegmentedDetector.detect (Detectors.swift:232-262): the bin of every pixel of the full grid, or -1. The bins live on
// the crop that holds the outer angle (2 ceil(outer / da) + 1 points with the grid's parity, at most n), whose
// pixels in FFT order are the full pattern's at the crop's source indices; alpha = sqrt(ax^2 + ay^2) and the
// azimuth atan2(ay, ax) - rotation reduced into [0, 2 pi), in float64 without contraction, as the reference
// rounds them.
std::vector<int> segment_map(const sk_grid& g, double ax, double ay, const sk_detector& d) {
#pragma clang fp contract(off)
  auto m = [&](int n, double da) {
    int k = int(2 * std::ceil(d.outer / da)) + 1;
    if (k % 2 != n % 2) k += 1;
    return std::min(k, n);
  };
  const int mx = m(g.nx, ax), my = m(g.ny, ay);
  const auto sx = crop_indices(g.nx, mx), sy = crop_indices(g.ny, my);
  const auto angx = fftfreq(mx, 1 / ax / double(mx)), angy = fftfreq(my, 1 / ay / double(my));
  std::vector<int> map((size_t)g.nx * g.ny, -1);
  const double two_pi = 2 * M_PI;
  for (int cx = 0; cx < mx; ++cx)
    for (int cy = 0; cy < my; ++cy) {
      const double sq = angx[cx] * angx[cx] + angy[cy] * angy[cy];
      const double alpha = std::sqrt(sq);
      if (!(alpha >= d.inner && alpha < d.outer)) continue;
      double phi = std::fmod(std::atan2(angy[cy], angx[cx]) - d.rotation, two_pi);
      if (phi < 0) phi += two_pi;
      const int r = (int)std::floor(double(d.bins_radial) * (alpha - d.inner) / (d.outer - d.inner));
      if (r >= d.bins_radial) continue;  // alpha just below outer can round to the last edge: dropped, as abTEM does
      const double f = std::floor(double(d.bins_azimuthal) * (phi / two_pi));
      const int az = (int)std::min(std::max(f, 0.0), double(d.bins_azimuthal - 1));
      map[(size_t)sx[cx] * g.ny + sy[cy]] = az + r * d.bins_azimuthal;
    }
  return map;
}

// ---- FFT pass helpers (full width: identity maps) ----------------------------------------------------------------
const int* identity_map(int n) {
  Context& c = ctx();
  auto it = c.identity.find(n);
  if (it != c.identity.end()) return it->second->p;
  auto b = std::make_unique<DevBuf<int>>();
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) v[i] = i;
  b->upload(v, c.stream);
  const int* p = b->p;
  c.identity[n] = std::move(b);
  return p;
}

// A pass launch; its ev_start / ev_stop (when set) time it: the fast passes attach them to the kernel's dispatch, the
// generic passes record them on the stream around their launches.
void row_pass(const RowArgs& a, int batch, bool fast, hipStream_t s) {
  if (fast && fast_row_supported(a.nx, a.ny, a.flags)) {
    launch_row_fast(a, batch, s);
    return;
  }
  if (fast && reg_row_supported(a)) {
    launch_row_reg(a, batch, s);
    return;
  }
  if (a.ev_start) SK_HIP(hipEventRecord(a.ev_start, s));
  launch_row_generic(a, batch, s);
  if (a.ev_stop) SK_HIP(hipEventRecord(a.ev_stop, s));
}

void col_pass(const ColArgs& a, int batch, bool fast, hipStream_t s) {
  if (fast && fast_col_supported(a.nx, a.nc, a.flags)) {
    launch_col_fast(a, batch, s);
    return;
  }
  if (fast && reg_col_supported(a)) {
    launch_col_reg(a, batch, s);
    return;
  }
  if (a.ev_start) SK_HIP(hipEventRecord(a.ev_start, s));
  launch_col_generic(a, batch, s);
  if (a.ev_stop) SK_HIP(hipEventRecord(a.ev_stop, s));
}

// Full 2-D transform of `batch` waves in place (buf: batch x nx x ny): forward unscaled (sign -1) or inverse with
// `inv_scale`, optionally multiplying the spectrum by the full-width P between the column transforms.
void fft2_batch(float2* buf, int nx, int ny, int batch, int sign, const float2* P, float scale, hipStream_t s) {
  RowArgs r{};
  r.nx = nx;
  r.ny = ny;
  r.nc = ny;
  r.fy = line_plan(ny, s).dev;
  r.keep = identity_map(ny);
  r.keepinv = r.keep;
  r.in_full = buf;
  r.out_full = buf;
  r.in_full_stride = r.out_full_stride = (long long)nx * ny;
  r.out_scale = 1.0f;
  ColArgs c{};
  c.nx = nx;
  c.ny = ny;
  c.nc = ny;
  c.fx = line_plan(nx, s).dev;
  c.keep = r.keep;
  c.G = buf;
  c.P = P;
  if (sign < 0) {
    r.flags = ROW_IN_FULL | ROW_FWD | ROW_OUT_FULL;
    row_pass(r, batch, false, s);
    c.flags = COL_IN_G | COL_FWD | COL_OUT | (P ? (COL_MUL_P | COL_INV) : 0);
    col_pass(c, batch, false, s);
    if (P) {
      r.flags = ROW_IN_FULL | ROW_INV | ROW_OUT_FULL;
      r.out_scale = scale;
      row_pass(r, batch, false, s);
    }
  } else {
    c.flags = COL_IN_G | COL_INV | COL_OUT;
    col_pass(c, batch, false, s);
    r.flags = ROW_IN_FULL | ROW_INV | ROW_OUT_FULL;
    r.out_scale = scale;
    row_pass(r, batch, false, s);
  }
}

// ---- potential preparation ---------------------------------------------------------------------------------------
std::unique_ptr<sk_potential> new_potential(const sk_grid* grid, int nslices, const double* dz, double energy,
                                            const double tilt[2]) {
  check_grid(grid);
  if (nslices < 1) fail(SK_ERR_INVALID, "a potential needs at least one slice");
  if (!(std::isfinite(energy) && energy > 0)) fail(SK_ERR_INVALID, "electron energy must be finite and positive");
  for (int j = 0; j < nslices; ++j)
    if (!(std::isfinite(dz[j]) && dz[j] > 0)) fail(SK_ERR_INVALID, "slice thickness must be finite and positive");
  if (tilt && !(std::isfinite(tilt[0]) && std::isfinite(tilt[1]))) fail(SK_ERR_INVALID, "invalid tilt");
  auto p = std::make_unique<sk_potential>();
  p->grid = *grid;
  p->nslices = nslices;
  p->energy = energy;
  p->lambda = wavelength(energy);
  p->sigma = sigma_of(energy);
  if (tilt) {
    p->tilt[0] = tilt[0];
    p->tilt[1] = tilt[1];
  }
  p->dz.assign(dz, dz + nslices);
  return p;
}

// transmissions and propagators from V (device, float64, nslices x N)
void finish_potential(sk_potential& p, const double* V_dev) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const int nx = p.grid.nx, ny = p.grid.ny;
  const size_t N = (size_t)nx * ny;
  kept_columns(p.grid, p.keep, p.keepinv, p.nkept, p.nc);
  p.keep_d.upload(p.keep, s);
  p.keepinv_d.upload(p.keepinv, s);
  p.aperture.alloc(N);
  launch_antialias(nx, ny, p.grid.extent_x, p.grid.extent_y, p.aperture.p, s);
  // t = IFFT2(FFT2(e^{i sigma V}) * A) (Multislice.swift:83-103), in chunks of slices
  p.t.alloc((size_t)p.nslices * N);
  DevBuf<float2> A;
  A.alloc(N);
  launch_aperture_as_complex(nx, ny, p.aperture.p, 1.0, A.p, s);
  for (int j = 0; j < p.nslices; ++j) launch_phase_to_complex(V_dev + j * N, p.sigma, p.t.p + j * N, N, s);
  const int chunk = std::max<int>(1, (int)std::min<size_t>(p.nslices, (size_t(512) << 20) / (N * sizeof(float2))));
  for (int j0 = 0; j0 < p.nslices; j0 += chunk) {
    const int cnt = std::min(chunk, p.nslices - j0);
    fft2_batch(p.t.p + (size_t)j0 * N, nx, ny, cnt, -1, A.p, float(1.0 / double(N)), s);
  }
  // propagators per distinct thickness (Multislice.swift:136, 145-148), compact, scaled by 1/N
  std::map<double, int> index;
  p.prop_of_slice.resize(p.nslices);
  for (int j = 0; j < p.nslices; ++j) {
    auto it = index.find(p.dz[j]);
    if (it == index.end()) {
      const int k = (int)p.prop_dz.size();
      index[p.dz[j]] = k;
      p.prop_dz.push_back(p.dz[j]);
      p.prop_of_slice[j] = k;
    } else {
      p.prop_of_slice[j] = it->second;
    }
  }
  p.props.alloc(p.prop_dz.size() * (size_t)nx * p.nc);
  for (size_t k = 0; k < p.prop_dz.size(); ++k)
    launch_propagator(nx, ny, p.grid.extent_x, p.grid.extent_y, p.lambda, p.prop_dz[k], p.tilt[0], p.tilt[1],
                      p.aperture.p, p.keep_d.p, p.nc, 1.0 / double(N), p.props.p + k * (size_t)nx * p.nc, s);
  SK_HIP(hipStreamSynchronize(s));
}

// ---- multislice driver -------------------------------------------------------------------------------------------
struct Workspace {
  DevBuf<float2> G, full, planes, rx, ry, spec;
  DevBuf<double> partial, acc, positions, tmp_re, tmp_im, norm, norm_partial;
  DevBuf<int> flag;
  std::vector<DevBuf<double>> pix;
  DevBuf<double> sqx, sqy;
  DevBuf<int> pos_dev;  // first position of a replayed batch
  DevBuf<LaneDesc> descs;          // lane pipeline: [steps + 1][lanes]
  DevBuf<int> step;                // lane pipeline: step counter
  DevBuf<float2> rx_all, ry_all;   // lane pipeline: ramps of every position
  DevBuf<double> acc_copies;       // lane pipeline with shared chunks: [lanes][npos][slots]
  std::vector<DevBuf<double>> pix_copies;  // and [lanes][npos][mx][my] per pixelated detector
  std::vector<DevBuf<int>> maps;
  std::vector<DevBuf<double>> weights;
};

Workspace& ws() {
  static Workspace* w = new Workspace;  // process lifetime (see ctx())
  return *w;
}

// Streams of a split scan (sk_options.streams = 0): the parts of each batch run their slice loops on two streams, so
// one part's row pass fills the tail of the other's column pass. Measured on the HAADF scan (4096 positions), two
// streams against one: 256 x 256 at 683 positions 4.30 M against 4.01 M position-slices/s; 512 x 512 at 128
// positions 963-989 k against 945-960 k at 171 positions for one stream, and 954-983 k at 171 for two. Three
// streams gain less, four lose. At 512 x 512 the split scan takes 3/4 of the cache-resident batch. Parts of fewer
// than 64 positions lose instead (the launches get too short): over 16 frozen-phonon configurations at 512 x 512,
// 24-48 positions run 2-5 % slower on two streams than on one, so the library splits batches of at least
// kMinSplitBatch positions only.
constexpr int kSplitStreams = 2;
constexpr int kMinSplitBatch = 128;
constexpr int kMaxStreams = 4;

void set_col_tile(Context& c, const sk_options* opt) {
  const int t = opt ? opt->col_tile : 0;
  if (t != 0 && t != 4 && t != 8 && t != 16) fail(SK_ERR_INVALID, "col_tile must be 0, 4, 8 or 16");
  c.col_tile = t;
  const int pr = opt ? opt->precision : 0;
  if (pr != 0 && pr != 1) fail(SK_ERR_INVALID, "precision must be 0 (float32) or 1 (split-BF16)");
  c.precision = pr;
  const int st = opt ? opt->streams : 0;
  if (st < 0 || st > kMaxStreams) fail(SK_ERR_INVALID, "streams must be 0 (the library's choice) to " +
                                                           std::to_string(kMaxStreams));
}

// the transform plan of length n for the slice loop under the running call's precision
LineFFT slice_plan(int n, hipStream_t s) {
  if (ctx().precision == 1) return mfma_line_plan(n, s);
  return line_plan(n, s).dev;
}

// Positions per launch. Every pass streams the batch's compact waves (nx x nc complex64 each) once in and once out,
// so a batch whose waves stay in the device's last-level cache (kCachedWaveBytes) runs from that cache rather than
// from HBM: at 512 x 512, 176 positions run 1.17 x faster per position than 1024, and 184 fall back to HBM rates.
// On a device with such a cache, the default takes the largest cache-resident batch when it holds at least
// kMinCachedBatch positions (enough to fill the GPU); otherwise a 4 GiB budget of at most 1024 positions. It then
// spreads `count` evenly over the batches, so no small remainder batch runs at a fraction of the GPU.
constexpr int kMinCachedBatch = 64;
// float64 intensities one batch of sk_hip_scan_intensities may hold (device, and each of two pinned host buffers)
constexpr size_t kIntensityBytes = size_t(256) << 20;
// device memory the lane pipeline may take for per-lane detector accumulators when lanes share positions
constexpr size_t kLaneCopyBytes = size_t(2) << 30;

int default_batch(const sk_potential& p, int count, const sk_options* opt, double cache_share = 1.0) {
  int b = opt && opt->batch > 0 ? opt->batch : 0;
  if (b > 0) return std::max(1, std::min(b, count));
  const size_t cached =
      ctx().big_llc ? (size_t)(cache_share * kCachedWaveBytes / ((double)p.grid.nx * p.nc * sizeof(float2))) : 0;
  if (cached >= (size_t)kMinCachedBatch)
    b = (int)std::min<size_t>(1024, cached);
  else
    b = (int)std::max<size_t>(1, std::min<size_t>(1024, (size_t(4) << 30) / ((size_t)p.grid.nx * p.nc * sizeof(float2))));
  count = std::max(1, count);
  const int batches = (count + b - 1) / b;
  return std::max(1, (count + batches - 1) / batches);
}

RowArgs row_args(const sk_potential& p, float2* G) {
  RowArgs r{};
  r.nx = p.grid.nx;
  r.ny = p.grid.ny;
  r.nc = p.nc;
  r.fy = slice_plan(p.grid.ny, ctx().stream);
  r.keep = p.keep_d.p;
  r.keepinv = p.keepinv_d.p;
  r.G = G;
  r.out_scale = 1.0f;
  r.in_full_stride = r.out_full_stride = r.plane_stride = (long long)p.grid.nx * p.grid.ny;
  // the kept set {iy : |ky| <= kc} is [0, hi] and [lo, ny) (fftfreq is symmetric about 0)
  int hi = -1;
  while (hi + 1 < p.nkept && p.keep[hi + 1] == hi + 1) ++hi;
  const int nneg = p.nkept - (hi + 1);
  r.kept_hi = hi;
  r.kept_lo = p.grid.ny - nneg;
  r.kept_shift = p.grid.ny - p.nkept;
  for (int c = hi + 1; c < p.nkept; ++c)
    if (p.keep[c] != r.kept_lo + (c - hi - 1)) fail(SK_ERR_DEVICE, "kept columns are not two ranges");
  return r;
}

ColArgs col_args(const sk_potential& p, float2* G) {
  ColArgs c{};
  c.nx = p.grid.nx;
  c.ny = p.grid.ny;
  c.nc = p.nc;
  c.fx = slice_plan(p.grid.nx, ctx().stream);
  c.keep = p.keep_d.p;
  c.G = G;
  c.col_tile = ctx().col_tile;  // 0: chosen per launch from the batch's size
  return c;
}

// Blocks per wave of the detecting column pass a scan of p0 launches: the count of its partial sums per detector
// slot and position.
int detect_blocks_per_wave(const sk_potential& p0, bool fast, int pipe, hipStream_t s) {
  const int nx = p0.grid.nx;
  if (fast && fast_col_supported(nx, p0.nc, COL_IN_G | COL_FWD | COL_MUL_P | COL_DETECT))
    return fast_col_blocks_per_wave(nx, p0.nc, pipe);
  ColArgs f = col_args(p0, nullptr);
  f.flags = COL_IN_G | COL_FWD | COL_MUL_P | COL_DETECT;
  if (fast && reg_col_supported(f)) return reg_col_blocks_per_wave(f);
  const int gcols = col_generic_lines(slice_plan(nx, s));
  return (p0.nc + gcols - 1) / gcols;
}

// Runs `batch` waves whose first row pass input is ready (full real-space waves in `full`, or compact G after a
// generated INIT); writes requested planes. planes_mask[j] = output index for plane j (or -1).
struct Timer {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;  // 0 row, 1 col
  size_t used = 0;
  void mark_start(hipStream_t s, int k) {
    if (!on) return;
    if (used + 2 > ev.size()) {
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        SK_HIP(hipEventCreate(&e));
        ev.push_back(e);
      }
      kind.resize(ev.size() / 2 + 1);
    }
    kind[used / 2] = k;
    SK_HIP(hipEventRecord(ev[used], s));
  }
  void mark_stop(hipStream_t s) {
    if (!on) return;
    SK_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
  }
  // the events that time the next launch of kind k, recorded by that launch itself (null, null when off)
  std::pair<hipEvent_t, hipEvent_t> take(int k) {
    if (!on) return {nullptr, nullptr};
    if (used + 2 > ev.size()) {
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        SK_HIP(hipEventCreate(&e));
        ev.push_back(e);
      }
      kind.resize(ev.size() / 2 + 1);
    }
    kind[used / 2] = k;
    used += 2;
    return {ev[used - 2], ev[used - 1]};
  }
  void harvest(sk_timing& t) {
    if (!on || !used) return;
    SK_HIP(hipEventSynchronize(ev[used - 1]));
    for (size_t i = 0; i < used; i += 2) {
      float ms = 0;
      SK_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      if (kind[i / 2] == 0) {
        t.row_ms += ms;
        t.row_launches++;
      } else {
        t.col_ms += ms;
        t.col_launches++;
      }
    }
    used = 0;
  }
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

void slice_loop(const sk_potential& p, int batch, float2* G, bool entry_full, float2* full, int first_slice_from_full,
                const int* plane_out, float2* planes, const ColArgs* final_col, bool fast, Timer& tm,
                hipStream_t stream = nullptr) {
  hipStream_t s = stream ? stream : ctx().stream;
  const size_t N = (size_t)p.grid.nx * p.grid.ny;
  RowArgs r = row_args(p, G);
  ColArgs c = col_args(p, G);
  for (int j = 0; j < p.nslices; ++j) {
    // row pass of slice j: (inverse of the previous state) * t_j * forward
    r.t = p.t.p + (size_t)j * N;
    if (j == 0 && entry_full) {
      r.flags = ROW_IN_FULL | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
      r.in_full = full;
    } else if (j == 0 && first_slice_from_full) {
      r.flags = ROW_IN_FULL | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
      r.in_full = full;
    } else {
      r.flags = ROW_IN_COMPACT | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
      if (j > 0 && plane_out && plane_out[j - 1] >= 0) {
        r.flags |= ROW_WRITE_PLANE;
        r.plane = planes + (size_t)plane_out[j - 1] * batch * N;
      }
    }
    std::tie(r.ev_start, r.ev_stop) = tm.take(0);
    row_pass(r, batch, fast, s);
    r.ev_start = r.ev_stop = nullptr;
    // column pass of slice j
    c.P = p.props.p + (size_t)p.prop_of_slice[j] * p.grid.nx * p.nc;
    if (j == p.nslices - 1 && final_col) {
      ColArgs f = *final_col;
      f.P = c.P;
      col_pass(f, batch, fast, s);
    } else {
      c.flags = COL_IN_G | COL_FWD | COL_MUL_P | COL_INV | COL_OUT;
      std::tie(c.ev_start, c.ev_stop) = tm.take(1);
      col_pass(c, batch, fast, s);
      c.ev_start = c.ev_stop = nullptr;
    }
    r.flags = 0;
    r.plane = nullptr;
  }
  if (!final_col) {
    // exit: the real-space wave after the last slice (and any plane still pending)
    const int q = plane_out ? plane_out[p.nslices - 1] : -1;
    if (q >= 0) {
      r.flags = ROW_IN_COMPACT | ROW_INV | ROW_WRITE_PLANE;
      r.plane = planes + (size_t)q * batch * N;
      row_pass(r, batch, fast, s);
    }
  }
}

// One batch of nb scan positions starting at b0 (+ *pos_dev when replayed from a graph) through potential P: the
// shift ramps, the first column pass (the probe, generated from the spectrum), every slice, the detector reduction of
// the last column pass and the sum of its partials into acc. G, rx, ry and det.partial hold nb positions; `full`
// (nb full-width waves) is needed only when the probe reaches past the bandlimit (fits = false).
void scan_batch(const sk_potential& P, const float2* spec, const double* positions, bool fits, bool fast, float2* G,
                float2* full, float2* rx, float2* ry, const DetPlan& det, int b0, const int* pos_dev, int nb,
                double* acc, Timer& t, hipStream_t s) {
  const sk_grid& g = P.grid;
  const int nx = g.nx, ny = g.ny;
  launch_ramps(nx, ny, g.extent_x / nx, g.extent_y / ny, positions, b0, pos_dev, nb, rx, ry, s);
  ColArgs init = col_args(P, fits ? G : full);
  init.spec = spec;
  init.rx = rx;
  init.ry = ry;
  init.flags = COL_GENERATE | COL_INV | COL_OUT;
  if (!fits) {
    init.nc = ny;
    init.keep = identity_map(ny);
  }
  col_pass(init, nb, fast && fits, s);
  ColArgs fin = col_args(P, G);
  fin.flags = COL_IN_G | COL_FWD | COL_MUL_P | COL_DETECT;
  fin.pos0 = b0;
  fin.pos_dev = pos_dev;
  fin.det = det;
  slice_loop(P, nb, G, false, full, fits ? 0 : 1, nullptr, nullptr, &fin, fast, t, s);
  launch_reduce_partials(det, nb, b0, pos_dev, acc, s);
}

// The lane pipeline of a scan through several potentials (frozen-phonon configurations of any slice counts).
// Positions are cut into chunks of W; lane l takes chunks l, l + L, ... and, for each, every potential in turn (the
// order rotated by l so that the lanes of one launch sit on different potentials). Every step advances each lane by
// one slice: one row launch, one column launch, one reduction of the lanes that finished an item, one increment of
// the device step counter. A lane's last slice of an item fuses the detectors with the probe of its next item.
// All potentials of one chunk run in one lane, in a fixed order, so the float64 accumulation is deterministic.
void run_pipeline(const sk_potential* const* pots, int npots, int npos, int W, int L, const sk_grid& g,
                  const DetPlan& dp, bool graph, Timer& tm, sk_timing& timing) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  Workspace& w = ws();
  const int nx = g.nx, ny = g.ny;
  const size_t N = (size_t)nx * ny;
  const sk_potential& p0 = *pots[0];
  const int B = L * W;
  struct Item {
    int pos, count, pot;
  };
  std::vector<std::vector<Item>> lane_items(L);
  const int chunks = (npos + W - 1) / W;
  // With no more lanes than chunks, every lane owns its chunks and runs all potentials over them (rotated, so that
  // the lanes hold different potentials at a time): no two lanes ever accumulate into one position. With more lanes
  // than chunks (a few positions over many potentials), the (chunk, potential) items go to the lanes longest first,
  // each to the lane with the fewest slices so far, and every lane accumulates into its own copy of the detector
  // accumulators; the copies are summed in lane order at the end.
  const bool copies = L > chunks;
  if (!copies) {
    for (int q = 0; q < chunks; ++q) {
      const int lane = q % L;
      for (int k = 0; k < npots; ++k)
        lane_items[lane].push_back({q * W, std::min(W, npos - q * W), (k + lane) % npots});
    }
  } else {
    std::vector<Item> items;
    for (int k = 0; k < npots; ++k)
      for (int q = 0; q < chunks; ++q) items.push_back({q * W, std::min(W, npos - q * W), k});
    std::stable_sort(items.begin(), items.end(),
                     [pots](const Item& a, const Item& b) { return pots[a.pot]->nslices > pots[b.pot]->nslices; });
    std::vector<long long> load(L, 0);
    for (const Item& it : items) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane_items[l].push_back(it);
      load[l] += pots[it.pot]->nslices;
    }
  }
  int nsteps = 0;
  for (auto& li : lane_items) {
    int t = 0;
    for (auto& it : li) t += pots[it.pot]->nslices;
    nsteps = std::max(nsteps, t);
  }
  std::vector<LaneDesc> descs((size_t)(nsteps + 1) * L, LaneDesc{});
  for (int l = 0; l < L; ++l) {
    const auto& li = lane_items[l];
    if (li.empty()) continue;
    const int acc_off = copies ? l * npos : 0;
    LaneDesc& pro = descs[l];
    pro.mode = LANE_INIT;
    pro.next_pos = li[0].pos;
    pro.next_count = li[0].count;
    pro.acc_off = acc_off;
    int st = 1;
    for (size_t m = 0; m < li.size(); ++m) {
      const sk_potential& P = *pots[li[m].pot];
      for (int j = 0; j < P.nslices; ++j, ++st) {
        LaneDesc& d = descs[(size_t)st * L + l];
        d.t = P.t.p + (size_t)j * N;
        d.P = P.props.p + (size_t)P.prop_of_slice[j] * nx * P.nc;
        d.pos = li[m].pos;
        d.count = li[m].count;
        d.acc_off = acc_off;
        if (j < P.nslices - 1) {
          d.mode = LANE_MID;
        } else if (m + 1 < li.size()) {
          d.mode = LANE_FINAL_INIT;
          d.next_pos = li[m + 1].pos;
          d.next_count = li[m + 1].count;
        } else {
          d.mode = LANE_FINAL;
        }
      }
    }
  }
  w.descs.upload(descs, s);
  // per descriptor row: the column launches it needs (1 mid slices, 2 items finishing, 4 items starting) and whether
  // any lane runs a slice (the row pass)
  std::vector<int> row_modes(nsteps + 1, 0);
  std::vector<char> row_busy(nsteps + 1, 0);
  for (int st = 0; st <= nsteps; ++st)
    for (int l = 0; l < L; ++l) {
      const LaneDesc& d = descs[(size_t)st * L + l];
      if (d.t) row_busy[st] = 1;
      if (d.mode == LANE_MID) row_modes[st] |= 1;
      if (d.mode == LANE_FINAL || d.mode == LANE_FINAL_INIT) row_modes[st] |= 2;
      if (d.mode == LANE_INIT || d.mode == LANE_FINAL_INIT) row_modes[st] |= 4;
    }
  w.step.ensure(1);
  SK_HIP(hipMemsetAsync(w.step.p, 0, sizeof(int), s));
  w.rx_all.ensure((size_t)npos * nx);
  w.ry_all.ensure((size_t)npos * ny);
  for (int q0 = 0; q0 < npos; q0 += 32768) {
    const int cnt = std::min(32768, npos - q0);
    launch_ramps(nx, ny, g.extent_x / nx, g.extent_y / ny, w.positions.p, q0, nullptr, cnt,
                 w.rx_all.p + (size_t)q0 * nx, w.ry_all.p + (size_t)q0 * ny, s);
  }
  DetPlan dpl = dp;  // where the lanes accumulate
  double* acc = w.acc.p;
  const int nslots = std::max(1, dp.nslots);
  if (copies) {
    const size_t n = (size_t)npos * nslots;
    w.acc_copies.ensure(n * L);
    SK_HIP(hipMemsetAsync(w.acc_copies.p, 0, sizeof(double) * n * L, s));
    acc = w.acc_copies.p;
    w.pix_copies.resize(std::max<size_t>(w.pix_copies.size(), (size_t)dp.npix));
    for (int q = 0; q < dp.npix; ++q) {
      const size_t m = (size_t)npos * dp.pix_stride[q];
      w.pix_copies[q].ensure(m * L);
      SK_HIP(hipMemsetAsync(w.pix_copies[q].p, 0, sizeof(double) * m * L, s));
      dpl.pix_acc[q] = w.pix_copies[q].p;
    }
  }
  PipeArgs pa{w.descs.p, w.step.p, L, W, w.rx_all.p, w.ry_all.p, 0};
  RowArgs r = row_args(p0, w.G.p);
  r.flags = ROW_IN_COMPACT | ROW_INV | ROW_MUL_T | ROW_FWD | ROW_OUT_COMPACT;
  r.pipe = pa;
  ColArgs cc = col_args(p0, w.G.p);
  cc.spec = w.spec.p;
  cc.det = dpl;
  cc.pipe = pa;
  auto step_fn = [&](Timer& t) {
    t.mark_start(s, 0);
    launch_row_pipe(r, B, s);
    t.mark_stop(s);
    t.mark_start(s, 1);
    launch_col_pipe(cc, B, s);
    t.mark_stop(s);
    launch_reduce_pipe(dpl, pa, B, acc, s);
    launch_step_increment(w.step.p, s);
  };
  // the lanes' copies into the scan's accumulators, in lane order
  auto fold_copies = [&] {
    if (!copies) return;
    launch_sum_copies(w.acc_copies.p, (long long)npos * nslots, L, w.acc.p, s);
    for (int q = 0; q < dp.npix; ++q)
      launch_sum_copies(w.pix_copies[q].p, (long long)npos * dp.pix_stride[q], L, dp.pix_acc[q], s);
  };
  if (!graph) {
    // the step goes to every launch by value, and each step launches only the kernels its lanes need
    r.pipe.step = cc.pipe.step = pa.step = nullptr;
    for (int st = 0; st <= nsteps; ++st) {  // row 0: the prologue, the probe of every lane's first item
      r.pipe.step_now = cc.pipe.step_now = pa.step_now = st;
      if (st > 0 && row_busy[st]) {
        tm.mark_start(s, 0);
        launch_row_pipe(r, B, s);
        tm.mark_stop(s);
      }
      const bool timed = (row_modes[st] & 1) != 0;  // the column time is that of steps with mid slices
      if (timed) tm.mark_start(s, 1);
      launch_col_pipe(cc, B, s, row_modes[st]);
      if (timed) tm.mark_stop(s);
      if (row_modes[st] & 2) launch_reduce_pipe(dpl, pa, B, acc, s);
      if ((st & 31) == 31) tm.harvest(timing);
    }
    tm.harvest(timing);
    fold_copies();
    return;
  }
  // prologue: the probe of every lane's first item (descriptor row 0)
  launch_col_pipe(cc, B, s);
  launch_step_increment(w.step.p, s);
  // hipGraph: K steps captured once; every replay reads its steps from the device counter
  const int K = std::min(nsteps, 32);
  Timer off;
  hipGraph_t gr;
  hipGraphExec_t ge = nullptr;
  SK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < K; ++k) step_fn(off);
  } catch (...) {
    hipGraph_t dropped;
    (void)hipStreamEndCapture(s, &dropped);
    if (dropped) (void)hipGraphDestroy(dropped);
    throw;
  }
  SK_HIP(hipStreamEndCapture(s, &gr));
  SK_HIP(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
  SK_HIP(hipGraphDestroy(gr));
  int st = 0;
  for (; st + K <= nsteps; st += K) SK_HIP(hipGraphLaunch(ge, s));
  for (; st < nsteps; ++st) step_fn(off);
  fold_copies();
  SK_HIP(hipStreamSynchronize(s));
  SK_HIP(hipGraphExecDestroy(ge));
}

}  // namespace

// Pieces of one device buffer, each aligned to 256 bytes.
struct Arena {
  unsigned char* base;
  size_t size, used;
  size_t aligned(size_t off) const { return (off + 255) & ~size_t(255); }
  template <class T>
  T* take(size_t count) {
    const size_t off = aligned(used), bytes = count * sizeof(T);
    if (off > size || bytes > size - off) fail(SK_ERR_MEMORY, "sk_hip_scan_noalloc: scratch exhausted");
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
  size_t left() const { return aligned(used) >= size ? 0 : size - aligned(used); }
};
// host staging of the annular detectors' pixel angles (static: the entry point allocates nothing)
constexpr int kNoallocMaxLength = 4096;
static double noalloc_sq[2 * kNoallocMaxLength];

int default_scan_batch(const sk_potential& p, int count) { return default_batch(p, count, nullptr); }

}  // namespace sk

using namespace sk;

// =================================================================================================================
// C ABI
// =================================================================================================================
extern "C" {

int sk_hip_available(void) {
  std::lock_guard<std::recursive_mutex> g(lock());
  try {
    require();
    return 1;
  } catch (const Error& e) {
    last_error() = e.what();
    return 0;
  }
}

const char* sk_hip_last_error(void) { return last_error().c_str(); }

#ifndef SK_BUILD_ID
#define SK_BUILD_ID "unversioned"
#endif
// the hash of the sources this library was compiled from, set by the build script
const char* sk_hip_build_id(void) { return SK_BUILD_ID; }

void sk_hip_set_timing(int enabled) {
  std::lock_guard<std::recursive_mutex> g(lock());
  ctx().timing = enabled != 0;
}

int sk_hip_last_timing(sk_timing* t) {
  if (!t) return SK_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> g(lock());
  *t = ctx().last;
  return SK_OK;
}

int sk_hip_potential_from_values(const sk_grid* grid, int nslices, const double* V, const double* dz, double energy,
                                 const double tilt[2], sk_potential** out) {
  return guarded([&] {
    Context& c = require();
    if (!out || !V || !dz) fail(SK_ERR_INVALID, "null argument");
    auto p = new_potential(grid, nslices, dz, energy, tilt);
    const size_t N = (size_t)grid->nx * grid->ny;
    DevBuf<double> Vd;
    Vd.upload(V, N * nslices, c.stream);
    finish_potential(*p, Vd.p);
    *out = p.release();
  });
}

int sk_hip_potential_from_atoms(const sk_grid* grid, int nslices, const double* dz, int natoms, const int* numbers,
                                const double* xy, const int* slice_of_atom, int pi_mode, double energy,
                                const double tilt[2], sk_potential** out) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    if (!out || !dz || (natoms > 0 && (!numbers || !xy || !slice_of_atom))) fail(SK_ERR_INVALID, "null argument");
    auto p = new_potential(grid, nslices, dz, energy, tilt);
    const int nx = grid->nx, ny = grid->ny;
    const size_t N = (size_t)nx * ny;
    const double sx = grid->extent_x / nx, sy = grid->extent_y / ny;
    // elements present, ascending (ProjectedPotential.swift:158)
    std::vector<int> zs;
    for (int a = 0; a < natoms; ++a) {
      if (numbers[a] < 1 || numbers[a] > 103 || !kLobato[numbers[a] - 1].present)
        fail(SK_ERR_INVALID, "element Z = " + std::to_string(numbers[a]) + " has no Lobato parameters");
      if (slice_of_atom[a] < 0 || slice_of_atom[a] >= nslices) fail(SK_ERR_INVALID, "slice of atom out of range");
      if (!(std::isfinite(xy[2 * a]) && std::isfinite(xy[2 * a + 1]))) fail(SK_ERR_INVALID, "non-finite position");
      zs.push_back(numbers[a]);
    }
    std::sort(zs.begin(), zs.end());
    zs.erase(std::unique(zs.begin(), zs.end()), zs.end());
    const int nz = std::max<int>(1, (int)zs.size());
    std::vector<int> zindex(104, -1);
    for (size_t e = 0; e < zs.size(); ++e) zindex[zs[e]] = (int)e;
    DevBuf<double> Vd;
    Vd.alloc(N * nslices);
    if (zs.empty()) {
      SK_HIP(hipMemsetAsync(Vd.p, 0, sizeof(double) * N * nslices, s));
    } else {
      DevBuf<int> zs_d;
      zs_d.upload(zs, s);
      DevBuf<float> K;
      K.alloc(N * zs.size());
      launch_lobato_kernels(nx, ny, grid->extent_x, grid->extent_y, zs_d.p, (int)zs.size(), pi_mode, c.lobato.p, K.p,
                            s);
      // bilinear deltas per (slice, element), merged per pixel in float64 in NumPy add.at order
      // (ProjectedPotential.swift:178-200): weight-major, then atom order
      std::vector<std::vector<int>> atoms_of(nslices * (size_t)nz);
      for (int a = 0; a < natoms; ++a) atoms_of[(size_t)slice_of_atom[a] * nz + zindex[numbers[a]]].push_back(a);
      const int chunk = std::max<int>(1, (int)std::min<size_t>(nslices, (size_t(1) << 30) / (N * nz * sizeof(float2))));
      std::vector<double> dense(N, 0.0);
      std::vector<char> touched(N, 0);
      DevBuf<float2> D, Sp;
      D.alloc((size_t)chunk * nz * N);
      Sp.alloc((size_t)chunk * N);
      for (int j0 = 0; j0 < nslices; j0 += chunk) {
        const int cnt = std::min(chunk, nslices - j0);
        std::vector<long long> idx;
        std::vector<int> idx32;
        std::vector<double> w;
        for (int j = j0; j < j0 + cnt; ++j)
          for (int e = 0; e < nz; ++e) {
            const auto& list = atoms_of[(size_t)j * nz + e];
            if (list.empty()) continue;
            std::vector<int> order;
            const size_t na = list.size();
            std::vector<int> rows(na), cols(na);
            std::vector<double> wt[4];
            for (auto& v : wt) v.resize(na);
            for (size_t q = 0; q < na; ++q) {
              const int a = list[q];
              const double px = xy[2 * a] / sx, py = xy[2 * a + 1] / sy;
              const int r = (int)std::floor(px), cc = (int)std::floor(py);
              const double x = px - r, y = py - cc, xyv = x * y;
              rows[q] = r;
              cols[q] = cc;
              wt[0][q] = 1 + xyv - y - x;
              wt[1][q] = x - xyv;
              wt[2][q] = y - xyv;
              wt[3][q] = xyv;
            }
            auto mod = [](int a, int m) {
              int r = a % m;
              return r < 0 ? r + m : r;
            };
            for (int k = 0; k < 4; ++k)
              for (size_t q = 0; q < na; ++q) {
                const int i = mod(rows[q] + ((k == 1 || k == 3) ? 1 : 0), nx);
                const int jj = mod(cols[q] + (k >= 2 ? 1 : 0), ny);
                const size_t pix = (size_t)i * ny + jj;
                if (!touched[pix]) {
                  touched[pix] = 1;
                  order.push_back((int)pix);
                }
                dense[pix] += wt[k][q];
              }
            const long long base = ((long long)(j - j0) * nz + e) * (long long)N;
            for (int pix : order) {
              idx.push_back(base + pix);
              w.push_back(dense[pix]);
              dense[pix] = 0.0;
              touched[pix] = 0;
            }
          }
        SK_HIP(hipMemsetAsync(D.p, 0, sizeof(float2) * (size_t)cnt * nz * N, s));
        if (!idx.empty()) {
          if ((size_t)cnt * nz * N > (size_t)INT32_MAX) fail(SK_ERR_INVALID, "potential chunk too large");
          idx32.assign(idx.begin(), idx.end());
          DevBuf<int> idx_d;
          DevBuf<double> w_d;
          idx_d.upload(idx32, s);
          w_d.upload(w, s);
          launch_scatter_deltas(idx_d.p, w_d.p, idx32.size(), D.p, s);
          SK_HIP(hipStreamSynchronize(s));
        }
        fft2_batch(D.p, nx, ny, cnt * nz, -1, nullptr, 1.0f, s);
        launch_combine_elements(D.p, K.p, cnt, nz, N, Sp.p, s);
        fft2_batch(Sp.p, nx, ny, cnt, +1, nullptr, float(1.0 / double(N)), s);
        launch_real_part(Sp.p, Vd.p + (size_t)j0 * N, (size_t)cnt * N, s);
        SK_HIP(hipStreamSynchronize(s));
      }
    }
    finish_potential(*p, Vd.p);
    p->V = std::move(Vd);
    p->has_values = true;
    *out = p.release();
  });
}

int sk_hip_potential_from_atoms_finite(const sk_grid* grid, int nslices, const double* dz, int natoms,
                                       const int* numbers, const double* xyz, int pi_mode, double energy,
                                       const double tilt[2], sk_potential** out) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    if (!out || !dz || (natoms > 0 && (!numbers || !xyz))) fail(SK_ERR_INVALID, "null argument");
    auto p = new_potential(grid, nslices, dz, energy, tilt);
    const int nx = grid->nx, ny = grid->ny;
    const size_t N = (size_t)nx * ny;
    const double sx = grid->extent_x / nx, sy = grid->extent_y / ny;
    std::vector<int> zs;
    for (int a = 0; a < natoms; ++a) {
      if (numbers[a] < 1 || numbers[a] > 103 || !kLobato[numbers[a] - 1].present)
        fail(SK_ERR_INVALID, "element Z = " + std::to_string(numbers[a]) + " has no Lobato parameters");
      for (int d = 0; d < 3; ++d)
        if (!std::isfinite(xyz[3 * a + d])) fail(SK_ERR_INVALID, "non-finite position");
      zs.push_back(numbers[a]);
    }
    std::sort(zs.begin(), zs.end());
    zs.erase(std::unique(zs.begin(), zs.end()), zs.end());
    std::vector<int> zindex(104, -1);
    for (size_t e = 0; e < zs.size(); ++e) zindex[zs[e]] = (int)e;
    // Lobato A, B per element; the z range beyond which an atom's share of a slice is below 1e-11 of its total:
    // c(k, T) decays as e^{-T kappa} with kappa >= min B
    std::vector<double> AB(10 * std::max<size_t>(1, zs.size()));
    double tcut = 0;
    const double kappa = 0.0208865737082965;
    for (size_t e = 0; e < zs.size(); ++e) {
      const LobatoRow& row = kLobato[zs[e] - 1];
      double bmin = 1e300;
      for (int i = 0; i < 5; ++i) {
        AB[10 * e + i] = M_PI * M_PI * row.a[i] / std::pow(row.b[i], 1.5) / kappa;
        AB[10 * e + 5 + i] = 2 * M_PI / std::sqrt(row.b[i]);
        bmin = std::min(bmin, AB[10 * e + 5 + i]);
      }
      tcut = std::max(tcut, 30.0 / bmin);
    }
    // atoms sorted by depth, with their bilinear corner and weights (ProjectedPotential.swift:178-200)
    std::vector<int> order(natoms);
    for (int a = 0; a < natoms; ++a) order[a] = a;
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return xyz[3 * u + 2] < xyz[3 * v + 2]; });
    std::vector<int> r(std::max(1, natoms)), cc(std::max(1, natoms)), el(std::max(1, natoms));
    std::vector<double> w(4 * (size_t)std::max(1, natoms)), zz(std::max(1, natoms));
    for (int k = 0; k < natoms; ++k) {
      const int a = order[k];
      const double px = xyz[3 * a] / sx, py = xyz[3 * a + 1] / sy;
      const double fr = std::floor(px), fc = std::floor(py);
      const double x = px - fr, y = py - fc, xyv = x * y;
      r[k] = (int)(((long long)fr % nx + nx) % nx);
      cc[k] = (int)(((long long)fc % ny + ny) % ny);
      w[4 * k] = 1 + xyv - y - x;
      w[4 * k + 1] = x - xyv;
      w[4 * k + 2] = y - xyv;
      w[4 * k + 3] = xyv;
      zz[k] = xyz[3 * a + 2];
      el[k] = zindex[numbers[a]];
    }
    std::vector<double> z0(nslices), z1(nslices);
    std::vector<int> lo(nslices), hi(nslices);
    double acc = 0;
    for (int j = 0; j < nslices; ++j) {
      z0[j] = acc;
      acc += dz[j];
      z1[j] = acc;
      lo[j] = int(std::lower_bound(zz.begin(), zz.begin() + natoms, z0[j] - tcut) - zz.begin());
      hi[j] = int(std::upper_bound(zz.begin(), zz.begin() + natoms, z1[j] + tcut) - zz.begin());
    }
    DevBuf<int> r_d, c_d, e_d, lo_d, hi_d;
    DevBuf<double> w_d, z_d, ab_d, z0_d, z1_d;
    r_d.upload(r, s);
    c_d.upload(cc, s);
    e_d.upload(el, s);
    w_d.upload(w, s);
    z_d.upload(zz, s);
    ab_d.upload(AB, s);
    FiniteAtoms at{r_d.p, c_d.p, w_d.p, z_d.p, e_d.p, ab_d.p};
    DevBuf<double> Vd;
    Vd.alloc(N * nslices);
    const int chunk = std::max<int>(1, (int)std::min<size_t>(nslices, (size_t(1) << 30) / (N * sizeof(float2))));
    DevBuf<float2> Sp;
    Sp.alloc((size_t)chunk * N);
    for (int j0 = 0; j0 < nslices; j0 += chunk) {
      const int cnt = std::min(chunk, nslices - j0);
      z0_d.upload(z0.data() + j0, cnt, s);
      z1_d.upload(z1.data() + j0, cnt, s);
      lo_d.upload(lo.data() + j0, cnt, s);
      hi_d.upload(hi.data() + j0, cnt, s);
      launch_finite_slices(nx, ny, grid->extent_x, grid->extent_y, pi_mode, at, z0_d.p, z1_d.p, lo_d.p, hi_d.p, cnt,
                           Sp.p, s);
      fft2_batch(Sp.p, nx, ny, cnt, +1, nullptr, float(1.0 / double(N)), s);
      launch_real_part(Sp.p, Vd.p + (size_t)j0 * N, (size_t)cnt * N, s);
      SK_HIP(hipStreamSynchronize(s));
    }
    finish_potential(*p, Vd.p);
    p->V = std::move(Vd);
    p->has_values = true;
    *out = p.release();
  });
}

void sk_hip_potential_release(sk_potential* p) {
  std::lock_guard<std::recursive_mutex> g(lock());
  delete p;
}

int sk_hip_potential_slice_count(const sk_potential* p) { return p ? p->nslices : SK_ERR_INVALID; }

int sk_hip_potential_grid(const sk_potential* p, sk_grid* grid) {
  if (!p || !grid) return SK_ERR_INVALID;
  *grid = p->grid;
  return SK_OK;
}

int sk_hip_potential_values(const sk_potential* p, double* V) {
  return guarded([&] {
    Context& c = require();
    if (!p || !V) fail(SK_ERR_INVALID, "null argument");
    if (!p->has_values) fail(SK_ERR_INVALID, "this potential was made from values, not built on the device");
    const size_t n = (size_t)p->grid.nx * p->grid.ny * p->nslices;
    SK_HIP(hipMemcpyAsync(V, p->V.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipStreamSynchronize(c.stream));
  });
}

int sk_hip_potential_transmissions(const sk_potential* p, double* t_re, double* t_im) {
  return guarded([&] {
    Context& c = require();
    if (!p || !t_re || !t_im) fail(SK_ERR_INVALID, "null argument");
    const size_t n = (size_t)p->grid.nx * p->grid.ny * p->nslices;
    DevBuf<double> re(n), im(n);
    launch_complex_to_split(p->t.p, re.p, im.p, n, 1.0, c.stream);
    SK_HIP(hipMemcpyAsync(t_re, re.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipMemcpyAsync(t_im, im.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SK_HIP(hipStreamSynchronize(c.stream));
  });
}

static void run_waves(const sk_potential* p, int count, const double* in_re, const double* in_im,
                      const std::vector<int>& planes, double* out_re, double* out_im, const sk_options* opt) {
  Context& c = require();
  hipStream_t s = c.stream;
  set_col_tile(c, opt);
  const size_t N = (size_t)p->grid.nx * p->grid.ny;
  const int nplanes = (int)planes.size();
  // plane j -> output slot (the incident plane -1 is copied on the host side)
  std::vector<int> plane_out(p->nslices, -1);
  int slot = 0;
  for (int q = 0; q < nplanes; ++q)
    if (planes[q] >= 0) plane_out[planes[q]] = slot++;
  const int nslot = slot;
  const int B = default_batch(*p, count, opt);
  const bool fast = (!opt || opt->fast_path != 0) && c.precision == 0;  // split-BF16 runs in the generic passes
  Workspace& w = ws();
  w.G.ensure((size_t)B * p->grid.nx * p->nc);
  w.full.ensure((size_t)B * N);
  w.planes.ensure((size_t)std::max(1, nslot) * B * N);
  w.tmp_re.ensure((size_t)B * N);
  w.tmp_im.ensure((size_t)B * N);
  Timer tm;
  for (int b0 = 0; b0 < count; b0 += B) {
    const int nb = std::min(B, count - b0);
    SK_HIP(hipMemcpyAsync(w.tmp_re.p, in_re + (size_t)b0 * N, nb * N * sizeof(double), hipMemcpyHostToDevice, s));
    SK_HIP(hipMemcpyAsync(w.tmp_im.p, in_im + (size_t)b0 * N, nb * N * sizeof(double), hipMemcpyHostToDevice, s));
    launch_real_to_complex(w.tmp_re.p, w.tmp_im.p, w.full.p, nb * N, s);
    slice_loop(*p, nb, w.G.p, true, w.full.p, 0, plane_out.data(), w.planes.p, nullptr, fast, tm);
    for (int q = 0, k = 0; q < nplanes; ++q) {
      double* ore = out_re + ((size_t)q * count + b0) * N;
      double* oim = out_im + ((size_t)q * count + b0) * N;
      if (planes[q] < 0) {
        std::memcpy(ore, in_re + (size_t)b0 * N, nb * N * sizeof(double));
        std::memcpy(oim, in_im + (size_t)b0 * N, nb * N * sizeof(double));
        continue;
      }
      launch_complex_to_split(w.planes.p + (size_t)k * nb * N, w.tmp_re.p, w.tmp_im.p, nb * N, 1.0, s);
      SK_HIP(hipMemcpyAsync(ore, w.tmp_re.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipMemcpyAsync(oim, w.tmp_im.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipStreamSynchronize(s));
      ++k;
    }
  }
  SK_HIP(hipStreamSynchronize(s));
}

int sk_hip_run(const sk_potential* p, const double* in_re, const double* in_im, const int* planes, int nplanes,
               double* out_re, double* out_im, const sk_options* opt) {
  return guarded([&] {
    if (!p || !in_re || !in_im || !planes || !out_re || !out_im || nplanes < 1) fail(SK_ERR_INVALID, "null argument");
    std::vector<int> pl(planes, planes + nplanes);
    for (int q = 0; q < nplanes; ++q) {
      if (pl[q] < -1 || pl[q] >= p->nslices || (q > 0 && pl[q] <= pl[q - 1]))
        fail(SK_ERR_INVALID, "invalid exit planes: afterSlices must be ascending, distinct, within -1 ..< " +
                                 std::to_string(p->nslices));
    }
    // out layout: nplanes x N for one wave
    run_waves(p, 1, in_re, in_im, pl, out_re, out_im, opt);
  });
}

int sk_hip_run_batch(const sk_potential* p, int count, const double* in_re, const double* in_im, double* out_re,
                     double* out_im, const sk_options* opt) {
  return guarded([&] {
    if (!p || count < 0) fail(SK_ERR_INVALID, "invalid argument");
    if (count == 0) return;
    if (!in_re || !in_im || !out_re || !out_im) fail(SK_ERR_INVALID, "null argument");
    std::vector<int> pl{p->nslices - 1};
    run_waves(p, count, in_re, in_im, pl, out_re, out_im, opt);
  });
}

// probe spectrum Psi0 / ||Psi0|| * extra on the device; returns whether its support fits the kept columns
// The probe's centred spectrum Psi_0 = A e^{-i chi} / ||A e^{-i chi}|| times `extra`, as complex64 in `out`
// (Probe.centredSpectrum, Probe.swift:113-136), with its float64 re / im left in tmp_re / tmp_im; norm and
// norm_partial (kSpectrumBlocks values) hold the float64 sum of squares.
constexpr int kSpectrumBlocks = 256;
static void spectrum_into(const sk_grid& g, const sk_probe& pr, double extra, double* tmp_re, double* tmp_im,
                          double* norm, double* norm_partial, float2* out, hipStream_t s) {
  const size_t N = (size_t)g.nx * g.ny;
  launch_probe_spectrum(g.nx, g.ny, g.extent_x, g.extent_y, wavelength(pr.energy), pr.semiangle_mrad, pr.soft,
                        pr.aberrations, tmp_re, tmp_im, s);
  launch_sum_squares(tmp_re, tmp_im, N, norm_partial, kSpectrumBlocks, norm, s);
  launch_normalise_spectrum(tmp_re, tmp_im, norm, extra, N, out, s);
}

static bool build_spectrum(const sk_grid& g, const sk_probe& pr, double extra, const std::vector<int>* keepinv,
                           float2* out) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  Workspace& w = ws();
  const size_t N = (size_t)g.nx * g.ny;
  w.tmp_re.ensure(N);
  w.tmp_im.ensure(N);
  w.norm.ensure(1);
  w.norm_partial.ensure(kSpectrumBlocks);
  spectrum_into(g, pr, extra, w.tmp_re.p, w.tmp_im.p, w.norm.p, w.norm_partial.p, out, s);
  if (!keepinv) return true;
  // support check: any non-zero amplitude on a pruned column?
  std::vector<double> re(N), im(N);
  SK_HIP(hipMemcpyAsync(re.data(), w.tmp_re.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
  SK_HIP(hipMemcpyAsync(im.data(), w.tmp_im.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
  SK_HIP(hipStreamSynchronize(s));
  for (int ix = 0; ix < g.nx; ++ix)
    for (int iy = 0; iy < g.ny; ++iy)
      if ((*keepinv)[iy] < 0 && (re[(size_t)ix * g.ny + iy] != 0.0 || im[(size_t)ix * g.ny + iy] != 0.0)) return false;
  return true;
}

static void check_probe(const sk_probe* pr) {
  if (!pr) fail(SK_ERR_INVALID, "no probe");
  if (!(std::isfinite(pr->energy) && pr->energy > 0)) fail(SK_ERR_INVALID, "electron energy must be finite and positive");
  if (!(pr->semiangle_mrad > 0) || std::isnan(pr->semiangle_mrad))
    fail(SK_ERR_INVALID, "semiangle cutoff must be finite and positive (or infinite)");
  for (double a : pr->aberrations)
    if (!std::isfinite(a)) fail(SK_ERR_INVALID, "invalid aberration: every coefficient must be finite");
  if (!(std::isfinite(pr->tilt[0]) && std::isfinite(pr->tilt[1]))) fail(SK_ERR_INVALID, "invalid tilt: the tilt must be finite");
}

// The arguments of a scan (STEM.scan's validation, STEM.swift:49-81): potentials on one grid, at the probe's energy
// and tilt, finite positions, an output per detector.
static void check_scan(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos,
                       const double* positions, const sk_detector* dets, int ndets, void* const* outputs) {
  if (!pots || npots < 1) fail(SK_ERR_INVALID, "invalid scan: no potential");
  check_probe(probe);
  if (npos < 1 || !positions) fail(SK_ERR_INVALID, "invalid scan: no positions");
  if (ndets < 0 || (ndets > 0 && (!dets || !outputs))) fail(SK_ERR_INVALID, "invalid detectors");
  if (!pots[0]) fail(SK_ERR_INVALID, "null potential");
  const sk_grid g = pots[0]->grid;
  for (int k = 0; k < npots; ++k) {
    if (!pots[k]) fail(SK_ERR_INVALID, "null potential");
    if (!same_grid(pots[k]->grid, g)) fail(SK_ERR_INVALID, "the waves and the potential are on different grids");
    if (pots[k]->energy != probe->energy) fail(SK_ERR_INVALID, "the probe's energy differs from the potential's");
    if (pots[k]->tilt[0] != probe->tilt[0] || pots[k]->tilt[1] != probe->tilt[1])
      fail(SK_ERR_INVALID, "invalid tilt: the probe's tilt differs from the potential's");
  }
  for (int q = 0; q < npos; ++q)
    if (!(std::isfinite(positions[2 * q]) && std::isfinite(positions[2 * q + 1])))
      fail(SK_ERR_INVALID, "invalid position");
}

int sk_hip_probes(const sk_grid* grid, const sk_probe* probe, int npos, const double* positions, double* out_re,
                  double* out_im) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    check_grid(grid);
    check_probe(probe);
    if (npos < 0 || (npos > 0 && (!positions || !out_re || !out_im))) fail(SK_ERR_INVALID, "invalid argument");
    for (int q = 0; q < npos; ++q)
      if (!(std::isfinite(positions[2 * q]) && std::isfinite(positions[2 * q + 1])))
        fail(SK_ERR_INVALID, "invalid position");
    const int nx = grid->nx, ny = grid->ny;
    const size_t N = (size_t)nx * ny;
    Workspace& w = ws();
    w.spec.ensure(N);
    build_spectrum(*grid, *probe, 1.0 / double(N), nullptr, w.spec.p);
    w.positions.upload(positions, 2 * (size_t)npos, s);
    const int B = std::max(1, std::min(npos, 64));
    w.rx.ensure((size_t)B * nx);
    w.ry.ensure((size_t)B * ny);
    w.full.ensure((size_t)B * N);
    DevBuf<double> re((size_t)B * N), im((size_t)B * N);
    for (int b0 = 0; b0 < npos; b0 += B) {
      const int nb = std::min(B, npos - b0);
      launch_ramps(nx, ny, grid->extent_x / nx, grid->extent_y / ny, w.positions.p, b0, nullptr, nb, w.rx.p, w.ry.p,
                   s);
      ColArgs cg{};
      cg.nx = nx;
      cg.ny = ny;
      cg.nc = ny;
      cg.fx = line_plan(nx, s).dev;
      cg.keep = identity_map(ny);
      cg.G = w.full.p;
      cg.spec = w.spec.p;
      cg.rx = w.rx.p;
      cg.ry = w.ry.p;
      cg.flags = COL_GENERATE | COL_INV | COL_OUT;
      col_pass(cg, nb, false, s);
      RowArgs r{};
      r.nx = nx;
      r.ny = ny;
      r.nc = ny;
      r.fy = line_plan(ny, s).dev;
      r.keep = cg.keep;
      r.keepinv = cg.keep;
      r.in_full = w.full.p;
      r.out_full = w.full.p;
      r.in_full_stride = r.out_full_stride = (long long)N;
      r.out_scale = 1.0f;
      r.flags = ROW_IN_FULL | ROW_INV | ROW_OUT_FULL;
      row_pass(r, nb, false, s);
      launch_complex_to_split(w.full.p, re.p, im.p, nb * N, 1.0, s);
      SK_HIP(hipMemcpyAsync(out_re + (size_t)b0 * N, re.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipMemcpyAsync(out_im + (size_t)b0 * N, im.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipStreamSynchronize(s));
    }
  });
}

static CtfParams ctf_params(const sk_ctf* ctf) {
  if (!ctf) fail(SK_ERR_INVALID, "no CTF");
  for (int k = 0; k < 25; ++k)
    if (!std::isfinite(ctf->aberrations[k])) fail(SK_ERR_INVALID, "invalid aberration: every coefficient must be finite");
  if (!(ctf->semiangle_mrad > 0)) fail(SK_ERR_INVALID, "invalid aperture: the semiangle must be positive");
  if (!(std::isfinite(ctf->focal_spread) && ctf->focal_spread >= 0))
    fail(SK_ERR_INVALID, "invalid aberration: focal spread must be finite and >= 0");
  if (!(std::isfinite(ctf->angular_spread) && ctf->angular_spread >= 0))
    fail(SK_ERR_INVALID, "invalid aberration: angular spread must be finite and >= 0");
  CtfParams c{};
  for (int k = 0; k < 25; ++k) c.ab[k] = ctf->aberrations[k];
  c.cutoff_mrad = ctf->semiangle_mrad;
  c.soft = ctf->soft;
  c.focal_spread = ctf->focal_spread;
  c.angular_spread = ctf->angular_spread;
  return c;
}

int sk_hip_ctf_kernel(const sk_grid* grid, double energy, const sk_ctf* ctf, double* out_re, double* out_im) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    check_grid(grid);
    if (!out_re || !out_im || !(std::isfinite(energy) && energy > 0)) fail(SK_ERR_INVALID, "invalid argument");
    const CtfParams cp = ctf_params(ctf);
    const size_t N = (size_t)grid->nx * grid->ny;
    DevBuf<double> re, im;
    re.alloc(N);
    im.alloc(N);
    launch_ctf_kernel(grid->nx, grid->ny, grid->extent_x, grid->extent_y, wavelength(energy), cp, 1.0f, nullptr,
                      re.p, im.p, s);
    SK_HIP(hipMemcpyAsync(out_re, re.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    SK_HIP(hipMemcpyAsync(out_im, im.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    SK_HIP(hipStreamSynchronize(s));
  });
}

int sk_hip_apply_ctf(const sk_grid* grid, double energy, const sk_ctf* ctf, int count, const double* in_re,
                     const double* in_im, double* out_re, double* out_im) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    check_grid(grid);
    if (count < 0 || (count > 0 && (!in_re || !in_im || !out_re || !out_im)) || !(std::isfinite(energy) && energy > 0))
      fail(SK_ERR_INVALID, "invalid argument");
    const CtfParams cp = ctf_params(ctf);
    if (count == 0) return;
    const int nx = grid->nx, ny = grid->ny;
    const size_t N = (size_t)nx * ny;
    // the kernel H (complex64); the transforms run unscaled and the inverse carries 1 / N
    DevBuf<float2> H, buf;
    H.alloc(N);
    launch_ctf_kernel(nx, ny, grid->extent_x, grid->extent_y, wavelength(energy), cp, 1.0f, H.p, nullptr, nullptr, s);
    const int B = (int)std::max<size_t>(1, std::min<size_t>(count, (size_t(1) << 30) / (N * sizeof(float2))));
    buf.alloc((size_t)B * N);
    DevBuf<double> re, im;
    re.alloc((size_t)B * N);
    im.alloc((size_t)B * N);
    for (int b0 = 0; b0 < count; b0 += B) {
      const int nb = std::min(B, count - b0);
      SK_HIP(hipMemcpyAsync(re.p, in_re + (size_t)b0 * N, nb * N * sizeof(double), hipMemcpyHostToDevice, s));
      SK_HIP(hipMemcpyAsync(im.p, in_im + (size_t)b0 * N, nb * N * sizeof(double), hipMemcpyHostToDevice, s));
      launch_real_to_complex(re.p, im.p, buf.p, nb * N, s);
      fft2_batch(buf.p, nx, ny, nb, -1, H.p, float(1.0 / double(N)), s);
      launch_complex_to_split(buf.p, re.p, im.p, nb * N, 1.0, s);
      SK_HIP(hipMemcpyAsync(out_re + (size_t)b0 * N, re.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipMemcpyAsync(out_im + (size_t)b0 * N, im.p, nb * N * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipStreamSynchronize(s));
    }
  });
}

// (destination -> source) map of a Fourier resampling from n to m points (WavesResample.swift:58-66): the lowest m
// frequencies of a crop, all n of a zero-pad, positive frequencies first; -1 for a destination no source fills
static std::vector<int> resample_sources(int n, int m) {
  std::vector<int> src(m, -1);
  if (m <= n) {
    if (m == 1) {
      src[0] = 0;
      return src;
    }
    const auto s = crop_indices(n, m);
    for (int d = 0; d < m; ++d) src[d] = s[d];
    return src;
  }
  if (n == 1) {
    src[0] = 0;
    return src;
  }
  const int positive = n % 2 == 0 ? n / 2 : n / 2 + 1;
  for (int i = 0; i < n; ++i) src[i < positive ? i : m - n + i] = i;
  return src;
}

int sk_hip_scan_waves(const sk_potential* p, const sk_probe* probe, int npos, const double* positions, int rows,
                      int cols, double* out_re, double* out_im, const sk_options* opt) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    set_col_tile(c, opt);
    if (!p || !positions || !out_re || !out_im || npos < 1) fail(SK_ERR_INVALID, "invalid waves scan");
    check_probe(probe);
    const sk_grid g = p->grid;
    const int nx = g.nx, ny = g.ny;
    const size_t N = (size_t)nx * ny;
    if (rows <= 0 || cols <= 0) {
      rows = nx;
      cols = ny;
    }
    if (rows < 2 || cols < 2) fail(SK_ERR_INVALID, "the waves detector needs at least 2 points per axis");
    const bool resample = rows != nx || cols != ny;
    const size_t M = (size_t)rows * cols;
    const bool fast = (!opt || opt->fast_path != 0) && c.precision == 0;
    const int B = default_batch(*p, npos, opt);
    Workspace& w = ws();
    w.G.ensure((size_t)B * nx * p->nc);
    w.full.ensure((size_t)B * N);
    w.planes.ensure((size_t)B * std::max(N, M));
    w.rx.ensure((size_t)B * nx);
    w.ry.ensure((size_t)B * ny);
    w.tmp_re.ensure((size_t)B * M);
    w.tmp_im.ensure((size_t)B * M);
    w.spec.ensure(N);
    const bool fits = build_spectrum(g, *probe, 1.0 / double(N), &p->keepinv, w.spec.p);
    w.positions.upload(positions, 2 * (size_t)npos, s);
    DevBuf<int> srcx, srcy;
    DevBuf<float2> spectra;
    if (resample) {
      srcx.upload(resample_sources(nx, rows), s);
      srcy.upload(resample_sources(ny, cols), s);
      spectra.alloc((size_t)B * M);
    }
    std::vector<int> plane_out(p->nslices, -1);
    plane_out[p->nslices - 1] = 0;
    Timer tm;
    for (int b0 = 0; b0 < npos; b0 += B) {
      const int nb = std::min(B, npos - b0);
      launch_ramps(nx, ny, g.extent_x / nx, g.extent_y / ny, w.positions.p, b0, nullptr, nb, w.rx.p, w.ry.p, s);
      ColArgs init = col_args(*p, fits ? w.G.p : w.full.p);
      init.spec = w.spec.p;
      init.rx = w.rx.p;
      init.ry = w.ry.p;
      init.flags = COL_GENERATE | COL_INV | COL_OUT;
      if (!fits) {
        init.nc = ny;
        init.keep = identity_map(ny);
      }
      col_pass(init, nb, fast && fits, s);
      // every slice, then the exit wave in real space (the probe's own normalisation: sum |psi|^2 = 1 / N)
      slice_loop(*p, nb, w.G.p, false, w.full.p, fits ? 0 : 1, plane_out.data(), w.planes.p, nullptr, fast, tm);
      const float2* result = w.planes.p;
      if (resample) {
        // WavesResample.swift:16-49: FFT, keep / pad the lowest frequencies, inverse FFT times (rows cols) / N:
        // the inverse runs unscaled and the whole factor is 1 / N
        fft2_batch(w.planes.p, nx, ny, nb, -1, nullptr, 1.0f, s);
        launch_resample_gather(w.planes.p, nx, ny, srcx.p, srcy.p, rows, cols, nb, spectra.p, s);
        fft2_batch(spectra.p, rows, cols, nb, +1, nullptr, float(1.0 / double(N)), s);
        result = spectra.p;
      }
      launch_complex_to_split(result, w.tmp_re.p, w.tmp_im.p, nb * M, 1.0, s);
      SK_HIP(hipMemcpyAsync(out_re + (size_t)b0 * M, w.tmp_re.p, nb * M * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipMemcpyAsync(out_im + (size_t)b0 * M, w.tmp_im.p, nb * M * sizeof(double), hipMemcpyDeviceToHost, s));
      SK_HIP(hipStreamSynchronize(s));
    }
  });
}

int sk_hip_scan_intensities(const sk_potential* p, const sk_probe* probe, int npos, const double* positions,
                            sk_deliver_fn deliver, void* user, const sk_options* opt) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    set_col_tile(c, opt);
    if (!p) fail(SK_ERR_INVALID, "invalid scan: no potential");
    check_probe(probe);
    if (npos < 1 || !positions) fail(SK_ERR_INVALID, "invalid scan: no positions");
    if (!deliver) fail(SK_ERR_INVALID, "no delivery callback");
    if (p->energy != probe->energy) fail(SK_ERR_INVALID, "the probe's energy differs from the potential's");
    if (p->tilt[0] != probe->tilt[0] || p->tilt[1] != probe->tilt[1])
      fail(SK_ERR_INVALID, "invalid tilt: the probe's tilt differs from the potential's");
    for (int q = 0; q < npos; ++q)
      if (!(std::isfinite(positions[2 * q]) && std::isfinite(positions[2 * q + 1])))
        fail(SK_ERR_INVALID, "invalid position");
    const sk_grid g = p->grid;
    const int nx = g.nx, ny = g.ny;
    const size_t N = (size_t)nx * ny;
    const bool fast = (!opt || opt->fast_path != 0) && c.precision == 0;
    // positions per launch: the scan's batch, within kIntensityBytes of float64 intensities per batch (the device
    // buffer and each of the two pinned host buffers hold one batch)
    int B = default_batch(*p, npos, opt);
    B = (int)std::max<size_t>(1, std::min<size_t>((size_t)B, kIntensityBytes / (N * sizeof(double))));
    Workspace& w = ws();
    w.G.ensure((size_t)B * nx * p->nc);
    w.rx.ensure((size_t)B * nx);
    w.ry.ensure((size_t)B * ny);
    w.spec.ensure(N);
    const bool fits = build_spectrum(g, *probe, 1.0 / double(N), &p->keepinv, w.spec.p);
    if (!fits) w.full.ensure((size_t)B * N);
    w.positions.upload(positions, 2 * (size_t)npos, s);
    DevBuf<double> inten((size_t)B * N);
    // two pinned host buffers: the caller's detectors read one batch while the device computes the next
    struct Pinned {
      double* h[2] = {nullptr, nullptr};
      ~Pinned() {
        for (auto x : h)
          if (x) (void)hipHostFree(x);
      }
    } pin;
    for (auto& x : pin.h) SK_HIP(hipHostMalloc(&x, sizeof(double) * (size_t)B * N, hipHostMallocDefault));
    hipEvent_t done[2];
    SK_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
    SK_HIP(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
    struct Events {
      hipEvent_t* e;
      ~Events() {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
      }
    } evs{done};
    Timer tm;
    int pend_first = -1, pend_count = 0, pend_buf = 0;
    auto hand_over = [&]() {
      if (pend_first < 0) return;
      SK_HIP(hipEventSynchronize(done[pend_buf]));
      if (deliver(user, pend_first, pend_count, pin.h[pend_buf]) != 0)
        fail(SK_ERR_INVALID, "the delivery callback stopped the scan");
      pend_first = -1;
    };
    int buf = 0;
    for (int b0 = 0; b0 < npos; b0 += B) {
      const int nb = std::min(B, npos - b0);
      launch_ramps(nx, ny, g.extent_x / nx, g.extent_y / ny, w.positions.p, b0, nullptr, nb, w.rx.p, w.ry.p, s);
      ColArgs init = col_args(*p, fits ? w.G.p : w.full.p);
      init.spec = w.spec.p;
      init.rx = w.rx.p;
      init.ry = w.ry.p;
      init.flags = COL_GENERATE | COL_INV | COL_OUT;
      if (!fits) {
        init.nc = ny;
        init.keep = identity_map(ny);
      }
      col_pass(init, nb, fast && fits, s);
      // the last slice's column pass stops after the propagator: G then holds FFT2(psi_exit) / N
      ColArgs fin = col_args(*p, w.G.p);
      fin.flags = COL_IN_G | COL_FWD | COL_MUL_P | COL_OUT;
      slice_loop(*p, nb, w.G.p, false, w.full.p, fits ? 0 : 1, nullptr, nullptr, &fin, fast, tm);
      launch_compact_intensity(w.G.p, nx, ny, p->nc, p->keepinv_d.p, nb, double(N) * double(N), inten.p, s);
      SK_HIP(hipMemcpyAsync(pin.h[buf], inten.p, sizeof(double) * (size_t)nb * N, hipMemcpyDeviceToHost, s));
      SK_HIP(hipEventRecord(done[buf], s));
      hand_over();  // the previous batch, while this one runs
      pend_first = b0;
      pend_count = nb;
      pend_buf = buf;
      buf ^= 1;
    }
    hand_over();
    SK_HIP(hipStreamSynchronize(s));
  });
}

int sk_hip_detector_shape(const sk_grid* grid, double energy, const sk_detector* det, int* shape) {
  int count = 0;
  int rc = guarded([&] {
    check_grid(grid);
    if (!det) fail(SK_ERR_INVALID, "no detector");
    int sh[2] = {0, 0};
    count = detector_count(*grid, wavelength(energy), *det, sh);
    if (shape) {
      shape[0] = sh[0];
      shape[1] = sh[1];
    }
  });
  return rc == SK_OK ? count : rc;
}

int sk_hip_scan(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos, const double* positions,
                const sk_detector* dets, int ndets, void* const* outputs, int outputs_on_device,
                const sk_options* opt) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    set_col_tile(c, opt);
    check_scan(pots, npots, probe, npos, positions, dets, ndets, outputs);
    const sk_potential& p0 = *pots[0];
    const sk_grid g = p0.grid;
    const int nx = g.nx, ny = g.ny;
    const size_t N = (size_t)nx * ny;
    const double lambda = wavelength(probe->energy);
    // ---- detector plan ----
    DetPlan dp{};
    Workspace& w = ws();
    std::vector<int> det_slot(ndets, -1), det_pix(ndets, -1), det_count(ndets, 0);
    std::vector<double> sqx, sqy;
    int nmask = 0, ncom = 0, npix = 0, nseg = 0;
    dp.ny = ny;
    w.maps.clear();
    w.weights.clear();
    w.pix.clear();
    w.maps.reserve(4 * kMaxCom + 2 * kMaxPix + kMaxSeg);
    w.weights.reserve(2 * kMaxCom);
    w.pix.reserve(kMaxPix);
    const double rxs = 1 / (double(nx) * (g.extent_x / nx)), rys = 1 / (double(ny) * (g.extent_y / ny));
    const double ax = rxs * lambda * 1e3, ay = rys * lambda * 1e3;
    for (int d = 0; d < ndets; ++d) {
      const sk_detector& D = dets[d];
      det_count[d] = detector_count(g, lambda, D, nullptr);
      if (!outputs[d]) fail(SK_ERR_INVALID, "null detector output");
      if (D.kind == SK_DET_ANNULAR) {
        if (dp.nslots >= kMaxSlots) fail(SK_ERR_INVALID, "too many detectors");
        const double o = annular_outer(g, lambda, D);
        if (sqx.empty()) {
          // the pixel angles (Detectors.swift:43-44, 124-131): fftfreq(n, 1 / d / n), squared in float64
          auto angx = fftfreq(nx, 1 / ax / double(nx)), angy = fftfreq(ny, 1 / ay / double(ny));
          sqx.resize(nx);
          sqy.resize(ny);
          for (int ix = 0; ix < nx; ++ix) sqx[ix] = angx[ix] * angx[ix];
          for (int iy = 0; iy < ny; ++iy) sqy[iy] = angy[iy] * angy[iy];
        }
        det_slot[d] = dp.nslots;
        dp.slot_kind[dp.nslots] = 0;
        dp.slot_arg[dp.nslots] = nmask;
        dp.slot_i2[dp.nslots] = D.inner * D.inner;
        dp.slot_o2[dp.nslots] = o * o;
        dp.nslots++;
        nmask++;
      } else if (D.kind == SK_DET_COM) {
        if (dp.nslots + 2 > kMaxSlots || ncom >= kMaxCom) fail(SK_ERR_INVALID, "too many centre-of-mass detectors");
        int mx, my;
        sk_detector region = D;
        region.crop_nx = region.crop_ny = 0;
        pixelated_size(g, region, mx, my);
        auto mapx = crop_map(nx, mx), mapy = crop_map(ny, my);
        const double per = D.units == 0 ? 1.0 : 1.0 / (lambda * 1e3);
        std::vector<double> wx(nx, 0.0), wy(ny, 0.0);
        for (int ix = 0; ix < nx; ++ix)
          if (mapx[ix] >= 0) wx[ix] = double(mapx[ix] - mx / 2) * ax * per;
        for (int iy = 0; iy < ny; ++iy)
          if (mapy[iy] >= 0) wy[iy] = double(mapy[iy] - my / 2) * ay * per;
        w.maps.emplace_back();
        w.maps.back().upload(mapx, s);
        dp.com_mapx[ncom] = w.maps.back().p;
        w.maps.emplace_back();
        w.maps.back().upload(mapy, s);
        dp.com_mapy[ncom] = w.maps.back().p;
        w.weights.emplace_back();
        w.weights.back().upload(wx, s);
        dp.com_wx[ncom] = w.weights.back().p;
        w.weights.emplace_back();
        w.weights.back().upload(wy, s);
        dp.com_wy[ncom] = w.weights.back().p;
        det_slot[d] = dp.nslots;
        dp.slot_kind[dp.nslots] = 1;
        dp.slot_arg[dp.nslots] = ncom;
        dp.slot_kind[dp.nslots + 1] = 2;
        dp.slot_arg[dp.nslots + 1] = ncom;
        dp.nslots += 2;
        ncom++;
      } else if (D.kind == SK_DET_SEGMENTED) {
        const int nb = D.bins_radial * D.bins_azimuthal;
        if (dp.nslots + nb > kMaxSlots || nseg >= kMaxSeg)
          fail(SK_ERR_INVALID, "too many detectors: the segmented detectors' bins and the other sums exceed " +
                                   std::to_string(kMaxSlots));
        w.maps.emplace_back();
        w.maps.back().upload(segment_map(g, ax, ay, D), s);
        dp.seg_map[nseg] = w.maps.back().p;
        det_slot[d] = dp.nslots;
        for (int b = 0; b < nb; ++b) {
          dp.slot_kind[dp.nslots + b] = 3;
          dp.slot_arg[dp.nslots + b] = nseg;
          dp.slot_bin[dp.nslots + b] = b;
        }
        dp.nslots += nb;
        nseg++;
      } else {
        if (npix >= kMaxPix) fail(SK_ERR_INVALID, "too many pixelated detectors");
        int mx, my;
        pixelated_size(g, D, mx, my);
        auto mapx = crop_map(nx, mx), mapy = crop_map(ny, my);
        w.maps.emplace_back();
        w.maps.back().upload(mapx, s);
        dp.pix_mapx[npix] = w.maps.back().p;
        w.maps.emplace_back();
        w.maps.back().upload(mapy, s);
        dp.pix_mapy[npix] = w.maps.back().p;
        const int cx = (int)crop_indices(nx, mx).size(), cy = (int)crop_indices(ny, my).size();
        dp.pix_my[npix] = cy;
        dp.pix_stride[npix] = (long long)cx * cy;
        w.pix.emplace_back();
        w.pix.back().alloc((size_t)npos * cx * cy);
        SK_HIP(hipMemsetAsync(w.pix.back().p, 0, sizeof(double) * (size_t)npos * cx * cy, s));
        dp.pix_acc[npix] = w.pix.back().p;
        det_pix[d] = npix;
        npix++;
      }
    }
    dp.npix = npix;
    if (!sqx.empty()) {
      w.sqx.upload(sqx, s);
      w.sqy.upload(sqy, s);
      dp.sqx = w.sqx.p;
      dp.sqy = w.sqy.p;
    } else {
      dp.sqx = dp.sqy = nullptr;
    }
    dp.scale = double(N) * double(N);  // |FFT2(psi_exit)|^2 of a probe (no 1/N): P' carries 1/N
    // ---- spectrum and positions ----
    w.spec.ensure(N);
    const bool fits = build_spectrum(g, *probe, 1.0 / double(N), &p0.keepinv, w.spec.p);
    w.positions.upload(positions, 2 * (size_t)npos, s);
    const int nslots = std::max(1, dp.nslots);
    w.acc.ensure((size_t)npos * nslots);
    SK_HIP(hipMemsetAsync(w.acc.p, 0, sizeof(double) * (size_t)npos * nslots, s));
    // ---- batches ----
    const bool fast = (!opt || opt->fast_path != 0) && c.precision == 0;  // split-BF16 runs in the generic passes
    int B = default_batch(p0, npos, opt);
    // lane pipeline: lanes of W positions, each on its own potential and slice at every step
    int pipe = opt ? opt->pipeline : 0;
    // The library's choice: lanes when there are several potentials and one potential's positions fill at most a
    // quarter of a launch. Lanes holding the same positions for different potentials each accumulate into their own
    // copy, so a launch carries up to a cache-sized batch of (position, potential) pairs. Measured at 512^2 over 16
    // configurations of 43-115 slices, lanes against one potential at a time: 1 position 3.44 x, 8 positions 2.50 x,
    // 16 positions 1.62 x, 32 positions 1.04 x, 48 positions 1.01 x, 64 positions 0.92 x, 144 positions 0.92 x.
    if (pipe < 0) pipe = (npots > 1 && 4 * npos <= default_batch(p0, 1 << 30, opt)) ? 1 : 0;
    int W = 0, L = 0;
    if (pipe) {
      const int q = pipe_positions_per_block();
      W = opt && opt->lane_waves > 0 ? opt->lane_waves
                                     : std::max(q, std::min(128, ((npos + 7) / 8 + q - 1) / q * q));
      const int chunks = (npos + W - 1) / W;
      // lanes: as many as a launch of the cache-sized batch holds, up to one per (chunk, potential) item; lanes
      // beyond the chunk count share chunks and need their own accumulator copies, within kLaneCopyBytes
      const int Bp = default_batch(p0, 1 << 30, opt);
      L = (int)std::max<long long>(1, std::min<long long>((long long)chunks * npots, std::max(Bp, W) / W));
      if (L > chunks) {
        double per_lane = (double)npos * std::max(1, dp.nslots);
        for (int k = 0; k < dp.npix; ++k) per_lane += (double)npos * dp.pix_stride[k];
        per_lane *= sizeof(double);
        const int fit = (int)std::min<double>(L, (double)kLaneCopyBytes / per_lane);
        L = std::max(std::min(L, chunks), fit);
      }
      if (!(fast && fits && pipeline_supported(nx, ny, p0.nc, W))) {
        if (opt && opt->pipeline > 0)
          fail(SK_ERR_INVALID, "the lane pipeline needs the fast passes (grid lengths 256 or 512), lanes of a "
                               "multiple of " + std::to_string(q) + " positions and a probe inside the bandlimit");
        pipe = 0;
      } else {
        B = L * W;
      }
    }
    // Tiles: a batch (positions per launch: a host chunk, or one graph replay) runs as parts of T positions, each
    // part through the whole slice loop. The caller's tile, or, when the caller names a batch larger than the
    // last-level cache holds, parts sized to the cache, so the waves stay in the cache between passes whatever the
    // batch (the default batch is itself cache-sized: one part).
    const bool user_batch = opt && opt->batch > 0, user_tile = opt && opt->tile > 0;
    if (opt && opt->tile < 0) fail(SK_ERR_INVALID, "tile must be 0 (the library's choice) or positive");
    int T = B;
    if (!pipe) {
      if (user_tile) T = std::min(opt->tile, B);
      else if (user_batch) T = std::min(B, default_batch(p0, B, nullptr));
    }
    // split scan: the parts of each tile run on their own streams, one potential at a time, through the fast
    // passes with the probe inside the bandlimit (with graph replay, each batch's graph forks into the streams and
    // joins them at its end). A timed scan (per-launch events) keeps the tile and runs it on one stream.
    const bool graph = opt && opt->use_graph;
    int split = 1;
    if (!pipe && fits && fast &&
        fast_col_supported(nx, p0.nc, COL_IN_G | COL_FWD | COL_MUL_P | COL_INV | COL_OUT)) {
      const int want = opt ? opt->streams : 0;
      if (want > 0) {
        split = want;
      } else {
        // at 512 x 512 the split scan takes 3/4 of the cache-resident batch
        const int Ts = (!user_tile && nx == 512 && ny == 512)
                           ? std::min(T, default_batch(p0, user_batch ? B : npos, nullptr, 0.75))
                           : T;
        if (Ts >= kMinSplitBatch) {
          split = kSplitStreams;
          T = Ts;
          if (!user_batch) B = T;
        }
      }
      split = std::max(1, std::min(split, T));
      if (c.timing) split = 1;
    }
    w.G.ensure((size_t)T * nx * p0.nc);
    w.rx.ensure((size_t)T * nx);
    w.ry.ensure((size_t)T * ny);
    if (!fits) w.full.ensure((size_t)T * N);
    const int bpw = detect_blocks_per_wave(p0, fast, pipe, s);
    dp.blocks_per_wave = bpw;
    w.partial.ensure((size_t)T * nslots * bpw);
    dp.partial = w.partial.p;
    Timer tm;
    tm.on = c.timing;
    sk_timing timing{};
    hipEvent_t e0, e1;
    SK_HIP(hipEventCreate(&e0));
    SK_HIP(hipEventCreate(&e1));
    SK_HIP(hipEventRecord(e0, s));
    // One batch of nb positions starting at b0 (+ *pos_dev when replayed from a graph) through potential P:
    // ramps, the first column pass (the probe, generated), every slice, the detector reduction.
    // `ss` and `off`: the stream and the first workspace position of a split scan's half (s and 0 otherwise).
    auto enqueue_batch = [&](const sk_potential& P, int b0, const int* pos_dev, int nb, Timer& t,
                             hipStream_t ss = nullptr, int off = 0) {
      DetPlan dph = dp;
      dph.partial = dp.partial + (size_t)off * nslots * bpw;
      scan_batch(P, w.spec.p, w.positions.p, fits, fast, w.G.p + (size_t)off * nx * p0.nc, w.full.p,
                 w.rx.p + (size_t)off * nx, w.ry.p + (size_t)off * ny, dph, b0, pos_dev, nb, w.acc.p, t,
                 ss ? ss : s);
    };
    // the streams of a split scan's parts: the scan's own, then side streams created on first use
    auto part_streams = [&](int k) {
      while ((int)c.side.size() < k - 1) {
        hipStream_t t;
        SK_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        c.side.push_back(t);
      }
      std::vector<hipStream_t> st(1, s);
      for (int q = 0; q < k - 1; ++q) st.push_back(c.side[q]);
      return st;
    };
    // One batch of nb positions starting at b0 through P: tiles of at most T positions, spread evenly, in order;
    // each tile's positions split over the streams `st` in parts of H = ceil(T / k), stream q always on the
    // workspace positions [q H, q H + H): a stream runs its parts of consecutive tiles one after another, and no two
    // streams ever share workspace, whatever the tile's size.
    // `only` >= 0: only stream `only`'s parts (a graph of that stream's share).
    auto enqueue_tiles = [&](const sk_potential& P, int b0, const int* pos_dev, int nb, Timer& t,
                             const std::vector<hipStream_t>& st, int only = -1) {
      const int tiles = (nb + T - 1) / T;
      const int Tp = (nb + tiles - 1) / tiles;
      const int k = (int)st.size();
      const int H = (T + k - 1) / k;
      for (int i = 0; i < tiles; ++i) {
        const int first = i * Tp, count = std::min(Tp, nb - first);
        if (count <= 0) break;
        for (int q = 0; q < k && q * H < count; ++q)
          if (only < 0 || q == only)
            enqueue_batch(P, b0 + first + q * H, pos_dev, std::min(H, count - q * H), t, st[q], q * H);
        if (k == 1) t.harvest(timing);
      }
    };
    // One batch through its slice loop launched kernel by kernel: on the scan's stream, or forked over the split
    // streams and joined back into it.
    auto run_batch_eager = [&](const sk_potential& P, int b0, int nb) {
      if (split <= 1) {
        enqueue_tiles(P, b0, nullptr, nb, tm, std::vector<hipStream_t>(1, s));
        return;
      }
      const std::vector<hipStream_t> st = part_streams(split);
      hipEvent_t fork;
      SK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
      SK_HIP(hipEventRecord(fork, s));
      for (int q = 1; q < split; ++q) SK_HIP(hipStreamWaitEvent(st[q], fork, 0));
      enqueue_tiles(P, b0, nullptr, nb, tm, st);
      std::vector<hipEvent_t> joins(split - 1);
      for (int q = 1; q < split; ++q) {
        SK_HIP(hipEventCreateWithFlags(&joins[q - 1], hipEventDisableTiming));
        SK_HIP(hipEventRecord(joins[q - 1], st[q]));
        SK_HIP(hipStreamWaitEvent(s, joins[q - 1], 0));
      }
      SK_HIP(hipEventDestroy(fork));
      for (auto e : joins) SK_HIP(hipEventDestroy(e));
    };
    if (pipe) {
      run_pipeline(pots, npots, npos, W, L, g, dp, graph, tm, timing);
    } else if (split > 1 && !graph) {
      // the parts of each tile, each through its own slice loop on its own stream
      const int k = split;
      const std::vector<hipStream_t> st = part_streams(k);
      hipEvent_t fork;
      SK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
      SK_HIP(hipEventRecord(fork, s));
      for (int q = 1; q < k; ++q) SK_HIP(hipStreamWaitEvent(st[q], fork, 0));
      for (int pk = 0; pk < npots; ++pk)
        for (int b0 = 0; b0 < npos; b0 += B) enqueue_tiles(*pots[pk], b0, nullptr, std::min(B, npos - b0), tm, st);
      std::vector<hipEvent_t> joins(k - 1);
      for (int q = 1; q < k; ++q) {
        SK_HIP(hipEventCreateWithFlags(&joins[q - 1], hipEventDisableTiming));
        SK_HIP(hipEventRecord(joins[q - 1], st[q]));
        SK_HIP(hipStreamWaitEvent(s, joins[q - 1], 0));
      }
      SK_HIP(hipEventDestroy(fork));
      for (auto e : joins) SK_HIP(hipEventDestroy(e));
    } else if (!graph) {
      const std::vector<hipStream_t> one(1, s);
      for (int k = 0; k < npots; ++k)
        for (int b0 = 0; b0 < npos; b0 += B) enqueue_tiles(*pots[k], b0, nullptr, std::min(B, npos - b0), tm, one);
    } else {
      // hipGraph replay: per potential, one set of graphs per batch shape (full, remainder), replayed for the
      // batches of that shape after the first; the first position of a replay is read from device memory, copied in
      // from pinned host memory before each launch. The first batch of a shape runs kernel by kernel while a second
      // host thread records (captures) the same slice loop on streams of its own and instantiates it, so that the
      // recording never leaves the GPU idle; a shape with a single batch is not captured at all. With the batch
      // split over streams, each stream's share is a graph of its own (a chain of launches on that stream),
      // launched on its stream between a fork and a join recorded outside the graphs: one graph holding both
      // shares as forked branches replayed them with less overlap than the two streams give kernel by kernel.
      line_plan(nx, s);
      line_plan(ny, s);
      identity_map(ny);
      const int nbatches = (npos + B - 1) / B;
      const int count_of[2] = {npos / B, npos % B ? 1 : 0};  // batches of the full and the remainder shape
      int* host_first = nullptr;
      SK_HIP(hipHostMalloc(&host_first, sizeof(int) * nbatches, hipHostMallocDefault));
      for (int i = 0; i < nbatches; ++i) host_first[i] = i * B;
      w.pos_dev.ensure(1);
      Timer off;
      // the fork and join events of split replays
      struct Events {
        std::vector<hipEvent_t> e;
        ~Events() {
          for (auto x : e) (void)hipEventDestroy(x);
        }
      } evs;
      for (int q = 0; q < split; ++q) {
        hipEvent_t x;
        SK_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        evs.e.push_back(x);
      }
      hipEvent_t fork_ev = evs.e[0];
      hipEvent_t* join_ev = evs.e.data() + 1;
      // the streams the replays run on: the scan's own, then the side streams (created before any capture begins)
      const std::vector<hipStream_t> rst = split > 1 ? part_streams(split) : std::vector<hipStream_t>(1, s);
      // the streams the graphs are recorded on: the capture runs on a second host thread while this one launches
      // the first batch on the scan's own streams, so neither waits for the other
      while ((int)c.capture.size() < split) {
        hipStream_t t;
        SK_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        c.capture.push_back(t);
      }
      const std::vector<hipStream_t> cst(c.capture.begin(), c.capture.begin() + split);
      (void)slice_plan(nx, s);  // every plan the recording looks up exists before the two threads share them
      (void)slice_plan(ny, s);
      int device = 0;
      SK_HIP(hipGetDevice(&device));
      // [shape][stream]: per potential; destroyed on the way out, also on errors
      std::vector<hipGraphExec_t> exec(2 * (size_t)split, nullptr);
      auto destroy_execs = [&exec]() {
        for (auto& e : exec) {
          if (e) (void)hipGraphExecDestroy(e);
          e = nullptr;
        }
      };
      try {
        for (int k = 0; k < npots; ++k) {
          bool seen[2] = {false, false};
          for (int i = 0; i < nbatches; ++i) {
            const int nb = std::min(B, npos - i * B);
            const int shape = nb == B ? 0 : 1;
            hipGraphExec_t* ge = exec.data() + (size_t)shape * split;
            if (!seen[shape]) {
              seen[shape] = true;
              // the capture: host-side recording of each stream's share of the slice loop on its capture stream,
              // then its instantiation; it launches nothing
              auto capture = [&]() {
                SK_HIP(hipSetDevice(device));
                for (int q = 0; q < split; ++q) {
                  hipGraph_t gr;
                  hipStream_t cs = cst[q];
                  SK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                  try {
                    enqueue_tiles(*pots[k], 0, w.pos_dev.p, nb, off, cst, split > 1 ? q : -1);
                  } catch (...) {
                    hipGraph_t dropped;
                    (void)hipStreamEndCapture(cs, &dropped);
                    if (dropped) (void)hipGraphDestroy(dropped);
                    throw;
                  }
                  SK_HIP(hipStreamEndCapture(cs, &gr));
                  const hipError_t inst = hipGraphInstantiate(&ge[q], gr, nullptr, nullptr, 0);
                  (void)hipGraphDestroy(gr);
                  SK_HIP(inst);
                }
              };
              std::exception_ptr failed;
              std::thread recorder;
              if (count_of[shape] >= 2)
                recorder = std::thread([&]() {
                  try {
                    capture();
                  } catch (...) {
                    failed = std::current_exception();
                  }
                });
              try {
                run_batch_eager(*pots[k], i * B, nb);
              } catch (...) {
                if (recorder.joinable()) recorder.join();
                throw;
              }
              if (recorder.joinable()) recorder.join();
              if (failed) std::rethrow_exception(failed);
              continue;
            }
            SK_HIP(hipMemcpyAsync(w.pos_dev.p, host_first + i, sizeof(int), hipMemcpyHostToDevice, s));
            if (split > 1) {
              SK_HIP(hipEventRecord(fork_ev, s));
              for (int q = 1; q < split; ++q) SK_HIP(hipStreamWaitEvent(rst[q], fork_ev, 0));
            }
            for (int q = 0; q < split; ++q) SK_HIP(hipGraphLaunch(ge[q], rst[q]));
            // joined before the next batch's copy of its first position, which the kernels above read
            for (int q = 1; q < split; ++q) {
              SK_HIP(hipEventRecord(join_ev[q - 1], rst[q]));
              SK_HIP(hipStreamWaitEvent(s, join_ev[q - 1], 0));
            }
          }
          SK_HIP(hipStreamSynchronize(s));
          destroy_execs();
        }
      } catch (...) {
        (void)hipStreamSynchronize(s);
        for (int q = 1; q < (int)rst.size(); ++q) (void)hipStreamSynchronize(rst[q]);
        destroy_execs();
        (void)hipHostFree(host_first);
        throw;
      }
      SK_HIP(hipStreamSynchronize(s));
      SK_HIP(hipHostFree(host_first));
    }
    SK_HIP(hipEventRecord(e1, s));
    // ---- outputs: averaged over the potentials (STEM.swift:146-150) ----
    const double inv = 1.0 / double(npots);
    std::vector<DevBuf<char>> staging;
    for (int d = 0; d < ndets; ++d) {
      const sk_detector& D = dets[d];
      const size_t cnt = (size_t)det_count[d] * npos;
      const size_t esz = D.out_f32 ? sizeof(float) : sizeof(double);
      void* dst = outputs[d];
      DevBuf<char> tmp;
      if (!outputs_on_device) {
        tmp.alloc(cnt * esz);
        dst = tmp.p;
      }
      if (D.kind == SK_DET_ANNULAR) {
        launch_gather_scalar(w.acc.p, nslots, det_slot[d], npos, inv, dst, 1, D.out_f32, s);
      } else if (D.kind == SK_DET_COM) {
        launch_gather_scalar(w.acc.p, nslots, det_slot[d], npos, inv, dst, 2, D.out_f32, s);
        char* dst1 = static_cast<char*>(dst) + esz;
        launch_gather_scalar(w.acc.p, nslots, det_slot[d] + 1, npos, inv, dst1, 2, D.out_f32, s);
      } else if (D.kind == SK_DET_SEGMENTED) {
        const int nb = D.bins_radial * D.bins_azimuthal;
        for (int b = 0; b < nb; ++b)
          launch_gather_scalar(w.acc.p, nslots, det_slot[d] + b, npos, inv, static_cast<char*>(dst) + b * esz, nb,
                               D.out_f32, s);
      } else {
        launch_finalise_outputs(w.pix[det_pix[d]].p, cnt, inv, dst, D.out_f32, s);
      }
      if (!outputs_on_device) {
        SK_HIP(hipMemcpyAsync(outputs[d], tmp.p, cnt * esz, hipMemcpyDeviceToHost, s));
        staging.push_back(std::move(tmp));
      }
    }
    SK_HIP(hipStreamSynchronize(s));
    float total = 0;
    SK_HIP(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // algorithmic bytes per launch: one complex64 read + write of every stored value of the batch
    timing.total_ms = total;
    if (timing.row_launches) timing.row_ms /= double(timing.row_launches);
    if (timing.col_launches) timing.col_ms /= double(timing.col_launches);
    timing.row_bytes = 2.0 * sizeof(float2) * double(T) * nx * p0.nc;
    timing.col_bytes = timing.row_bytes;
    timing.batch = T;
    timing.group = B;
    c.last = timing;
  });
}

// ---- the scan without allocation ------------------------------------------------------------------------------

int sk_hip_scan_noalloc(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos,
                        const double* positions, const sk_detector* dets, int ndets, void* const* outputs,
                        int outputs_on_device, const sk_options* opt) {
  return guarded([&] {
    Context& c = require();
    hipStream_t s = c.stream;
    if (!c.scratch_dev) fail(SK_ERR_DEVICE, "sk_hip_scan_noalloc: no device address for the scratch array: " +
                                                c.scratch_why);
    set_col_tile(c, opt);
    if (c.precision != 0) fail(SK_ERR_INVALID, "sk_hip_scan_noalloc runs the float32 passes only (precision 0)");
    if (opt && (opt->use_graph || opt->pipeline > 0 || opt->streams > 1 || opt->tile > 0))
      fail(SK_ERR_INVALID, "sk_hip_scan_noalloc runs its batches on one stream, without graph replay, lanes or tiles");
    check_scan(pots, npots, probe, npos, positions, dets, ndets, outputs);
    if (ndets < 1) fail(SK_ERR_INVALID, "sk_hip_scan_noalloc: no detector");
    const sk_potential& p0 = *pots[0];
    const sk_grid g = p0.grid;
    const int nx = g.nx, ny = g.ny;
    if (nx > kNoallocMaxLength || ny > kNoallocMaxLength)
      fail(SK_ERR_MEMORY, "sk_hip_scan_noalloc: grid lengths above " + std::to_string(kNoallocMaxLength) +
                              " exceed its scratch; use sk_hip_scan");
    const size_t N = (size_t)nx * ny;
    const double lambda = wavelength(probe->energy);
    // ---- detectors: annular sums (Detectors.swift:91-136), one slot each ----
    DetPlan dp{};
    dp.ny = ny;
    if (ndets > kMaxSlots) fail(SK_ERR_INVALID, "too many detectors");
    for (int d = 0; d < ndets; ++d) {
      const sk_detector& D = dets[d];
      if (D.kind != SK_DET_ANNULAR)
        fail(SK_ERR_INVALID, "sk_hip_scan_noalloc reduces annular detectors only; use sk_hip_scan");
      (void)detector_count(g, lambda, D, nullptr);  // validates
      if (!outputs[d]) fail(SK_ERR_INVALID, "null detector output");
      const double o = annular_outer(g, lambda, D);
      dp.slot_kind[d] = 0;
      dp.slot_arg[d] = d;
      dp.slot_i2[d] = D.inner * D.inner;
      dp.slot_o2[d] = o * o;
    }
    dp.nslots = ndets;
    dp.npix = 0;
    dp.scale = double(N) * double(N);
    const int nslots = ndets;
    const double rxs = 1 / (double(nx) * (g.extent_x / nx)), rys = 1 / (double(ny) * (g.extent_y / ny));
    const double ax = rxs * lambda * 1e3, ay = rys * lambda * 1e3;
    for (int ix = 0; ix < nx; ++ix) {
      const double a = fftfreq_at(ix, nx, 1 / ax / double(nx));
      noalloc_sq[ix] = a * a;
    }
    for (int iy = 0; iy < ny; ++iy) {
      const double a = fftfreq_at(iy, ny, 1 / ay / double(ny));
      noalloc_sq[nx + iy] = a * a;
    }
    // ---- the scratch: fixed part, then the batches' waves ----
    const bool fast = !opt || opt->fast_path != 0;
    const int bpw = detect_blocks_per_wave(p0, fast, 0, s);
    dp.blocks_per_wave = bpw;
    Arena a{c.scratch_dev, kNoallocScratchBytes, 0};
    const size_t fixed = N * sizeof(float2) + (2 + (size_t)nslots) * npos * sizeof(double) +
                         (outputs_on_device ? 0 : (size_t)ndets * npos * sizeof(double)) +
                         (size_t)(nx + ny) * sizeof(double) + (2 + kSpectrumBlocks) * sizeof(double) + 8 * 256;
    const size_t per_pos = (size_t)nx * p0.nc * sizeof(float2) + (size_t)(nx + ny) * sizeof(float2) +
                           (size_t)nslots * bpw * sizeof(double);
    const size_t need = fixed + std::max(per_pos + 4 * 256, 2 * N * sizeof(double));
    if (need > kNoallocScratchBytes)
      fail(SK_ERR_MEMORY, "sk_hip_scan_noalloc: " + std::to_string(npos) + " positions on a " + std::to_string(nx) +
                              " x " + std::to_string(ny) + " grid need at least " + std::to_string(need) +
                              " bytes of scratch; it holds " + std::to_string(kNoallocScratchBytes) +
                              " (use sk_hip_scan)");
    float2* spec = a.take<float2>(N);
    double* pos_d = a.take<double>(2 * (size_t)npos);
    double* acc = a.take<double>((size_t)npos * nslots);
    double* stage = outputs_on_device ? nullptr : a.take<double>((size_t)ndets * npos);
    double* sq = a.take<double>((size_t)nx + ny);
    double* norm = a.take<double>(1);
    int* flag = reinterpret_cast<int*>(a.take<double>(1));
    double* norm_partial = a.take<double>(kSpectrumBlocks);
    // the rest holds T positions' compact waves, ramps and partial sums (and, before the first batch, the spectrum's
    // float64 temporaries, which are dead once the spectrum is built)
    const size_t rest = a.left();
    const int tmax = (int)std::min<size_t>((size_t)npos, (rest - 4 * 256) / per_pos);
    const int T = std::max(1, std::min(tmax, default_batch(p0, npos, opt)));
    unsigned char* region = c.scratch_dev + a.aligned(a.used);
    double* tmp_re = reinterpret_cast<double*>(region);
    double* tmp_im = tmp_re + N;
    float2* G = a.take<float2>((size_t)T * nx * p0.nc);
    float2* rx = a.take<float2>((size_t)T * nx);
    float2* ry = a.take<float2>((size_t)T * ny);
    dp.partial = a.take<double>((size_t)T * nslots * bpw);
    // ---- inputs ----
    SK_HIP(hipMemcpyAsync(pos_d, positions, 2 * sizeof(double) * (size_t)npos, hipMemcpyHostToDevice, s));
    SK_HIP(hipMemcpyAsync(sq, noalloc_sq, sizeof(double) * ((size_t)nx + ny), hipMemcpyHostToDevice, s));
    dp.sqx = sq;
    dp.sqy = sq + nx;
    launch_fill_zero(acc, sizeof(double) * (size_t)npos * nslots, s);
    launch_fill_zero(flag, sizeof(int), s);
    spectrum_into(g, *probe, 1.0 / double(N), tmp_re, tmp_im, norm, norm_partial, spec, s);
    launch_pruned_support(tmp_re, tmp_im, p0.keepinv_d.p, nx, ny, flag, s);
    int outside = 0;
    SK_HIP(hipMemcpyAsync(&outside, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    SK_HIP(hipStreamSynchronize(s));
    if (outside)
      fail(SK_ERR_INVALID, "sk_hip_scan_noalloc: the probe reaches past the bandlimit (it needs full-width waves); "
                           "use sk_hip_scan");
    // ---- the batches: potentials outer, positions inner, each batch through every slice ----
    Timer off;
    for (int k = 0; k < npots; ++k)
      for (int b0 = 0; b0 < npos; b0 += T)
        scan_batch(*pots[k], spec, pos_d, true, fast, G, nullptr, rx, ry, dp, b0, nullptr, std::min(T, npos - b0),
                   acc, off, s);
    // ---- outputs: averaged over the potentials (STEM.swift:146-150) ----
    const double inv = 1.0 / double(npots);
    for (int d = 0; d < ndets; ++d) {
      const size_t esz = dets[d].out_f32 ? sizeof(float) : sizeof(double);
      void* dst = outputs_on_device ? outputs[d] : static_cast<void*>(stage + (size_t)d * npos);
      launch_gather_scalar(acc, nslots, d, npos, inv, dst, 1, dets[d].out_f32, s);
      if (!outputs_on_device)
        SK_HIP(hipMemcpyAsync(outputs[d], dst, esz * (size_t)npos, hipMemcpyDeviceToHost, s));
    }
    SK_HIP(hipStreamSynchronize(s));
    sk_timing t{};
    t.batch = T;
    t.group = T;
    c.last = t;
  });
}

}  // extern "C"
