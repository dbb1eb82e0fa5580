// sk_internal.h — shared definitions of the MI355X multislice library (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/simulato_hip.h"

namespace sk {

// ---- errors ----------------------------------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& m) { throw Error(code, m); }

// The C ABI's process-wide lock (one call at a time, as the reference's runLock) and the calling thread's last error
// message (sk_api.hip).
std::recursive_mutex& api_lock();
std::string& last_error();

// Runs f under the lock, turning a thrown Error into its status code and message.
template <class F>
int guarded_call(F&& f) {
  std::lock_guard<std::recursive_mutex> g(api_lock());
  try {
    f();
    return SK_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return SK_ERR_DEVICE;
  }
}

#define SK_HIP(call)                                                                                         \
  do {                                                                                                       \
    hipError_t e_ = (call);                                                                                  \
    if (e_ != hipSuccess)                                                                                    \
      ::sk::fail(e_ == hipErrorOutOfMemory ? SK_ERR_MEMORY : SK_ERR_DEVICE,                                  \
                 std::string(#call) + ": " + hipGetErrorString(e_));                                         \
  } while (0)

// ---- device buffers ----------------------------------------------------------------------------------------------
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    SK_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // copies from (pageable) host memory and waits, so the host array may be released on return
  void upload(const T* h, size_t count, hipStream_t s) {
    ensure(count);
    if (count) {
      SK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      SK_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// ---- 1-D transform plan (device view) -------------------------------------------------------------------------------
constexpr int kMaxStages = 20;

// A 1-D complex transform of logical length n, done in LDS as a Stockham transform of length m (= n, or the
// power-of-two Bluestein length).
struct LineFFT {
  int n, m, nst, blue;
  int radix[kMaxStages];
  const float2* roots;  // m entries, e^{-2 pi i k / m}
  const float2* chirp;  // n entries, e^{-pi i k^2 / n} (Bluestein)
  const float2* khat;   // m entries, FFT_m of the conjugated chirp, times 1/m (Bluestein)
  // split-BF16 MFMA line transform of length m (m = 256, 384, 512 or 1024; sk_mfma.h): 0 = the float32 Stockham
  // stages above
  int use_mfma;
  const void* mtab16;   // DFT-16 block matrices in MFMA B-operand order, bfloat16 hi/lo
  const void* mtabn2;   // DFT-(m/16) block matrices, same order
  const float2* mtw;    // e^{-2 pi i j2 k1 / m}, [k1 16][j2 m/16]
};

// Host-owned tables of one length.
struct LinePlan {
  LineFFT dev{};
  DevBuf<float2> roots, chirp, khat;
};

// Returns the shared plan of length n (built once per process and device).
const LinePlan& line_plan(int n, hipStream_t s);
// The plan of length n with the split-BF16 MFMA line transform, or fails: n = 256, 384, 512 or 1024, or a length
// done by Bluestein's algorithm whose power-of-two length m is 256, 512 or 1024 (65 <= n <= 512).
const LineFFT& mfma_line_plan(int n, hipStream_t s);
bool mfma_supported(int n);

// ---- lane pipeline ----------------------------------------------------------------------------------------------
// The scan's waves are split into lanes of lane_waves positions. At every step each lane advances one slice of its
// current item (a chunk of positions through one potential), so one launch serves lanes on different potentials and
// slices. The host writes one LaneDesc per lane per step; the kernels read row `*step` of the table.
enum LaneMode : int { LANE_IDLE = 0, LANE_MID = 1, LANE_FINAL = 2, LANE_FINAL_INIT = 3, LANE_INIT = 4 };

struct LaneDesc {
  const float2* t;  // row pass: transmission of (potential, slice); null: the lane is idle
  const float2* P;  // column pass: propagator of (potential, slice)
  int mode;         // column pass: LaneMode
  int pos;          // first position of the current item (detector accumulators)
  int count;        // positions of the current item (<= lane_waves)
  int next_pos;     // FINAL_INIT / INIT: first position of the lane's next item
  int next_count;   // its positions
  int acc_off;      // added to positions where the lane accumulates its detectors (its own accumulator copy)
};

struct PipeArgs {
  const LaneDesc* descs;  // [steps][lanes]; null: not a pipelined launch
  const int* step;        // device step counter (replayed graphs); null: the step is step_now
  int lanes, lane_waves;
  const float2* rx_all;   // [npos][nx] shift ramps of every position
  const float2* ry_all;   // [npos][ny]
  int step_now;           // the step of this launch when there is no device counter
};

// the descriptor row of a pipelined launch
__device__ __forceinline__ int pipe_step(const PipeArgs& p) { return p.step ? *p.step : p.step_now; }

// ---- pass descriptions --------------------------------------------------------------------------------------------
// Row pass (along y, contiguous): flags
enum RowFlags : int {
  ROW_IN_COMPACT = 1,   // input: compact columns of G (zero at the pruned ky)
  ROW_IN_FULL = 2,      // input: a full-width buffer
  ROW_INV = 4,          // apply the unscaled inverse transform along y first
  ROW_WRITE_PLANE = 8,  // write the line after the inverse (real space) to `plane`
  ROW_MUL_T = 16,       // multiply by the transmission t(x, y)
  ROW_FWD = 32,         // apply the forward transform along y
  ROW_OUT_COMPACT = 64, // store the kept columns into G
  ROW_OUT_FULL = 128    // store the full line into `out_full`
};

struct RowArgs {
  int nx, ny, nc, flags;
  LineFFT fy;
  const int* keep;      // nc: kept iy (or -1 past the kept count)
  const int* keepinv;   // ny: compact index of iy or -1
  float2* G;            // compact waves [batch][nx][nc]
  const float2* in_full;  // [batch][nx][ny]
  float2* out_full;       // [batch][nx][ny]
  float2* plane;          // [batch][nx][ny]
  const float2* t;        // [nx][ny]
  float out_scale;        // applied to full / plane outputs
  long long in_full_stride, out_full_stride, plane_stride;
  // the kept set as two ranges: iy in [0, kept_hi] -> c = iy; iy in [kept_lo, ny) -> c = iy - kept_shift
  int kept_hi, kept_lo, kept_shift;
  PipeArgs pipe;          // lane pipeline (fast passes only); pipe.descs null otherwise
  hipEvent_t ev_start, ev_stop;  // host side: events that time this launch (null: untimed)
};

enum ColFlags : int {
  COL_IN_G = 1,        // input: G columns
  COL_GENERATE = 2,    // input: probe spectrum x ramps
  COL_FWD = 4,         // forward transform along x
  COL_MUL_P = 8,       // multiply by the compact propagator
  COL_DETECT = 16,     // detectors on |.|^2 (final pass)
  COL_INV = 32,        // unscaled inverse transform along x
  COL_OUT = 64         // store back into G
};

constexpr int kMaxSlots = 32;
constexpr int kMaxSeg = 4;
constexpr int kMaxPix = 4;
constexpr int kMaxCom = 4;

// Detector reduction plan of the final column pass.
struct DetPlan {
  int nslots;                   // scalar sums per wave
  int slot_kind[kMaxSlots];     // 0: annular, 1: COM x, 2: COM y, 3: a segmented detector's bin
  int slot_arg[kMaxSlots];      // COM index (COM slots), segmented detector index (bin slots)
  int slot_bin[kMaxSlots];      // bin slots: the bin (radial * azimuthal bins + azimuthal)
  const int* seg_map[kMaxSeg];  // [nx][ny] bin of each pixel, or -1 (outside the bins)
  int ny;
  double slot_i2[kMaxSlots];    // annular: inner^2, outer^2 [mrad^2], half-open inner^2 <= a2 < outer^2
  double slot_o2[kMaxSlots];
  const double* sqx;            // [nx] angle_x^2 of the pattern's pixels [mrad^2] (float64, host-rounded)
  const double* sqy;            // [ny]
  const int* com_mapx[kMaxCom];   // [nx] crop output index or -1
  const int* com_mapy[kMaxCom];   // [ny]
  const double* com_wx[kMaxCom];  // [nx] angle of the crop row (mrad or 1/A)
  const double* com_wy[kMaxCom];  // [ny]
  int npix;
  const int* pix_mapx[kMaxPix];
  const int* pix_mapy[kMaxPix];
  int pix_my[kMaxPix];
  double* pix_acc[kMaxPix];       // [npos][mx][my] accumulators
  long long pix_stride[kMaxPix];  // mx * my
  double* partial;                // [batch][nslots][blocks per wave]
  int blocks_per_wave;
  double scale;                   // intensity scale (N^2 for a probe)
};

struct ColArgs {
  int nx, ny, nc, flags;
  LineFFT fx;
  const int* keep;        // nc: the iy of each compact column (or -1)
  float2* G;              // [batch][nx][nc]
  const float2* P;        // [nx][nc]
  const float2* spec;     // [nx][ny] probe spectrum / N
  const float2* rx;       // [batch][nx]
  const float2* ry;       // [batch][ny]
  int pos0;               // position index of wave 0 (detectors)
  const int* pos_dev;     // optional device offset added to pos0 (a replayed graph reads its batch here)
  int col_tile;           // 512-point fast column pass: columns per block, 8 or 16 (0 = the default)
  DetPlan det;
  PipeArgs pipe;          // lane pipeline (fast passes only); pipe.descs null otherwise
  hipEvent_t ev_start, ev_stop;  // host side: events that time this launch (null: untimed)
};

__device__ __forceinline__ long long first_position(int pos0, const int* pos_dev) {
  return (long long)pos0 + (pos_dev ? *pos_dev : 0);
}

// launchers (sk_passes.hip)
void launch_row_generic(const RowArgs& a, int batch, hipStream_t s);
void launch_col_generic(const ColArgs& a, int batch, hipStream_t s);
void launch_reduce_partials(const DetPlan& d, int batch, int pos0, const int* pos_dev, double* acc, hipStream_t s);
int col_generic_lines(const LineFFT& p);  // columns per block of the generic column pass of plan p

// fast passes (sk_fast.hip): return false when the length has no fast kernel
bool fast_row_supported(int nx, int ny, int flags);
bool fast_col_supported(int nx, int nc, int flags);
void launch_row_fast(const RowArgs& a, int batch, hipStream_t s);
void launch_col_fast(const ColArgs& a, int batch, hipStream_t s);
// Bytes of compact waves (batch x nx x nc complex64) a batch may hold and still stay in the last-level cache between
// passes: the 256 MiB Infinity Cache of an MI355X less room for t, P and the partial sums (measured at 512 x 512:
// 176 positions, 242 MiB, run from the cache; 184 positions, 253 MiB, at HBM rates).
constexpr size_t kCachedWaveBytes = size_t(240) << 20;
// columns per block the fast column pass uses for a request of 0, 8 or 16 and a batch of `wave_bytes` compact waves
int fast_col_tile(int nx, int requested, double wave_bytes);
int fast_col_blocks_per_wave(int nx, int nc, int pipeline);  // blocks of the detecting column pass per wave
// register-resident passes for lines of 384 and 1024 points and Bluestein's lengths 257..512 on 1024 points
// (sk_reg.hip): the row modes of the fast passes, and the middle slices' column pass
bool reg_row_supported(const RowArgs& a);
bool reg_col_supported(const ColArgs& a);
int reg_col_blocks_per_wave(const ColArgs& a);  // blocks per wave of the register column pass (its detector partials)
void launch_row_reg(const RowArgs& a, int batch, hipStream_t s);
void launch_col_reg(const ColArgs& a, int batch, hipStream_t s);
// lane pipeline launches (sk_fast.hip): the row pass and the column pass of one step, the partial-sum reduction of
// the lanes that finished an item, and the step counter's increment
bool pipeline_supported(int nx, int ny, int nc, int lane_waves);
void launch_row_pipe(const RowArgs& a, int batch, hipStream_t s);
// modes: a mask of the lanes served at 512 (1 mid slices, 2 items finishing, 4 items starting); 0 launches nothing
void launch_col_pipe(const ColArgs& a, int batch, hipStream_t s, int modes = 7);
void launch_reduce_pipe(const DetPlan& d, const PipeArgs& p, int batch, double* acc, hipStream_t s);
void launch_step_increment(int* step, hipStream_t s);
// out[e] += sum over c = 0 .. copies-1, in that order, of in[c * n + e]
void launch_sum_copies(const double* in, long long n, int copies, double* out, hipStream_t s);
int pipe_positions_per_block();
// positions per launch the scan takes for `count` positions through p when the caller names no batch (sk_api.hip)
int default_scan_batch(const sk_potential& p, int count);

// elementwise kernels (sk_prepare.hip)
void launch_phase_to_complex(const double* V, double sigma, float2* out, size_t n, hipStream_t s);
void launch_real_to_complex(const double* re, const double* im, float2* out, size_t n, hipStream_t s);
void launch_complex_to_split(const float2* in, double* re, double* im, size_t n, double scale, hipStream_t s);
// float64 |G|^2 * scale of compact spectra on the full grid, zero at the pruned columns
void launch_compact_intensity(const float2* G, int nx, int ny, int nc, const int* keepinv, size_t batch, double scale,
                              double* out, hipStream_t s);
// the CTF's parameters on the device (ContrastTransferFunction.swift:36-71)
struct CtfParams {
  double ab[25];
  double cutoff_mrad;  // +inf: no aperture
  int soft;
  double focal_spread;
  double angular_spread;
};
void launch_ctf_kernel(int nx, int ny, double ex, double ey, double lambda, const CtfParams& c, float scale,
                       float2* out, double* re, double* im, hipStream_t s);
void launch_resample_gather(const float2* in, int nx, int ny, const int* srcx, const int* srcy, int rows, int cols,
                            size_t batch, float2* out, hipStream_t s);
void launch_antialias(int nx, int ny, double ex, double ey, double* out, hipStream_t s);
void launch_propagator(int nx, int ny, double ex, double ey, double lambda, double dz, double tx, double ty,
                       const double* aperture, const int* keep, int nc, double scale, float2* out, hipStream_t s);
void launch_aperture_as_complex(int nx, int ny, const double* aperture, double scale, float2* out, hipStream_t s);
void launch_probe_spectrum(int nx, int ny, double ex, double ey, double lambda, double cutoff, int soft,
                           const double* ab, double* re, double* im, hipStream_t s);
void launch_sum_squares(const double* re, const double* im, size_t n, double* partial, int blocks, double* out,
                        hipStream_t s);
void launch_normalise_spectrum(const double* re, const double* im, const double* norm2, double extra,
                               size_t n, float2* out, hipStream_t s);
void launch_ramps(int nx, int ny, double sx, double sy, const double* positions, int pos0, const int* pos_dev, int count,
                  float2* rx, float2* ry, hipStream_t s);
void launch_scatter_deltas(const int* idx, const double* w, size_t n, float2* out, hipStream_t s);
void launch_lobato_kernels(int nx, int ny, double ex, double ey, const int* zs, int nz, int pi_mode,
                           const double* table, float* out, hipStream_t s);
void launch_combine_elements(const float2* D, const float* K, int nslices, int nz, size_t N, float2* out,
                             hipStream_t s);
void launch_real_part(const float2* in, double* out, size_t n, hipStream_t s);
// finite projection: atoms sorted by z, with their bilinear corner (r, c), weights w[4], depth z, element index e,
// and per element the Lobato A[5], B[5]
struct FiniteAtoms {
  const int* r;
  const int* c;
  const double* w;
  const double* z;
  const int* e;
  const double* AB;
};
void launch_finite_slices(int nx, int ny, double ex, double ey, int pi_mode, const FiniteAtoms& at, const double* z0,
                          const double* z1, const int* lo, const int* hi, int cnt, float2* out, hipStream_t s);
void launch_finalise_outputs(const double* acc, size_t n, double inv, void* out, int f32, hipStream_t s);
void launch_gather_scalar(const double* acc, int nslots, int slot, int npos, double inv, void* out, int stride,
                          int f32, hipStream_t s);
// clears `bytes` (a multiple of 4) of device memory with a kernel
void launch_fill_zero(void* p, size_t bytes, hipStream_t s);
// *flag = 1 when the float64 spectrum [nx][ny] is non-zero on a column with keepinv[iy] < 0 (*flag cleared first)
void launch_pruned_support(const double* re, const double* im, const int* keepinv, int nx, int ny, int* flag,
                           hipStream_t s);

// ---- device complex helpers ------------------------------------------------------------------------------------
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

}  // namespace sk
