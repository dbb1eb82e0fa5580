/*
 * simulato_hip.h — the C ABI of the MI355X (gfx950) multislice STEM library.
 *
 * A `.hip` backend of SimulatoKit sits beside `.cpu` and `.metal` in `Backend` (Sources/SimulatoMultislice/
 * Multislice.swift:43-46) and reaches the GPU only through these entry points, which replace the Metal engine's
 * internal interface (Sources/SimulatoMultislice/MetalEngine.swift, MetalMultislice.swift). Plain C types only.
 *
 * Conventions (as the reference's Grid/Waves, Sources/SimulatoCore/Grid.swift:7-91, Fields.swift:8-29):
 *   - arrays are row-major [ix * ny + iy], x the first index;
 *   - host complex arrays are split re/im float64 (the layout of `ComplexField`);
 *   - forward FFT unscaled e^{-2 pi i}, inverse scaled 1/(nx ny) (FFT2D.swift:198-215);
 *   - every entry point returns SK_OK (0) or a negative status; the message is in sk_hip_last_error()
 *     (per thread), as SimulatoError.metalUnavailable / .metalFailure carry theirs (SimulatoError.swift:22-23).
 *   - host arrays are caller-owned and read or written only during the call; device memory is owned by handles.
 *   - calls are serialised by a process-wide lock (the reference's `runLock`, MetalEngine.swift:51).
 */
#ifndef SIMULATO_HIP_H
#define SIMULATO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum sk_status {
  SK_OK = 0,
  SK_ERR_UNAVAILABLE = -1, /* no HIP device (SimulatoError.metalUnavailable) */
  SK_ERR_INVALID = -2,     /* a refused input (the SimulatoError cases thrown before work) */
  SK_ERR_DEVICE = -3,      /* a HIP runtime failure (SimulatoError.metalFailure) */
  SK_ERR_MEMORY = -4       /* a device allocation failed */
};

/* An nx x ny lateral grid over extent_x x extent_y Angstrom (Grid.swift:8-27). */
typedef struct sk_grid {
  int nx, ny;
  double extent_x, extent_y;
} sk_grid;

/* A device-resident prepared potential: per slice the bandlimited transmission function (complex64) and the
 * propagators of its distinct slice thicknesses, for one beam energy and tilt. Immutable after creation; may be
 * shared between calls (MetalPreparedPotential, MetalEngine.swift:277-332). */
typedef struct sk_potential sk_potential;

/* The probe (Sources/SimulatoProbe/Probe.swift:12-69, Aberrations.swift:10-60). aberrations[] in the order
 * C10 C12 phi12 C21 phi21 C23 phi23 C30 C32 phi32 C34 phi34 C41 phi41 C43 phi43 C45 phi45 C50 C52 phi52 C54 phi54
 * C56 phi56 (Angstrom, rad). semiangle_mrad may be +infinity (no aperture). */
typedef struct sk_probe {
  double energy;          /* eV */
  double semiangle_mrad;  /* aperture cutoff */
  int soft;               /* abTEM's soft edge (default 1) */
  double aberrations[25];
  double tilt[2];         /* mrad, enters the propagators only */
} sk_probe;

/* Detector kinds (Sources/SimulatoDetector/Detectors.swift:91-263). */
enum sk_detector_kind { SK_DET_ANNULAR = 0, SK_DET_PIXELATED = 1, SK_DET_COM = 2, SK_DET_SEGMENTED = 3 };
enum sk_max_angle { SK_ANGLE_CUTOFF = 0, SK_ANGLE_VALID = 1, SK_ANGLE_FULL = 2 };

typedef struct sk_detector {
  int kind;              /* sk_detector_kind */
  double inner;          /* annular, segmented: inner angle, mrad */
  double outer;          /* annular: outer angle, mrad; < 0 = abTEM's default (floor of the cutoff angle); segmented:
                            outer angle, mrad (> inner) */
  int max_angle;         /* pixelated / COM region: sk_max_angle */
  int units;             /* COM: 0 = mrad, 1 = 1/Angstrom */
  int crop_nx, crop_ny;  /* pixelated: a numeric centred crop of the fftshifted pattern (0, 0 = by max_angle) */
  int out_f32;           /* 1: this detector's output array is float32 instead of float64 */
  int bins_radial;       /* segmented (SegmentedDetector, Detectors.swift:207-263): radial bins between inner and outer */
  int bins_azimuthal;    /* segmented: azimuthal bins; values per position [radial][azimuthal] */
  double rotation;       /* segmented: the azimuth's origin, rad */
} sk_detector;

/* Execution knobs (ExecutionOptions.swift:21-49). None changes which numbers are computed, only their rounding. */
typedef struct sk_options {
  int batch;        /* positions per launch (sk_hip_scan: per host chunk, or per graph replay with use_graph; the
                       other entry points: waves resident per kernel launch); 0 = the library's choice for the grid */
  int use_graph;    /* capture the slice loop in hipGraphs and replay them: per potential, one graph per batch shape
                       and stream, recorded while the first batch of that shape runs kernel by kernel and replayed
                       for the rest (a shape with one batch runs kernel by kernel only) */
  int fast_path;    /* 1 (default): register-resident FFT passes where the length has one; 0: generic LDS passes */
  int pipeline;     /* scan: 1 = lanes of positions advance one slice per launch, each through its own potential
                       and slice (potentials of different slice counts share every launch); 0 = one potential at a
                       time; -1 = the library's choice (lanes when npots > 1 and the positions fill at most a
                       quarter of a launch, else one potential at a time) */
  int lane_waves;   /* pipeline: positions per lane (0 = the library's choice) */
  int col_tile;     /* columns per block of the column pass: 8 or 16 for 512- and 384-point columns, 4 or 8 for
                       1024-point columns and Bluestein's lengths on 1024 points (0 = the library's choice) */
  int precision;    /* transforms of the slice loop: 0 = float32 (default); 1 = split-BF16 on the matrix cores
                       (bfloat16 hi/lo pairs, three products, float32 accumulation; grid lengths 256,
                       384, 512 and 1024, and the Bluestein lengths 65-512 that have a prime factor
                       above 13) */
  int streams;      /* scan, one potential at a time, fast passes, probe inside the bandlimit: the parts of each
                       batch run their slice loops on this many streams (1-4; with use_graph, each stream's share
                       of a batch is a graph of its own); 0 = the library's choice (2 for batches of at least 128 positions,
                       with 3/4 of the cache-resident batch at 512 x 512, else 1); the lane pipeline and other
                       scans use one stream */
  int tile;         /* sk_hip_scan, one potential at a time: positions each kernel launch carries, a part of the
                       batch; 0 = the library's choice: the whole batch when its waves fit the device's last-level
                       cache, else parts that do (on an MI355X, 128-171 positions at 512 x 512), so a batch larger
                       than the cache still runs from the cache between passes; tile = batch streams the batch's
                       waves through HBM */
} sk_options;

/* --- availability (MetalEngine.require / Multislice.metalIsAvailable, MetalEngine.swift:63-77,
 *     MetalMultislice.swift:13-21) ------------------------------------------------------------------------- */
int sk_hip_available(void);
const char* sk_hip_last_error(void);
/* Build-only (no counterpart in the reference): the SHA-256 of the library's sources (simulatokit_amd/csrc and
 * include) that this binary was compiled from, so that a loader can refuse a binary built from other sources. */
const char* sk_hip_build_id(void);

/* --- potentials --------------------------------------------------------------------------------------------- */

/* From projected potentials V [nslices x nx x ny] in V*Angstrom and thicknesses dz [nslices]
 * (MetalMultislice.prepare(_: ProjectedPotential, ...), MetalEngine.swift:310-320; Multislice.transmissionFunctions
 * and .propagator, Multislice.swift:83-130). The transmission phase sigma*V and the propagator phase
 * -pi*lambda*k^2*dz are evaluated in float64 on the device and rounded to complex64. */
int sk_hip_potential_from_values(const sk_grid* grid, int nslices, const double* V, const double* dz,
                                 double energy, const double tilt[2], sk_potential** out);

/* From atoms already wrapped and sliced by the caller (Potential.init, Sources/SimulatoMultislice/Potential.swift:
 * 251-265): numbers Z [natoms], positions xy [natoms x 2] in Angstrom, slice_of_atom [natoms]. The slices are built
 * on the device (MetalMultislice.prepare(_: Potential, ...), MetalEngine.swift:254-308; SliceBuilder,
 * ProjectedPotential.swift:109-202): bilinear deltas merged per pixel in float64 on the host, FFT, times the Lobato
 * f(k^2)/S(k) of each element, inverse FFT, real part. pi_mode 0 = exact pi, 1 = abTEM's float32 pi. */
int sk_hip_potential_from_atoms(const sk_grid* grid, int nslices, const double* dz, int natoms, const int* numbers,
                                const double* xy, const int* slice_of_atom, int pi_mode, double energy,
                                const double tilt[2], sk_potential** out);

/* Build-only (abTEM's Potential(projection="finite"); SimulatoKit has none, docs/parity.md:58): as
 * sk_hip_potential_from_atoms, but slice j holds the 3-D Lobato potential of every atom integrated over the slice's
 * own depth range [z_j, z_j + dz_j] only (z from the entrance surface, not periodic), so an atom near a boundary
 * feeds both neighbours. xyz [natoms x 3] in Angstrom, wrapped into the box. Built on the device in float64 per
 * frequency; with every atom deep inside its slice it equals the infinite projection. */
int sk_hip_potential_from_atoms_finite(const sk_grid* grid, int nslices, const double* dz, int natoms,
                                       const int* numbers, const double* xyz, int pi_mode, double energy,
                                       const double tilt[2], sk_potential** out);

void sk_hip_potential_release(sk_potential* p);
int sk_hip_potential_slice_count(const sk_potential* p);
/* The grid a potential was prepared on (ProjectedPotential.grid, ProjectedPotential.swift:13-40). */
int sk_hip_potential_grid(const sk_potential* p, sk_grid* grid);

/* Test hooks (Multislice.gpuPotentialSlices, MetalMultislice.swift:104-107): V of every slice as built on the
 * device [nslices x nx x ny] (atoms form only), and the transmission functions as stored (complex64 widened). */
int sk_hip_potential_values(const sk_potential* p, double* V);
int sk_hip_potential_transmissions(const sk_potential* p, double* t_re, double* t_im);

/* --- multislice ------------------------------------------------------------------------------------------------ */

/* One wave through the potential (MetalMultislice.run, MetalEngine.swift:458-473; Multislice.run,
 * Multislice.swift:52-80). planes: ascending zero-based slice indices, -1 = the incident wave (ExitPlanes,
 * Multislice.swift:13-40). out_re/out_im: nplanes x nx x ny. */
int sk_hip_run(const sk_potential* p, const double* in_re, const double* in_im, const int* planes, int nplanes,
               double* out_re, double* out_im, const sk_options* opt);

/* Exit waves after the last slice of `count` incident waves (MetalMultislice.runBatch, MetalEngine.swift:476-500;
 * Multislice.runBatch, MetalMultislice.swift:52-83). in/out: count x nx x ny, input order kept. */
int sk_hip_run_batch(const sk_potential* p, int count, const double* in_re, const double* in_im, double* out_re,
                     double* out_im, const sk_options* opt);

/* --- HRTEM imaging -------------------------------------------------------------------------------------------- */

/* An objective lens (ContrastTransferFunction, ContrastTransferFunction.swift:36-71). */
typedef struct sk_ctf {
  double aberrations[25];  /* sk_probe order */
  double semiangle_mrad;   /* objective aperture; +inf (or <= 0 is refused) for none */
  int soft;                /* soft aperture edge */
  double focal_spread;     /* Angstrom, 1/e half-width Delta of the temporal envelope; 0 for none */
  double angular_spread;   /* mrad, beta of the spatial envelope; 0 for none */
} sk_ctf;

/* The CTF kernel H on a grid (ContrastTransferFunction.kernel, ContrastTransferFunction.swift:84-111), evaluated in
 * float64 on the device: out_re / out_im nx x ny, FFT order. */
int sk_hip_ctf_kernel(const sk_grid* grid, double energy, const sk_ctf* ctf, double* out_re, double* out_im);

/* `count` waves with the CTF applied (Waves.applying, ContrastTransferFunction.swift:140-153): IFFT(FFT(psi) H);
 * the intensity of the result is the HRTEM image. in/out: count x nx x ny split float64 (may alias). */
int sk_hip_apply_ctf(const sk_grid* grid, double energy, const sk_ctf* ctf, int count, const double* in_re,
                     const double* in_im, double* out_re, double* out_im);

/* --- probes ------------------------------------------------------------------------------------------------- */

/* The probes at `npos` positions [npos x 2] Angstrom as the scan builds them (Probe.build(at:),
 * Probe.swift:72-109; Multislice.gpuProbes, MetalMultislice.swift:110-113): out npos x nx x ny, real space,
 * sum_k |Psi(k)|^2 = 1. */
int sk_hip_probes(const sk_grid* grid, const sk_probe* probe, int npos, const double* positions, double* out_re,
                  double* out_im);

/* --- STEM scan ----------------------------------------------------------------------------------------------- */

/* STEM.scan through `npots` potentials (one per frozen-phonon configuration; slice counts may differ)
 * (STEM.swift:49-81, 120-150; Multislice.metalScanIntensities, MetalEngine.swift:583-676): every position's probe
 * through every potential, each detector reduced on the device from |FFT(psi_exit)|^2 (float64 accumulation),
 * summed in potential order and divided by npots. outputs[d] receives npos x count_d values (count: 1 annular,
 * 2 COM (x, y), mx*my pixelated), float64 or float32 per sk_detector.out_f32; outputs_on_device = 1 when the
 * output pointers are device memory (the results then never leave HBM). */
int sk_hip_scan(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos, const double* positions,
                const sk_detector* dets, int ndets, void* const* outputs, int outputs_on_device,
                const sk_options* opt);

/* Build-only, with no counterpart in the reference: sk_hip_scan for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The same scan as sk_hip_scan
 * with annular detectors only, run on one stream without graph replay, lanes or tiles; it calls no hipMalloc,
 * hipFree, hipHostMalloc or allocating library routine. Its device memory is one static 64 MiB array per device,
 * held until the process ends: the prepared potentials, the positions and the outputs stay the caller's, and the
 * array holds the probe spectrum, the positions, the detector sums (npos x ndets float64, plus npos x ndets more for
 * host outputs) and the waves of as many positions per launch as the rest holds (at 512 x 512 about 40). A scan whose
 * fixed part leaves no room for one position, or a grid length above 4096, is refused with SK_ERR_MEMORY and a
 * message naming the limit. Every call on a device shares that device's array, which keeps the last call's contents
 * (the call clears what must start at zero): calls that overlap on one device race. The potentials must have been
 * prepared by this library (which builds the transform tables this entry point reads). The sums, per position, equal
 * sk_hip_scan's with the same positions per launch. */
int sk_hip_scan_noalloc(const sk_potential* const* pots, int npots, const sk_probe* probe, int npos,
                        const double* positions, const sk_detector* dets, int ndets, void* const* outputs,
                        int outputs_on_device, const sk_options* opt);

/* The complex exit wave of every position's probe through one potential (WavesDetector, STEM.scan(probe:potentials:
 * scan:detector:), WavesDetector.swift:107-215; abTEM's WavesDetector(gpts)): rows x cols > 0 resamples each wave in
 * Fourier space to that many points over the same extent (Waves.resampled, WavesResample.swift:16-49); 0 keeps the
 * simulation grid. out_re / out_im: npos x rows x cols, position order kept. */
int sk_hip_scan_waves(const sk_potential* p, const sk_probe* probe, int npos, const double* positions, int rows,
                      int cols, double* out_re, double* out_im, const sk_options* opt);

/* A caller's detector (any conformer of the open `Detector` protocol, Detectors.swift:84-88) receives the diffraction
 * intensities |FFT2(psi_exit)|^2 of `count` consecutive positions starting at `first`: count x nx x ny float64 values,
 * each position's in FFT order [ix * ny + iy] (DiffractionIntensity(grid:beam:values:), Detectors.swift:18-36). The
 * array is valid during the call only. Return 0 to go on; any other value stops the scan, which then returns
 * SK_ERR_INVALID. The callback runs on the calling thread and must not call into this library. */
typedef int (*sk_deliver_fn)(void* user, int first, int count, const double* intensities);

/* The diffraction intensity of every position's probe through one potential, delivered batch by batch in position
 * order (Multislice.metalScanIntensities -> MetalMultislice.scanIntensities, MetalEngine.swift:583-676). The STEM
 * driver calls it once per potential and reduces the intensities with the caller's detectors on the host, summed in
 * potential order (STEM.metalScan / accumulate / average, STEM.swift:120-150); the built-in detector kinds are reduced
 * on the device by sk_hip_scan instead. The next batch computes while `deliver` reads the current one. */
int sk_hip_scan_intensities(const sk_potential* p, const sk_probe* probe, int npos, const double* positions,
                            sk_deliver_fn deliver, void* user, const sk_options* opt);

/* The values each detector returns per position (Detector.shape(for:beam:), Detectors.swift:84-88): writes the
 * shape (up to 2 ints) and returns the count, or a negative status for a refused detector. */
int sk_hip_detector_shape(const sk_grid* grid, double energy, const sk_detector* det, int* shape);

/* --- measurement --------------------------------------------------------------------------------------------- */

/* Device time of the dominant kernels of the last sk_hip_scan on the calling thread, measured with HIP events on
 * the library's stream: mean milliseconds per launch of the row pass and of the column pass, their launch counts,
 * the bytes of the stored compact waves each launch reads and writes once, and the positions per launch. */
typedef struct sk_timing {
  double row_ms, col_ms;
  long long row_launches, col_launches;
  double row_bytes, col_bytes;
  double total_ms;
  long long batch;   /* positions per kernel launch (the tile) */
  long long group;   /* positions per launch (the batch: a host chunk, or one graph replay) */
} sk_timing;
int sk_hip_last_timing(sk_timing* t);
void sk_hip_set_timing(int enabled);

/* --- autotuning ---------------------------------------------------------------------------------------------- */

/* Build-only (no counterpart in the reference; it sets the knobs of ExecutionOptions, ExecutionOptions.swift:21-49):
 * times sk_hip_scan over `npos` positions spread over the potential's extent for each candidate execution plan
 * (positions per launch, column tile width at 512, hipGraph replay) on this device and returns the fastest. The
 * candidates change rounding only, never which numbers are computed. npos <= 0 = 4096. */
#define SK_TUNE_MAX 32
typedef struct sk_tune_report {
  sk_options best;
  double best_rate;                   /* position-slices per second of `best` */
  int ncandidates;
  sk_options candidates[SK_TUNE_MAX];
  double rates[SK_TUNE_MAX];          /* position-slices per second of each candidate */
  char device[64];                    /* the device's name and compute-unit count, e.g. "gfx950 256 CU" */
} sk_tune_report;
int sk_hip_autotune(const sk_potential* const* pots, int npots, const sk_probe* probe, const sk_detector* dets,
                    int ndets, int npos, sk_tune_report* report);
/* Build-only: the device string sk_tune_report.device holds (the key under which a caller may keep tuned plans). */
int sk_hip_device_name(char* name, int len);

#ifdef __cplusplus
}
#endif
#endif /* SIMULATO_HIP_H */
