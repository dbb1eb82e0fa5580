"""ctypes binding of the float64 CPU oracle (oracle/simulato_oracle.cpp). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the float64 reference
the HIP library is tested against; the product package never imports it. Arrays are NumPy, row-major ``[ix, iy]``; complex arrays are
complex128. Each wrapper names the reference function it restates (paths relative to the SimulatoKit checkout).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_wavelength.restype = C.c_double
        L.orc_wavelength.argtypes = [C.c_double]
        L.orc_sigma.restype = C.c_double
        L.orc_sigma.argtypes = [C.c_double]
        L.orc_last_scan_seconds.restype = C.c_double
        L.orc_configuration_seed.restype = C.c_ulonglong
        L.orc_configuration_seed.argtypes = [C.c_ulonglong, C.c_longlong]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())


# ---- constants, grid ---------------------------------------------------------------------------------------
def wavelength(energy):  # ElectronBeam.swift:129-134
    return lib().orc_wavelength(float(energy))


def sigma(energy):  # ElectronBeam.swift:115-121
    return lib().orc_sigma(float(energy))


def fftfreq(n, d):  # Grid.swift:82-86
    out = np.empty(n)
    lib().orc_fftfreq(C.c_int(n), C.c_double(d), _d(out))
    return out


def lobato_f(z, k2, pi_mode=0):  # LobatoParameterization.swift:45-63; pi_mode 1 = abTEM float32 pi
    k2 = np.ascontiguousarray(k2, dtype=np.float64)
    out = np.empty_like(k2)
    _check(lib().orc_lobato_f(C.c_int(z), C.c_int(pi_mode), _d(k2), C.c_longlong(k2.size), _d(out)))
    return out


# ---- FFT ---------------------------------------------------------------------------------------------------
def fft1(x, inverse=False):
    """Unscaled 1-D DFT along the last axis (sign -1 forward, +1 inverse without 1/n)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    lib().orc_fft1(C.c_int(n), C.c_longlong(a.size // n), C.c_int(1 if inverse else -1),
                   a.ctypes.data_as(_dp))
    return a


def fft2(x, inverse=False):  # FFT2D.swift:202-215 (inverse scaled by 1/(nx ny))
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    nx, ny = a.shape
    lib().orc_fft2(C.c_int(nx), C.c_int(ny), C.c_int(1 if inverse else -1), a.ctypes.data_as(_dp))
    return a


def antialias(nx, ny, ex, ey):  # Antialias.swift:18-39
    out = np.empty((nx, ny))
    lib().orc_antialias(C.c_int(nx), C.c_int(ny), C.c_double(ex), C.c_double(ey), _d(out))
    return out


# ---- slicing, potential, transmission, propagator ---------------------------------------------------------------
def slice_atoms(xyz, box, thicknesses):  # Slicing.swift:74-104
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3).copy()
    box = np.ascontiguousarray(box, dtype=np.float64)
    th = np.ascontiguousarray(thicknesses, dtype=np.float64)
    s = np.empty(len(xyz), dtype=np.int32)
    lib().orc_slice_atoms(C.c_int(len(xyz)), _d(xyz), _d(box), C.c_int(len(th)), _d(th), _i(s))
    return xyz, s


def potential(nx, ny, numbers, xyz, box, thicknesses, pi_mode=0):
    """Potential.build (Potential.swift:285-293): V [nslices, nx, ny] and the slice of every atom."""
    numbers = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    box = np.ascontiguousarray(box, dtype=np.float64)
    th = np.ascontiguousarray(thicknesses, dtype=np.float64)
    V = np.empty((len(th), nx, ny))
    s = np.empty(len(numbers), dtype=np.int32)
    _check(lib().orc_potential(C.c_int(nx), C.c_int(ny), C.c_double(box[0]), C.c_double(box[1]),
                               C.c_int(len(numbers)), _i(numbers), _d(xyz), _d(box), C.c_int(len(th)), _d(th),
                               C.c_int(pi_mode), _d(V), _i(s)))
    return V, s


def potential_finite(nx, ny, numbers, xyz, box, thicknesses, pi_mode=0):
    """The finite projection (abTEM's Potential(projection="finite"); none in the reference): V [nslices, nx, ny],
    each slice the 3-D Lobato potential integrated over its own z range only."""
    numbers = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    box = np.ascontiguousarray(box, dtype=np.float64)
    th = np.ascontiguousarray(thicknesses, dtype=np.float64)
    V = np.empty((len(th), nx, ny))
    _check(lib().orc_potential_finite(C.c_int(nx), C.c_int(ny), C.c_double(box[0]), C.c_double(box[1]),
                                      C.c_int(len(numbers)), _i(numbers), _d(xyz), _d(box), C.c_int(len(th)),
                                      _d(th), C.c_int(pi_mode), _d(V)))
    return V


def transmission(V, extent, energy):  # Multislice.swift:83-103
    V = np.ascontiguousarray(V, dtype=np.float64)
    ns, nx, ny = V.shape
    t = np.empty((ns, nx, ny), dtype=np.complex128)
    lib().orc_transmission(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]), C.c_int(ns),
                           _d(V), C.c_double(energy), t.ctypes.data_as(_dp))
    return t


def propagator(nx, ny, extent, energy, dz, tilt=(0.0, 0.0)):  # Multislice.swift:108-130
    out = np.empty((nx, ny), dtype=np.complex128)
    lib().orc_propagator(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                         C.c_double(energy), C.c_double(dz), C.c_double(tilt[0]), C.c_double(tilt[1]),
                         out.ctypes.data_as(_dp))
    return out


# ---- probe ---------------------------------------------------------------------------------------------------
ABERRATION_NAMES = ("C10", "C12", "phi12", "C21", "phi21", "C23", "phi23", "C30", "C32", "phi32", "C34", "phi34",
                    "C41", "phi41", "C43", "phi43", "C45", "phi45", "C50", "C52", "phi52", "C54", "phi54", "C56",
                    "phi56")


def aberration_vector(ab=None):
    v = np.zeros(25)
    for k, x in (ab or {}).items():
        v[ABERRATION_NAMES.index(k)] = x
    return v


def probe(nx, ny, extent, energy, cutoff_mrad, position, aberrations=None, soft=True):  # Probe.swift:72-109
    out = np.empty((nx, ny), dtype=np.complex128)
    ab = aberration_vector(aberrations)
    lib().orc_probe(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]), C.c_double(energy),
                    C.c_double(cutoff_mrad), C.c_int(1 if soft else 0), _d(ab), C.c_double(position[0]),
                    C.c_double(position[1]), out.ctypes.data_as(_dp))
    return out


def probe_spectrum(nx, ny, extent, energy, cutoff_mrad, aberrations=None, soft=True):  # Probe.swift:113-136
    out = np.empty((nx, ny), dtype=np.complex128)
    ab = aberration_vector(aberrations)
    lib().orc_probe_spectrum(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                             C.c_double(energy), C.c_double(cutoff_mrad), C.c_int(1 if soft else 0), _d(ab),
                             out.ctypes.data_as(_dp))
    return out


# ---- multislice ----------------------------------------------------------------------------------------------
def multislice(V, dz, extent, energy, incident, planes=None, tilt=(0.0, 0.0)):
    """Multislice.run (Multislice.swift:52-59, 132-157): waves at the planes (default: the last slice)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    ns, nx, ny = V.shape
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, dtype=np.float64), (ns,)))
    planes = np.ascontiguousarray([ns - 1] if planes is None else planes, dtype=np.int32)
    inc = np.ascontiguousarray(incident, dtype=np.complex128)
    out = np.empty((len(planes), nx, ny), dtype=np.complex128)
    lib().orc_multislice(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                         C.c_double(energy), C.c_double(tilt[0]), C.c_double(tilt[1]), C.c_int(ns), _d(V), _d(dz),
                         inc.ctypes.data_as(_dp), _i(planes), C.c_int(len(planes)), out.ctypes.data_as(_dp))
    return out


def multislice_batch(V, dz, extent, energy, incidents, tilt=(0.0, 0.0), threads=8):
    """Multislice.runBatch(.cpu) (MetalMultislice.swift:28-60): exit waves after the last slice."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    ns, nx, ny = V.shape
    dz = np.ascontiguousarray(np.broadcast_to(np.asarray(dz, dtype=np.float64), (ns,)))
    inc = np.ascontiguousarray(incidents, dtype=np.complex128)
    out = np.empty_like(inc)
    lib().orc_multislice_batch(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                               C.c_double(energy), C.c_double(tilt[0]), C.c_double(tilt[1]), C.c_int(ns), _d(V),
                               _d(dz), C.c_int(len(inc)), inc.ctypes.data_as(_dp), C.c_int(threads),
                               out.ctypes.data_as(_dp))
    return out


def diffraction_intensity(wave, values_norm=False):  # Detectors.swift:28-36
    w = np.ascontiguousarray(wave, dtype=np.complex128)
    nx, ny = w.shape
    out = np.empty((nx, ny))
    lib().orc_diffraction_intensity(C.c_int(nx), C.c_int(ny), w.ctypes.data_as(_dp), C.c_int(int(values_norm)),
                                    _d(out))
    return out


def diffraction_pattern(wave, extent, values_norm=False, max_angle="cutoff"):  # Waves.swift:62-86
    w = np.ascontiguousarray(wave, dtype=np.complex128)
    nx, ny = w.shape
    mode = {"cutoff": 0, "full": 2}[max_angle]
    mx, my = C.c_int(), C.c_int()
    lib().orc_diffraction_pattern(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                                  w.ctypes.data_as(_dp), C.c_int(int(values_norm)), C.c_int(mode), C.byref(mx),
                                  C.byref(my), None)
    out = np.empty((mx.value, my.value))
    lib().orc_diffraction_pattern(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                                  w.ctypes.data_as(_dp), C.c_int(int(values_norm)), C.c_int(mode), C.byref(mx),
                                  C.byref(my), _d(out))
    return out


# ---- detectors -----------------------------------------------------------------------------------------------
MAX_ANGLE = {"cutoff": 0, "valid": 1, "full": 2}


def det_annular(inner=0.0, outer=None):  # Detectors.swift:91-136
    return ("annular", float(inner), -1.0 if outer is None else float(outer), 0, 0, 1, 1, 0.0)


def det_pixelated(max_angle="valid"):  # Detectors.swift:140-167
    return ("pixelated", 0.0, 0.0, MAX_ANGLE[max_angle], 0, 1, 1, 0.0)


def det_com(max_angle="valid", units="inverse_angstroms"):  # Detectors.swift:172-206
    return ("com", 0.0, 0.0, MAX_ANGLE[max_angle], 0 if units == "mrad" else 1, 1, 1, 0.0)


def det_segmented(radial, azimuthal, inner, outer, rotation=0.0):  # Detectors.swift:207-263
    return ("segmented", float(inner), float(outer), 0, 0, int(radial), int(azimuthal), float(rotation))


_KIND = {"annular": 0, "pixelated": 1, "com": 2, "segmented": 3}


def _spec(d):
    return np.array([_KIND[d[0]], d[1], d[2], d[3], d[4], d[5], d[6], d[7]], dtype=np.float64)


def pixelated_size(nx, ny, extent, max_angle):
    mx, my = C.c_int(), C.c_int()
    lib().orc_pixelated_size(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                             C.c_int(MAX_ANGLE[max_angle]), C.byref(mx), C.byref(my))
    return mx.value, my.value


def detector_shape(nx, ny, extent, d):
    if d[0] == "annular":
        return ()
    if d[0] == "com":
        return (2,)
    if d[0] == "segmented":
        return (d[5], d[6])
    return pixelated_size(nx, ny, extent, {0: "cutoff", 1: "valid", 2: "full"}[d[3]])


def detect(P, extent, energy, d):
    P = np.ascontiguousarray(P, dtype=np.float64)
    nx, ny = P.shape
    shape = detector_shape(nx, ny, extent, d)
    out = np.empty(int(np.prod(shape)) if shape else 1)
    spec = _spec(d)
    _check(lib().orc_detect(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                            C.c_double(energy), _d(spec), _d(P), _d(out)))
    return out.reshape(shape) if shape else out[0]


def stem_scan(potentials, dz_list, extent, energy, cutoff_mrad, positions, detectors, aberrations=None,
              soft=True, tilt=(0.0, 0.0), threads=8):
    """STEM.scan(.cpu) (STEM.swift:49-56, 95-150). potentials: list of V [ns, nx, ny]; dz_list: per potential,
    the slice thicknesses. Returns one array per detector, [npos, *shape], averaged over the potentials."""
    Vs = [np.ascontiguousarray(v, dtype=np.float64) for v in potentials]
    nx, ny = Vs[0].shape[1:]
    nsl = np.array([v.shape[0] for v in Vs], dtype=np.int32)
    V = np.ascontiguousarray(np.concatenate([v.reshape(-1) for v in Vs]))
    dz = np.ascontiguousarray(np.concatenate([np.broadcast_to(np.asarray(d, dtype=np.float64), (v.shape[0],))
                                              for d, v in zip(dz_list, Vs)]))
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
    specs = np.ascontiguousarray(np.concatenate([_spec(d) for d in detectors]))
    shapes = [detector_shape(nx, ny, extent, d) for d in detectors]
    counts = [int(np.prod(s)) if s else 1 for s in shapes]
    out = np.empty(sum(counts) * len(pos))
    ab = aberration_vector(aberrations)
    _check(lib().orc_stem_scan(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]),
                               C.c_double(energy), C.c_double(cutoff_mrad), C.c_int(1 if soft else 0), _d(ab),
                               C.c_double(tilt[0]), C.c_double(tilt[1]), C.c_int(len(Vs)), _i(nsl), _d(V), _d(dz),
                               C.c_int(len(pos)), _d(pos), C.c_int(len(detectors)), _d(specs), C.c_int(threads),
                               _d(out)))
    res, off = [], 0
    for s, c in zip(shapes, counts):
        res.append(out[off:off + c * len(pos)].reshape((len(pos),) + tuple(s)))
        off += c * len(pos)
    return res


def last_scan_seconds():
    """Wall time of the last stem_scan's position loop (after the transmission functions were prepared)."""
    return lib().orc_last_scan_seconds()


# ---- frozen phonons --------------------------------------------------------------------------------------------
def gaussian_normals(seed, count):  # FrozenPhonons.swift:14-58
    out = np.empty(count)
    lib().orc_gaussian_normals(C.c_ulonglong(seed), C.c_longlong(count), _d(out))
    return out


def xoshiro_stream(seed, count):
    out = np.empty(count, dtype=np.uint64)
    lib().orc_xoshiro_stream(C.c_ulonglong(seed), C.c_longlong(count), out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return out


def xoshiro_from_state(state, count):
    st = np.ascontiguousarray(state, dtype=np.uint64)
    out = np.empty(count, dtype=np.uint64)
    lib().orc_xoshiro_from_state(st.ctypes.data_as(C.POINTER(C.c_ulonglong)), C.c_longlong(count),
                                 out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return out


def splitmix_stream(seed, count):
    out = np.empty(count, dtype=np.uint64)
    lib().orc_splitmix_stream(C.c_ulonglong(seed), C.c_longlong(count),
                              out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return out


def configuration_seed(seed, c):  # FrozenPhonons.swift:122-130
    return lib().orc_configuration_seed(C.c_ulonglong(seed), C.c_longlong(c))


def frozen_displacements(n_atoms_numbers, sigmas, configurations, seed, directions=(True, True, True)):
    """FrozenPhonons(structure:sigmas:configurations:seed:) (FrozenPhonons.swift:95-115): [configs, atoms, 3]."""
    numbers = np.asarray(n_atoms_numbers)
    out = np.zeros((configurations, len(numbers), 3))
    for c in range(configurations):
        g = gaussian_normals(configuration_seed(seed, c), 3 * len(numbers)).reshape(-1, 3)
        s = np.array([sigmas[int(z)] for z in numbers])
        for ax in range(3):
            if directions[ax]:
                out[c, :, ax] = s * g[:, ax]
    return out


# ---- exit waves over a scan (WavesDetector.swift:33-217, WavesResample.swift:16-66) ----------------------------
def crop_indices(n, m):  # Waves.swift:90-96
    if m >= n:
        return list(range(n))
    low = m // 2 + 1 if m % 2 == 1 else m // 2
    return list(range(low)) + list(range(n - m // 2, n))


def resample_map(n, m):  # WavesResample.swift:58-66: (source, destination) of every surviving frequency
    if m <= n:
        return [(0, 0)] if m == 1 else [(s, d) for d, s in enumerate(crop_indices(n, m))]
    if n == 1:
        return [(0, 0)]
    positive = n // 2 if n % 2 == 0 else n // 2 + 1
    return [(i, i if i < positive else m - n + i) for i in range(n)]


def resample_wave(w, rows, cols):  # Waves.resampled (WavesResample.swift:16-49)
    nx, ny = w.shape
    F = fft2(np.ascontiguousarray(w, dtype=complex))
    out = np.zeros((rows, cols), dtype=complex)
    mx, my = resample_map(nx, rows), resample_map(ny, cols)
    for sx, dx in mx:
        for sy, dy in my:
            out[dx, dy] = F[sx, sy]
    return fft2(out, inverse=True) * (rows * cols / (nx * ny))


def scan_waves(V, dz, extent, energy, cutoff_mrad, positions, gpts=None):
    """STEM.scan(..., detector: WavesDetector) on the CPU path (WavesDetector.swift:107-215): the exit wave of every
    position's probe, resampled to gpts = (rows, cols) when given; [npos, rows, cols] complex."""
    nx, ny = V.shape[1:]
    probes = [probe(nx, ny, extent, energy, cutoff_mrad, tuple(p)) for p in positions]
    exits = multislice_batch(V, dz, extent, energy, probes)
    if gpts is None:
        return np.stack(exits)
    return np.stack([resample_wave(e, gpts[0], gpts[1]) for e in exits])


# ---- contrast transfer function (ContrastTransferFunction.swift:36-153) ----------------------------------------
def ctf_kernel(nx, ny, extent, energy, aberrations=None, cutoff_mrad=np.inf, soft=True, focal_spread=0.0,
               angular_spread=0.0):
    out = np.empty((nx, ny), dtype=complex)
    lib().orc_ctf_kernel(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]), C.c_double(energy),
                         _d(aberration_vector(aberrations)), C.c_double(cutoff_mrad), C.c_int(1 if soft else 0),
                         C.c_double(focal_spread), C.c_double(angular_spread), out.ctypes.data_as(_dp))
    return out


def apply_ctf(wave, extent, energy, aberrations=None, cutoff_mrad=np.inf, soft=True, focal_spread=0.0,
              angular_spread=0.0):
    w = np.array(wave, dtype=complex, copy=True, order="C")
    nx, ny = w.shape
    lib().orc_apply_ctf(C.c_int(nx), C.c_int(ny), C.c_double(extent[0]), C.c_double(extent[1]), C.c_double(energy),
                        _d(aberration_vector(aberrations)), C.c_double(cutoff_mrad), C.c_int(1 if soft else 0),
                        C.c_double(focal_spread), C.c_double(angular_spread), w.ctypes.data_as(_dp))
    return w
