"""Detectors and the STEM driver of the `.hip` backend.

Mirrors Sources/SimulatoDetector/{Detectors,STEM}.swift. The detectors are reduced inside the final column pass on
the device (float64 accumulation), so no diffraction pattern leaves HBM unless a pixelated detector asks for it.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native as N
from .core import ElectronBeam, ExecutionOptions, Grid, Normalization, SimulatoError, Waves, fftfreq
from .multislice import AnyPotential, PreparedPotential, _options, _prepared, _raise, _sk_grid
from .probe import GridScan, LineScan, Probe

MAX_ANGLE = {"cutoff": 0, "valid": 1, "full": 2}


@dataclass(frozen=True)
class AnnularDetector:
    """Sum of P over inner^2 <= alpha^2 < outer^2 [mrad]; outer None = floor of the cutoff angle
    (Detectors.swift:91-136)."""

    inner: float = 0.0
    outer: Optional[float] = None

    def __post_init__(self):
        if not (math.isfinite(self.inner) and self.inner >= 0):
            raise SimulatoError("invalidDetector", "invalid detector: inner must be finite and >= 0")
        if self.outer is not None and not (math.isfinite(self.outer) and self.outer >= self.inner):
            raise SimulatoError("invalidDetector", f"invalid detector: outer must be finite and >= inner, got {self.outer}")

    def descriptor(self) -> N.SkDetector:
        return N.SkDetector(0, self.inner, -1.0 if self.outer is None else self.outer, 0, 0, 0, 0, 0)


@dataclass(frozen=True)
class PixelatedDetector:
    """The pattern cropped and fftshifted (Detectors.swift:140-167): max_angle "cutoff", "valid" or "full"; or a
    numeric crop (mx, my) of abTEM's fft_crop, which the reference does not offer; dtype "float64" or "float32"."""

    max_angle: str = "valid"
    crop: Optional[Tuple[int, int]] = None
    dtype: str = "float64"

    def descriptor(self) -> N.SkDetector:
        cx, cy = self.crop if self.crop else (0, 0)
        return N.SkDetector(1, 0.0, 0.0, MAX_ANGLE[self.max_angle], 0, int(cx), int(cy),
                            1 if self.dtype == "float32" else 0)


@dataclass(frozen=True)
class CenterOfMassDetector:
    """(sum P alpha_x, sum P alpha_y) over the pixelated region, not divided by sum P; units "mrad" or
    "inverse_angstroms" (Detectors.swift:172-206)."""

    max_angle: str = "valid"
    units: str = "inverse_angstroms"

    def descriptor(self) -> N.SkDetector:
        return N.SkDetector(2, 0.0, 0.0, MAX_ANGLE[self.max_angle], 0 if self.units == "mrad" else 1, 0, 0, 0)


@dataclass(frozen=True)
class SegmentedDetector:
    """radial_bins x azimuthal_bins bins between inner and outer [mrad], the azimuth measured from `rotation` [rad],
    summed in float64 (abTEM's SegmentedDetector / polar binning; Detectors.swift:207-263). Values per position
    [radial][azimuthal]."""

    radial_bins: int
    azimuthal_bins: int
    inner: float
    outer: float
    rotation: float = 0.0

    def __post_init__(self):
        if self.radial_bins < 1 or self.azimuthal_bins < 1:
            raise SimulatoError("invalidDetector", "bins must be at least 1")
        if not (math.isfinite(self.inner) and math.isfinite(self.outer) and self.inner >= 0
                and self.outer > self.inner and math.isfinite(self.rotation)):
            raise SimulatoError("invalidDetector", "need 0 <= inner < outer, finite, and a finite rotation")

    def descriptor(self) -> N.SkDetector:
        d = N.SkDetector(3, float(self.inner), float(self.outer), 0, 0, 0, 0, 0)
        d.bins_radial, d.bins_azimuthal, d.rotation = int(self.radial_bins), int(self.azimuthal_bins), \
            float(self.rotation)
        return d


BUILTIN_DETECTORS = (AnnularDetector, PixelatedDetector, CenterOfMassDetector, SegmentedDetector)
Detector = Union[AnnularDetector, PixelatedDetector, CenterOfMassDetector, SegmentedDetector]


@dataclass(frozen=True)
class DiffractionIntensity:
    """The full, unshifted diffraction intensity |FFT2(psi)|^2 of one exit wave, in FFT order [ix * ny + iy]: what
    every detector reads (DiffractionIntensity, Detectors.swift:11-81). A caller's own detector receives these."""

    grid: Grid
    beam: ElectronBeam
    values: np.ndarray  # nx * ny float64

    def __post_init__(self):
        v = np.asarray(self.values, dtype=np.float64).reshape(-1)
        if v.size != self.grid.count:
            raise ValueError(f"DiffractionIntensity: {self.grid.nx} x {self.grid.ny} values required")
        object.__setattr__(self, "values", v)

    @property
    def pattern(self) -> np.ndarray:
        """The values as an [nx, ny] array (FFT order, not shifted)."""
        return self.values.reshape(self.grid.nx, self.grid.ny)

    @property
    def angular_sampling(self) -> Tuple[float, float]:
        """mrad per pixel per axis: (1 / extent) * lambda * 1e3 (Detectors.swift:39-41)."""
        lam = self.beam.wavelength
        return (self.grid.reciprocal_sampling_x * lam * 1e3, self.grid.reciprocal_sampling_y * lam * 1e3)

    @staticmethod
    def angles(m: int, d: float) -> np.ndarray:
        """The angles [mrad] of the pixels of an m-point axis at angular sampling d: fftfreq(m, 1 / d / m)
        (Detectors.swift:43-44)."""
        return fftfreq(m, 1 / d / float(m))

    @property
    def cutoff_angles(self) -> Tuple[float, float]:
        """floor(n'/2) * d per axis, n' the antialias cutoff points with the grid's parity (Detectors.swift:46-67)."""
        kcut = 2.0 / 3.0 / max(self.grid.sampling_x, self.grid.sampling_y)

        def size(n, extent):
            f = 10.0 ** 7
            m = int(math.floor(round(kcut * extent * f) / f))
            return m + 1 if m % 2 != n % 2 else m

        mx = size(self.grid.nx, self.grid.nx * self.grid.sampling_x)
        my = size(self.grid.ny, self.grid.ny * self.grid.sampling_y)
        ax, ay = self.angular_sampling
        return ((mx // 2) * ax, (my // 2) * ay)


def is_builtin(d) -> bool:
    return isinstance(d, BUILTIN_DETECTORS)


def _check_custom(d):
    """A caller's detector follows the reference's open Detector protocol (Detectors.swift:84-88): shape(grid, beam)
    gives the shape of one position's values, detect(intensity) the values."""
    if not (callable(getattr(d, "shape", None)) and callable(getattr(d, "detect", None))):
        raise SimulatoError("invalidDetector", f"invalid detector: {type(d).__name__} has no shape(grid, beam) and "
                            "detect(intensity)")


@dataclass
class ScanMeasurement:
    """One detector's values over a scan, row-major (scan x, scan y, the detector's shape) (STEM.swift:14-43)."""

    scan_shape: Tuple[int, ...]
    value_shape: Tuple[int, ...]
    values: np.ndarray

    def __getitem__(self, ix_iy):
        ix, iy = ix_iy
        return self.values.reshape(self.scan_shape + self.value_shape)[ix, iy]

    def image(self) -> Optional[np.ndarray]:
        if self.value_shape != () or len(self.scan_shape) != 2:
            return None
        return self.values.reshape(self.scan_shape)


def detector_shape(d: Detector, grid, energy: float) -> Tuple[int, ...]:
    desc = d.descriptor()
    shape = np.zeros(2, dtype=np.int32)
    rc = N.lib().sk_hip_detector_shape(C.byref(_sk_grid(grid)), energy, C.byref(desc), N.iptr(shape))
    if rc < 0:
        _raise(rc, "detector")
    if isinstance(d, AnnularDetector):
        return ()
    if isinstance(d, CenterOfMassDetector):
        return (2,)
    if isinstance(d, SegmentedDetector):
        return (int(d.radial_bins), int(d.azimuthal_bins))
    return (int(shape[0]), int(shape[1]))


@dataclass(frozen=True)
class WavesDetector:
    """The complex exit wave (abTEM's WavesDetector; WavesDetector.swift:33-65): on the simulation grid, or with rows
    and columns resampled in Fourier space to that many points over the same extent."""

    rows: Optional[int] = None
    columns: Optional[int] = None

    def __post_init__(self):
        if (self.rows is None) != (self.columns is None):
            raise SimulatoError("invalidDetector", "give both rows and columns, or neither")
        if self.rows is not None and (self.rows < 2 or self.columns < 2):
            raise SimulatoError("unsupportedGridLength", "the waves detector needs at least 2 points per axis")


@dataclass
class ScanWaves:
    """The exit waves of a scan, [positions..., rows, columns] complex128 (ScanWaves, WavesDetector.swift:67-105)."""

    scan_shape: Tuple[int, ...]
    grid: Grid
    beam: ElectronBeam
    values: np.ndarray

    @property
    def rows(self) -> int:
        return self.grid.nx

    @property
    def columns(self) -> int:
        return self.grid.ny

    def __getitem__(self, ix_iy) -> Waves:
        ix, iy = ix_iy
        field = self.values.reshape(self.scan_shape + (self.rows, self.columns))[ix, iy]
        return Waves(field, self.grid, self.beam, Normalization.RECIPROCAL_SPACE)  # a probe's exit wave


class STEM:
    @staticmethod
    def scan(probe: Probe, potentials: Sequence[AnyPotential], scan: Union[GridScan, np.ndarray],
             detectors: Sequence, options: Optional[ExecutionOptions] = None,
             device_outputs: Optional[Sequence[int]] = None) -> List[ScanMeasurement]:
        """Every position's probe through every potential, each detector averaged over the potentials
        (STEM.scan, STEM.swift:49-81, 120-150). `scan` is a GridScan or an [npos, 2] array of positions (a line).
        The built-in detectors are reduced on the device (sk_hip_scan); any other object with shape(grid, beam) and
        detect(intensity) methods (the reference's open Detector protocol) reads each position's
        DiffractionIntensity on the host, delivered by sk_hip_scan_intensities, as the reference's `.metal` scan does
        for every detector. `device_outputs`: optional device pointers (one per detector, built-in detectors only)
        that receive the values instead of host arrays; the returned measurements then hold None."""
        if not potentials:
            raise SimulatoError("invalidScan", "invalid scan: no potential")
        if isinstance(scan, (GridScan, LineScan)):
            positions, scan_shape = scan.positions, tuple(scan.shape)
        else:
            positions = np.asarray(scan, dtype=np.float64).reshape(-1, 2)
            scan_shape = (len(positions),)
        for p in potentials:
            if not p.grid.matches(probe.grid):
                raise SimulatoError("gridMismatch", "the waves and the potential are on different grids")
        custom = [k for k, d in enumerate(detectors) if not is_builtin(d)]
        for k in custom:
            _check_custom(detectors[k])
        if custom:
            if device_outputs is not None:
                raise SimulatoError("invalidDetector", "device outputs take the built-in detectors only")
            builtin = [k for k in range(len(detectors)) if k not in set(custom)]
            potentials = [_prepared(p, probe.beam, probe.tilt) for p in potentials]  # built once for both paths
            out: List[Optional[ScanMeasurement]] = [None] * len(detectors)
            if builtin:
                for k, m in zip(builtin, STEM.scan(probe, potentials, scan, [detectors[k] for k in builtin],
                                                   options)):
                    out[k] = m
            for k, m in zip(custom, STEM._scan_intensities(probe, potentials, positions, scan_shape,
                                                           [detectors[k] for k in custom], options)):
                out[k] = m
            return out
        prepared = [_prepared(p, probe.beam, probe.tilt) for p in potentials]
        handles = (C.c_void_p * len(prepared))(*[p.handle.value for p in prepared])
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        descs = (N.SkDetector * max(1, len(detectors)))(*[d.descriptor() for d in detectors])
        shapes = [detector_shape(d, probe.grid, probe.beam.energy) for d in detectors]
        outs, ptrs = [], []
        for d, s in zip(detectors, shapes):
            dt = np.float32 if getattr(d, "dtype", "float64") == "float32" else np.float64
            a = np.empty((len(pos),) + s, dtype=dt)
            outs.append(a)
            ptrs.append(a.ctypes.data)
        on_device = device_outputs is not None
        if on_device:
            ptrs = list(device_outputs)
        ptr_arr = (C.c_void_p * max(1, len(ptrs)))(*ptrs)
        opt = _options(options)
        _raise(N.lib().sk_hip_scan(handles, len(prepared), C.byref(probe.descriptor()), len(pos), N.dptr(pos), descs,
                                   len(detectors), ptr_arr, 1 if on_device else 0, C.byref(opt)), "STEM.scan")
        return [ScanMeasurement(tuple(scan_shape), s, None if on_device else a.reshape(-1))
                for s, a in zip(shapes, outs)]

    @staticmethod
    def _scan_intensities(probe: Probe, potentials: Sequence[AnyPotential], positions: np.ndarray,
                          scan_shape: Tuple[int, ...], detectors: Sequence, options: Optional[ExecutionOptions]
                          ) -> List[ScanMeasurement]:
        """The caller's detectors over every position and potential, from the intensities the device delivers,
        summed in potential order and divided by the potential count (STEM.metalScan / accumulate / average,
        STEM.swift:120-150)."""
        grid, beam = probe.grid, probe.beam
        shapes = [tuple(int(v) for v in d.shape(grid, beam)) for d in detectors]
        counts = [int(np.prod(s)) if s else 1 for s in shapes]
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        npos = len(pos)
        sums = [np.zeros(npos * c, dtype=np.float64) for c in counts]
        prepared = [_prepared(p, beam, probe.tilt) for p in potentials]
        n = grid.count
        failure: List[BaseException] = []

        def deliver(_user, first, count, ptr):
            try:
                block = np.ctypeslib.as_array(ptr, shape=(count * n,)).reshape(count, n)
                for j in range(count):
                    inten = DiffractionIntensity(grid, beam, block[j].copy())
                    q = first + j
                    for d, det in enumerate(detectors):
                        v = np.asarray(det.detect(inten), dtype=np.float64).reshape(-1)
                        if v.size != counts[d]:
                            raise SimulatoError("invalidDetector", "a detector returned the wrong number of values")
                        sums[d][q * counts[d]:(q + 1) * counts[d]] += v
                return 0
            except BaseException as e:  # handed back after the library has returned
                failure.append(e)
                return 1

        cb = N.DELIVER_FN(deliver)
        opt = _options(options)
        for prep in prepared:
            rc = N.lib().sk_hip_scan_intensities(prep.handle, C.byref(probe.descriptor()), npos, N.dptr(pos), cb,
                                                 None, C.byref(opt))
            if failure:
                raise failure[0]
            _raise(rc, "STEM.scan")
        inv = 1.0 / len(prepared)
        return [ScanMeasurement(tuple(scan_shape), s, v * inv) for s, v in zip(shapes, sums)]

    @staticmethod
    def scan_waves(probe: Probe, potential: AnyPotential, scan: Union[GridScan, np.ndarray],
                   detector: WavesDetector = WavesDetector(), options: Optional[ExecutionOptions] = None) -> ScanWaves:
        """The exit wave of every position's probe through one potential (STEM.scan(probe:potentials:scan:detector:),
        WavesDetector.swift:107-124), resampled when the detector has rows and columns."""
        if isinstance(potential, (list, tuple)):
            if len(potential) != 1:
                raise SimulatoError("severalPotentials", "the waves detector takes one potential; scan the "
                                    "configurations with STEM.scan_frozen_phonon_waves")
            potential = potential[0]
        if not potential.grid.matches(probe.grid):
            raise SimulatoError("gridMismatch", "the waves and the potential are on different grids")
        if isinstance(scan, (GridScan, LineScan)):
            positions, scan_shape = scan.positions, tuple(scan.shape)
        else:
            positions = np.asarray(scan, dtype=np.float64).reshape(-1, 2)
            scan_shape = (len(positions),)
        prepared = _prepared(potential, probe.beam, probe.tilt)
        rows = detector.rows or probe.grid.nx
        cols = detector.columns or probe.grid.ny
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        re = np.empty((len(pos), rows, cols))
        im = np.empty_like(re)
        opt = _options(options)
        _raise(N.lib().sk_hip_scan_waves(prepared.handle, C.byref(probe.descriptor()), len(pos), N.dptr(pos),
                                         int(detector.rows or 0), int(detector.columns or 0), N.dptr(re), N.dptr(im),
                                         C.byref(opt)), "STEM.scan_waves")
        grid = Grid(rows, cols, probe.grid.extent_x, probe.grid.extent_y)
        return ScanWaves(scan_shape, grid, probe.beam, re + 1j * im)

    @staticmethod
    def scan_frozen_phonon_waves(probe: Probe, potentials: Sequence[AnyPotential], scan,
                                 detector: WavesDetector = WavesDetector(),
                                 options: Optional[ExecutionOptions] = None) -> List[ScanWaves]:
        """One ScanWaves per configuration (STEM.scanFrozenPhonons, WavesDetector.swift:143-181): complex waves are
        not averaged over configurations."""
        if not potentials:
            raise SimulatoError("invalidScan", "invalid scan: no potential")
        return [STEM.scan_waves(probe, p, scan, detector, options) for p in potentials]


def last_timing() -> dict:
    """Device timing of the last scan (see sk_hip_last_timing)."""
    t = N.SkTiming()
    N.lib().sk_hip_last_timing(C.byref(t))
    return {f: getattr(t, f) for f, _ in N.SkTiming._fields_}


def set_timing(enabled: bool):
    N.lib().sk_hip_set_timing(1 if enabled else 0)


__all__ = ["AnnularDetector", "PixelatedDetector", "CenterOfMassDetector", "SegmentedDetector", "ScanMeasurement",
           "DiffractionIntensity", "BUILTIN_DETECTORS", "is_builtin",
           "STEM", "WavesDetector", "ScanWaves", "last_timing", "set_timing", "detector_shape"]
