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
from .core import ElectronBeam, ExecutionOptions, Grid, Normalization, SimulatoError, Waves
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


Detector = Union[AnnularDetector, PixelatedDetector, CenterOfMassDetector, SegmentedDetector]


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
             detectors: Sequence[Detector], options: Optional[ExecutionOptions] = None,
             device_outputs: Optional[Sequence[int]] = None) -> List[ScanMeasurement]:
        """Every position's probe through every potential, each detector averaged over the potentials
        (STEM.scan, STEM.swift:49-81, 120-150). `scan` is a GridScan or an [npos, 2] array of positions (a line).
        `device_outputs`: optional device pointers (one per detector) that receive the values instead of host
        arrays; the returned measurements then hold None."""
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
           "STEM", "WavesDetector", "ScanWaves", "last_timing", "set_timing", "detector_shape"]
