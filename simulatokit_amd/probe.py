"""Probe, aperture, aberrations and grid scans of the `.hip` backend.

Mirrors Sources/SimulatoProbe/{Probe,Aberrations,GridScan}.swift. Probes are built on the device
(include/simulato_hip.h: sk_hip_probes), as the reference's Metal path builds them (MetalEngine.swift:539-553).
"""
from __future__ import annotations

import ctypes as C
import copy
import math
from dataclasses import dataclass, fields
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .core import ElectronBeam, Grid, Normalization, SimulatoError, Waves, as_float_pair

ABERRATION_NAMES = ("C10", "C12", "phi12", "C21", "phi21", "C23", "phi23", "C30", "C32", "phi32", "C34", "phi34",
                    "C41", "phi41", "C43", "phi43", "C45", "phi45", "C50", "C52", "phi52", "C54", "phi54", "C56",
                    "phi56")


@dataclass
class Aberrations:
    """Polar aberration coefficients, Angstrom and rad; defocus = -C10 (Aberrations.swift:10-60)."""

    C10: float = 0.0
    C12: float = 0.0
    phi12: float = 0.0
    C21: float = 0.0
    phi21: float = 0.0
    C23: float = 0.0
    phi23: float = 0.0
    C30: float = 0.0
    C32: float = 0.0
    phi32: float = 0.0
    C34: float = 0.0
    phi34: float = 0.0
    C41: float = 0.0
    phi41: float = 0.0
    C43: float = 0.0
    phi43: float = 0.0
    C45: float = 0.0
    phi45: float = 0.0
    C50: float = 0.0
    C52: float = 0.0
    phi52: float = 0.0
    C54: float = 0.0
    phi54: float = 0.0
    C56: float = 0.0
    phi56: float = 0.0

    @property
    def defocus(self) -> float:
        return -self.C10

    @defocus.setter
    def defocus(self, v: float):
        self.C10 = -v

    @property
    def Cs(self) -> float:
        return self.C30

    @Cs.setter
    def Cs(self, v: float):
        self.C30 = v

    def vector(self) -> List[float]:
        return [float(getattr(self, n)) for n in ABERRATION_NAMES]

    @property
    def is_finite(self) -> bool:
        return all(math.isfinite(getattr(self, f.name)) for f in fields(self))


@dataclass(frozen=True)
class Aperture:
    """Semiangle cutoff [mrad] (infinity: none) with abTEM's one-pixel soft edge (Probe.swift:12-49)."""

    semiangle_cutoff: float
    soft: bool = True

    def __post_init__(self):
        if not (self.semiangle_cutoff > 0) or math.isnan(self.semiangle_cutoff):
            raise SimulatoError("invalidAperture", f"semiangle cutoff must be finite and positive (or infinite), got {self.semiangle_cutoff} mrad")


class Probe:
    """A convergent probe (Probe.swift:52-157)."""

    def __init__(self, beam: ElectronBeam, grid: Grid, aperture: Aperture, aberrations: Aberrations = None,
                 tilt=(0.0, 0.0)):
        aberrations = aberrations or Aberrations()
        if not aberrations.is_finite:
            raise SimulatoError("invalidAberration", "invalid aberration: every coefficient must be finite")
        tilt = as_float_pair(tilt)
        if not (math.isfinite(tilt[0]) and math.isfinite(tilt[1])):
            raise SimulatoError("invalidTilt", "invalid tilt: the tilt must be finite")
        self.beam, self.grid, self.aperture, self.aberrations, self.tilt = beam, grid, aperture, aberrations, tilt

    def descriptor(self) -> N.SkProbe:
        p = N.SkProbe()
        p.energy = self.beam.energy
        p.semiangle_mrad = self.aperture.semiangle_cutoff
        p.soft = 1 if self.aperture.soft else 0
        for i, v in enumerate(self.aberrations.vector()):
            p.aberrations[i] = v
        p.tilt[0], p.tilt[1] = self.tilt
        return p

    def build_many(self, positions: Sequence[Sequence[float]]) -> np.ndarray:
        """Real-space probes at the positions [npos, nx, ny] (complex128), sum_k |Psi(k)|^2 = 1."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 2))
        if not np.isfinite(pos).all():
            raise SimulatoError("invalidPosition", "the probe position must be finite")
        g = self.grid
        re = np.empty((len(pos), g.nx, g.ny))
        im = np.empty_like(re)
        from .multislice import _raise, _sk_grid
        _raise(N.lib().sk_hip_probes(C.byref(_sk_grid(g)), C.byref(self.descriptor()), len(pos), N.dptr(pos),
                                     N.dptr(re), N.dptr(im)), "Probe.build")
        return re + 1j * im

    def build(self, at: Sequence[float]) -> Waves:
        """The probe centred at `at` (x, y) [Angstrom] (Probe.build(at:), Probe.swift:72-109)."""
        return Waves(self.build_many([at])[0], self.grid, self.beam, Normalization.RECIPROCAL_SPACE, self.tilt)


class GridScan:
    """A rectangular grid of probe positions (GridScan.swift:10-67), index ix * ny + iy."""

    def __init__(self, start: Sequence[float], end: Sequence[float], nx: int, ny: int, endpoint: bool = False):
        if nx < 1 or ny < 1:
            raise SimulatoError("invalidScan", f"invalid scan: gpts must be at least 1, got {nx} x {ny}")
        start, end = as_float_pair(start), as_float_pair(end)
        if not all(math.isfinite(v) for v in start + end):
            raise SimulatoError("invalidScan", "invalid scan: start and end must be finite")
        if not (end[0] - start[0] > 0 or end[1] - start[1] > 0):
            raise SimulatoError("invalidScan", "invalid scan: scan extent must be positive on at least one axis (abTEM's rule)")
        self.start, self.end, self.nx, self.ny, self.endpoint = start, end, int(nx), int(ny), endpoint

    @staticmethod
    def with_sampling(start, end, sampling_x: float, sampling_y: float, endpoint: bool = False) -> "GridScan":
        """ceil(extent / sampling) points per axis (GridScan.swift:251-265)."""
        if not (math.isfinite(sampling_x) and math.isfinite(sampling_y) and sampling_x > 0 and sampling_y > 0):
            raise SimulatoError("invalidScan", f"invalid scan: sampling must be finite and positive, got ({sampling_x}, {sampling_y}) A")
        start, end = as_float_pair(start), as_float_pair(end)
        ex, ey = end[0] - start[0], end[1] - start[1]
        if ex < 0 or ey < 0:
            raise SimulatoError("invalidScan", "invalid scan: end must not lie before start")
        px = math.ceil(ex / sampling_x) if ex > 0 else 0
        py = math.ceil(ey / sampling_y) if ey > 0 else 0
        extra = 1 if endpoint else 0
        return GridScan(start, end, max(px + (extra if ex > 0 else 1), 1), max(py + (extra if ey > 0 else 1), 1),
                        endpoint)

    @staticmethod
    def linspace(a: float, b: float, n: int, endpoint: bool) -> np.ndarray:
        """NumPy's linspace: i * step + start (GridScan.swift:269-276)."""
        div = n - 1 if endpoint else n
        if div <= 0:
            return np.array([a])
        step = (b - a) / float(div)
        out = np.arange(n, dtype=np.float64) * step + a
        if endpoint and n > 1:
            out[-1] = b
        return out

    @property
    def shape(self) -> Tuple[int, int]:
        return (self.nx, self.ny)

    @property
    def positions(self) -> np.ndarray:
        xs = self.linspace(self.start[0], self.end[0], self.nx, self.endpoint)
        ys = self.linspace(self.start[1], self.end[1], self.ny, self.endpoint)
        return np.stack(np.meshgrid(xs, ys, indexing="ij"), axis=-1).reshape(-1, 2)


def _check(rc: int, what: str):
    from .multislice import _raise
    _raise(rc, what)


def _grid_desc(g: Grid):
    from .multislice import _sk_grid
    return _sk_grid(g)


class ContrastTransferFunction:
    """The objective lens of HRTEM imaging (abTEM's CTF; ContrastTransferFunction.swift:36-135): the aberration phase
    e^{-i chi}, the spatial (angular spread beta, mrad) and temporal (focal spread Delta, Angstrom) envelopes of the
    quasi-coherent approximation, and the objective aperture (semiangle mrad; inf for none)."""

    def __init__(self, aberrations: Optional[Aberrations] = None, semiangle_cutoff: float = math.inf,
                 soft_aperture: bool = True, focal_spread: float = 0.0, angular_spread: float = 0.0):
        self.aberrations = aberrations or Aberrations()
        if not all(math.isfinite(v) for v in self.aberrations.vector()):
            raise SimulatoError("invalidAberration", "every coefficient must be finite")
        if not (semiangle_cutoff > 0):
            raise SimulatoError("invalidAperture", f"invalid aperture: {semiangle_cutoff}")
        if not (math.isfinite(focal_spread) and focal_spread >= 0):
            raise SimulatoError("invalidAberration", f"focal spread must be finite and >= 0 A, got {focal_spread}")
        if not (math.isfinite(angular_spread) and angular_spread >= 0):
            raise SimulatoError("invalidAberration", f"angular spread must be finite and >= 0 mrad, got {angular_spread}")
        self.semiangle_cutoff, self.soft_aperture = float(semiangle_cutoff), bool(soft_aperture)
        self.focal_spread, self.angular_spread = float(focal_spread), float(angular_spread)

    def descriptor(self) -> N.SkCtf:
        d = N.SkCtf()
        for k, v in enumerate(self.aberrations.vector()):
            d.aberrations[k] = v
        d.semiangle_mrad, d.soft = self.semiangle_cutoff, 1 if self.soft_aperture else 0
        d.focal_spread, d.angular_spread = self.focal_spread, self.angular_spread
        return d

    def kernel(self, grid: Grid, energy: float) -> np.ndarray:
        """H on the grid in FFT order, evaluated in float64 on the device (ContrastTransferFunction.swift:84-111)."""
        re = np.empty((grid.nx, grid.ny))
        im = np.empty_like(re)
        _check(N.lib().sk_hip_ctf_kernel(C.byref(_grid_desc(grid)), energy, C.byref(self.descriptor()), N.dptr(re),
                                         N.dptr(im)), "ctf kernel")
        return re + 1j * im

    @staticmethod
    def scherzer_defocus(Cs: float, energy: float, factor: float = 1.5) -> float:
        """sign(Cs) sqrt(factor |Cs| lambda) [Angstrom] (ContrastTransferFunction.swift:113-120)."""
        if not (math.isfinite(factor) and factor > 0):
            raise SimulatoError("invalidAberration", "the Scherzer factor must be finite and positive")
        lam = ElectronBeam(energy).wavelength
        s = 1.0 if Cs > 0 else (-1.0 if Cs < 0 else 0.0)
        return s * math.sqrt(factor * abs(Cs) * lam)

    @staticmethod
    def point_resolution(Cs: float, energy: float, divisor: float = 6.0) -> float:
        """(lambda^3 |Cs| / divisor)^(1/4) [Angstrom] (ContrastTransferFunction.swift:122-127)."""
        if not (math.isfinite(divisor) and divisor > 0):
            raise SimulatoError("invalidAberration", "the divisor must be finite and positive")
        lam = ElectronBeam(energy).wavelength
        return math.pow(math.pow(lam, 3) * abs(Cs) / divisor, 1.0 / 4)

    @staticmethod
    def nyquist_sampling(semiangle_cutoff: float, energy: float, oversampling: float = 4.0) -> float:
        """1 / (oversampling alpha_c / lambda 1e-3) [Angstrom] (ContrastTransferFunction.swift:129-135)."""
        if not (math.isfinite(semiangle_cutoff) and semiangle_cutoff > 0):
            raise SimulatoError("invalidAperture", f"invalid aperture: {semiangle_cutoff}")
        if not (math.isfinite(oversampling) and oversampling > 0):
            raise SimulatoError("invalidAperture", f"invalid oversampling: {oversampling}")
        lam = ElectronBeam(energy).wavelength
        return 1 / (oversampling * semiangle_cutoff / lam * 1e-3)


def apply_ctf(waves: Waves, ctf: ContrastTransferFunction) -> Waves:
    """The waves with the CTF applied, IFFT(FFT(psi) H) on the device (Waves.applying,
    ContrastTransferFunction.swift:140-153); the intensity of the result is the HRTEM image."""
    re = np.ascontiguousarray(waves.field.real)
    im = np.ascontiguousarray(waves.field.imag)
    ore, oim = np.empty_like(re), np.empty_like(im)
    _check(N.lib().sk_hip_apply_ctf(C.byref(_grid_desc(waves.grid)), waves.beam.energy, C.byref(ctf.descriptor()), 1,
                                    N.dptr(re), N.dptr(im), N.dptr(ore), N.dptr(oim)), "apply_ctf")
    return Waves(ore + 1j * oim, waves.grid, waves.beam, waves.normalization, waves.tilt)


class LineScan:
    """Probe positions along a straight line (abTEM's LineScan; LineScan.swift:26-198): `gpts` points from `start` to
    `end` [Angstrom], the last at `end` with `endpoint`. A given `sampling` wins over `gpts`: the count becomes
    ceil(length / sampling) and the realised sampling length / (count - 1) with `endpoint` and more than one point,
    else length / count (abTEM's _adjust_gpts / _adjust_sampling)."""

    def __init__(self, start=(0.0, 0.0), end=None, gpts: Optional[int] = None, sampling: Optional[float] = None,
                 endpoint: bool = True, fractional: bool = False, extent=None):
        if end is None:
            raise SimulatoError("invalidScan", "invalid scan: a line scan needs an end point")
        s, e = as_float_pair(start), as_float_pair(end)
        if fractional:
            if extent is None:
                raise SimulatoError("invalidScan", "provide the potential's extent for fractional coordinates")
            ex = as_float_pair(extent)
            s, e = (ex[0] * s[0], ex[1] * s[1]), (ex[0] * e[0], ex[1] * e[1])
        if not all(math.isfinite(v) for v in s + e):
            raise SimulatoError("invalidScan", "start and end must be finite")
        if gpts is not None and not (1 <= gpts < 1_000_000_000):
            raise SimulatoError("invalidScan", f"gpts must be at least 1 and below 1e9, got {gpts}")
        if sampling is not None and not (math.isfinite(sampling) and sampling > 0):
            raise SimulatoError("invalidScan", f"sampling must be finite and positive, got {sampling} A")
        if gpts is None and sampling is None:
            raise SimulatoError("invalidScan", "provide gpts or sampling for a line scan")
        self.start, self.end, self.endpoint = s, e, bool(endpoint)
        self.gpts = int(gpts) if gpts is not None else 0
        self.sampling = float(sampling) if sampling is not None else 0.0
        if sampling is not None:
            self._adjust_gpts()
        self._adjust_sampling()

    @property
    def extent(self) -> float:
        dx, dy = self.end[0] - self.start[0], self.end[1] - self.start[1]
        return math.sqrt(dx * dx + dy * dy)

    @property
    def direction(self) -> Tuple[float, float]:
        L = self.extent
        return ((self.end[0] - self.start[0]) / L, (self.end[1] - self.start[1]) / L) if L > 0 else (math.nan, math.nan)

    @property
    def angle(self) -> float:
        d = self.direction
        return math.atan2(d[1], d[0])

    @property
    def shape(self) -> Tuple[int]:
        return (self.gpts,)

    def _adjust_gpts(self):
        n = math.ceil(self.extent / self.sampling)
        if not (1 <= n < 1e9):
            raise SimulatoError("invalidScan", f"the line of {self.extent} A at sampling {self.sampling} A has {n} positions")
        self.gpts = int(n)
        self._adjust_sampling()

    def _adjust_sampling(self):
        self.sampling = self.extent / (self.gpts - 1) if self.endpoint and self.gpts > 1 else self.extent / self.gpts

    def adding_margin(self, before: float, after: Optional[float] = None) -> "LineScan":
        """The line with `before` added at the start and `after` at the end along its direction (abTEM add_margin);
        the count is recomputed from the sampling after each end moves, as abTEM's setters do."""
        after = before if after is None else after
        if not (math.isfinite(before) and math.isfinite(after)):
            raise SimulatoError("invalidScan", "the margin must be finite")
        if not self.extent > 0:
            raise SimulatoError("invalidScan", "a line of zero length has no direction for a margin")
        line = copy.copy(self)
        d = self.direction
        line.start = (self.start[0] - d[0] * before, self.start[1] - d[1] * before)
        line._adjust_gpts()
        line.end = (self.end[0] + d[0] * after, self.end[1] + d[1] * after)
        line._adjust_gpts()
        return line

    @staticmethod
    def at_position(center, extent: float = 1.0, angle: float = 0.0, gpts: Optional[int] = None,
                    sampling: Optional[float] = None, endpoint: bool = True) -> "LineScan":
        """A line of `extent` centred on `center` at `angle` degrees to the first axis (abTEM at_position)."""
        a = angle * (math.pi / 180)
        c = as_float_pair(center)
        dx, dy = math.cos(a), math.sin(a)
        return LineScan((c[0] - extent / 2 * dx, c[1] - extent / 2 * dy), (c[0] + extent / 2 * dx, c[1] + extent / 2 * dy),
                        gpts=gpts, sampling=sampling, endpoint=endpoint)

    @property
    def positions(self) -> np.ndarray:
        xs = GridScan.linspace(self.start[0], self.end[0], self.gpts, self.endpoint)
        ys = GridScan.linspace(self.start[1], self.end[1], self.gpts, self.endpoint)
        return np.stack([np.broadcast_to(xs, (self.gpts,)), np.broadcast_to(ys, (self.gpts,))], axis=-1)


__all__ = ["Aberrations", "Aperture", "Probe", "GridScan", "LineScan", "ContrastTransferFunction", "apply_ctf",
           "ABERRATION_NAMES"]
