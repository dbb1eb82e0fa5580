"""Grid, beam, structure, waves, options and errors: the host side of SimulatoCore that the `.hip` path needs.

Mirrors Sources/SimulatoCore/{Grid,ElectronBeam,PhysicalConstants,AtomicStructure,Waves,SimulatoError}.swift and
Sources/SimulatoMultislice/ExecutionOptions.swift. Arrays are NumPy, row-major ``[ix, iy]`` (x first).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np


class SimulatoError(Exception):
    """A refusal (SimulatoError.swift:7-57). ``kind`` names the reference's case, e.g. "invalidEnergy"."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


# ---- constants (PhysicalConstants.swift:8-31) ---------------------------------------------------------------------
PLANCK = 6.62607004e-34
ELECTRON_MASS = 9.10938356e-31
SPEED_OF_LIGHT = 299792458.0
ELEMENTARY_CHARGE = 1.6021766208e-19
KAPPA = 0.0208865737082965
_ASE_KG = 6.0221408585491615e+26
_ASE_COULOMB = 6.241509125883258e+18
_ASE_SECOND = 98226947884640.62
_ASE_JOULE = 6.241509125883258e+18


def fftfreq(n: int, d: float) -> np.ndarray:
    """NumPy's fftfreq, bit for bit (Grid.swift:82-86)."""
    val = 1.0 / (float(n) * d)
    positive = (n - 1) // 2 + 1
    i = np.arange(n)
    return np.where(i < positive, i, i - n).astype(np.float64) * val


@dataclass(frozen=True)
class Grid:
    """nx x ny points over extent_x x extent_y Angstrom (Grid.swift:8-27)."""

    nx: int
    ny: int
    extent_x: float
    extent_y: float

    def __post_init__(self):
        if self.nx < 2:
            raise SimulatoError("unsupportedGridLength", f"grid length {self.nx} is not supported: lengths must be at least 2")
        if self.ny < 2:
            raise SimulatoError("unsupportedGridLength", f"grid length {self.ny} is not supported: lengths must be at least 2")
        if not (math.isfinite(self.extent_x) and math.isfinite(self.extent_y) and self.extent_x > 0 and self.extent_y > 0):
            raise SimulatoError("invalidExtent", f"grid extent must be finite and positive, got ({self.extent_x}, {self.extent_y}) A")

    @staticmethod
    def from_sampling(extent_x: float, extent_y: float, sampling_x: float, sampling_y: float) -> "Grid":
        """abTEM's grid from a target sampling: ceil(extent / sampling) points (Grid.swift:32-42)."""
        if not (math.isfinite(sampling_x) and math.isfinite(sampling_y) and sampling_x > 0 and sampling_y > 0):
            raise SimulatoError("invalidSampling", f"grid sampling must be finite and positive, got ({sampling_x}, {sampling_y}) A")
        if not (math.isfinite(extent_x) and math.isfinite(extent_y) and extent_x > 0 and extent_y > 0):
            raise SimulatoError("invalidExtent", f"grid extent must be finite and positive, got ({extent_x}, {extent_y}) A")
        px, py = math.ceil(extent_x / sampling_x), math.ceil(extent_y / sampling_y)
        if px >= 1e9 or py >= 1e9:
            raise SimulatoError("invalidSampling", "grid sampling gives too many points")
        return Grid(int(px), int(py), extent_x, extent_y)

    @property
    def sampling_x(self) -> float:
        return self.extent_x / self.nx

    @property
    def sampling_y(self) -> float:
        return self.extent_y / self.ny

    @property
    def reciprocal_sampling_x(self) -> float:  # Grid.swift:52-53
        return 1 / (float(self.nx) * self.sampling_x)

    @property
    def reciprocal_sampling_y(self) -> float:
        return 1 / (float(self.ny) * self.sampling_y)

    @property
    def count(self) -> int:
        return self.nx * self.ny

    @property
    def frequencies_x(self) -> np.ndarray:
        return fftfreq(self.nx, self.sampling_x)

    @property
    def frequencies_y(self) -> np.ndarray:
        return fftfreq(self.ny, self.sampling_y)

    def matches(self, other: "Grid", relative_tolerance: float = 1e-12) -> bool:  # Grid.swift:73-79
        def close(a, b):
            return a == b or abs(a - b) <= relative_tolerance * max(abs(a), abs(b))
        return (self.nx == other.nx and self.ny == other.ny and close(self.extent_x, other.extent_x)
                and close(self.extent_y, other.extent_y))


@dataclass(frozen=True)
class ElectronBeam:
    """A beam of kinetic energy [eV] (ElectronBeam.swift:10-44)."""

    energy: float

    def __post_init__(self):
        if not (math.isfinite(self.energy) and self.energy > 0):
            raise SimulatoError("invalidEnergy", f"electron energy must be finite and positive, got {self.energy} eV")

    @property
    def wavelength(self) -> float:
        c2 = SPEED_OF_LIGHT * SPEED_OF_LIGHT
        return (PLANCK * SPEED_OF_LIGHT / math.sqrt(self.energy * (2 * ELECTRON_MASS * c2 / ELEMENTARY_CHARGE
                                                                   + self.energy)) / ELEMENTARY_CHARGE * 1.0e10)

    @property
    def interaction_parameter(self) -> float:
        mass = (1 + ELEMENTARY_CHARGE * self.energy / (ELECTRON_MASS * (SPEED_OF_LIGHT * SPEED_OF_LIGHT))) \
            * ELECTRON_MASS
        hsj = PLANCK * _ASE_SECOND * _ASE_JOULE
        return 2 * math.pi * mass * _ASE_KG * ELEMENTARY_CHARGE * _ASE_COULOMB * self.wavelength / (hsj * hsj)


@dataclass
class AtomicStructure:
    """An orthogonal box and atoms in Cartesian Angstrom (AtomicStructure.swift:8-55)."""

    box: Tuple[float, float, float]
    numbers: np.ndarray
    positions: np.ndarray

    def __post_init__(self):
        self.box = tuple(float(b) for b in self.box)
        self.numbers = np.asarray(self.numbers, dtype=np.int32).reshape(-1)
        self.positions = np.asarray(self.positions, dtype=np.float64).reshape(-1, 3)
        if not all(math.isfinite(b) and b > 0 for b in self.box):
            raise SimulatoError("invalidBox", f"the box must be orthogonal with finite positive sides, got {self.box} A")
        if len(self.numbers) != len(self.positions):
            raise SimulatoError("invalidPosition", "one position per atom")
        bad = ~np.isfinite(self.positions).all(axis=1)
        if bad.any():
            i = int(np.argmax(bad))
            raise SimulatoError("invalidPosition", f"atom {i} has a non-finite position")

    def repeated(self, nx: int, ny: int, nz: int) -> "AtomicStructure":
        """ASE's repeat order: x outermost, then y, then z, atoms innermost (AtomicStructure.swift:40-54)."""
        if min(nx, ny, nz) < 1:
            raise ValueError("repeat counts must be positive")
        shifts = np.array([(mx * self.box[0], my * self.box[1], mz * self.box[2])
                           for mx in range(nx) for my in range(ny) for mz in range(nz)])
        pos = (shifts[:, None, :] + self.positions[None, :, :]).reshape(-1, 3)
        num = np.tile(self.numbers, len(shifts))
        return AtomicStructure((self.box[0] * nx, self.box[1] * ny, self.box[2] * nz), num, pos)


class Normalization:
    """Waves.Normalization (Waves.swift:10-15)."""

    VALUES = "values"  # a plane wave: the pattern is computed from psi / N
    RECIPROCAL_SPACE = "reciprocal_space"  # a probe: sum_k |Psi(k)|^2 = 1


@dataclass
class Waves:
    """A wave on a grid with its beam, normalisation and tilt [mrad] (Waves.swift:8-97)."""

    field: np.ndarray
    grid: Grid
    beam: ElectronBeam
    normalization: str = Normalization.VALUES
    tilt: Tuple[float, float] = (0.0, 0.0)

    def __post_init__(self):
        self.field = np.asarray(self.field, dtype=np.complex128)
        if self.field.shape != (self.grid.nx, self.grid.ny):
            raise ValueError("Waves: the field must match the grid")

    @staticmethod
    def plane_wave(beam: ElectronBeam, grid: Grid, tilt=(0.0, 0.0)) -> "Waves":
        return Waves(np.ones((grid.nx, grid.ny), dtype=np.complex128), grid, beam, Normalization.VALUES, tuple(tilt))

    def intensity(self) -> np.ndarray:
        return self.field.real ** 2 + self.field.imag ** 2

    def apply_ctf(self, ctf) -> "Waves":
        """The waves with a ContrastTransferFunction applied on the device (Waves.applying,
        ContrastTransferFunction.swift:140-153)."""
        from .probe import apply_ctf
        return apply_ctf(self, ctf)

    @property
    def total_intensity(self) -> float:
        return float(self.intensity().sum())


@dataclass
class ExecutionOptions:
    """How the `.hip` path executes (ExecutionOptions.swift:21-49). None of these changes which numbers are
    computed, only their rounding."""

    batch: int = 0          # positions per launch (a scan's host chunk, or one graph replay with use_graph); 0 =
                            # the library's choice
    use_graph: bool = False  # replay the slice loop from captured hipGraphs (the first batch of a shape is captured)
    fast_path: bool = True   # register-resident FFT passes where the grid length has one
    pipeline: int = -1       # scan: 1 lanes of positions each on its own potential and slice per launch, 0 one
                             # potential at a time, -1 the library's choice (lanes for a few positions over several
                             # potentials, else one potential at a time)
    lane_waves: int = 0      # positions per lane (0 = the library's choice)
    col_tile: int = 0        # columns per column-pass block: 8 or 16 (512, 384 points), 4 or 8 (1024); 0 = library's
    precision: str = "float32"  # the slice loop's transforms: "float32", or "split_bf16" (matrix cores)
    streams: int = 0         # scan one potential at a time through the fast passes: streams the parts of each batch
                             # run on (1-4), 0 the library's choice (2 for batches of at least 128 positions)
    tile: int = 0            # scan one potential at a time: positions per kernel launch within a batch; 0 = the
                             # library's choice (parts that stay in the last-level cache); tile = batch streams the
                             # batch through HBM


@dataclass(frozen=True)
class HipAccuracy:
    """The agreement of `.hip` (complex64) with the float64 path that the tests hold it to, relative to the array's
    largest magnitude (the reference registers the same bounds for its float32 backend, ExecutionOptions.swift:
    55-94)."""

    wave: float = 1e-4
    intensity: float = 2e-4
    pattern: float = 2e-4
    detector: float = 2e-4
    detector_floor: float = 1e-5
    total_intensity_per_slice: float = 1.5e-6
    slice_limit: int = 120
    grid_limit: int = 1024

    def detector_bound(self, image_maximum: float, incident_intensity: float = 1.0) -> float:
        return self.detector * max(image_maximum, self.detector_floor * incident_intensity)


REGISTERED = HipAccuracy()


class Backend:
    """The backend this package provides: the MI355X HIP library (beside the reference's `.cpu` and `.metal`)."""

    HIP = "hip"


def as_float_pair(v: Optional[Sequence[float]]) -> Tuple[float, float]:
    if v is None:
        return (0.0, 0.0)
    return (float(v[0]), float(v[1]))


__all__ = ["SimulatoError", "Grid", "ElectronBeam", "AtomicStructure", "Waves", "Normalization", "ExecutionOptions",
           "HipAccuracy", "REGISTERED", "Backend", "fftfreq", "KAPPA", "field"]
