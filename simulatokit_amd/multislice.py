"""Potentials, slicing, frozen phonons and the multislice entry points of the `.hip` backend.

Mirrors Sources/SimulatoMultislice/{Slicing,Potential,ProjectedPotential,LobatoParameterization,FrozenPhonons,
Multislice,MetalMultislice}.swift. The arithmetic runs in the HIP library (include/simulato_hip.h); this module
holds the host-side bookkeeping the reference's Swift host code does (slicing, wrapping, validation) and hands the
device plain arrays.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Union

import numpy as np

from . import _native as N
from .core import (AtomicStructure, ElectronBeam, ExecutionOptions, Grid, Normalization, SimulatoError, Waves,
                   as_float_pair)

MAXIMUM_SLICE_COUNT = 1_000_000


def _raise(rc: int, what: str = ""):
    if rc == N.SK_OK:
        return
    msg = N.last_error()
    kind = {N.SK_ERR_UNAVAILABLE: "hipUnavailable", N.SK_ERR_INVALID: "invalidInput",
            N.SK_ERR_MEMORY: "hipFailure", N.SK_ERR_DEVICE: "hipFailure"}.get(rc, "hipFailure")
    if "exit planes" in msg:
        kind = "invalidExitPlanes"
    elif "different grids" in msg:
        kind = "gridMismatch"
    elif "invalid detector" in msg:
        kind = "invalidDetector"
    elif "invalid scan" in msg:
        kind = "invalidScan"
    elif "tilt" in msg:
        kind = "invalidTilt"
    raise SimulatoError(kind, (what + ": " if what else "") + msg)


def hip_is_available() -> bool:
    """Multislice.metalIsAvailable's twin (MetalMultislice.swift:18-21)."""
    try:
        return N.lib().sk_hip_available() == 1
    except RuntimeError:
        return False


def _options(opt: Optional[ExecutionOptions]) -> N.SkOptions:
    o = opt or ExecutionOptions()
    s = N.SkOptions()
    s.batch = int(o.batch)
    s.use_graph = 1 if o.use_graph else 0
    s.fast_path = 1 if o.fast_path else 0
    s.pipeline = int(o.pipeline)
    s.lane_waves = int(o.lane_waves)
    s.col_tile = int(o.col_tile)
    if o.precision not in ("float32", "split_bf16"):
        raise ValueError(f"precision must be 'float32' or 'split_bf16', got {o.precision!r}")
    s.precision = 1 if o.precision == "split_bf16" else 0
    s.streams = int(o.streams)
    s.tile = int(o.tile)
    return s


# ---- slicing (Slicing.swift:12-105) --------------------------------------------------------------------------
class Slicing:
    @staticmethod
    def thicknesses(depth: float, requested: float, maximum_count: int = MAXIMUM_SLICE_COUNT) -> List[float]:
        """n = ceil(Lz / s) equal slices of Lz / n (Slicing.swift:17-25)."""
        if not (math.isfinite(requested) and requested > 0):
            raise SimulatoError("invalidSliceThickness", f"slice thickness must be finite and positive, got {requested} A")
        n = math.ceil(depth / requested)
        if not (n >= 1 and n <= min(maximum_count, 1 << 53)):
            raise SimulatoError("invalidSliceThickness", f"slice thickness must be finite and positive, got {requested} A")
        return [depth / n] * int(n)

    @staticmethod
    def validated(thicknesses: Sequence[float], depth: float, maximum_count: int = MAXIMUM_SLICE_COUNT) -> List[float]:
        """A list of thicknesses summing to the depth within NumPy's isclose (Slicing.swift:29-38)."""
        t = [float(x) for x in thicknesses]
        if not t or len(t) > maximum_count:
            raise SimulatoError("invalidSliceThickness", "slice thickness must be finite and positive, got 0 A")
        for s in t:
            if not (math.isfinite(s) and s > 0):
                raise SimulatoError("invalidSliceThickness", f"slice thickness must be finite and positive, got {s} A")
        total = 0.0
        for s in t:
            total += s
        if not abs(total - depth) <= 1e-8 + 1e-5 * abs(depth):
            raise SimulatoError("sliceThicknessesDoNotSumToBox", f"slice thicknesses sum to {total} A, the box is {depth} A deep")
        return t

    @staticmethod
    def crystal_slice_thicknesses(structure: AtomicStructure, tolerance: float = 0.2) -> List[float]:
        """Thicknesses that follow the structure's atomic planes (abTEM's crystal_slice_thicknesses; Slicing.swift:
        51-70): every atom's z as given (not wrapped), with 0 and the box depth, falls in bin floor(z / tolerance);
        the thicknesses are the gaps between the centres (b + 1/2) * tolerance of the occupied bins, bottom up. They
        must sum to the depth within NumPy's isclose, as abTEM asserts."""
        if not (math.isfinite(tolerance) and tolerance > 0):
            raise SimulatoError("invalidSliceThickness", f"slice thickness must be finite and positive, got {tolerance} A")
        depth = float(structure.box[2])
        bins = set()
        for z in list(np.asarray(structure.positions, dtype=np.float64)[:, 2]) + [0.0, depth]:
            q = float(z) / tolerance
            if not (math.isfinite(q) and abs(math.floor(q)) <= float(1 << 53)):
                raise SimulatoError("invalidSliceThickness",
                                    f"slice thickness must be finite and positive, got {tolerance} A")
            bins.add(int(math.floor(q)))
        centres = [(b + 0.5) * tolerance for b in sorted(bins)]
        out = [hi - lo for lo, hi in zip(centres[:-1], centres[1:])]
        total = 0.0
        for s in out:
            total += s
        if not out or not abs(total - depth) <= 1e-8 + 1e-5 * abs(depth):
            raise SimulatoError("sliceThicknessesDoNotSumToBox",
                                f"slice thicknesses sum to {total} A, the box is {depth} A deep")
        return out

    @staticmethod
    def wrapped(structure: AtomicStructure) -> np.ndarray:
        """Fractional coordinates mod 1 (ASE wrap), z within 1e-10 of the top face set to 0 (Slicing.swift:74-88)."""
        box = np.array(structure.box)
        f = np.fmod(structure.positions / box, 1.0)
        f = np.where(f < 0, f + 1.0, f)
        f = np.where(f == 1.0, 0.0, f)
        p = f * box
        p[:, 2] = np.where(p[:, 2] > box[2] - 1e-10, 0.0, p[:, 2])
        return p

    @staticmethod
    def slice_indices(z: np.ndarray, thicknesses: Sequence[float]) -> np.ndarray:
        """e_{j-1} <= z < e_j with e_j = sum_{i<=j} s_i - 1e-12; past the last edge: the last slice (Slicing.swift:92-104)."""
        edges = []
        acc = 0.0
        for s in thicknesses:
            acc += s
            edges.append(acc)
        edges = np.array(edges) - 1e-12
        idx = np.searchsorted(edges, np.asarray(z), side="right")
        return np.minimum(idx, len(thicknesses) - 1).astype(np.int32)


# ---- parameterisation (LobatoParameterization.swift:12-75) --------------------------------------------------------
_LOBATO_PRESENT = None


def _lobato_present() -> np.ndarray:
    global _LOBATO_PRESENT
    if _LOBATO_PRESENT is None:
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "lobato_table.inc")
        present = np.zeros(104, dtype=bool)
        z = 0
        with open(path) as fh:
            for line in fh:
                if line.startswith("{"):
                    z += 1
                    present[z] = line.startswith("{1,")
        _LOBATO_PRESENT = present
    return _LOBATO_PRESENT


@dataclass(frozen=True)
class LobatoParameterization:
    """pi_constant: "exact" (default) or "abtem_float32" (abTEM's float32 pi, which its fixtures use)."""

    pi_constant: str = "exact"

    @staticmethod
    def supports(z: int) -> bool:
        return 1 <= z <= 103 and bool(_lobato_present()[z])

    @property
    def pi_mode(self) -> int:
        return 1 if self.pi_constant == "abtem_float32" else 0


# ---- potentials ---------------------------------------------------------------------------------------------------
class PreparedPotential:
    """A potential resident on the device for one beam and tilt (sk_potential; MetalPreparedPotential's twin)."""

    def __init__(self, handle: int, grid: Grid, beam: ElectronBeam, tilt, thicknesses):
        self._h = C.c_void_p(handle)
        self.grid = grid
        self.beam = beam
        self.tilt = tilt
        self.slice_thicknesses = list(thicknesses)

    @property
    def handle(self):
        return self._h

    @property
    def slice_count(self) -> int:
        return len(self.slice_thicknesses)

    def transmissions(self) -> np.ndarray:
        """The transmission functions as stored on the device (complex64 widened), [slices, nx, ny]."""
        n = self.slice_count * self.grid.count
        re, im = np.empty(n), np.empty(n)
        _raise(N.lib().sk_hip_potential_transmissions(self._h, N.dptr(re), N.dptr(im)), "transmissions")
        return (re + 1j * im).reshape(self.slice_count, self.grid.nx, self.grid.ny)

    def values(self) -> np.ndarray:
        """V of every slice as built on the device (atoms form), [slices, nx, ny]."""
        V = np.empty(self.slice_count * self.grid.count)
        _raise(N.lib().sk_hip_potential_values(self._h, N.dptr(V)), "values")
        return V.reshape(self.slice_count, self.grid.nx, self.grid.ny)

    def release(self):
        import sys
        if sys.is_finalizing():
            return  # the process is ending; the device memory goes with it
        if self._h is not None and self._h.value:
            N.lib().sk_hip_potential_release(self._h)
        self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _sk_grid(g: Grid) -> N.SkGrid:
    return N.SkGrid(g.nx, g.ny, g.extent_x, g.extent_y)


class ProjectedPotential:
    """V [V*Angstrom] per slice on a grid (ProjectedPotential.swift:12-106), here from values given by the caller."""

    def __init__(self, grid: Grid, values: np.ndarray, thicknesses: Sequence[float]):
        self.grid = grid
        self.values = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, grid.nx, grid.ny)
        self.slice_thicknesses = [float(x) for x in thicknesses]
        if len(self.slice_thicknesses) != self.values.shape[0]:
            raise ValueError("one thickness per slice")

    @property
    def slice_count(self) -> int:
        return len(self.slice_thicknesses)

    @property
    def thickness(self) -> float:
        return float(sum(self.slice_thicknesses))

    def prepare(self, beam: ElectronBeam, tilt=(0.0, 0.0)) -> PreparedPotential:
        h = C.c_void_p()
        dz = np.ascontiguousarray(self.slice_thicknesses, dtype=np.float64)
        t = np.ascontiguousarray(as_float_pair(tilt), dtype=np.float64)
        _raise(N.lib().sk_hip_potential_from_values(C.byref(_sk_grid(self.grid)), self.slice_count,
                                                    N.dptr(self.values), N.dptr(dz), beam.energy, N.dptr(t),
                                                    C.byref(h)), "prepare")
        return PreparedPotential(h.value, self.grid, beam, as_float_pair(tilt), self.slice_thicknesses)


class Potential:
    """A structure sliced on a grid, built slice by slice on the device (Potential.swift:12-92)."""

    def __init__(self, structure: AtomicStructure, nx: Optional[int] = None, ny: Optional[int] = None,
                 slice_thickness: Optional[float] = None, slice_thicknesses: Optional[Sequence[float]] = None,
                 sampling: Optional[Union[float, Sequence[float]]] = None,
                 parameterization: LobatoParameterization = LobatoParameterization(),
                 maximum_slice_count: int = MAXIMUM_SLICE_COUNT, projection: str = "infinite"):
        """`projection`: "infinite" (the reference's: each atom's whole potential in the slice that holds it) or
        "finite" (abTEM's projection="finite", which the reference lacks: each slice integrates every atom's 3-D
        potential over its own depth range only)."""
        if projection not in ("infinite", "finite"):
            raise ValueError(f"projection must be 'infinite' or 'finite', got {projection!r}")
        self.projection = projection
        if sampling is not None:
            sx, sy = (sampling, sampling) if np.isscalar(sampling) else sampling
            g = Grid.from_sampling(structure.box[0], structure.box[1], float(sx), float(sy))
            nx, ny = g.nx, g.ny
        self.grid = Grid(int(nx), int(ny), structure.box[0], structure.box[1])
        if slice_thicknesses is None:
            if slice_thickness is None:
                raise ValueError("give slice_thickness or slice_thicknesses")
            slice_thicknesses = Slicing.thicknesses(structure.box[2], slice_thickness, maximum_slice_count)
        self.slice_thicknesses = Slicing.validated(slice_thicknesses, structure.box[2], maximum_slice_count)
        for z in np.unique(structure.numbers):
            if not LobatoParameterization.supports(int(z)):
                raise SimulatoError("unsupportedElement", f"element Z = {int(z)} has no Lobato parameters (Z = 1 ... 103 are tabulated)")
        self.parameterization = parameterization
        self.numbers = structure.numbers.copy()
        self.positions = Slicing.wrapped(structure)
        self.slice_of_atom = Slicing.slice_indices(self.positions[:, 2], self.slice_thicknesses)

    @property
    def slice_count(self) -> int:
        return len(self.slice_thicknesses)

    @property
    def thickness(self) -> float:
        return float(sum(self.slice_thicknesses))

    def prepare(self, beam: ElectronBeam, tilt=(0.0, 0.0)) -> PreparedPotential:
        h = C.c_void_p()
        dz = np.ascontiguousarray(self.slice_thicknesses, dtype=np.float64)
        t = np.ascontiguousarray(as_float_pair(tilt), dtype=np.float64)
        num = np.ascontiguousarray(self.numbers, dtype=np.int32)
        xy = np.ascontiguousarray(self.positions[:, :2], dtype=np.float64)
        soa = np.ascontiguousarray(self.slice_of_atom, dtype=np.int32)
        if self.projection == "finite":
            xyz = np.ascontiguousarray(self.positions, dtype=np.float64)
            _raise(N.lib().sk_hip_potential_from_atoms_finite(
                C.byref(_sk_grid(self.grid)), self.slice_count, N.dptr(dz), len(num), N.iptr(num), N.dptr(xyz),
                self.parameterization.pi_mode, beam.energy, N.dptr(t), C.byref(h)), "prepare")
            return PreparedPotential(h.value, self.grid, beam, as_float_pair(tilt), self.slice_thicknesses)
        _raise(N.lib().sk_hip_potential_from_atoms(C.byref(_sk_grid(self.grid)), self.slice_count, N.dptr(dz),
                                                   len(num), N.iptr(num), N.dptr(xy), N.iptr(soa),
                                                   self.parameterization.pi_mode, beam.energy, N.dptr(t),
                                                   C.byref(h)), "prepare")
        return PreparedPotential(h.value, self.grid, beam, as_float_pair(tilt), self.slice_thicknesses)

    def build(self, beam: Optional[ElectronBeam] = None) -> ProjectedPotential:
        """V of every slice as the device builds it (Multislice.gpuPotentialSlices, MetalMultislice.swift:104-107)."""
        prepared = self.prepare(beam or ElectronBeam(200e3))
        V = prepared.values()
        prepared.release()
        return ProjectedPotential(self.grid, V, self.slice_thicknesses)


AnyPotential = Union[Potential, ProjectedPotential, PreparedPotential]


def _prepared(potential: AnyPotential, beam: ElectronBeam, tilt) -> PreparedPotential:
    if isinstance(potential, PreparedPotential):
        if potential.beam.energy != beam.energy or tuple(potential.tilt) != tuple(as_float_pair(tilt)):
            raise SimulatoError("invalidTilt", "the prepared potential was made for another beam or tilt")
        return potential
    return potential.prepare(beam, tilt)


# ---- exit planes (Multislice.swift:13-40) ----------------------------------------------------------------------------
class ExitPlanes:
    def __init__(self, kind: str, arg=None):
        self.kind, self.arg = kind, arg

    @staticmethod
    def last():
        return ExitPlanes("last")

    @staticmethod
    def every(m: int):
        return ExitPlanes("every", m)

    @staticmethod
    def after_slices(planes: Sequence[int]):
        return ExitPlanes("after", list(planes))

    def planes(self, slice_count: int) -> List[int]:
        n = slice_count
        if self.kind == "last":
            return [n - 1]
        if self.kind == "every":
            m = self.arg
            if m < 1:
                raise SimulatoError("invalidExitPlanes", f"invalid exit planes: every({m}): the step must be at least 1")
            if m >= n:
                return [n - 1]
            p = list(range(m - 1, n, m))
            if p[-1] != n - 1:
                p.append(n - 1)
            return [-1] + p
        lst = self.arg
        if lst != sorted(lst) or len(set(lst)) != len(lst) or not all(-1 <= x < n for x in lst):
            raise SimulatoError("invalidExitPlanes", f"invalid exit planes: afterSlices must be ascending, distinct, within -1 ..< {n}")
        return lst if lst and lst[-1] == n - 1 else lst + [n - 1]


# ---- the multislice entry points ----------------------------------------------------------------------------------
class Multislice:
    """Multislice.run / runBatch on the `.hip` backend (Multislice.swift:49-80, MetalMultislice.swift:49-101)."""

    @staticmethod
    def run(incident: Waves, through: AnyPotential, exit_planes: ExitPlanes = ExitPlanes.last(),
            options: Optional[ExecutionOptions] = None) -> List[Waves]:
        if not incident.grid.matches(through.grid):
            raise SimulatoError("gridMismatch", "the waves and the potential are on different grids")
        prepared = _prepared(through, incident.beam, incident.tilt)
        planes = exit_planes.planes(prepared.slice_count)
        g = prepared.grid
        re = np.ascontiguousarray(incident.field.real)
        im = np.ascontiguousarray(incident.field.imag)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        ore = np.empty((len(planes), g.nx, g.ny))
        oim = np.empty_like(ore)
        opt = _options(options)
        _raise(N.lib().sk_hip_run(prepared.handle, N.dptr(re), N.dptr(im), N.iptr(pl), len(planes), N.dptr(ore),
                                  N.dptr(oim), C.byref(opt)), "Multislice.run")
        return [Waves(ore[q] + 1j * oim[q], incident.grid, incident.beam, incident.normalization, incident.tilt)
                for q in range(len(planes))]

    @staticmethod
    def run_batch(incidents: Sequence[Waves], through: AnyPotential,
                  options: Optional[ExecutionOptions] = None) -> List[Waves]:
        if not incidents:
            return []
        first = incidents[0]
        for w in incidents:
            if not w.grid.matches(through.grid):
                raise SimulatoError("gridMismatch", "the waves and the potential are on different grids")
            if tuple(w.tilt) != tuple(first.tilt):
                raise SimulatoError("invalidTilt", "invalid tilt: the waves of one .hip batch must share one tilt; run each tilt separately")
        prepared = _prepared(through, first.beam, first.tilt)
        g = prepared.grid
        stack = np.stack([w.field for w in incidents])
        re = np.ascontiguousarray(stack.real)
        im = np.ascontiguousarray(stack.imag)
        ore = np.empty_like(re)
        oim = np.empty_like(im)
        opt = _options(options)
        _raise(N.lib().sk_hip_run_batch(prepared.handle, len(incidents), N.dptr(re), N.dptr(im), N.dptr(ore),
                                        N.dptr(oim), C.byref(opt)), "Multislice.runBatch")
        return [Waves(ore[q] + 1j * oim[q], w.grid, w.beam, w.normalization, w.tilt) for q, w in enumerate(incidents)]


# ---- frozen phonons (FrozenPhonons.swift:14-156) ---------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _splitmix_finalise(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


class GaussianGenerator:
    """SplitMix64 seeds xoshiro256**; Box-Muller, both outputs used (FrozenPhonons.swift:14-58)."""

    def __init__(self, seed: int):
        x = seed & _M64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & _M64
            s.append(_splitmix_finalise(x))
        self._s = s
        self._spare = None

    def next_uint64(self) -> int:
        s = self._s

        def rotl(v, k):
            return ((v << k) | (v >> (64 - k))) & _M64

        result = (rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return result

    def next_normal(self) -> float:
        if self._spare is not None:
            z, self._spare = self._spare, None
            return z
        u1 = (float(self.next_uint64() >> 11) + 1) * 2.0 ** -53
        u2 = float(self.next_uint64() >> 11) * 2.0 ** -53
        r = math.sqrt(-2 * math.log(u1))
        theta = 2 * math.pi * u2
        self._spare = r * math.sin(theta)
        return r * math.cos(theta)

    def normals(self, count: int) -> np.ndarray:
        return np.array([self.next_normal() for _ in range(count)])


class FrozenPhonons:
    """Configurations of one structure, atoms displaced by Gaussians of sigma_Z (FrozenPhonons.swift:61-156)."""

    def __init__(self, structure: AtomicStructure, displacements: Optional[Sequence[np.ndarray]] = None,
                 sigmas: Optional[Dict[int, float]] = None, configurations: int = 1, seed: int = 0,
                 directions: str = "xyz"):
        self.structure = structure
        if displacements is None:
            if configurations < 1:
                raise SimulatoError("invalidFrozenPhonons", "invalid frozen phonons: configurations must be at least 1")
            sigmas = sigmas or {}
            for z in np.unique(structure.numbers):
                if int(z) not in sigmas:
                    raise SimulatoError("invalidFrozenPhonons", f"invalid frozen phonons: sigma must be given for every element; Z = {int(z)} has none")
            for z, s in sigmas.items():
                if not (math.isfinite(s) and s >= 0):
                    raise SimulatoError("invalidFrozenPhonons", f"invalid frozen phonons: sigma of Z = {z} must be finite and non-negative, got {s}")
            s_of = np.array([sigmas[int(z)] for z in structure.numbers])
            mask = np.array([a in directions for a in "xyz"], dtype=np.float64)
            displacements = []
            for c in range(configurations):
                g = GaussianGenerator(self.configuration_seed(seed, c))
                r = g.normals(3 * len(structure.numbers)).reshape(-1, 3)
                displacements.append(s_of[:, None] * r * mask[None, :])
        displacements = [np.asarray(d, dtype=np.float64).reshape(-1, 3) for d in displacements]
        if not displacements:
            raise SimulatoError("invalidFrozenPhonons", "invalid frozen phonons: no configuration")
        for c, d in enumerate(displacements):
            if len(d) != len(structure.numbers):
                raise SimulatoError("invalidFrozenPhonons", f"invalid frozen phonons: configuration {c} has {len(d)} displacements for {len(structure.numbers)} atoms")
            if not np.isfinite(d).all():
                raise SimulatoError("invalidFrozenPhonons", f"invalid frozen phonons: configuration {c} has a non-finite displacement")
        self.displacements = displacements

    @staticmethod
    def configuration_seed(seed: int, c: int) -> int:
        """mix(mix(seed) + c), mix = SplitMix64's finaliser of v + golden gamma (FrozenPhonons.swift:122-130)."""
        def mix(v):
            return _splitmix_finalise((v + 0x9E3779B97F4A7C15) & _M64)
        return mix((mix(seed & _M64) + c) & _M64)

    @property
    def count(self) -> int:
        return len(self.displacements)

    def configuration(self, c: int) -> AtomicStructure:
        return AtomicStructure(self.structure.box, self.structure.numbers,
                               self.structure.positions + self.displacements[c])

    def lazy_potentials(self, nx: int, ny: int, slice_thickness: float,
                        parameterization: LobatoParameterization = LobatoParameterization()) -> List[Potential]:
        return [Potential(self.configuration(c), nx, ny, slice_thickness=slice_thickness,
                          parameterization=parameterization) for c in range(self.count)]


__all__ = ["Slicing", "LobatoParameterization", "Potential", "ProjectedPotential", "PreparedPotential", "ExitPlanes",
           "Multislice", "FrozenPhonons", "GaussianGenerator", "hip_is_available", "Normalization"]
