"""The synthetic workloads the benchmarks and tests run: seeded structures and scan settings of the five
configurations this project measures (perovskite HRTEM, HAADF scan, frozen-phonon 4D-STEM, precision study,
surrogate training cells). Every structure is built from its parameters and a seed; nothing is read from disk.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .core import AtomicStructure

SRTIO3_A = 3.905


def srtio3_cell(a: float = SRTIO3_A) -> AtomicStructure:
    """Cubic SrTiO3: Sr (0,0,0), Ti (1/2,1/2,1/2), O at the face centres (1/2,1/2,0), (1/2,0,1/2), (0,1/2,1/2), in
    ASE's order as the reference's fixture generator builds it (tools/pins/potential-multislice/pins.py:265-269)."""
    frac = np.array([(0, 0, 0), (0.5, 0.5, 0.5), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)])
    return AtomicStructure((a, a, a), [38, 22, 8, 8, 8], frac * a)


def srtio3(nx: int, ny: int, nz: int) -> AtomicStructure:
    return srtio3_cell().repeated(nx, ny, nz)


@dataclass
class HaadfScan:
    """The HAADF scan workload: SrTiO3 tiled (tx, ty, tz), n x n grid, slices of `slice_thickness`, 200 keV probe with
    a soft `semiangle` aperture, a scan_n x scan_n GridScan over one unit cell, annular detector inner..outer.
    `batch`: positions per launch (0: the library's choice)."""

    tiles: Sequence[int] = (8, 8, 20)
    n: int = 512
    slice_thickness: float = 2.0
    energy: float = 200e3
    semiangle: float = 20.0
    scan_n: int = 64
    inner: float = 50.0
    outer: float = 200.0
    batch: int = 0

    @property
    def structure(self) -> AtomicStructure:
        return srtio3(*self.tiles)

    @property
    def name(self) -> str:
        t = "x".join(str(x) for x in self.tiles)
        return f"srtio3_{t}_{self.n}sq_haadf_{self.scan_n}x{self.scan_n}"


def random_cell(seed: int, extent: float, thickness: float, natoms: int, elements: Sequence[int],
                min_separation: float = 0.0, max_tries: int = 200000) -> AtomicStructure:
    """Uniform random atoms in an orthogonal extent x extent x thickness box, Z drawn uniformly from `elements`; with
    `min_separation` > 0, positions are drawn by rejection so that no two atoms (periodic in x, y) are closer."""
    rng = np.random.default_rng(seed)
    box = np.array([extent, extent, thickness])
    numbers = rng.choice(np.asarray(elements), natoms)
    if min_separation <= 0:
        return AtomicStructure(tuple(box), numbers, rng.uniform(0, 1, (natoms, 3)) * box)
    pos: List[np.ndarray] = []
    tries = 0
    while len(pos) < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random_cell: could not place the atoms with the requested separation")
        p = rng.uniform(0, 1, 3) * box
        if pos:
            d = np.asarray(pos) - p
            d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
            if (np.einsum("ij,ij->i", d, d) < min_separation ** 2).any():
                continue
        pos.append(p)
    return AtomicStructure(tuple(box), numbers, np.asarray(pos))


def uneven_frozen_phonon_stack(seed: int = 20261202, extent: float = 40.0, natoms: int = 2000,
                               elements=(8, 14, 22, 38, 79), configurations: int = 16, t_min: float = 40.0,
                               t_max: float = 120.0, sigma: float = 0.08):
    """The frozen-phonon workload with uneven slice counts: one master cell of depth t_max with `natoms` atoms
    (uniform positions, seeded), displaced per configuration as FrozenPhonons(seed:) does (sigma on every axis), and a
    thickness T_c = t_min + (t_max - t_min) u_c per configuration (u_c from seed + 1); configuration c keeps the
    displaced atoms with 0 <= z < T_c in a box of depth T_c. Returns (structures, T)."""
    from .multislice import FrozenPhonons

    master = random_cell(seed, extent, t_max, natoms, elements)
    fp = FrozenPhonons(master, sigmas={int(z): sigma for z in np.unique(master.numbers)},
                       configurations=configurations, seed=seed)
    T = t_min + (t_max - t_min) * np.random.default_rng(seed + 1).uniform(0, 1, configurations)
    out = []
    for c in range(configurations):
        moved = fp.configuration(c)
        z = moved.positions[:, 2]
        keep = (z >= 0) & (z < T[c])
        out.append(AtomicStructure((extent, extent, float(T[c])), moved.numbers[keep], moved.positions[keep]))
    return out, T


def slice_count(depth: float, requested: float) -> int:
    return math.ceil(depth / requested)
