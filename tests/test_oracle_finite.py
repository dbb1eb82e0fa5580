"""Pins of the oracle's finite projection (abTEM's Potential(projection="finite"); SimulatoKit has none, so no
reference fixture exists: the pins are the mathematics).

- Against independent quadrature: a slice that holds no part of the atom's own z range carries, in Fourier space,
  F(k) = integral over its z range of the 2-D transform at height t of the 3-D Lobato potential. Here that transform
  is a numerical Hankel transform of V3(sqrt(rho^2 + t^2)) (scipy) integrated over t by Gauss-Legendre, with V3 from
  the Lobato table's real-space form; the oracle uses the closed form.
- Telescoping: over a stack that contains the potential's whole z range, the finite slices sum to the infinite ones.
- Continuity: an atom moved across a slice boundary changes the slices continuously, although its own slice flips.
- The thick-slab limit: one slice much thicker than the potential's range is the infinite projection.
"""
import math
import os
import re

import numpy as np
import pytest
from scipy import integrate, special

from oracle import oracle as O

KAPPA = 0.0208865737082965
TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "simulatokit_amd", "csrc",
                     "lobato_table.inc")


def lobato_ab(z):
    rows = [line for line in open(TABLE) if line.startswith("{")]
    nums = [float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", rows[z - 1].split("//")[0])]
    a, b = np.array(nums[1:6]), np.array(nums[6:11])
    return math.pi ** 2 * a / b ** 1.5 / KAPPA, 2 * math.pi / np.sqrt(b)


def v3(r, A, B):
    """The Lobato 3-D potential sum_i A_i (2/(B_i r) + 1) e^{-B_i r}."""
    return float(np.sum(A * (2 / (B * r) + 1) * np.exp(-B * r)))


def slab_transform(k, t0, t1, A, B, nodes=24):
    """integral_{t0}^{t1} dt 2 pi integral_0^inf rho J0(2 pi k rho) V3(sqrt(rho^2 + t^2)) d rho, numerically."""
    x, w = np.polynomial.legendre.leggauss(nodes)
    ts = 0.5 * (t1 - t0) * x + 0.5 * (t1 + t0)
    total = 0.0
    for t, wt in zip(ts, w):
        f = lambda rho: rho * special.j0(2 * math.pi * k * rho) * v3(math.hypot(rho, t), A, B)
        val, _ = integrate.quad(f, 0, 60.0, limit=400, epsabs=1e-13, epsrel=1e-11)
        total += wt * 2 * math.pi * val
    return 0.5 * (t1 - t0) * total


@pytest.mark.parametrize("z", [14, 79])
def test_slice_away_from_the_atom_equals_the_quadrature(z):
    n, L = 64, 16.0
    box = (L, L, 20.0)
    x0, y0, za = 4 * L / n, 8 * L / n, 10.3  # on a grid point: the own slice's bilinear delta is one pixel
    V = O.potential_finite(n, n, [z], [[x0, y0, za]], box, [4.0] * 5)
    A, B = lobato_ab(z)
    dx = L / n
    for j, (t0, t1) in ((0, (0 - za, 4 - za)), (4, (16 - za, 20 - za))):
        # an atom on a grid point is one delta; the slice's transform then is F(k) / S(k), S = sin(s)/s dx dy with
        # s = |k| dx (the compensation every projection in the slices carries, ProjectedPotential.swift:130-143)
        Fk = np.fft.fft2(V[j]) * dx * dx
        for mx, my in ((0, 0), (1, 0), (3, 2), (5, 7)):
            kx, ky = mx / L, my / L
            k = math.hypot(kx, ky)
            s = k * dx
            Fk_mk = Fk[mx, my] * (math.sin(s) / s if s > 0 else 1.0)
            want = slab_transform(k, t0, t1, A, B) * np.exp(-2j * math.pi * (kx * x0 + ky * y0))
            assert abs(Fk_mk - want) <= 1e-7 * abs(want) + 1e-12, (j, mx, my, Fk_mk, want)


def test_finite_slices_sum_to_the_infinite_slices_over_a_deep_stack():
    rng = np.random.default_rng(7)
    n, L, depth = 48, 12.0, 70.0
    numbers = [8, 14, 38, 79, 22]
    xyz = np.column_stack([rng.uniform(0, L, 5), rng.uniform(0, L, 5), rng.uniform(30, 40, 5)])
    th = [2.0] * 35
    Vf = O.potential_finite(n, n, numbers, xyz, (L, L, depth), th)
    Vi, _ = O.potential(n, n, numbers, xyz, (L, L, depth), th)
    scale = np.abs(Vi.sum(0)).max()
    assert np.abs(Vf.sum(0) - Vi.sum(0)).max() <= 1e-9 * scale
    # the slices themselves differ (the finite projection is not the infinite one slice by slice)
    assert np.abs(Vf - Vi).max() > 1e-3 * scale


def test_an_atom_crossing_a_slice_boundary_changes_the_slices_continuously():
    n, L = 32, 8.0
    th = [2.0] * 4
    a = O.potential_finite(n, n, [29], [[3.0, 4.0, 4.0 - 1e-7]], (L, L, 8.0), th)
    b = O.potential_finite(n, n, [29], [[3.0, 4.0, 4.0 + 1e-7]], (L, L, 8.0), th)
    scale = np.abs(a).max()
    assert np.abs(a - b).max() <= 1e-5 * scale
    # the infinite projection jumps: the whole atom moves from slice 1 to slice 2
    ia, _ = O.potential(n, n, [29], [[3.0, 4.0, 4.0 - 1e-7]], (L, L, 8.0), th)
    ib, _ = O.potential(n, n, [29], [[3.0, 4.0, 4.0 + 1e-7]], (L, L, 8.0), th)
    assert np.abs(ia - ib).max() > 0.5 * np.abs(ia).max()


def test_one_thick_slab_is_the_infinite_projection():
    n, L = 32, 10.0
    xyz = [[2.0, 3.0, 30.0], [7.5, 6.0, 31.0]]
    Vf = O.potential_finite(n, n, [8, 79], xyz, (L, L, 60.0), [60.0])
    Vi, _ = O.potential(n, n, [8, 79], xyz, (L, L, 60.0), [60.0])
    assert np.abs(Vf - Vi).max() <= 1e-9 * np.abs(Vi).max()
