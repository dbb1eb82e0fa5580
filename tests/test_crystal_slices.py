"""Slicing.crystal_slice_thicknesses against abTEM 1.0.10's crystal_slice_thicknesses.

The expected thicknesses are the numbers the reference's own test holds inline
(Tests/SimulatoMultisliceTests/CrystalSliceThicknessTests.swift:27-39, printed by abTEM), at its rtol of 1e-7. The
refusals are the ones that test lists (:42-58).
"""
import math

import numpy as np
import pytest

import simulatokit_amd as sk


def planted(z, depth):
    return sk.AtomicStructure((4.0, 4.0, depth), [6] * len(z), [[1.0, 2.0, v] for v in z])


def aluminium():
    h = 2.025
    return sk.AtomicStructure((4.05, 4.05, 4.05), [13] * 4, [[0, 0, 0], [0, h, h], [h, 0, h], [h, h, 0]])


def close(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-7 * abs(y), (x, y)


def test_planted_structures_match_abtem():
    four = planted([0.0, 0.6, 1.0, 2.05, 3.99], 4)
    close(sk.Slicing.crystal_slice_thicknesses(four),
          [0.4, 0.6000000000000001, 1.0, 1.8000000000000003, 0.20000000000000018])
    close(sk.Slicing.crystal_slice_thicknesses(four, tolerance=0.5), [0.5, 0.5, 1.0, 1.5, 0.5])
    # 0.3 / 0.2 and 0.6 / 0.2 fall just below 1.5 and 3 in float64: bins 1 and 2, as NumPy's floor gives
    six = planted([0.3, 0.6, 1.8, 3.0, 4.2, 5.4], 6)
    close(sk.Slicing.crystal_slice_thicknesses(six, tolerance=0.2),
          [0.20000000000000004, 0.19999999999999996, 1.4000000000000001, 1.2, 1.1999999999999997,
           1.2000000000000002, 0.6000000000000005])


def test_aluminium_at_a_commensurate_tolerance():
    al = aluminium()
    t = sk.Slicing.crystal_slice_thicknesses(al, tolerance=0.45)
    close(t, [1.7999999999999998, 2.2500000000000004])
    p = sk.Potential(al, 16, 16, slice_thicknesses=t)
    assert p.slice_count == 2
    # the atoms at z = 0 go to the first slice and those at z = 2.025 to the second
    assert sorted(set(p.slice_of_atom.tolist())) == [0, 1]


def test_refuses_what_abtem_refuses():
    six = planted([0.3, 0.6, 1.8, 3.0, 4.2, 5.4], 6)
    with pytest.raises(sk.SimulatoError):
        sk.Slicing.crystal_slice_thicknesses(six, tolerance=0.7)
    with pytest.raises(sk.SimulatoError):
        sk.Slicing.crystal_slice_thicknesses(aluminium())
    with pytest.raises(sk.SimulatoError):
        sk.Slicing.crystal_slice_thicknesses(planted([-0.3, 1.0], 4))
    with pytest.raises(sk.SimulatoError):
        sk.Slicing.crystal_slice_thicknesses(planted([0.05], 0.1))


@pytest.mark.parametrize("t", [0.0, -0.2, math.nan, math.inf, 1e-300])
def test_refuses_a_bad_tolerance(t):
    with pytest.raises(sk.SimulatoError):
        sk.Slicing.crystal_slice_thicknesses(planted([0.0, 1.0], 4), tolerance=t)


def test_commensurate_layers_give_the_layer_spacing():
    # layers every 0.5 A in a 4 A box at a tolerance that divides the spacing: every slice is one spacing, save the
    # half bins at the two faces, and the thicknesses sum to the depth
    z = np.arange(0.0, 4.0, 0.5) + 0.01
    t = sk.Slicing.crystal_slice_thicknesses(planted(list(z), 4.0), tolerance=0.1)
    assert abs(sum(t) - 4.0) < 1e-12
    assert np.allclose(t[1:-1], 0.5, atol=1e-12)
