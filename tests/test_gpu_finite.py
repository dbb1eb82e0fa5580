"""The finite projection on the GPU (abTEM's Potential(projection="finite"); SimulatoKit has none, so the parity is
against the oracle, whose finite projection is pinned by quadrature and invariants in test_oracle_finite.py).

- Slices, transmission functions and exit waves equal the oracle's on a small perovskite stack (atoms on and near
  slice boundaries feed both neighbours).
- At the HRTEM grid (256^2 over SrTiO3 4 x 4 cells), slices and the exit wave of a three-cell stack equal the oracle's.
- At the HRTEM grid (256^2, SrTiO3 4 x 4 x 10, 800 atoms) with 30 Angstrom of vacuum above and below, the finite
  slices sum to the infinite ones (the property holds whatever the size: every atom's z range lies inside the stack).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import srtio3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


def test_finite_slices_and_exit_wave_equal_the_oracle():
    st = srtio3(2, 2, 3)  # 7.81 x 7.81 x 11.7 Angstrom, 60 atoms; O and Ti planes sit on slice boundaries
    pot = sk.Potential(st, 96, 96, slice_thickness=1.9525, projection="finite")
    beam = sk.ElectronBeam(200e3)
    prepared = pot.prepare(beam)
    V = prepared.values()
    Vo = O.potential_finite(96, 96, st.numbers, pot.positions, st.box, pot.slice_thicknesses)
    print("finite V rel max vs oracle", rel_max(V, Vo))
    assert rel_max(V, Vo) <= 1e-5
    Vi, _ = O.potential(96, 96, st.numbers, pot.positions, st.box, pot.slice_thicknesses)
    assert rel_max(V, Vi) > 1e-3  # the projection differs slice by slice
    t = prepared.transmissions()
    to = O.transmission(Vo, (st.box[0], st.box[1]), 200e3)
    assert rel_max(t, to) <= 1e-5
    w = sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), prepared)[0]
    want = O.multislice(Vo, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((96, 96), complex))[0]
    assert rel_max(w.field, want) <= sk.REGISTERED.wave


def test_finite_slices_and_exit_wave_at_the_hrtem_grid_equal_the_oracle():
    # the HRTEM configuration's grid and lateral cell (256^2 over SrTiO3 4 x 4 cells, 1.9525 A slices, planes on the
    # slice boundaries), three unit cells deep (240 atoms, 6 slices) so that the oracle's per-atom quadrature
    # finishes in about 15 s
    st = srtio3(4, 4, 3)
    pot = sk.Potential(st, 256, 256, slice_thickness=1.9525, projection="finite")
    assert pot.slice_count == 6
    beam = sk.ElectronBeam(200e3)
    prepared = pot.prepare(beam)
    V = prepared.values()
    Vo = O.potential_finite(256, 256, st.numbers, pot.positions, st.box, pot.slice_thicknesses)
    print("256^2 finite V rel max vs oracle", rel_max(V, Vo))
    assert rel_max(V, Vo) <= 1e-5
    w = sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), prepared)[0]
    want = O.multislice(Vo, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((256, 256), complex))[0]
    print("256^2 finite exit wave rel max vs oracle", rel_max(w.field, want))
    assert rel_max(w.field, want) <= sk.REGISTERED.wave
    prepared.release()


def test_finite_slices_sum_to_the_infinite_slices_at_the_hrtem_grid():
    st = srtio3(4, 4, 10)
    pad = 30.0
    box = (st.box[0], st.box[1], st.box[2] + 2 * pad)
    xyz = st.positions + np.array([0.0, 0.0, pad])
    padded = sk.AtomicStructure(box, st.numbers, xyz)
    fin = sk.Potential(padded, 256, 256, slice_thickness=2.0, projection="finite")
    inf = sk.Potential(padded, 256, 256, slice_thickness=2.0)
    assert fin.slice_count == 50
    beam = sk.ElectronBeam(200e3)
    Vf = fin.prepare(beam).values()
    Vi = inf.prepare(beam).values()
    scale = np.abs(Vi.sum(0)).max()
    err = np.abs(Vf.sum(0) - Vi.sum(0)).max() / scale
    print("sum of finite slices vs sum of infinite slices, rel max", err)
    assert err <= 1e-5
    # the first 12 Angstrom lie over 18 Angstrom from every atom, beyond the reach of the slowest Lobato term
    # (e^{-B T} < 1e-13 for Sr's smallest B = 1.72 / Angstrom)
    assert np.abs(Vf[:6]).max() <= 1e-12 * scale
