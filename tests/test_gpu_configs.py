"""The frozen-phonon 4D-STEM workload and the precision study on the GPU, at sizes the float64 oracle finishes in
seconds, against the oracle.

- Frozen phonons with uneven slice counts: configurations of one seeded random cell, each cut to its own thickness
  (so their slice counts differ), displaced as FrozenPhonons(seed:) does; a scan into a numeric 128 x 128 float32
  crop of the pattern and the centre of mass, through the lane pipeline (every launch mixes configurations at
  different slices) and one configuration at a time; both against the oracle's average over the configurations.
- Precision study: the HAADF geometry (SrTiO3 8 x 8 x 20, 40 slices; 20 slices at 509) on 256^2 (register passes),
  384^2 (generic mixed radix 3 x 2^7) and 509^2 (Bluestein) grids; annular values and cutoff patterns against the
  oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import SRTIO3_A, srtio3, uneven_frozen_phonon_stack

pytestmark = pytest.mark.gpu

ACC = sk.REGISTERED


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


def centred_crop(full_shifted, m):
    """The m x m centred crop of an fftshifted even-sized pattern: the frequencies -m/2 .. m/2 - 1 of each axis
    (Waves.cropIndices, Waves.swift:90-96, then fftshift)."""
    n = full_shifted.shape[-1]
    lo = n // 2 - m // 2
    return full_shifted[..., lo:lo + m, lo:lo + m]


def test_frozen_phonons_with_uneven_slice_counts_equal_the_oracle():
    structures, T = uneven_frozen_phonon_stack(seed=20261202, extent=20.0, natoms=150, configurations=3,
                                               t_min=6.0, t_max=14.0)
    n = 256
    pots = [sk.Potential(st, n, n, slice_thickness=1.0) for st in structures]
    counts = [p.slice_count for p in pots]
    assert len(set(counts)) == 3, counts
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    positions = np.array([[5.0, 5.0], [12.5, 7.25], [3.3, 16.1], [10.0, 10.0]])
    dets = [sk.PixelatedDetector(crop=(128, 128), dtype="float32"), sk.CenterOfMassDetector("cutoff", "mrad")]
    prepared = [p.prepare(probe.beam) for p in pots]
    lanes = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=1, lane_waves=8))
    one = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=0))
    assert lanes[0].values.dtype == np.float32 and lanes[0].value_shape == (128, 128)
    for x, y in zip(one, lanes):
        assert rel_max(y.values.astype(np.float64), x.values.astype(np.float64)) <= 1e-6
    Vs, ths = [], []
    for st, p in zip(structures, pots):
        V, _ = O.potential(n, n, st.numbers, st.positions, st.box, p.slice_thicknesses)
        Vs.append(V)
        ths.append(p.slice_thicknesses)
    full, com = O.stem_scan(Vs, ths, (20.0, 20.0), 200e3, 20.0, positions,
                            [O.det_pixelated("full"), O.det_com("cutoff", "mrad")])
    want = centred_crop(full.reshape(len(positions), n, n), 128).reshape(len(positions), -1)
    got = lanes[0].values.reshape(len(positions), -1).astype(np.float64)
    err = rel_max(got, want)
    print("frozen phonons: pattern rel max", err, "rel L2", rel_l2(got, want))
    assert err <= ACC.pattern
    c = com.reshape(len(positions), 2)
    assert np.abs(lanes[1].values.reshape(-1, 2) - c).max() <= ACC.detector_bound(np.abs(c).max())


@pytest.mark.parametrize("n,tz", [(256, 20), (384, 20), (509, 10)])
def test_precision_study_grid_equals_the_oracle(n, tz):
    # 509 (Bluestein on the host too) at 20 slices keeps the float64 oracle near 20 s
    st = srtio3(8, 8, tz)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    assert pot.slice_count == 2 * tz
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 2, 2).positions
    dets = [sk.AnnularDetector(30, 100), sk.PixelatedDetector("cutoff")]
    got = sk.STEM.scan(probe, [pot.prepare(probe.beam)], positions, dets, sk.ExecutionOptions(batch=4))
    V, _ = O.potential(n, n, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    haadf, pat = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 20.0, positions,
                             [O.det_annular(30, 100), O.det_pixelated("cutoff")])
    e_haadf = rel_l2(got[0].values, haadf.reshape(-1))
    e_pat = rel_l2(got[1].values, pat.reshape(-1))
    print(f"n={n}: HAADF rel L2 {e_haadf:.3e}, pattern rel L2 {e_pat:.3e}, "
          f"pattern rel max {rel_max(got[1].values, pat.reshape(-1)):.3e}")
    assert e_haadf <= 1e-4
    assert e_pat <= 1e-4
    assert rel_max(got[1].values, pat.reshape(-1)) <= ACC.pattern
