"""The 512-point fast passes against the float64 oracle, on an MI355X.

The 512 x 512 grid is the HAADF workload's; its column pass has two block shapes (8 or 16 columns per block) and the
lane pipeline has its own column kernel, so each runs here on the same inputs: exit waves of a plane wave, and a scan
with annular, centre-of-mass and pixelated detectors. Bounds as for the other grids (complex64 vs float64).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import SRTIO3_A, srtio3

pytestmark = pytest.mark.gpu

ACC = sk.REGISTERED


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


@pytest.fixture(scope="module")
def stack():
    st = srtio3(4, 4, 2)
    pot = sk.Potential(st, 512, 512, slice_thickness=2.0)
    V, _ = O.potential(512, 512, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    return st, pot, V


@pytest.mark.parametrize("col_tile", [8, 16])
def test_plane_wave_exit_512_equals_the_oracle(stack, col_tile):
    st, pot, V = stack
    beam = sk.ElectronBeam(200e3)
    opt = sk.ExecutionOptions(col_tile=col_tile)
    w = sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), pot, options=opt)[0]
    want = O.multislice(V, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((512, 512), complex))[0]
    assert rel_max(w.field, want) <= ACC.wave


def test_col_tile_choices_are_refused_outside_8_and_16(stack):
    _, pot, _ = stack
    with pytest.raises(sk.SimulatoError):
        sk.Multislice.run(sk.Waves.plane_wave(sk.ElectronBeam(200e3), pot.grid), pot,
                          options=sk.ExecutionOptions(col_tile=4))


@pytest.fixture(scope="module")
def scan_case(stack):
    st, pot, V = stack
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(21.4))
    scan = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 3, 4)
    dets = [sk.AnnularDetector(50, 200), sk.CenterOfMassDetector("cutoff", "mrad"), sk.PixelatedDetector("cutoff")]
    want = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 21.4, scan.positions,
                       [O.det_annular(50, 200), O.det_com("cutoff", "mrad"), O.det_pixelated("cutoff")])
    return pot.prepare(probe.beam), probe, scan, dets, want


@pytest.mark.parametrize("opt", [dict(col_tile=8), dict(col_tile=16), dict(pipeline=1, lane_waves=8),
                                 dict(col_tile=8, use_graph=True, batch=5)],
                         ids=["tile8", "tile16", "pipeline", "graph"])
def test_scan_512_equals_the_oracle(scan_case, opt):
    prepared, probe, scan, dets, want = scan_case
    got = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(**opt))
    for m, wv in zip(got, want):
        wv = wv.reshape(-1)
        assert np.abs(m.values - wv).max() <= ACC.detector_bound(np.abs(wv).max())


@pytest.mark.parametrize("batch", [1, 5, 7, 12])
def test_scan_512_balanced_spans_equal_the_oracle(scan_case, batch):
    """The 8-column pass cuts the (tile pair, position) units of a launch into equal spans, one per pair of blocks;
    batches of 5 and 7 give spans that cross from one tile pair to the next (P reloaded mid-span), batch 1 gives
    spans of one unit. The 16-column blocks run the same launches as the control."""
    prepared, probe, scan, dets, want = scan_case
    got = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=batch, col_tile=8))
    ctl = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=batch, col_tile=16))
    for m, c, wv in zip(got, ctl, want):
        wv = wv.reshape(-1)
        assert np.abs(m.values - wv).max() <= ACC.detector_bound(np.abs(wv).max())
        assert np.abs(m.values - c.values).max() <= 1e-6 * np.abs(wv).max()


@pytest.mark.parametrize("tiles", [(4, 2, 2), (2, 4, 2)], ids=["x_long", "y_long"])
def test_scan_512_anisotropic_sampling_equals_the_oracle(tiles):
    """Unequal sampling in x and y. With the cell twice as long in x, the antialias cutoff follows the coarser x
    sampling, so only |ky| up to a third of the y Nyquist frequency is kept: the kept set is narrower than the 2/3
    band the band-pruned row kernel assumes, and the plain row kernel runs. With the cell longer in y, the kept set
    is the full 2/3 band in y and the column pass's zero rows of P reach further in x. Both run through the column
    pass's P-band pruning."""
    st = srtio3(*tiles)
    pot = sk.Potential(st, 512, 512, slice_thickness=2.0)
    V, _ = O.potential(512, 512, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    scan = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 2, 2)
    dets = [sk.AnnularDetector(50, 200), sk.PixelatedDetector("cutoff")]
    want = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 20.0, scan.positions,
                       [O.det_annular(50, 200), O.det_pixelated("cutoff")])
    got = sk.STEM.scan(probe, [pot.prepare(probe.beam)], scan, dets)
    for m, wv in zip(got, want):
        wv = wv.reshape(-1)
        assert np.abs(m.values - wv).max() <= ACC.detector_bound(np.abs(wv).max())


@pytest.mark.parametrize("npos, nz", [(200, (1, 2, 3)), (24, (1, 2, 3, 1, 2, 3, 2, 1))],
                         ids=["lanes_own_chunks", "lanes_share_chunks"])
def test_lanes_at_512_equal_one_potential_at_a_time(npos, nz):
    """The lanes' 512-point column launches (mid slices, items finishing, items starting) over several potentials of
    uneven slice counts, where lanes run several items in a row, so a lane's next probe is generated into the
    storage its detectors have just read. 200 positions: lanes of 32 own their chunks. 24 positions over 8
    potentials: 24 items on 21 lanes of 8 that share chunks, each with its own accumulators. Each (position,
    potential) wave is the same float32 computation as one potential at a time, so only the float64 order of the
    sum over potentials may differ."""
    pots = [sk.Potential(srtio3(4, 4, k), 512, 512, slice_thickness=2.0) for k in nz]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    rng = np.random.default_rng(7)
    positions = rng.uniform(0.0, 4 * SRTIO3_A, (npos, 2))
    dets = [sk.AnnularDetector(50, 200), sk.CenterOfMassDetector("cutoff", "mrad"), sk.PixelatedDetector("cutoff")]
    prepared = [p.prepare(probe.beam) for p in pots]
    one = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=0))
    lanes = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=1))
    for x, y in zip(one, lanes):
        assert rel_max(y.values, x.values) <= 1e-12
