"""The measured workloads on the GPU, at their real sizes.

- HRTEM (SrTiO3 4 x 4 x 10, 256 x 256, 20 slices, 200 keV plane wave): exit-wave intensity against the float64
  oracle.
- HAADF scan (SrTiO3 8 x 8 x 20, 512 x 512, 40 slices, 64 x 64 positions, 50-200 mrad): the full scan, checked by
  properties that hold whatever its size (the 4mm symmetry of SrTiO3 [001] maps the image onto its transpose and its
  mirrors; the detector never sees more than the incident intensity) and against the oracle at sample positions.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import HaadfScan, SRTIO3_A, srtio3

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


def test_hrtem_exit_intensity_equals_the_oracle():
    st = srtio3(4, 4, 10)
    pot = sk.Potential(st, 256, 256, slice_thickness=2.0)
    assert pot.slice_count == 20
    beam = sk.ElectronBeam(200e3)
    w = sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), pot)[0]
    V, _ = O.potential(256, 256, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    want = O.multislice(V, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((256, 256), complex))[0]
    I, Io = w.intensity(), np.abs(want) ** 2
    print("HRTEM intensity rel L2", rel_l2(I, Io), "rel max", rel_max(I, Io))
    assert rel_l2(I, Io) <= TOL
    assert rel_max(I, Io) <= TOL


@pytest.fixture(scope="module")
def haadf_scan():
    w = HaadfScan()
    st = w.structure
    pot = sk.Potential(st, w.n, w.n, slice_thickness=w.slice_thickness)
    probe = sk.Probe(sk.ElectronBeam(w.energy), pot.grid, sk.Aperture(w.semiangle))
    scan = sk.GridScan((0.0, 0.0), (SRTIO3_A, SRTIO3_A), w.scan_n, w.scan_n)
    prepared = pot.prepare(probe.beam)
    dets = [sk.AnnularDetector(w.inner, w.outer), sk.AnnularDetector(0, None)]
    haadf, total = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=w.batch))
    return w, st, pot, scan, haadf, total


def test_full_haadf_scan_has_the_crystal_symmetry(haadf_scan):
    w, _, _, _, haadf, total = haadf_scan
    img = haadf.image()
    assert img.shape == (64, 64)
    n = w.scan_n
    mirror = img[(n - np.arange(n)) % n, :]
    scale = np.abs(img).max()
    print("transpose", np.abs(img - img.T).max() / scale, "mirror", np.abs(img - mirror).max() / scale)
    assert np.abs(img - img.T).max() <= TOL * scale
    assert np.abs(img - mirror).max() <= TOL * scale
    # the annulus never holds more than the bandlimited total, which never exceeds the incident intensity 1
    assert (img >= 0).all() and (img <= total.image() + 1e-12).all() and (total.image() <= 1 + 1e-5).all()


def test_full_haadf_scan_equals_the_oracle_at_sample_positions(haadf_scan):
    w, st, pot, scan, haadf, _ = haadf_scan
    pick = [0, 1, 33, 64 * 17 + 5, 64 * 32 + 32, 4095]
    pos = scan.positions[pick]
    V, _ = O.potential(w.n, w.n, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    want = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), w.energy, w.semiangle, pos,
                       [O.det_annular(w.inner, w.outer)], threads=6)[0]
    got = haadf.values[pick]
    print("HAADF rel err at samples", np.abs(got - want) / np.abs(want))
    assert rel_max(got, want) <= TOL


def test_default_batch_is_sized_to_the_infinity_cache(haadf_scan):
    """The library's positions per launch: the largest batch whose compact waves (nx x 352 complex64 at 512^2) take at
    most 240 MiB, spread evenly over the scan: 4,096 positions -> 24 launches of at most 171 on one stream. The
    default split scan (two streams) takes 3/4 of that at 512^2: 32 launches of 128, 64 on each stream. A caller's
    batch of 1024 runs as tiles of that size (1024 positions per launch, 128 per kernel launch), and with tile = 1024
    as whole 1024-position kernel launches streaming from HBM; the chunking changes no number the scan computes."""
    w, st, pot, scan, haadf, _ = haadf_scan
    probe = sk.Probe(sk.ElectronBeam(w.energy), pot.grid, sk.Aperture(w.semiangle))
    prepared = pot.prepare(probe.beam)
    det = [sk.AnnularDetector(w.inner, w.outer)]
    sk.set_timing(True)
    try:
        sk.STEM.scan(probe, [prepared], scan, det)
        assert sk.last_timing()["batch"] == 128
        one = sk.STEM.scan(probe, [prepared], scan, det, sk.ExecutionOptions(streams=1))[0]
        assert sk.last_timing()["batch"] == 171
        tiled = sk.STEM.scan(probe, [prepared], scan, det, sk.ExecutionOptions(batch=1024))[0]
        assert (sk.last_timing()["batch"], sk.last_timing()["group"]) == (128, 1024)
        streamed = sk.STEM.scan(probe, [prepared], scan, det, sk.ExecutionOptions(batch=1024, tile=1024))[0]
        assert (sk.last_timing()["batch"], sk.last_timing()["group"]) == (1024, 1024)
    finally:
        sk.set_timing(False)
    assert rel_max(tiled.values, haadf.values) <= 1e-12
    assert rel_max(streamed.values, haadf.values) <= 1e-12
    assert rel_max(one.values, haadf.values) <= 1e-12
    # 1024 positions per graph replay, in cache-sized tiles on two streams, and whole from HBM
    for o in (sk.ExecutionOptions(batch=1024, use_graph=True), sk.ExecutionOptions(batch=1024, tile=1024,
                                                                                   use_graph=True)):
        g = sk.STEM.scan(probe, [prepared], scan, det, o)[0]
        assert rel_max(g.values, haadf.values) <= 1e-12


@pytest.mark.parametrize("n", [256, 512])
def test_split_scans_equal_one_stream(n):
    """Scans whose batches are split over 2, 3 and 4 streams compute the same numbers as one stream, bit for bit
    (every position runs the same kernels; only the launches it shares differ), for every detector kind, including
    a batch that does not divide evenly and a last batch smaller than the streams' parts."""
    st = srtio3(4, 4, 6)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 7, 9).positions
    dets = [sk.AnnularDetector(30, 90), sk.PixelatedDetector(crop=(32, 32)), sk.CenterOfMassDetector(),
            sk.SegmentedDetector(2, 4, 10, 40)]
    prepared = pot.prepare(probe.beam)
    want = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(streams=1, batch=20))
    for streams in (2, 3, 4):
        got = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(streams=streams, batch=20))
        for x, y in zip(want, got):
            assert np.array_equal(np.asarray(x.values), np.asarray(y.values)), (streams, type(x))
    dflt = sk.STEM.scan(probe, [prepared], positions, dets)
    for x, y in zip(want, dflt):
        assert np.array_equal(np.asarray(x.values), np.asarray(y.values))


@pytest.mark.parametrize("graph", [False, True])
def test_split_scans_over_several_potentials_equal_one_stream(graph):
    """Several potentials of different slice counts scanned one at a time (pipeline=0), each batch split over 2 and 3
    streams, with and without graph replay (which captures each potential's batches anew): the same numbers as one
    stream, bit for bit. Each part accumulates its positions' sums on its own stream, in potential order."""
    pots = [sk.Potential(srtio3(4, 4, tz), 256, 256, slice_thickness=2.0) for tz in (2, 3, 5)]
    assert [p.slice_count for p in pots] == [4, 6, 10]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 7, 9).positions
    dets = [sk.AnnularDetector(30, 90), sk.PixelatedDetector(crop=(32, 32)), sk.CenterOfMassDetector()]
    prepared = [p.prepare(probe.beam) for p in pots]
    want = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(streams=1, batch=20, pipeline=0))
    for streams in (2, 3):
        got = sk.STEM.scan(probe, prepared, positions, dets,
                           sk.ExecutionOptions(streams=streams, batch=20, pipeline=0, use_graph=graph))
        for x, y in zip(want, got):
            assert np.array_equal(np.asarray(x.values), np.asarray(y.values)), (streams, graph, type(x))


def test_streams_outside_one_to_four_are_refused():
    st = srtio3(2, 2, 2)
    pot = sk.Potential(st, 256, 256, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    for bad in (-1, 5):
        with pytest.raises(sk.SimulatoError):
            sk.STEM.scan(probe, [pot], [(1.0, 1.0)], [sk.AnnularDetector(30, 90)], sk.ExecutionOptions(streams=bad))
