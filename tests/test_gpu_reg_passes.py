"""The register-resident passes for lines of 384 and 1024 points and for Bluestein's lengths 257..512 (on 1024
points) against the generic LDS passes (ExecutionOptions(fast_path=False)) and the float64 oracle, at every
column-pass width (ExecutionOptions.col_tile: 8 or 16 columns per block at 384 points, 4 or 8 at 1024).

Both paths compute the same numbers in float32 with different operation orders, so they agree to float32 rounding
over the slice loop; both are held to the oracle at the scan's float32 bound.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import SRTIO3_A, srtio3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


# (nx, ny, tiles, tz): 384 = 8 8 6 on both axes; 1024; 509 by Bluestein on 1024 points; mixed axes (the column
# pass's tiles then span the other length's kept columns)
GRIDS = [(384, 384, 8, 5), (1024, 1024, 8, 3), (509, 509, 8, 5), (509, 384, 8, 4), (300, 1024, 4, 3),
         (1024, 384, 8, 3), (384, 1024, 8, 3)]


@pytest.mark.parametrize("nx,ny,tiles,tz", GRIDS)
def test_register_passes_scan_as_the_generic_passes_and_the_oracle(nx, ny, tiles, tz):
    st = srtio3(tiles, tiles, tz)
    pot = sk.Potential(st, nx, ny, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 3, 3).positions[[0, 4, 8]]
    prepared = pot.prepare(probe.beam)
    outer = 100.0  # inside the largest simulated angle of every grid here (about 150 mrad at 384 x 384 over 31 A)
    # every kind of the detecting pass: an annular slot, pixelated accumulators, two centre-of-mass slots and the
    # bins of a segmented detector
    dets = [sk.AnnularDetector(40, outer), sk.PixelatedDetector("cutoff"), sk.CenterOfMassDetector("cutoff", "mrad"),
            sk.SegmentedDetector(2, 4, 10, 40, 0.3)]
    reg = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions())
    gen = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(fast_path=False))
    widths = (8, 16) if nx == 384 else ((4, 8) if nx in (509, 1024) else ())
    for tc in widths:  # every column-pass width agrees with the default one to float32 rounding
        other = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(col_tile=tc))
        for k in range(len(dets)):
            assert rel_max(other[k].values, reg[k].values) <= 2e-5, (tc, k)
    V, _ = O.potential(nx, ny, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    want = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 20.0, positions,
                       [O.det_annular(40, outer), O.det_pixelated("cutoff"), O.det_com("cutoff", "mrad"),
                        O.det_segmented(2, 4, 10, 40, 0.3)])
    for k, name in enumerate(("annular", "pattern", "COM", "segmented")):
        e_rg = rel_max(reg[k].values, gen[k].values)
        e_ro = rel_l2(reg[k].values, np.asarray(want[k]).reshape(-1))
        e_go = rel_l2(gen[k].values, np.asarray(want[k]).reshape(-1))
        print(f"{nx}x{ny} {name}: register vs generic {e_rg:.2e}; vs oracle: register {e_ro:.2e}, generic {e_go:.2e}")
        assert e_rg <= 2e-5
        assert e_ro <= 1e-4
    prepared.release()


@pytest.mark.parametrize("nx,ny,tiles,tz", [(384, 384, 4, 3), (509, 509, 4, 3), (1024, 1024, 2, 2)])
def test_register_passes_exit_planes_as_the_generic_passes(nx, ny, tiles, tz):
    # a plane wave with every exit plane: the entry, middle (with plane), and exit row modes
    st = srtio3(tiles, tiles, tz)
    pot = sk.Potential(st, nx, ny, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    prepared = pot.prepare(beam)
    wave = sk.Waves.plane_wave(beam, pot.grid)
    reg = sk.Multislice.run(wave, prepared, exit_planes=sk.ExitPlanes.every(1))
    gen = sk.Multislice.run(wave, prepared, exit_planes=sk.ExitPlanes.every(1),
                            options=sk.ExecutionOptions(fast_path=False))
    assert len(reg) == len(gen) == pot.slice_count + 1
    worst = max(rel_max(a.field, b.field) for a, b in zip(reg, gen))
    print(f"{nx}x{ny}: {len(reg)} planes, register vs generic rel max {worst:.2e}")
    assert worst <= 2e-5
    prepared.release()


def test_column_widths_the_length_has_not_are_refused():
    st = srtio3(4, 4, 2)
    pot = sk.Potential(st, 384, 384, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    with pytest.raises(sk.SimulatoError):
        sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), pot, options=sk.ExecutionOptions(col_tile=4))


@pytest.mark.parametrize("n", [384, 509, 1024])
def test_register_passes_replayed_from_graphs_give_the_same_bits(n):
    # 13 positions in batches of 4: the first batch of each shape runs kernel by kernel while its graph is captured,
    # the next two replay it (the batch's first position read from device memory), the remainder of 1 runs once
    st = srtio3(4, 4, 2)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 4, 4).positions[:13]
    prepared = pot.prepare(probe.beam)
    dets = [sk.AnnularDetector(40, 100), sk.CenterOfMassDetector("cutoff", "mrad"), sk.PixelatedDetector("cutoff")]
    eager = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(batch=4))
    for streams in (1, 2):
        graph = sk.STEM.scan(probe, [prepared], positions, dets,
                             sk.ExecutionOptions(batch=4, use_graph=True, streams=streams))
        for k, (x, y) in enumerate(zip(eager, graph)):
            assert np.array_equal(x.values, y.values), (n, streams, k)
    prepared.release()
