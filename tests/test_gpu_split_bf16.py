"""The split-BF16 matrix-core transforms (ExecutionOptions(precision="split_bf16")) against the float64 oracle.

The slice loop's line transforms run as two dense DFT stages on v_mfma_f32_32x32x16_bf16 (bfloat16 hi/lo pairs,
three products, float32 accumulation), for transform lengths 16 n2 with n2 = 16, 24, 32, 64: grid lengths 256, 384,
512, 1024, and Bluestein lengths whose 2 n - 1 point convolution fits 256, 512 or 1024 points (509 among them).
The precision study (the HAADF geometry at 4096 positions per launch, grid lengths 256, 384 and 509) reports their
error next to the float32 transforms'; the bounds here are the study's (the matrix-core path carries about 2^-16
relative per product, where float32 carries 2^-24).
"""
import time

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


@pytest.mark.parametrize("n,cells,tz", [(256, 4, 10), (384, 4, 10), (509, 4, 10), (1024, 2, 4)])
def test_split_bf16_plane_wave_exit_equals_the_oracle(n, cells, tz):
    # 256, 384 and 1024: direct transforms (n2 = 16, 24, 64); 509: Bluestein through two 1024-point transforms
    st = srtio3(cells, cells, tz)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    prepared = pot.prepare(beam)
    wave = sk.Waves.plane_wave(beam, pot.grid)
    got = sk.Multislice.run(wave, prepared, options=sk.ExecutionOptions(precision="split_bf16"))[0]
    f32 = sk.Multislice.run(wave, prepared)[0]
    V, _ = O.potential(n, n, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    want = O.multislice(V, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((n, n), complex))[0]
    e_bf, e_32 = rel_max(got.field, want), rel_max(f32.field, want)
    print(f"n={n}: exit wave rel max: split-BF16 {e_bf:.3e}, float32 {e_32:.3e}")
    assert e_bf <= 1e-3
    assert e_32 <= 1e-4


@pytest.mark.parametrize("n,tz,outer", [(256, 20, 100.0), (384, 20, 150.0), (509, 10, 200.0)])
def test_precision_study_at_4096_positions(n, tz, outer):
    """The precision study: SrTiO3 8 x 8 cells, 2 A slices (40 slices; 20 at 509, which keeps the float64 oracle
    near 20 s), 200 keV probe with a 20 mrad aperture, a 64 x 64 scan over one cell as ONE launch of 4096
    positions, annular 50 mrad to `outer` (200 mrad, or less where the grid's largest simulated angle, about
    103 mrad at 256 and 154 mrad at 384, is smaller: the detector refuses a larger outer angle). Float32 and split-BF16 transforms; the annular values and cutoff patterns of a
    sample of positions against the float64 oracle, and the sample's annular values of the 4096-position launch
    against a launch of the sample alone."""
    st = srtio3(8, 8, tz)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    assert pot.slice_count == 2 * tz
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 64, 64).positions
    assert len(positions) == 4096
    sample = np.array([0, 1387, 2730, 4095])
    prepared = pot.prepare(probe.beam)
    V, _ = O.potential(n, n, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    haadf, pat = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 20.0, positions[sample],
                             [O.det_annular(50, outer), O.det_pixelated("cutoff")])
    rows = []
    for name, prec in (("float32", "float32"), ("split-BF16", "split_bf16")):
        opt = sk.ExecutionOptions(batch=4096, tile=4096, precision=prec)
        t0 = time.perf_counter()
        full = sk.STEM.scan(probe, [prepared], positions, [sk.AnnularDetector(50, outer)], opt)[0].values
        wall = time.perf_counter() - t0
        part = sk.STEM.scan(probe, [prepared], positions[sample],
                            [sk.AnnularDetector(50, outer), sk.PixelatedDetector("cutoff")],
                            sk.ExecutionOptions(batch=len(sample), precision=prec))
        assert full.shape == (4096,)
        assert rel_max(full[sample], part[0].values) <= 1e-12
        e_h, e_p = rel_l2(full[sample], haadf.reshape(-1)), rel_l2(part[1].values, pat.reshape(-1))
        rows.append((e_h, e_p))
        print(f"n={n} {name}: HAADF rel L2 {e_h:.3e}, pattern rel L2 {e_p:.3e}; 4096 positions x "
              f"{pot.slice_count} slices in {wall:.3f} s ({4096 * pot.slice_count / wall:,.0f} position-slices/s, "
              f"wall clock of STEM.scan)")
    assert rows[0][0] <= 1e-4 and rows[0][1] <= 1e-4
    assert rows[1][0] <= 1e-3 and rows[1][1] <= 1e-3


@pytest.mark.parametrize("n,tz", [(512, 10)])
def test_split_bf16_precision_study(n, tz):
    st = srtio3(8, 8, tz)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = sk.GridScan((0, 0), (SRTIO3_A, SRTIO3_A), 2, 2).positions
    dets = [sk.AnnularDetector(30, 100), sk.PixelatedDetector("cutoff")]
    prepared = pot.prepare(probe.beam)
    bf = sk.STEM.scan(probe, [prepared], positions, dets, sk.ExecutionOptions(precision="split_bf16"))
    f32 = sk.STEM.scan(probe, [prepared], positions, dets)
    V, _ = O.potential(n, n, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    haadf, pat = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), 200e3, 20.0, positions,
                             [O.det_annular(30, 100), O.det_pixelated("cutoff")])
    rows = []
    for name, m in (("split-BF16", bf), ("float32", f32)):
        rows.append((name, rel_l2(m[0].values, haadf.reshape(-1)), rel_l2(m[1].values, pat.reshape(-1))))
        print(f"n={n} {name}: annular rel L2 {rows[-1][1]:.3e}, pattern rel L2 {rows[-1][2]:.3e}")
    assert rows[0][1] <= 1e-3 and rows[0][2] <= 1e-3
    assert rows[1][1] <= 1e-4 and rows[1][2] <= 1e-4


def test_split_bf16_refuses_other_grid_lengths():
    st = srtio3(2, 2, 2)
    pot = sk.Potential(st, 96, 96, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    with pytest.raises(sk.SimulatoError):
        sk.Multislice.run(sk.Waves.plane_wave(beam, pot.grid), pot,
                          options=sk.ExecutionOptions(precision="split_bf16"))
