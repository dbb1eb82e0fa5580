"""The split-BF16 matrix-core transforms (ExecutionOptions(precision="split_bf16")) against the float64 oracle.

The slice loop's line transforms run as two dense DFT stages on v_mfma_f32_32x32x16_bf16 (bfloat16 hi/lo pairs,
three products, float32 accumulation). The precision study of the HAADF geometry reports their error next to the
float32 transforms'; the bounds here are the study's (the matrix-core path carries about 2^-16 relative per product,
where float32 carries 2^-24).
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


def test_split_bf16_plane_wave_exit_equals_the_oracle():
    st = srtio3(4, 4, 10)
    pot = sk.Potential(st, 256, 256, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    prepared = pot.prepare(beam)
    wave = sk.Waves.plane_wave(beam, pot.grid)
    got = sk.Multislice.run(wave, prepared, options=sk.ExecutionOptions(precision="split_bf16"))[0]
    f32 = sk.Multislice.run(wave, prepared)[0]
    V, _ = O.potential(256, 256, st.numbers, st.positions, st.box, pot.slice_thicknesses)
    want = O.multislice(V, pot.slice_thicknesses, (st.box[0], st.box[1]), 200e3, np.ones((256, 256), complex))[0]
    e_bf, e_32 = rel_max(got.field, want), rel_max(f32.field, want)
    print(f"exit wave rel max: split-BF16 {e_bf:.3e}, float32 {e_32:.3e}")
    assert e_bf <= 1e-3
    assert e_32 <= 1e-4


@pytest.mark.parametrize("n,tz", [(256, 20), (512, 10)])
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
