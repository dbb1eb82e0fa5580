"""WavesDetector / ScanWaves against abTEM 1.0.10 (WavesDetectorTests.swift): the complex exit waves of a 3 x 2 scan
through aluminium on the simulation grid and Fourier-resampled to five other grids. The oracle at assert_allclose's
float64 defaults (rtol 1e-7, atol 1e-12 of the maximum); the HIP path at the registered float32 wave bound."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

import simulatokit_amd as sk

W = golden("waves")


def setup_case():
    m = W.meta
    st = sk.AtomicStructure(m["box"], [a[0] for a in m["atoms"]], [a[1] for a in m["atoms"]])
    return m, st


def scan_positions(m):
    s = m["scan"]
    return sk.GridScan(s["start"], s["end"], *s["gpts"])


def allclose_ratio(got, want):
    atol = 1e-12 * np.abs(want).max()
    r = np.abs(got.real - want.real) / (atol + 1e-7 * np.abs(want.real))
    i = np.abs(got.imag - want.imag) / (atol + 1e-7 * np.abs(want.imag))
    return float(max(r.max(), i.max()))


@pytest.mark.parametrize("run", W.meta["runs"], ids=lambda r: r["name"])
def test_oracle_exit_waves_equal_abtem(run):
    m, st = setup_case()
    n = m["gpts"]
    th = [m["sliceThickness"]] * int(round(m["box"][2] / m["sliceThickness"]))
    V, _ = O.potential(n, n, st.numbers, st.positions, m["box"], th, pi_mode=1)
    scan = scan_positions(m)
    got = O.scan_waves(V, th, (m["box"][0], m["box"][1]), m["energy"], m["semiangle"], scan.positions, run["gpts"])
    want = W.array(run["name"]).reshape(got.shape)
    assert list(scan.shape) + list(got.shape[1:]) == run["shape"]
    assert allclose_ratio(got, want) <= 1.0


def test_resample_map_pads_and_crops_like_abtem():
    assert O.resample_map(4, 3) == [(0, 0), (1, 1), (3, 2)]
    assert O.resample_map(3, 5) == [(0, 0), (1, 1), (2, 4)]
    assert O.resample_map(4, 6) == [(0, 0), (1, 1), (2, 4), (3, 5)]
    w = np.random.default_rng(0).normal(size=(12, 10)) + 0j
    assert np.allclose(O.resample_wave(w, 12, 10), w)  # same grid: identity


@pytest.mark.gpu
@pytest.mark.parametrize("run", W.meta["runs"], ids=lambda r: r["name"])
def test_hip_exit_waves_equal_abtem(run):
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())
    m, st = setup_case()
    n = m["gpts"]
    pot = sk.Potential(st, n, n, slice_thickness=m["sliceThickness"],
                       parameterization=sk.LobatoParameterization("abtem_float32"))
    probe = sk.Probe(sk.ElectronBeam(m["energy"]), pot.grid, sk.Aperture(m["semiangle"]))
    det = sk.WavesDetector() if run["gpts"] is None else sk.WavesDetector(*run["gpts"])
    out = sk.STEM.scan_waves(probe, pot, scan_positions(m), det)
    assert list(out.scan_shape) + [out.rows, out.columns] == run["shape"]
    want = W.array(run["name"]).reshape(out.values.shape)
    err = np.abs(out.values - want).max() / np.abs(want).max()
    print(run["name"], "rel max", err)
    assert err <= sk.REGISTERED.wave
    w = out[2, 1]
    assert w.grid.nx == out.rows and w.grid.ny == out.columns
