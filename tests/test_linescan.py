"""LineScan against abTEM 1.0.10's fixtures (LineScanTests.swift: construction, points, sampling, endpoint,
add_margin, at_position, fractional), and the oracle's STEM scan over the fixture's two lines (Al, 64^2, annular
50-150 mrad) against abTEM's values."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden, scaled_max

import simulatokit_amd as sk

LS = golden("linescan")


def make(c):
    a = c["args"]
    gpts = a.get("gpts")
    endpoint = a.get("endpoint", True)
    if a.get("atPosition"):
        return sk.LineScan.at_position(a["center"], a.get("extent", 1.0), a.get("angle", 0.0), gpts=gpts,
                                       sampling=a.get("sampling"), endpoint=endpoint)
    scan = sk.LineScan(a.get("start", (0.0, 0.0)), a["end"], gpts=gpts, sampling=a.get("sampling"), endpoint=endpoint,
                       fractional=a.get("fractional", False), extent=a.get("extent2"))
    if "margin" in a:
        scan = scan.adding_margin(*a["margin"])
    return scan


@pytest.mark.parametrize("c", LS.meta["lineScans"], ids=lambda c: c["label"])
def test_line_scan_construction_equals_abtem(c):
    scan = make(c)
    assert scan.gpts == c["gpts"]
    assert scan.endpoint == c["endpoint"]
    np.testing.assert_allclose(list(scan.start) + list(scan.end), c["start"] + c["end"], rtol=1e-14)
    assert scan.sampling == c["sampling"]
    assert scan.extent == c["extent"]
    if scan.extent > 0:
        assert scan.angle == c["angle"]
    want = LS.array(f"line.{c['label']}.positions").reshape(-1, 2)
    np.testing.assert_allclose(scan.positions, want, rtol=1e-7, atol=1e-12 * np.abs(want).max())
    assert scan.shape == (scan.gpts,)


def test_line_scan_refusals():
    with pytest.raises(sk.SimulatoError):
        sk.LineScan((0, 0), (1, 1))  # neither gpts nor sampling
    with pytest.raises(sk.SimulatoError):
        sk.LineScan((0, 0), (1, 1), gpts=0)
    with pytest.raises(sk.SimulatoError):
        sk.LineScan((0, 0), (1, 1), sampling=-1.0)
    with pytest.raises(sk.SimulatoError):
        sk.LineScan((0, 0), (0, 0), gpts=3).adding_margin(0.5)
    with pytest.raises(sk.SimulatoError):
        sk.LineScan((0, 0), (1, 1), gpts=3, fractional=True)


def line_stem(spec):
    m = LS.meta
    numbers = [a[0] for a in m["atoms"]]
    xyz = np.array([a[1] for a in m["atoms"]])
    n = m["gpts"]
    th = [m["sliceThickness"]] * int(round(m["box"][2] / m["sliceThickness"]))
    V, _ = O.potential(n, n, numbers, xyz, m["box"], th, pi_mode=1)
    scan = sk.LineScan(spec["start"], spec["end"], gpts=spec["gpts"], endpoint=spec.get("endpoint", True))
    return V, th, scan


@pytest.mark.parametrize("spec", LS.meta["scans"], ids=lambda s: s["label"])
def test_oracle_stem_over_a_line_equals_abtem(spec):
    m = LS.meta
    V, th, scan = line_stem(spec)
    np.testing.assert_allclose(scan.positions, LS.array(f"scan.{spec['label']}.positions").reshape(-1, 2), rtol=1e-12)
    got = O.stem_scan([V], [th], (m["box"][0], m["box"][1]), m["energy"], m["semiangle"], scan.positions,
                      [O.det_annular(50, 150)])[0]
    want = LS.array(f"scan.{spec['label']}.adf").reshape(-1)
    assert scaled_max(got.reshape(-1), want, np.abs(want).max()) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("spec", LS.meta["scans"], ids=lambda s: s["label"])
def test_hip_stem_over_a_line_equals_abtem(spec):
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())
    m = LS.meta
    st = sk.AtomicStructure(m["box"], [a[0] for a in m["atoms"]], [a[1] for a in m["atoms"]])
    pot = sk.Potential(st, m["gpts"], m["gpts"], slice_thickness=m["sliceThickness"],
                       parameterization=sk.LobatoParameterization("abtem_float32"))
    probe = sk.Probe(sk.ElectronBeam(m["energy"]), pot.grid, sk.Aperture(m["semiangle"]))
    scan = sk.LineScan(spec["start"], spec["end"], gpts=spec["gpts"], endpoint=spec.get("endpoint", True))
    got = sk.STEM.scan(probe, [pot], scan, [sk.AnnularDetector(50, 150)])[0]
    assert got.scan_shape == (spec["gpts"],)
    want = LS.array(f"scan.{spec['label']}.adf").reshape(-1)
    assert np.abs(got.values - want).max() <= sk.REGISTERED.detector_bound(np.abs(want).max())
