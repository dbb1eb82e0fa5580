"""A caller's own detector on the `.hip` backend: STEM.scan hands every position's diffraction intensity to any object
with shape(grid, beam) and detect(intensity), as the reference's `.metal` scan does for its open `Detector` protocol
(Detectors.swift:84-88; STEM.metalScan, STEM.swift:120-150; MetalMultislice.scanIntensities, MetalEngine.swift:
583-676), through sk_hip_scan_intensities.

The HIP path is held against the float64 oracle: the oracle's probe through the oracle's potential, its
|FFT2(psi_exit)|^2 (orc_diffraction_intensity), reduced by the same Python detector, averaged over the same
frozen-phonon configurations, within the registered detector bound 2e-4 x the largest value.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_max

import simulatokit_amd as sk
from simulatokit_amd.detector import DiffractionIntensity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


class RadialProfile:
    """The intensity summed over rings `width` mrad wide, from the centre out to `rings` rings."""

    def __init__(self, rings=8, width=12.5):
        self.rings, self.width = rings, width

    def shape(self, grid, beam):
        return (self.rings,)

    def detect(self, inten):
        g = inten.grid
        ax, ay = inten.angular_sampling
        a = np.sqrt(DiffractionIntensity.angles(g.nx, ax)[:, None] ** 2 +
                    DiffractionIntensity.angles(g.ny, ay)[None, :] ** 2)
        ring = np.floor(a / self.width).astype(np.int64).reshape(-1)
        keep = ring < self.rings
        return np.bincount(ring[keep], weights=inten.values[keep], minlength=self.rings)


class PythonAnnular:
    """The annular sum inner^2 <= alpha^2 < outer^2 written as a caller's detector."""

    def __init__(self, inner, outer):
        self.inner, self.outer = inner, outer

    def shape(self, grid, beam):
        return ()

    def detect(self, inten):
        g = inten.grid
        ax, ay = inten.angular_sampling
        a2 = (DiffractionIntensity.angles(g.nx, ax) ** 2)[:, None] + (DiffractionIntensity.angles(g.ny, ay) ** 2)[None, :]
        m = (a2 >= self.inner ** 2) & (a2 < self.outer ** 2)
        return [float(inten.pattern[m].sum())]


def perovskite(cells=(2, 2, 2)):
    a = 3.905
    cell = sk.AtomicStructure((a, a, a), [38, 22, 8, 8, 8],
                              [[0, 0, 0], [a / 2, a / 2, a / 2], [a / 2, a / 2, 0], [a / 2, 0, a / 2], [0, a / 2, a / 2]])
    return cell.repeated(*cells)


def oracle_custom(structures, n, dz, positions, detector, energy=200e3, semiangle=20.0):
    """The oracle's scan with a host detector: per configuration V, then per position the probe, multislice and
    |FFT2 psi|^2, summed in configuration order and divided by the count."""
    grid = sk.Grid(n, n, structures[0].box[0], structures[0].box[1])
    beam = sk.ElectronBeam(energy)
    extent = (grid.extent_x, grid.extent_y)
    sums = None
    for st in structures:
        thick = [st.box[2] / math.ceil(st.box[2] / dz)] * math.ceil(st.box[2] / dz)
        V, _ = O.potential(n, n, st.numbers, st.positions, st.box, thick)
        probes = [O.probe(n, n, extent, energy, semiangle, p) for p in positions]
        exits = O.multislice_batch(V, thick, extent, energy, np.array(probes))
        vals = np.array([np.asarray(detector.detect(DiffractionIntensity(grid, beam, O.diffraction_intensity(e))),
                                    dtype=np.float64).reshape(-1) for e in exits])
        sums = vals if sums is None else sums + vals
    return sums / len(structures)


def test_radial_profile_over_frozen_phonons_equals_the_oracle():
    st = perovskite()
    fp = sk.FrozenPhonons(st, sigmas={38: 0.08, 22: 0.08, 8: 0.08}, configurations=2, seed=7)
    confs = [fp.configuration(c) for c in range(fp.count)]
    pots = [sk.Potential(c, 128, 128, slice_thickness=2.0) for c in confs]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    positions = np.array([[0.0, 0.0], [1.2, 3.1], [3.905, 1.9525], [6.5, 7.0]])
    det = RadialProfile()
    got = sk.STEM.scan(probe, pots, positions, [det])[0]
    assert got.value_shape == (8,) and got.values.shape == (4 * 8,)
    want = oracle_custom(confs, 128, 2.0, positions, det)
    err = rel_max(got.values.reshape(4, 8), want)
    print(f"radial profile over 2 configurations, 4 positions: rel max vs float64 oracle {err:.2e}", flush=True)
    assert err <= 2e-4
    # every probe carries unit intensity; the rings out to 100 mrad hold most of it
    assert np.all(want.sum(axis=1) <= 1.0 + 1e-9) and np.all(got.values.reshape(4, 8).sum(axis=1) > 0.5)


@pytest.mark.parametrize("n", [256, 384, 509, 512, 1024])
def test_radial_profile_through_the_register_and_fast_passes_equals_the_oracle(n):
    # the intensity scan's slice loop runs the fast passes (256, 512) or the register passes (384, 1024, and 509 by
    # Bluestein on 1024 points) for its first and middle column passes, and the library's other passes for the last
    st = perovskite((2, 2, 1))
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = np.array([[0.0, 0.0], [1.2, 3.1]])
    det = RadialProfile()
    got = sk.STEM.scan(probe, [pot], positions, [det])[0]
    want = oracle_custom([st], n, 2.0, positions, det)
    err = rel_max(got.values.reshape(2, 8), want)
    print(f"{n}x{n} radial profile, 2 positions: rel max vs float64 oracle {err:.2e}", flush=True)
    assert err <= 2e-4


def test_a_python_annular_detector_equals_the_builtin_one():
    st = perovskite()
    pot = sk.Potential(st, 128, 128, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    scan = sk.GridScan((0, 0), (3.905, 3.905), 3, 3)
    builtin = sk.AnnularDetector(50, 120)
    mixed = sk.STEM.scan(probe, [pot], scan, [PythonAnnular(50, 120), builtin, RadialProfile(4, 25.0)])
    alone = sk.STEM.scan(probe, [pot], scan, [builtin])[0]
    assert [m.value_shape for m in mixed] == [(), (), (4,)]
    assert mixed[0].scan_shape == (3, 3)
    # the built-in detector's numbers do not depend on the other detectors in the list
    assert np.array_equal(mixed[1].values, alone.values)
    # the same pixels summed on the host from the delivered intensities, in another order
    assert rel_max(mixed[0].values, alone.values) <= 1e-9
    assert mixed[0].image().shape == (3, 3)


def test_a_failing_detector_stops_the_scan_and_the_library_stays_usable():
    st = perovskite((1, 1, 2))
    pot = sk.Potential(st, 64, 64, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))

    class Broken:
        def shape(self, grid, beam):
            return ()

        def detect(self, inten):
            raise ValueError("broken detector")

    class WrongCount(Broken):
        def detect(self, inten):
            return [1.0, 2.0]

    with pytest.raises(ValueError, match="broken detector"):
        sk.STEM.scan(probe, [pot], np.zeros((3, 2)), [Broken()])
    with pytest.raises(sk.SimulatoError, match="wrong number of values"):
        sk.STEM.scan(probe, [pot], np.zeros((3, 2)), [WrongCount()])
    ok = sk.STEM.scan(probe, [pot], np.zeros((3, 2)), [PythonAnnular(0, 20)])[0]
    assert np.all(np.isfinite(ok.values)) and np.all(ok.values > 0)


def test_intensities_of_many_batches_arrive_in_position_order():
    # 300 positions at 64^2 in batches of 64 (five deliveries, the last partial) against one position per batch
    st = perovskite((1, 1, 1))
    pot = sk.Potential(st, 64, 64, slice_thickness=1.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    positions = np.random.default_rng(3).uniform(0, 3.905, (300, 2))

    class Total:
        def shape(self, grid, beam):
            return ()

        def detect(self, inten):
            return [inten.values.sum()]

    total = sk.STEM.scan(probe, [pot], positions, [Total()], sk.ExecutionOptions(batch=64))[0]
    one = sk.STEM.scan(probe, [pot], positions[::37], [Total()], sk.ExecutionOptions(batch=1))[0]
    assert rel_max(total.values[::37], one.values) <= 1e-12
    # a unit probe loses only what the thin cell scatters past the bandlimit
    assert np.all(total.values <= 1.0 + 1e-5) and np.all(total.values > 0.9)
