"""The frozen-phonon 4D-STEM workload at its own size, on an MI355X.

The workload: a seeded 40 x 40 A master cell 120 A deep with 2,000 atoms (Z from {8, 14, 22, 38, 79}, uniform
positions), displaced per configuration as FrozenPhonons(seed:) does (sigma 0.08 A on every axis), and cut per
configuration to its own thickness T_c uniform in 40-120 A, so the 16 configurations have 42-115 slices of 1 A; a
512^2 grid, a 200 keV probe with a 20 mrad soft aperture, a 128 x 128 scan over the cell, a 128 x 128 float32 crop of
the pattern and the centre of mass, averaged over the configurations in potential order (STEM.swift:120-150,
FrozenPhonons.swift:61-137).

- Parity: all 16 configurations at two scan positions, through the library's default plan (the lane pipeline at two
  positions), the lanes with a small lane, and one configuration at a time, against the float64 oracle's scan
  through its own potentials (the registered bounds: 2e-4 of the largest pattern value; detectors 2e-4 of the
  largest value).
- At full size (16 configurations x 16,384 positions, 20.3 M position-slices, outputs left in device memory): the
  lane pipeline and one configuration at a time agree (float64 centre of mass to 1e-12; the float32 patterns to one
  rounding), every pattern is finite and non-negative, and no pattern holds more than the unit intensity of the
  incident probe.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, rel_max

import simulatokit_amd as sk
from simulatokit_amd.workloads import uneven_frozen_phonon_stack

pytestmark = pytest.mark.gpu

ACC = sk.REGISTERED
N = 512
EXTENT = 40.0


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


@pytest.fixture(scope="module")
def workload():
    structures, T = uneven_frozen_phonon_stack()  # 40 x 40 A, 2,000 atoms, 16 configurations, 40-120 A
    pots = [sk.Potential(st, N, N, slice_thickness=1.0) for st in structures]
    counts = [p.slice_count for p in pots]
    assert len(structures) == 16 and min(counts) <= 45 and max(counts) >= 110, counts
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    prepared = [p.prepare(probe.beam) for p in pots]
    scan = sk.GridScan((0, 0), (EXTENT, EXTENT), 128, 128)
    dets = [sk.PixelatedDetector(crop=(128, 128), dtype="float32"), sk.CenterOfMassDetector()]
    print(f"config 3: slices per configuration {counts} ({sum(counts)} in all)", flush=True)
    return structures, pots, prepared, probe, scan, dets


def centred_crop(full_shifted, m):
    """The m x m centred crop of an fftshifted even-sized pattern (Waves.cropIndices, Waves.swift:90-96)."""
    n = full_shifted.shape[-1]
    lo = n // 2 - m // 2
    return full_shifted[..., lo:lo + m, lo:lo + m]


def test_config3_at_two_positions_equals_the_oracle(workload):
    structures, pots, prepared, probe, scan, dets = workload
    picks = [0, 77 * 128 + 51]  # the scan's first position and one inside the cell
    positions = scan.positions[picks]
    t0 = time.perf_counter()

    def oracle_one(c):
        st, p = structures[c], pots[c]
        V, _ = O.potential(N, N, st.numbers, st.positions, st.box, p.slice_thicknesses)
        full, com = O.stem_scan([V], [p.slice_thicknesses], (EXTENT, EXTENT), 200e3, 20.0, positions,
                                [O.det_pixelated("full"), O.det_com()], threads=2)
        return full, com

    with ThreadPoolExecutor(max_workers=8) as pool:
        parts = list(pool.map(oracle_one, range(len(structures))))
    full = sum(p[0] for p in parts) / len(parts)  # in potential order, as the reference sums
    com = sum(p[1] for p in parts) / len(parts)
    print(f"oracle: 16 configurations at 2 positions in {time.perf_counter() - t0:.1f} s", flush=True)
    want = centred_crop(full.reshape(2, N, N), 128).reshape(2, -1)
    want_com = com.reshape(2, 2)
    plans = {"default": sk.ExecutionOptions(), "lanes of 8": sk.ExecutionOptions(pipeline=1, lane_waves=8),
             "one configuration at a time": sk.ExecutionOptions(pipeline=0)}
    results = {}
    for name, opt in plans.items():
        got = sk.STEM.scan(probe, prepared, positions, dets, opt)
        pat = got[0].values.reshape(2, -1).astype(np.float64)
        e_max, e_l2 = rel_max(pat, want), rel_l2(pat, want)
        e_com = np.abs(got[1].values.reshape(2, 2) - want_com).max() / np.abs(want_com).max()
        print(f"{name}: pattern rel max {e_max:.2e}, rel L2 {e_l2:.2e}; COM rel max {e_com:.2e}", flush=True)
        assert got[0].values.dtype == np.float32 and got[0].value_shape == (128, 128)
        assert e_max <= ACC.pattern
        assert e_l2 <= 1e-4
        assert e_com <= ACC.detector
        results[name] = got
    # every plan computes the same float32 waves and sums them in potential order
    for name in ("lanes of 8", "one configuration at a time"):
        assert rel_max(results[name][1].values, results["default"][1].values) <= 1e-12


def test_config3_at_full_size_lanes_equal_one_configuration_at_a_time(workload):
    import torch

    structures, pots, prepared, probe, scan, dets = workload
    npos = scan.positions.shape[0]
    assert npos == 16384
    slices = sum(p.slice_count for p in pots)
    out = {}
    for name, opt in (("lanes", sk.ExecutionOptions(pipeline=1)), ("one at a time", sk.ExecutionOptions(pipeline=0))):
        pat = torch.empty((npos, 128, 128), dtype=torch.float32, device="cuda:0")
        com = torch.empty((npos, 2), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sk.STEM.scan(probe, prepared, scan, dets, opt, device_outputs=[pat.data_ptr(), com.data_ptr()])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(f"{name}: {npos} positions x {slices} slices in {dt:.1f} s, "
              f"{npos * slices / dt:,.0f} position-slices/s", flush=True)
        out[name] = (pat, com)
    (pa, ca), (pb, cb) = out["lanes"], out["one at a time"]
    assert float((ca - cb).abs().max() / cb.abs().max()) <= 1e-12
    assert float((pa.double() - pb.double()).abs().max() / pb.abs().max()) <= 1e-6
    for pat in (pa, pb):
        assert bool(torch.isfinite(pat).all()) and float(pat.min()) >= 0.0
        sums = pat.double().sum(dim=(1, 2))
        # a unit probe: the crop holds part of the unit intensity that stays inside the bandlimit
        assert float(sums.max()) <= 1.0 + 1e-6 and float(sums.min()) > 0.0
    assert bool(torch.isfinite(ca).all())
