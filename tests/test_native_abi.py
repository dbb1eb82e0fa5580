"""The C-ABI library and the host-side logic, without a GPU.

- the library builds for gfx950 and exports every entry point include/simulato_hip.h declares;
- with no device (or SK_HIP_HIDE_DEVICE) every entry point refuses with SK_ERR_UNAVAILABLE, never a fallback;
- the host mirror's slicing, scan positions, exit planes and frozen-phonon generator agree with the oracle and the
  reference's fixtures (these are pure host computations, so they are checked here).
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT, golden

import simulatokit_amd as sk
from simulatokit_amd import _native as N


def header_symbols():
    text = ""
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if name.endswith(".h"):
            with open(os.path.join(ROOT, "include", name)) as fh:
                text += fh.read()
    return sorted(set(re.findall(r"\b(sk_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(N.LIB_PATH), "run __graft_entry__.build() first"
    lib = C.CDLL(N.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(N.EXPORTED) == sorted(syms)


def test_library_carries_the_hash_of_its_sources():
    from simulatokit_amd.build import source_hash
    lib = C.CDLL(N.LIB_PATH)
    lib.sk_hip_build_id.restype = C.c_char_p
    assert lib.sk_hip_build_id().decode() == source_hash()


_STALE_SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
import simulatokit_amd.build as B
B.source_hash = lambda: "0" * 64  # sources that differ from the ones the binary was built from
from simulatokit_amd import _native as N
try:
    N.lib()
except RuntimeError as e:
    print("refused:", e)
"""


def test_a_binary_built_from_other_sources_is_refused():
    res = subprocess.run([sys.executable, "-c", _STALE_SCRIPT, ROOT], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert res.stdout.startswith("refused:") and "simulatokit_amd.build" in res.stdout, res.stdout


def test_library_is_built_for_gfx950():
    out = subprocess.run(["bash", "-c", f"strings {N.LIB_PATH} | grep -c gfx950"], capture_output=True, text=True)
    assert int(out.stdout.strip() or 0) > 0


_UNAVAILABLE_SCRIPT = r"""
import ctypes as C, os, sys
sys.path.insert(0, sys.argv[1])
from simulatokit_amd import _native as N
lib = N.lib()
print(lib.sk_hip_available())
print(lib.sk_hip_last_error().decode())
g = N.SkGrid(64, 64, 10.0, 10.0)
V = (C.c_double * (64 * 64))()
dz = (C.c_double * 1)(2.0)
h = C.c_void_p()
print(lib.sk_hip_potential_from_values(C.byref(g), 1, V, dz, 200e3, None, C.byref(h)))
"""


def test_hidden_device_refuses_with_unavailable():
    env = dict(os.environ, SK_HIP_HIDE_DEVICE="1")
    res = subprocess.run([sys.executable, "-c", _UNAVAILABLE_SCRIPT, ROOT], capture_output=True, text=True, env=env,
                         timeout=120)
    lines = res.stdout.strip().splitlines()
    assert res.returncode == 0, res.stderr
    assert lines[0] == "0"
    assert "unavailable" in lines[1]
    assert int(lines[2]) == N.SK_ERR_UNAVAILABLE


def test_python_entry_points_raise_when_the_device_is_hidden():
    env = dict(os.environ, SK_HIP_HIDE_DEVICE="1")
    script = r"""
import sys; sys.path.insert(0, sys.argv[1])
import numpy as np
import simulatokit_amd as sk
assert not sk.hip_is_available()
g = sk.Grid(32, 32, 8.0, 8.0)
pp = sk.ProjectedPotential(g, np.zeros((1, 32, 32)), [2.0])
try:
    sk.Multislice.run(sk.Waves.plane_wave(sk.ElectronBeam(200e3), g), pp)
except sk.SimulatoError as e:
    print(e.kind)
"""
    res = subprocess.run([sys.executable, "-c", script, ROOT], capture_output=True, text=True, env=env, timeout=120)
    assert res.stdout.strip() == "hipUnavailable", res.stdout + res.stderr


# ---- host logic against the oracle and the fixtures ----------------------------------------------------------------
def test_slicing_matches_the_oracle_and_the_fixtures():
    S1 = golden("slice1")
    for cid in ("A1", "A2", "B1", "D1"):
        c = S1.meta["cases"][cid]
        st = sk.AtomicStructure(c["box"], c["numbers"], c["positions"])
        p = sk.Potential(st, *c["gpts"], slice_thickness=c["slice_thickness_requested"])
        assert list(p.slice_of_atom) == c["slice_of_atom"]
        assert np.allclose(p.slice_thicknesses, c["slice_thickness"], rtol=0, atol=1e-15)
        xyz, s = O.slice_atoms(np.array(c["positions"]), c["box"], p.slice_thicknesses)
        assert np.array_equal(p.positions, xyz)
        assert np.array_equal(p.slice_of_atom, s)


def test_wrapping_top_face_and_negative_coordinates():
    st = sk.AtomicStructure((4.0, 4.0, 4.0), [13, 13, 13], [[-1.0, 5.0, 4.0 - 1e-11], [4.0, 0.0, 1.0], [1, 1, 8.5]])
    xyz, s = O.slice_atoms(st.positions, st.box, [2.0, 2.0])
    w = sk.Slicing.wrapped(st)
    assert np.array_equal(w, xyz)
    assert np.array_equal(sk.Slicing.slice_indices(w[:, 2], [2.0, 2.0]), s)


def test_grid_scan_positions_equal_abtem():
    S2 = golden("slice2")
    for cid in ("S1", "S2", "S3"):
        sc = S2.meta["cases"][cid]["scan"]
        g = sk.GridScan(sc["start"], sc["end"], *sc["gpts"])
        assert np.array_equal(g.positions.reshape(-1), S2.array(f"{cid}.scan_positions").reshape(-1))


def test_grid_scan_with_sampling_gives_abtems_points():
    S = golden("slice4scan")
    for s in S.meta["scans"]:
        smp = s["sampling"]
        sx, sy = (smp, smp) if np.isscalar(smp) else smp
        g = sk.GridScan.with_sampling(s["start"], s["end"], sx, sy)
        assert [g.nx, g.ny] == s["gpts"]
    g = sk.GridScan.with_sampling((0, 0), (10, 10), 1, 1, endpoint=True)
    assert (g.nx, g.ny) == (11, 11)
    with pytest.raises(sk.SimulatoError):
        sk.GridScan.with_sampling((2, 0), (1, 1), 1, 1)


def test_exit_planes():
    assert sk.ExitPlanes.last().planes(5) == [4]
    assert sk.ExitPlanes.every(2).planes(5) == [-1, 1, 3, 4]
    assert sk.ExitPlanes.every(7).planes(5) == [4]
    assert sk.ExitPlanes.after_slices([-1, 2]).planes(5) == [-1, 2, 4]
    with pytest.raises(sk.SimulatoError):
        sk.ExitPlanes.after_slices([3, 1]).planes(5)
    with pytest.raises(sk.SimulatoError):
        sk.ExitPlanes.every(0).planes(5)


def test_gaussian_generator_agrees_with_the_oracle():
    g = sk.GaussianGenerator(12345)
    ours = g.normals(1001)
    assert np.array_equal(ours, O.gaussian_normals(12345, 1001))
    for seed in (0, 7, 2**63 + 5):
        for c in (0, 1, 15):
            assert sk.FrozenPhonons.configuration_seed(seed, c) == O.configuration_seed(seed, c)


def test_frozen_phonon_displacements_agree_with_the_oracle():
    st = sk.AtomicStructure((5.0, 5.0, 5.0), [8, 38, 8, 22], np.random.default_rng(0).uniform(0, 5, (4, 3)))
    fp = sk.FrozenPhonons(st, sigmas={8: 0.07, 38: 0.08, 22: 0.06}, configurations=3, seed=99)
    want = O.frozen_displacements(st.numbers, {8: 0.07, 38: 0.08, 22: 0.06}, 3, 99)
    for c in range(3):
        assert np.array_equal(fp.displacements[c], want[c])
    with pytest.raises(sk.SimulatoError):
        sk.FrozenPhonons(st, sigmas={8: 0.1}, configurations=2, seed=1)


def test_refusals_before_work():
    with pytest.raises(sk.SimulatoError) as e:
        sk.ElectronBeam(-1)
    assert e.value.kind == "invalidEnergy"
    with pytest.raises(sk.SimulatoError):
        sk.Grid(1, 8, 1.0, 1.0)
    with pytest.raises(sk.SimulatoError):
        sk.Aperture(0.0)
    with pytest.raises(sk.SimulatoError):
        sk.AnnularDetector(inner=10, outer=5)
    st = sk.AtomicStructure((4.0, 4.0, 4.0), [200], [[0, 0, 0]])
    with pytest.raises(sk.SimulatoError) as e:
        sk.Potential(st, 16, 16, slice_thickness=2.0)
    assert e.value.kind == "unsupportedElement"
    st = sk.AtomicStructure((4.0, 4.0, 4.0), [13], [[0, 0, 0]])
    with pytest.raises(sk.SimulatoError) as e:
        sk.Potential(st, 16, 16, slice_thicknesses=[1.0, 1.0])
    assert e.value.kind == "sliceThicknessesDoNotSumToBox"


def test_an_object_that_is_no_detector_is_refused_before_work():
    g = sk.Grid(32, 32, 8.0, 8.0)
    pp = sk.ProjectedPotential(g, np.zeros((1, 32, 32)), [2.0])
    probe = sk.Probe(sk.ElectronBeam(200e3), g, sk.Aperture(20.0))
    with pytest.raises(sk.SimulatoError) as e:
        sk.STEM.scan(probe, [pp], np.zeros((1, 2)), [object()])
    assert e.value.kind == "invalidDetector"


def test_diffraction_intensity_angles_follow_the_detector_rules():
    from simulatokit_amd.detector import DiffractionIntensity, detector_shape
    beam = sk.ElectronBeam(200e3)
    for g in (sk.Grid(512, 512, 31.24, 31.24), sk.Grid(90, 75, 12.0, 10.0), sk.Grid(128, 96, 7.81, 5.2)):
        inten = DiffractionIntensity(g, beam, np.zeros(g.count))
        ax, ay = inten.angular_sampling
        assert ax == g.reciprocal_sampling_x * beam.wavelength * 1e3
        # the cutoff angles are floor(m / 2) pixels of the cutoff-sized pattern the library crops to
        mx, my = detector_shape(sk.PixelatedDetector("cutoff"), g, beam.energy)
        assert inten.cutoff_angles == ((mx // 2) * ax, (my // 2) * ay)
        assert np.array_equal(DiffractionIntensity.angles(g.nx, ax), sk.core.fftfreq(g.nx, 1 / ax / g.nx))
    with pytest.raises(ValueError):
        DiffractionIntensity(sk.Grid(8, 8, 1.0, 1.0), beam, np.zeros(63))
