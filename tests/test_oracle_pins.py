"""The float64 oracle against the reference's abTEM fixtures, at the reference's own bounds.

Mirrors Tests/SimulatoKitParityTests/Multislice/{BeamAndScatteringTests,PotentialPinsTests,MultislicePinsTests,
GridSamplingAndTiltTests}.swift and Detector/{STEMPinsTests,FrozenPhononsTests,ScanPatternsAndNoiseTests}.swift.
The fixtures are abTEM 1.0.10 float64 (tests/golden/make_golden.py copies them); nothing here runs the reference.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden, rel_diff, rel_max, scaled_max

S1 = golden("slice1")
PI = {"abtem": 1, "exact": 0}


def n_slices(depth, requested):
    """Slicing.thicknesses (Slicing.swift:17-25)."""
    n = math.ceil(depth / requested)
    return [depth / n] * n


def case_potential(meta, cid, pi):
    c = meta["cases"][cid]
    box = c["box"]
    nx, ny = c["gpts"]
    th = n_slices(box[2], c["slice_thickness_requested"])
    numbers = np.array(c["numbers"], dtype=np.int32)
    xyz = np.array(c["positions"], dtype=np.float64).reshape(-1, 3)
    V, s = O.potential(nx, ny, numbers, xyz, box, th, pi_mode=PI[pi])
    return V, th, s, (box[0], box[1])


# ---- BeamAndScatteringTests ---------------------------------------------------------------------------------
def test_wavelength_and_sigma_equal_abtem_to_the_bit():
    for b in S1.meta["beam"]:
        assert O.wavelength(b["energy"]) == b["wavelength"]
        assert O.sigma(b["energy"]) == b["sigma"]


def test_kirkland_wavelength_and_sigma_at_200kev():
    # BeamAndScatteringTests.swift:29-34 (textbook values, Kirkland 2010)
    assert abs(O.wavelength(200e3) - 0.02508) < 0.5e-5
    assert abs(O.sigma(200e3) - 7.288e-4) < 0.5e-7


@pytest.mark.parametrize("pi", ["abtem", "exact"])
def test_scattering_factor_equals_abtem(pi):
    z_of = {"H": 1, "C": 6, "O": 8, "Al": 13, "Si": 14, "Ti": 22, "Sr": 38, "Au": 79, "U": 92}
    k2 = np.array(S1.meta["f_k2"])
    for sym, entry in S1.meta["f"].items():
        want = np.array(entry[pi])
        got = O.lobato_f(z_of[sym], k2, pi_mode=PI[pi])
        assert rel_max(got, want) <= 1e-14, sym


# ---- PotentialPinsTests -------------------------------------------------------------------------------------
@pytest.mark.parametrize("cid", ["A1", "A2", "B1", "C1", "D1"])
def test_slicing_equals_abtem(cid):
    c = S1.meta["cases"][cid]
    th = n_slices(c["box"][2], c["slice_thickness_requested"])
    assert np.allclose(th, c["slice_thickness"], rtol=0, atol=1e-15)
    if c["numbers"]:
        _, s = O.slice_atoms(np.array(c["positions"]), c["box"], th)
        assert list(s) == c["slice_of_atom"]


@pytest.mark.parametrize("cid", ["A1", "A2", "B1", "D1"])
def test_potential_and_transmission_equal_abtem(cid):
    c = S1.meta["cases"][cid]
    V, th, _, extent = case_potential(S1.meta, cid, "abtem")
    for j in range(c["V_stored_slices"]):
        assert rel_max(V[j], S1.array(f"{cid}.V.{j}")) <= 1e-11, f"{cid} V slice {j}"
    for j in range(len(th)):
        assert rel_diff(V[j].sum(), c["V_sum"][j]) <= 1e-11
        assert rel_diff(np.abs(V[j]).max(), c["V_max_abs"][j]) <= 1e-11
    t = O.transmission(V[:1], extent, c["energy"])
    assert rel_max(t[0], S1.array(f"{cid}.t.0")) <= 1e-11


@pytest.mark.parametrize("cid", ["A1", "A2", "B1", "D1"])
def test_k0_identity_sum_v_dxdy_equals_sum_f0(cid):
    # PotentialPinsTests.swift:60-77: sum(V) dx dy = sum over atoms of f(0)
    c = S1.meta["cases"][cid]
    V, th, s, extent = case_potential(S1.meta, cid, "exact")
    nx, ny = c["gpts"]
    dxdy = (extent[0] / nx) * (extent[1] / ny)
    for j in range(len(th)):
        zs = [z for z, sl in zip(c["numbers"], s) if sl == j]
        want = sum(O.lobato_f(z, np.array([0.0]))[0] for z in zs)
        assert rel_diff(V[j].sum() * dxdy, want) <= 1e-10


# ---- MultislicePinsTests --------------------------------------------------------------------------------------
CASE_BEAMS = [("A1", "plane"), ("A2", "plane"), ("B1", "plane"), ("B2", "probe"), ("C1", "plane"),
              ("C1", "probe"), ("D1", "plane"), ("D1", "probe")]


def incident(meta, cid, label, extent, gpts):
    b = meta["cases"][cid]["beams"][label]
    nx, ny = gpts
    if b["kind"] == "plane":
        return np.ones((nx, ny), dtype=np.complex128), True
    return O.probe(nx, ny, extent, meta["cases"][cid]["energy"], b["cutoff_mrad"], b["position"],
                   aberrations=b["aberrations"]), False


@pytest.mark.parametrize("cid,label", [cb for cb in CASE_BEAMS if cb[1] == "probe"])
def test_incident_probe_equals_abtem(cid, label):
    pc = "B1" if cid == "B2" else cid
    c = S1.meta["cases"][pc]
    psi, _ = incident(S1.meta, cid, label, (c["box"][0], c["box"][1]), c["gpts"])
    assert rel_max(psi, S1.array(f"{cid}/{label}.incident")) <= 1e-12


@pytest.mark.parametrize("pi,bound", [("abtem", 1e-10), ("exact", 1e-6)])
@pytest.mark.parametrize("cid,label", CASE_BEAMS)
def test_exit_waves_patterns_and_scalars_equal_abtem(cid, label, pi, bound):
    pc = "B1" if cid == "B2" else cid
    c = S1.meta["cases"][cid]
    V, th, _, extent = case_potential(S1.meta, pc, pi)
    psi0, values_norm = incident(S1.meta, cid, label, extent, c["gpts"])
    planes = c["exit_planes"]
    waves = O.multislice(V, th, extent, c["energy"], psi0, planes=planes)
    b = c["beams"][label]
    for q, w in enumerate(waves):
        pattern = O.diffraction_pattern(w, extent, values_norm=values_norm, max_angle="cutoff")
        assert list(pattern.shape) == b["pattern_gpts"]
        assert rel_diff(float(np.sum(np.abs(w) ** 2)), b["total_intensity"][q]) <= bound
        assert rel_diff(pattern.sum(), b["pattern_sum"][q]) <= bound
        assert rel_diff(np.abs(w).max(), b["exit_max_abs"][q]) <= bound
        mx, my = pattern.shape
        assert abs(pattern[mx // 2, my // 2] - b["pattern_center"][q]) <= bound * pattern.max()
        if q in b["stored_planes"]:
            assert rel_max(w, S1.array(f"{cid}/{label}.exit.{q}")) <= bound
        if q in b["stored_pattern_planes"]:
            assert rel_max(pattern, S1.array(f"{cid}/{label}.pattern.{q}")) <= bound


# ---- TruthTests ---------------------------------------------------------------------------------------------
def test_plane_wave_through_vacuum_stays_one():
    # TruthTests.swift:47-56
    V = np.zeros((5, 64, 64))
    w = O.multislice(V, [2.0] * 5, (10.0, 10.0), 200e3, np.ones((64, 64), dtype=complex))[0]
    assert np.abs(w - 1).max() <= 1e-14


def test_probe_through_vacuum_is_the_defocused_probe():
    # TruthTests.swift:58-79: after T of vacuum the probe equals the probe with C10 + T
    n, L = 96, 12.0
    ab = {"C10": -50.0, "C30": 1e4, "C12": 20.0, "phi12": 0.3}
    psi = O.probe(n, n, (L, L), 200e3, 20.0, (6.0, 6.0), aberrations=ab)
    dz = [2.0, 2.5, 1.5, 2.0]
    waves = O.multislice(np.zeros((4, n, n)), dz, (L, L), 200e3, psi, planes=[0, 1, 2, 3])
    for q, w in enumerate(waves):
        shifted = dict(ab, C10=ab["C10"] + sum(dz[:q + 1]))
        want = O.probe(n, n, (L, L), 200e3, 20.0, (6.0, 6.0), aberrations=shifted)
        assert rel_max(w, want) <= 1e-12, q


def test_thickness_series_is_consistent_monotone_and_conserves_intensity():
    # TruthTests.swift:190-215
    c = S1.meta["cases"]["B1"]
    V, th, _, extent = case_potential(S1.meta, "B1", "exact")
    planes = [-1] + list(range(1, len(th), 2))  # ExitPlanes.every(2)
    series = O.multislice(V, th, extent, c["energy"], np.ones((256, 256), dtype=complex), planes=planes)
    assert len(series) == 11
    numbers = np.array(c["numbers"], dtype=np.int32)
    xyz = np.array(c["positions"]).reshape(-1, 3)
    for cells in (1, 3, 7):
        depth = cells * 4.05
        keep = xyz[:, 2] < depth - 1e-9
        Vp, _ = O.potential(256, 256, numbers[keep], xyz[keep], [c["box"][0], c["box"][1], depth],
                            n_slices(depth, 2.025))
        w = O.multislice(Vp, n_slices(depth, 2.025), extent, c["energy"], np.ones((256, 256), dtype=complex))[0]
        assert rel_max(w, series[cells]) <= 1e-13, cells
    n = 256 * 256
    totals = [float((np.abs(w) ** 2).sum()) for w in series]
    for q in range(1, len(series)):
        assert totals[q] <= totals[q - 1] * (1 + 1e-13)
    for q, w in enumerate(series):
        pattern = O.diffraction_pattern(w, extent, values_norm=True, max_angle="full")
        assert rel_diff(pattern.sum(), totals[q] / n) <= 1e-12
        assert pattern[128, 128] <= pattern.sum() * (1 + 1e-12)


def test_fft_matches_the_direct_dft_at_awkward_lengths():
    # SimulatoCoreTests.swift:21-52 (vs a direct O(n^2) DFT)
    rng = np.random.default_rng(1)
    for n in (2, 3, 5, 7, 11, 13, 17, 30, 75, 90, 97, 120, 509):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        k = np.arange(n)
        ang = (np.outer(k, k) % n) * (2 * np.pi / n)
        direct = (np.cos(ang) - 1j * np.sin(ang)) @ x
        assert rel_max(O.fft1(x[None])[0], direct) <= 1e-12, n
        back = O.fft1(O.fft1(x[None]), inverse=True)[0] / n
        assert rel_max(back, x) <= 1e-14, n


# ---- GridSamplingAndTiltTests (90 x 75, mixed radix on both axes, tilt) ---------------------------------------
@pytest.mark.parametrize("run", ["plane0", "planeT", "probe0", "probeT"])
def test_exit_waves_on_an_odd_grid_with_tilt_equal_abtem(run):
    S4 = golden("slice4")
    m = S4.meta
    nx, ny = m["gpts"]
    box = m["box"]
    th = n_slices(box[2], m["sliceThickness"])
    numbers = np.array([a[0] for a in m["atoms"]], dtype=np.int32)
    xyz = np.array([a[1] for a in m["atoms"]])
    V, _ = O.potential(nx, ny, numbers, xyz, box, th)
    tilt = m["runs"][run]["tilt"]
    extent = (box[0], box[1])
    if run.startswith("plane"):
        psi0 = np.ones((nx, ny), dtype=complex)
    else:
        psi0 = O.probe(nx, ny, extent, m["energy"], 20.0, m["probePosition"])
    w = O.multislice(V, th, extent, m["energy"], psi0, tilt=tilt)[0]
    assert rel_max(w, S4.array(run)) <= 1e-7


# ---- STEMPinsTests ------------------------------------------------------------------------------------------
S2 = golden("slice2")


def stem_case(cid, pi):
    c = S2.meta["cases"][cid]
    box = c["box"]
    nx, ny = c["gpts"]
    th = n_slices(box[2], c["slice_thickness_requested"])
    numbers = np.array(c["numbers"], dtype=np.int32)
    xyz = np.array(c["positions"], dtype=np.float64).reshape(-1, 3)
    nconf = max(1, len(c["seeds"]))
    Vs = []
    for k in range(nconf):
        d = S2.array(f"{cid}.displacement.{k}") if len(numbers) else np.zeros((0, 3))
        V, _ = O.potential(nx, ny, numbers, xyz + d, box, th, pi_mode=PI[pi])
        Vs.append(V)
    return c, Vs, th, (box[0], box[1])


def grid_positions(scan):
    """GridScan.positions (GridScan.swift:52-66): linspace per axis, endpoint false, index ix * ny + iy."""
    (x0, y0), (x1, y1), (gx, gy) = scan["start"], scan["end"], scan["gpts"]
    xs = np.arange(gx) * ((x1 - x0) / gx) + x0
    ys = np.arange(gy) * ((y1 - y0) / gy) + y0
    return np.array([(x, y) for x in xs for y in ys])


DETS = {"BF": O.det_annular(0, 20), "ADF": O.det_annular(30, 60), "HAADF": O.det_annular(70, 130),
        "annular_default_outer": O.det_annular(0, None), "pixelated_cutoff": O.det_pixelated("cutoff"),
        "pixelated_valid": O.det_pixelated("valid"), "pixelated_full": O.det_pixelated("full"),
        "DPC": O.det_segmented(2, 4, 0, 20), "DPC_rotated": O.det_segmented(2, 4, 0, 20, 0.3)}


@pytest.mark.parametrize("cid", ["S1", "S2", "S3"])
def test_scan_positions_equal_abtem(cid):
    c = S2.meta["cases"][cid]
    p = grid_positions(c["scan"])
    assert np.array_equal(p.reshape(-1), S2.array(f"{cid}.scan_positions").reshape(-1))


@pytest.mark.parametrize("pi", ["abtem", "exact"])
@pytest.mark.parametrize("cid", ["S1", "S2", "S3"])
def test_detectors_equal_abtem(cid, pi):
    c, Vs, th, extent = stem_case(cid, pi)
    pos = grid_positions(c["scan"])
    names = list(DETS)
    res = O.stem_scan(Vs, [th] * len(Vs), extent, c["energy"], c["cutoff_mrad"], pos, [DETS[n] for n in names])
    tight = pi == "abtem"
    for name, m in zip(names, res):
        entry = c["detectors"][name]
        if entry["kind"] == "annular":
            assert scaled_max(m.reshape(-1), S2.array(f"{cid}.{name}").reshape(-1), entry["scale"]) <= \
                (1e-10 if tight else 1e-6), name
        elif entry["kind"] == "segmented":
            # STEMPinsTests.swift:43-47: against the float64 bin sums, and abTEM's float32 sums to 1e-6
            assert entry["params"] == {"inner": 0.0, "outer": 20.0, "nbins_radial": 2, "nbins_azimuthal": 4,
                                       "rotation": 0.3 if name == "DPC_rotated" else 0.0}
            assert scaled_max(m.reshape(-1), S2.array(f"{cid}.{name}.float64").reshape(-1), entry["scale"]) <= \
                (1e-10 if tight else 1e-6), name
            assert scaled_max(m.reshape(-1), S2.array(f"{cid}.{name}").reshape(-1), entry["scale"]) <= 1e-6, name
        else:
            sums = np.array(entry["sums"]).reshape(-1)
            mine = m.reshape(len(pos), -1).sum(axis=1)
            assert scaled_max(mine, sums, np.abs(sums).max()) <= (1e-10 if tight else 1e-6), name
            sy = c["scan"]["gpts"][1]
            for ix, iy in entry["stored_positions"]:
                want = S2.array(f"{cid}.{name}.position.{ix}.{iy}")
                assert rel_max(m[ix * sy + iy], want) <= (1e-10 if tight else 1e-6), (name, ix, iy)


@pytest.mark.parametrize("sym,z", [("C", 6), ("Si", 14), ("Cu", 29), ("Au", 79)])
def test_single_atom_haadf_equals_abtem(sym, z):
    # STEMPinsTests.swift:62-76
    want = S2.meta["cases"]["S4"]["HAADF"][sym]
    V, _ = O.potential(128, 128, [z], np.array([[4.0, 4.0, 1.0]]), [8.0, 8.0, 2.0], [2.0], pi_mode=1)
    psi = O.probe(128, 128, (8.0, 8.0), 200e3, 25.0, (4.0, 4.0))
    w = O.multislice(V, [2.0], (8.0, 8.0), 200e3, psi)[0]
    P = O.diffraction_intensity(w)
    got = O.detect(P, (8.0, 8.0), 200e3, O.det_annular(70, 130))
    assert abs(got - want) / want <= 1e-10


def test_parallel_batch_equals_sequential_runs_bit_for_bit():
    # STEMPinsTests.swift:78-93
    c, Vs, th, extent = stem_case("S2", "exact")
    pos = grid_positions(c["scan"])[:6]
    inc = np.stack([O.probe(128, 128, extent, c["energy"], 20.0, p) for p in pos])
    batch = O.multislice_batch(Vs[0], th, extent, c["energy"], inc, threads=4)
    for q in range(len(pos)):
        single = O.multislice(Vs[0], th, extent, c["energy"], inc[q])[0]
        assert np.array_equal(batch[q], single)


# ---- FrozenPhononsTests ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("pi,bound", [("abtem", 1e-10), ("exact", 1e-6)])
def test_frozen_phonon_mean_pattern_equals_abtem(pi, bound):
    c = S2.meta["cases"]["F1"]
    box = c["box"]
    nx, ny = c["gpts"]
    th = n_slices(box[2], c["slice_thickness_requested"])
    numbers = np.array(c["numbers"], dtype=np.int32)
    xyz = np.array(c["positions"]).reshape(-1, 3)
    mean = None
    for k in range(c["num_configs"]):
        V, _ = O.potential(nx, ny, numbers, xyz + S2.array(f"F1.displacement.{k}"), box, th, pi_mode=PI[pi])
        if k == 0:
            assert rel_max(V[0], S2.array("F1.V.config0.slice0")) <= (1e-11 if pi == "abtem" else 1e-7)
        w = O.multislice(V, th, (box[0], box[1]), c["energy"], np.ones((nx, ny), dtype=complex))[0]
        total = float((np.abs(w) ** 2).sum())
        assert rel_diff(total, c["per_config_total"][k]) <= bound
        p = O.diffraction_pattern(w, (box[0], box[1]), values_norm=True)
        mean = p if mean is None else mean + p
    mean = mean / c["num_configs"]
    assert rel_max(mean, S2.array("F1.pattern_mean")) <= bound


# ---- ScanPatternsAndNoiseTests (pixelated cutoff crop and centre of mass over a 6 x 5 scan) ---------------------
def test_pixelated_and_centre_of_mass_over_a_scan_equal_abtem():
    S = golden("slice4scan")
    m = S.meta
    n = m["gpts"]
    box = m["box"]
    th = n_slices(box[2], m["sliceThickness"])
    numbers = np.array([a[0] for a in m["atoms"]], dtype=np.int32)
    xyz = np.array([a[1] for a in m["atoms"]])
    V, _ = O.potential(n, n, numbers, xyz, box, th)
    pos = grid_positions(m["scan"])
    dets = [O.det_pixelated("cutoff"), O.det_com("cutoff", "mrad"), O.det_com("cutoff", "inverse_angstroms")]
    dp, com_mrad, com_inv = O.stem_scan([V], [th], (box[0], box[1]), m["energy"], m["semiangle"], pos, dets)
    assert rel_max(dp.reshape(-1), S.array("dp").reshape(-1)) <= 1e-6
    for got, name in ((com_mrad, "com_mrad"), (com_inv, "com_inv")):
        want = S.array(name).reshape(-1)
        want = np.stack([want.real, want.imag], axis=1)
        assert rel_max(got.reshape(-1), want.reshape(-1)) <= 1e-5, name


# ---- detector truths (DetectorTruthTests.swift:23-57) ---------------------------------------------------------
def test_detector_truths_parseval_bf_plus_df_half_open_annulus():
    rng = np.random.default_rng(3)
    n, L, E = 64, 10.0, 200e3
    psi = O.probe(n, n, (L, L), E, 20.0, (5.0, 5.0))
    P = O.diffraction_intensity(psi)
    assert abs(P.sum() - 1.0) <= 1e-12  # Parseval for a normalised probe
    full = O.detect(P, (L, L), E, O.det_annular(0, None))
    inner = O.detect(P, (L, L), E, O.det_annular(0, 40))
    outer = O.detect(P, (L, L), E, O.det_annular(40, None))
    assert abs(inner + outer - full) <= 1e-15 + 1e-14 * full
    # a pixel exactly on the boundary angle belongs to the outer annulus only
    dalpha = (1 / L) * O.wavelength(E) * 1e3
    Q = np.zeros((n, n))
    Q[3, 0] = 1.0
    edge = 3 * dalpha
    assert O.detect(Q, (L, L), E, O.det_annular(0, edge)) == 0.0
    assert O.detect(Q, (L, L), E, O.det_annular(edge, 2 * edge)) == 1.0
    with pytest.raises(ValueError):
        O.detect(P, (L, L), E, O.det_annular(0, 1e4))
    _ = rng


# ---- GaussianGenerator (FrozenPhonons.swift:14-58) --------------------------------------------------------------
def test_splitmix64_and_xoshiro_known_answers():
    # SplitMix64 from seed 1234567 (Vigna's reference implementation, splitmix64.c)
    assert [int(v) for v in O.splitmix_stream(1234567, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]
    # xoshiro256** from the state (1, 2, 3, 4) (Blackman and Vigna's xoshiro256starstar.c)
    assert [int(v) for v in O.xoshiro_from_state([1, 2, 3, 4], 4)] == [11520, 0, 1509978240, 1215971899390074240]
    # and seeded by SplitMix64(0): the state words are SplitMix64(0)'s first four outputs
    sm = [int(v) for v in O.splitmix_stream(0, 4)]
    s = sm[:]
    mask = (1 << 64) - 1

    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & mask

    want = []
    for _ in range(6):
        want.append((rotl((s[1] * 5) & mask, 7) * 9) & mask)
        t = (s[1] << 17) & mask
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    assert [int(v) for v in O.xoshiro_stream(0, 6)] == want


def test_gaussian_generator_statistics():
    # FrozenPhononsTests.swift:35-84 (truth 1)
    z = O.gaussian_normals(1, 1_000_000)
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1) < 5e-3
    assert abs(np.mean(z ** 4) - 3) < 3e-2
    assert abs(np.mean(z[1:] * z[:-1])) < 5e-3
    assert O.configuration_seed(7, 0) != O.configuration_seed(8, 0)
    assert O.configuration_seed(7, 1) != O.configuration_seed(8, 0)
