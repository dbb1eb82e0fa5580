"""The HIP path against the float64 oracle and the reference's abTEM fixtures, on an MI355X.

Bounds (complex64 on the device vs float64): relative max |a - b| / max |b| <= 1e-4 for waves (the reference
registers the same for its float32 backend, ExecutionOptions.swift:73-94); detectors within 2e-4 x max(image max,
1e-5 x incident). Mirrors MetalGPUTests.swift / MetalTests.swift / STEMPinsTests.swift:95-168, with `.hip` in
place of `.metal` and the oracle in place of `.cpu`.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden, rel_l2, rel_max, scaled_max

import simulatokit_amd as sk

pytestmark = pytest.mark.gpu

ACC = sk.REGISTERED


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


def n_slices(depth, requested):
    n = math.ceil(depth / requested)
    return [depth / n] * n


def structure_of(c):
    return sk.AtomicStructure(c["box"], c["numbers"], c["positions"])


# ---- potentials and transmissions ---------------------------------------------------------------------------------
@pytest.mark.parametrize("cid", ["A1", "A2", "B1", "D1"])
def test_device_potential_equals_the_oracle_and_abtem(cid):
    S1 = golden("slice1")
    c = S1.meta["cases"][cid]
    st = structure_of(c)
    pot = sk.Potential(st, *c["gpts"], slice_thickness=c["slice_thickness_requested"],
                       parameterization=sk.LobatoParameterization("abtem_float32"))
    prepared = pot.prepare(sk.ElectronBeam(c["energy"]))
    V = prepared.values()
    Vo, _ = O.potential(*c["gpts"], np.array(c["numbers"], dtype=np.int32), np.array(c["positions"]), c["box"],
                        pot.slice_thicknesses, pi_mode=1)
    print(cid, "V rel max vs oracle", rel_max(V, Vo))
    assert rel_max(V, Vo) <= 1e-5
    for j in range(c["V_stored_slices"]):
        assert rel_max(V[j], S1.array(f"{cid}.V.{j}")) <= 1e-5
    t = prepared.transmissions()
    to = O.transmission(Vo[:2], (c["box"][0], c["box"][1]), c["energy"])
    print(cid, "t rel max vs oracle", rel_max(t[:2], to))
    assert rel_max(t[:2], to) <= 1e-5
    assert rel_max(t[0], S1.array(f"{cid}.t.0")) <= 1e-5


# ---- multislice: plane waves and probes ------------------------------------------------------------------------
CASE_BEAMS = [("A1", "plane"), ("A2", "plane"), ("B1", "plane"), ("B2", "probe"), ("C1", "plane"),
              ("C1", "probe"), ("D1", "plane"), ("D1", "probe")]


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("cid,label", CASE_BEAMS)
def test_exit_waves_equal_the_oracle_and_abtem(cid, label, fast):
    S1 = golden("slice1")
    pc = "B1" if cid == "B2" else cid
    c = S1.meta["cases"][pc]
    cb = S1.meta["cases"][cid]
    b = cb["beams"][label]
    beam = sk.ElectronBeam(cb["energy"])
    pot = sk.Potential(structure_of(c), *c["gpts"], slice_thickness=c["slice_thickness_requested"],
                       parameterization=sk.LobatoParameterization("abtem_float32"))
    grid = pot.grid
    if b["kind"] == "plane":
        incident = sk.Waves.plane_wave(beam, grid)
    else:
        ab = sk.Aberrations(**b["aberrations"])
        probe = sk.Probe(beam, grid, sk.Aperture(b["cutoff_mrad"]), ab)
        incident = probe.build(b["position"])
        assert rel_max(incident.field, S1.array(f"{cid}/{label}.incident")) <= 1e-5
    planes = cb["exit_planes"]
    waves = sk.Multislice.run(incident, pot, sk.ExitPlanes.after_slices(planes),
                              sk.ExecutionOptions(fast_path=fast))
    V, _ = O.potential(*c["gpts"], np.array(c["numbers"], dtype=np.int32), np.array(c["positions"]), c["box"],
                       pot.slice_thicknesses, pi_mode=1)
    want = O.multislice(V, pot.slice_thicknesses, (grid.extent_x, grid.extent_y), cb["energy"], incident.field,
                        planes=planes)
    for q, w in enumerate(waves):
        err = rel_max(w.field, want[q])
        assert err <= ACC.wave, (q, err)
        if q in b["stored_planes"]:
            assert rel_max(w.field, S1.array(f"{cid}/{label}.exit.{q}")) <= ACC.wave
        # sum |psi|^2 per slice (registered 1.5e-6 per slice)
        tot, tot_o = w.total_intensity, float((np.abs(want[q]) ** 2).sum())
        assert abs(tot - tot_o) / tot_o <= ACC.total_intensity_per_slice * max(1, planes[q] + 1)


@pytest.mark.parametrize("run", ["plane0", "planeT", "probe0", "probeT"])
def test_odd_grid_and_tilt_equal_abtem(run):
    S4 = golden("slice4")
    m = S4.meta
    nx, ny = m["gpts"]
    st = sk.AtomicStructure(m["box"], [a[0] for a in m["atoms"]], [a[1] for a in m["atoms"]])
    pot = sk.Potential(st, nx, ny, slice_thickness=m["sliceThickness"])
    beam = sk.ElectronBeam(m["energy"])
    tilt = m["runs"][run]["tilt"]
    if run.startswith("plane"):
        inc = sk.Waves.plane_wave(beam, pot.grid, tilt)
    else:
        inc = sk.Probe(beam, pot.grid, sk.Aperture(20.0), tilt=tilt).build(m["probePosition"])
    w = sk.Multislice.run(inc, pot)[0]
    assert tuple(w.tilt) == tuple(tilt)
    assert rel_max(w.field, S4.array(run)) <= ACC.wave


@pytest.mark.parametrize("nx,ny", [(67, 53), (131, 137), (1021, 1019)])
def test_bluestein_grid_equals_the_oracle(nx, ny):
    # both lengths prime, so both axes take Bluestein's algorithm on the device: convolution lengths 256 and 128
    # (67 x 53), 512 (131 x 137) and 2048 (1021 x 1019), at about 0.14 A sampling
    rng = np.random.default_rng(5)
    L = (nx * 9.0 / 67, ny * 7.5 / 53)
    numbers = rng.choice([8, 14, 38], 30)
    xyz = rng.uniform(0, 1, (30, 3)) * np.array([L[0], L[1], 6.0])
    st = sk.AtomicStructure((L[0], L[1], 6.0), numbers, xyz)
    pot = sk.Potential(st, nx, ny, slice_thickness=2.0)
    beam = sk.ElectronBeam(200e3)
    inc = sk.Probe(beam, pot.grid, sk.Aperture(25.0)).build((4.0, 3.0))
    w = sk.Multislice.run(inc, pot)[0]
    V, _ = O.potential(nx, ny, numbers, xyz, st.box, pot.slice_thicknesses)
    want = O.multislice(V, pot.slice_thicknesses, L, 200e3, O.probe(nx, ny, L, 200e3, 25.0, (4.0, 3.0)))[0]
    assert rel_max(w.field, want) <= ACC.wave


def test_run_batch_equals_single_runs_and_the_oracle():
    S2 = golden("slice2")
    c = S2.meta["cases"]["S2"]
    pot = sk.Potential(structure_of(c), *c["gpts"], slice_thickness=c["slice_thickness_requested"])
    beam = sk.ElectronBeam(c["energy"])
    probe = sk.Probe(beam, pot.grid, sk.Aperture(c["cutoff_mrad"]))
    pos = sk.GridScan(c["scan"]["start"], c["scan"]["end"], *c["scan"]["gpts"]).positions[:7]
    inc = [probe.build(p) for p in pos]
    prepared = pot.prepare(beam)
    batch = sk.Multislice.run_batch(inc, prepared, sk.ExecutionOptions(batch=3))
    for q in range(len(pos)):
        single = sk.Multislice.run(inc[q], prepared)[0]
        assert np.array_equal(batch[q].field, single.field), q  # bit for bit, whatever the batch
    V, _ = O.potential(*c["gpts"], np.array(c["numbers"], dtype=np.int32), np.array(c["positions"]), c["box"],
                       pot.slice_thicknesses)
    want = O.multislice_batch(V, pot.slice_thicknesses, (pot.grid.extent_x, pot.grid.extent_y), c["energy"],
                              np.stack([w.field for w in inc]))
    for q in range(len(pos)):
        assert rel_max(batch[q].field, want[q]) <= ACC.wave


# ---- STEM scans --------------------------------------------------------------------------------------------------
DETS = {"BF": sk.AnnularDetector(0, 20), "ADF": sk.AnnularDetector(30, 60), "HAADF": sk.AnnularDetector(70, 130),
        "annular_default_outer": sk.AnnularDetector(0, None), "pixelated_cutoff": sk.PixelatedDetector("cutoff"),
        "pixelated_valid": sk.PixelatedDetector("valid"), "pixelated_full": sk.PixelatedDetector("full"),
        "DPC": sk.SegmentedDetector(2, 4, 0, 20), "DPC_rotated": sk.SegmentedDetector(2, 4, 0, 20, 0.3)}


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("cid", ["S1", "S2", "S3"])
def test_stem_detectors_equal_abtem(cid, fast):
    S2 = golden("slice2")
    c = S2.meta["cases"][cid]
    st = structure_of(c)
    nconf = max(1, len(c["seeds"]))
    disp = [S2.array(f"{cid}.displacement.{k}") if len(c["numbers"]) else np.zeros((0, 3)) for k in range(nconf)]
    fp = sk.FrozenPhonons(st, displacements=disp)
    pots = fp.lazy_potentials(*c["gpts"], c["slice_thickness_requested"])
    beam = sk.ElectronBeam(c["energy"])
    probe = sk.Probe(beam, pots[0].grid, sk.Aperture(c["cutoff_mrad"]))
    scan = sk.GridScan(c["scan"]["start"], c["scan"]["end"], *c["scan"]["gpts"])
    names = list(DETS)
    ms = sk.STEM.scan(probe, pots, scan, [DETS[n] for n in names], sk.ExecutionOptions(fast_path=fast))
    for name, m in zip(names, ms):
        entry = c["detectors"][name]
        if entry["kind"] in ("annular", "segmented"):
            # segmented: the float64 bin sums (STEMPinsTests.swift:43-47)
            key = f"{cid}.{name}" + (".float64" if entry["kind"] == "segmented" else "")
            want = S2.array(key).reshape(-1)
            bound = ACC.detector_bound(np.abs(want).max())
            assert np.abs(m.values - want).max() <= bound, name
        else:
            sums = np.array(entry["sums"]).reshape(-1)
            mine = m.values.reshape(len(sums), -1).sum(axis=1)
            assert scaled_max(mine, sums, np.abs(sums).max()) <= ACC.detector, name
            for ix, iy in entry["stored_positions"]:
                want = S2.array(f"{cid}.{name}.position.{ix}.{iy}")
                assert rel_max(m[ix, iy], want) <= ACC.pattern, (name, ix, iy)


@pytest.mark.parametrize("sym,z", [("C", 6), ("Si", 14), ("Cu", 29), ("Au", 79)])
def test_single_atom_haadf_equals_abtem(sym, z):
    S2 = golden("slice2")
    want = S2.meta["cases"]["S4"]["HAADF"][sym]
    st = sk.AtomicStructure((8.0, 8.0, 2.0), [z], [[4.0, 4.0, 1.0]])
    pot = sk.Potential(st, 128, 128, slice_thickness=2.0, parameterization=sk.LobatoParameterization("abtem_float32"))
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(25.0))
    got = sk.STEM.scan(probe, [pot], np.array([[4.0, 4.0]]), [sk.AnnularDetector(70, 130)])[0].values[0]
    assert abs(got - want) / want <= ACC.detector


def test_pixelated_and_centre_of_mass_equal_abtem():
    S = golden("slice4scan")
    m = S.meta
    st = sk.AtomicStructure(m["box"], [a[0] for a in m["atoms"]], [a[1] for a in m["atoms"]])
    pot = sk.Potential(st, m["gpts"], m["gpts"], slice_thickness=m["sliceThickness"])
    probe = sk.Probe(sk.ElectronBeam(m["energy"]), pot.grid, sk.Aperture(m["semiangle"]))
    scan = sk.GridScan(m["scan"]["start"], m["scan"]["end"], *m["scan"]["gpts"])
    dp, cm, ci = sk.STEM.scan(probe, [pot], scan, [sk.PixelatedDetector("cutoff"),
                                                   sk.CenterOfMassDetector("cutoff", "mrad"),
                                                   sk.CenterOfMassDetector("cutoff", "inverse_angstroms")])
    assert dp.value_shape == (42, 42)
    assert rel_max(dp.values, S.array("dp").reshape(-1)) <= ACC.pattern
    for got, name in ((cm, "com_mrad"), (ci, "com_inv")):
        w = S.array(name).reshape(-1)
        assert rel_max(got.values, np.stack([w.real, w.imag], 1).reshape(-1)) <= 1e-4, name


def test_lane_pipeline_with_uneven_slice_counts():
    # three potentials of 2, 3 and 5 unit cells (4, 6, 10 slices) mixed in every launch of the lane pipeline
    a = 3.905
    cell = sk.AtomicStructure((a, a, a), [38, 22, 8, 8, 8],
                              [[0, 0, 0], [a / 2, a / 2, a / 2], [a / 2, a / 2, 0], [a / 2, 0, a / 2], [0, a / 2, a / 2]])
    pots = [sk.Potential(cell.repeated(3, 3, nz), 256, 256, slice_thickness=2.0) for nz in (2, 3, 5)]
    assert [p.slice_count for p in pots] == [4, 6, 10]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    scan = sk.GridScan((0, 0), (a, a), 6, 7)
    dets = [sk.AnnularDetector(50, 150), sk.CenterOfMassDetector("cutoff", "mrad"), sk.PixelatedDetector("valid")]
    prepared = [p.prepare(probe.beam) for p in pots]
    seq = sk.STEM.scan(probe, prepared, scan, dets, sk.ExecutionOptions(pipeline=0))
    pipe = sk.STEM.scan(probe, prepared, scan, dets, sk.ExecutionOptions(pipeline=1, lane_waves=8))
    for x, y in zip(seq, pipe):  # the same float32 waves; only the float64 order of the potential sum may differ
        assert rel_max(y.values, x.values) <= 1e-12
    graph = sk.STEM.scan(probe, prepared, scan, dets, sk.ExecutionOptions(pipeline=1, lane_waves=8, use_graph=True))
    for x, y in zip(pipe, graph):
        assert np.array_equal(x.values, y.values)
    Vs, ths = [], []
    for p in pots:
        st = cell.repeated(3, 3, {4: 2, 6: 3, 10: 5}[p.slice_count])
        V, _ = O.potential(256, 256, st.numbers, st.positions, st.box, p.slice_thicknesses)
        Vs.append(V)
        ths.append(p.slice_thicknesses)
    want = O.stem_scan(Vs, ths, (3 * a, 3 * a), 200e3, 20.0, scan.positions,
                       [O.det_annular(50, 150), O.det_com("cutoff", "mrad"), O.det_pixelated("valid")])
    for got, wv in zip(pipe, want):
        wv = wv.reshape(-1)
        assert np.abs(got.values - wv).max() <= ACC.detector_bound(np.abs(wv).max())


def test_scan_is_deterministic_and_batch_independent():
    st = sk.AtomicStructure((3.905, 3.905, 3.905), [38, 22, 8, 8, 8],
                            [[0, 0, 0], [1.9525, 1.9525, 1.9525], [1.9525, 1.9525, 0], [1.9525, 0, 1.9525],
                             [0, 1.9525, 1.9525]]).repeated(4, 4, 4)
    pot = sk.Potential(st, 256, 256, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    scan = sk.GridScan((0, 0), (3.905, 3.905), 8, 8)
    prepared = pot.prepare(probe.beam)
    dets = [sk.AnnularDetector(50, 200), sk.CenterOfMassDetector()]
    a = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=64))
    b = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=64))
    c = sk.STEM.scan(probe, [prepared], scan, dets, sk.ExecutionOptions(batch=5))
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x.values, y.values)
        assert np.array_equal(x.values, z.values)
    # the slice loop replayed from captured hipGraphs gives the same bits: the first batch of a shape runs kernel by
    # kernel while its graph is captured, the later batches of that shape replay it (batch 5: 12 full batches and a
    # remainder of 4, run once, uncaptured; batch 16: 4 batches; batch 64: one batch, never captured), on one stream
    # (the library's choice for batches under 128) and with each batch's graph forked into parts on 2 and 3 streams
    for batch, streams in ((64, 0), (5, 0), (16, 0), (64, 2), (16, 2), (5, 3)):
        g = sk.STEM.scan(probe, [prepared], scan, dets,
                         sk.ExecutionOptions(batch=batch, use_graph=True, streams=streams))
        for x, y in zip(a, g):
            assert np.array_equal(x.values, y.values), (batch, streams)
    # a batch run as tiles (uneven: 64 positions in tiles of at most 20, spread as 4 x 16), with and without graph
    # replay, on one and two streams
    for graph, streams in ((False, 1), (False, 2), (True, 1), (True, 2)):
        t = sk.STEM.scan(probe, [prepared], scan, dets,
                         sk.ExecutionOptions(batch=64, tile=20, use_graph=graph, streams=streams))
        for x, y in zip(a, t):
            assert np.array_equal(x.values, y.values), ("tiles", graph, streams)


def test_lane_pipeline_single_position_over_many_potentials():
    """A CBED-style scan: one probe position over five potentials of different slice counts. The lanes hold the same
    position for different potentials at once, each accumulating into its own copy of the detector sums, which are
    added in lane order at the end; the per-potential waves are the same float32 numbers as one potential at a
    time, so only the float64 order of the sum over potentials may differ."""
    a = 3.905
    cell = sk.AtomicStructure((a, a, a), [38, 22, 8, 8, 8],
                              [[0, 0, 0], [a / 2, a / 2, a / 2], [a / 2, a / 2, 0], [a / 2, 0, a / 2], [0, a / 2, a / 2]])
    pots = [sk.Potential(cell.repeated(3, 3, nz), 256, 256, slice_thickness=2.0) for nz in (1, 2, 3, 4, 6)]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    positions = np.array([[a / 3, a / 4]])
    dets = [sk.AnnularDetector(50, 150), sk.CenterOfMassDetector("cutoff", "mrad"), sk.PixelatedDetector("valid")]
    prepared = [p.prepare(probe.beam) for p in pots]
    seq = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=0))
    for opt in (sk.ExecutionOptions(pipeline=1), sk.ExecutionOptions(pipeline=-1),
                sk.ExecutionOptions(pipeline=1, use_graph=True)):
        lanes = sk.STEM.scan(probe, prepared, positions, dets, opt)
        for x, y in zip(seq, lanes):
            assert rel_max(y.values, x.values) <= 1e-12
    # the same twice in a row: the accumulator copies start from zero on every scan
    again = sk.STEM.scan(probe, prepared, positions, dets, sk.ExecutionOptions(pipeline=1))
    for x, y in zip(lanes, again):
        assert np.array_equal(x.values, y.values)
