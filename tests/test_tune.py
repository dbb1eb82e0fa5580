"""The execution-plan autotuner: its cache on the host, and on the GPU that the plan it picks is the fastest it timed
and scans to the same numbers as the library's default plan."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import simulatokit_amd as sk
from simulatokit_amd import tune as T
from tests.conftest import rel_max

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cache_round_trip_and_key(tmp_path):
    g = sk.Grid(512, 512, 31.24, 31.24)
    key = T.tune_key("gfx950 256 CU", g, [40], [sk.AnnularDetector(50, 200)], 4096)
    assert key == "gfx950 256 CU|512x512|S=40|P<=4096|AnnularDetector"
    # a plan tuned for a large scan is not reused for a one-position scan, nor for one of 4097 positions
    assert T.tune_key("gfx950 256 CU", g, [40], [sk.AnnularDetector(50, 200)], 1) != key
    assert T.tune_key("gfx950 256 CU", g, [40], [sk.AnnularDetector(50, 200)], 4097) != key
    assert T.tune_key("gfx950 256 CU", g, [40], [sk.AnnularDetector(50, 200)], 3000) == key
    assert [T.positions_bucket(n) for n in (0, 1, 2, 3, 64, 65, 4096)] == [1, 1, 2, 4, 64, 128, 4096]
    path = str(tmp_path / "sub" / "tune.json")
    res = T.TuneResult(sk.ExecutionOptions(batch=1024, col_tile=8), 7.5e5, "gfx950 256 CU",
                       [{"batch": 1024, "rate": 7.5e5}], key)
    T.store_cache(path, key, res)
    T.store_cache(path, key + "x", res)
    data = T.load_cache(path)
    assert set(data) == {key, key + "x"}
    assert sk.ExecutionOptions(**data[key]["options"]) == res.options
    with open(path, "w") as f:
        f.write("not json")
    assert T.load_cache(path) == {}
    assert T.load_cache(str(tmp_path / "missing.json")) == {}


def test_autotune_refuses_without_a_device():
    script = r"""
import sys; sys.path.insert(0, sys.argv[1])
import numpy as np
import simulatokit_amd as sk
g = sk.Grid(32, 32, 8.0, 8.0)
pp = sk.ProjectedPotential(g, np.zeros((1, 32, 32)), [2.0])
probe = sk.Probe(sk.ElectronBeam(200e3), g, sk.Aperture(20.0))
try:
    sk.autotune(probe, [pp], [sk.AnnularDetector(0, 20)], cache=None)
except sk.SimulatoError as e:
    print(e.kind)
"""
    env = dict(os.environ, SK_HIP_HIDE_DEVICE="1")
    res = subprocess.run([sys.executable, "-c", script, ROOT], capture_output=True, text=True, env=env, timeout=120)
    assert res.stdout.strip() == "hipUnavailable", res.stdout + res.stderr


@pytest.mark.gpu
def test_autotune_picks_its_fastest_plan_and_scans_the_same_numbers(tmp_path):
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())
    a = 3.905
    cell = sk.AtomicStructure((a, a, a), [38, 22, 8, 8, 8],
                              [[0, 0, 0], [a / 2, a / 2, a / 2], [a / 2, a / 2, 0], [a / 2, 0, a / 2], [0, a / 2, a / 2]])
    pot = sk.Potential(cell.repeated(4, 4, 2), 256, 256, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    dets = [sk.AnnularDetector(50, 150), sk.CenterOfMassDetector()]
    cache = str(tmp_path / "tune.json")
    res = sk.autotune(probe, [pot], dets, positions=1024, cache=cache)
    assert not res.from_cache
    assert len(res.candidates) >= 2
    assert res.rate == max(c["rate"] for c in res.candidates)
    assert res.options.batch in (0, 1024)  # 0: the library's cache-sized batch (two launches of 512 here)
    # the default batch is timed split over the library's streams and on one stream
    assert {c["streams"] for c in res.candidates if c["batch"] == 0} == {0, 1}
    assert res.device.startswith("gfx950")
    again = sk.autotune(probe, [pot], dets, positions=1024, cache=cache)
    assert again.from_cache and again.options == res.options
    assert json.load(open(cache))[res.key]["rate"] == res.rate
    # another position count is another entry, tuned afresh
    small = sk.autotune(probe, [pot], dets, positions=8, cache=cache)
    assert not small.from_cache and small.key != res.key
    assert set(json.load(open(cache))) == {res.key, small.key}
    scan = sk.GridScan((0, 0), (a, a), 8, 8)
    want = sk.STEM.scan(probe, [pot], scan, dets)
    got = sk.STEM.scan(probe, [pot], scan, dets, res.options)
    for x, y in zip(want, got):
        assert rel_max(y.values, x.values) <= 1e-12
