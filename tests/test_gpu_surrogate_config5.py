"""The trained HAADF surrogate on held-out cells of the surrogate workload at its own size (30 x 30 Angstrom random
cells of 200-800 atoms, thickness 20-100 Angstrom, Z in {6, 8, 14, 22, 38, 79}; 512 x 512 waves, 2 Angstrom slices,
200 keV, 20 mrad, annular 50-200 mrad, a 96 x 96 scan), with the committed weights, judged at the real 2 %
tolerance by plain multislice on 5 % of the positions.

The judge's decision is checked both ways: its spot values are plain multislice at those positions, it falls back
exactly when the spot error exceeds the tolerance, the fallback delivers the plain multislice image, and at a
tolerance just above the spot error it delivers the network's image. The network's own error on each cell is held
to the value the committed weights reached, and the test prints what the surrogate gains over plain multislice run
with the library's default plan.
"""
import os
import time

import numpy as np
import pytest

import simulatokit_amd as sk
from simulatokit_amd import surrogate as SG

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(SG.__file__), "data", "haadf_surrogate.npz")
# held-out cells (seeds after the 2,000 training cells): the thinnest and the thickest of the first held-out ones,
# with the network's relative L2 error the committed weights reach on them (bfloat16 weights, float32 arithmetic:
# deterministic up to the order of float32 sums)
HELD_OUT = {20263254: 0.04076, 20263225: 0.36506}  # seed -> network error (20.8 A, 225 atoms; 97.5 A, 764 atoms)
TOL = 0.02


@pytest.fixture(scope="module")
def model():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())
    return SG.HaadfSurrogate.from_file(WEIGHTS)


@pytest.mark.parametrize("seed", sorted(HELD_OUT))
def test_judged_surrogate_on_a_held_out_cell(model, seed):
    work = SG.SurrogateWorkload()
    st = work.cell(seed)
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    truth = work.haadf(st)
    plain_s = time.perf_counter() - t0
    js = SG.judged_scan(model, work, st, fraction=0.05, tolerance=TOL, seed=seed)
    npos = work.scan_n ** 2
    assert len(js.checked) == round(0.05 * npos) and len(np.unique(js.checked)) == len(js.checked)
    # the judge's values are plain multislice at those positions
    spot = work.haadf(st, work.scan().positions[js.checked])
    assert np.abs(spot - truth.reshape(-1)[js.checked]).max() <= 1e-9 * np.abs(truth).max()
    pred = js.surrogate.reshape(-1)[js.checked]
    assert abs(js.spot_error - np.linalg.norm(pred - spot) / np.linalg.norm(spot)) <= 1e-6
    # it falls back exactly when the spot error exceeds the tolerance, and then delivers plain multislice
    assert js.fallback == (js.spot_error > TOL)
    if js.fallback:
        assert np.array_equal(js.image, truth)
    else:
        assert js.image is js.surrogate
    # just above its spot error the judge accepts the network's image
    accepted = SG.judged_scan(model, work, st, fraction=0.05, tolerance=js.spot_error * 1.01 + 1e-12, seed=seed)
    assert not accepted.fallback and accepted.image is accepted.surrogate
    err = float(np.linalg.norm(js.surrogate - truth) / np.linalg.norm(truth))
    gain = plain_s / (accepted.seconds["surrogate"] + accepted.seconds["judge"])
    print(f"cell {seed} ({st.box[2]:.0f} A, {len(st.numbers)} atoms): network error {err:.4f}, spot error "
          f"{js.spot_error:.4f}, fallback {js.fallback}; plain {plain_s * 1e3:.0f} ms, surrogate + judge "
          f"{(accepted.seconds['surrogate'] + accepted.seconds['judge']) * 1e3:.0f} ms ({gain:.1f} x when accepted)",
          flush=True)
    # the committed weights' accuracy on this cell, and the judge's estimate of it
    assert err <= HELD_OUT[seed] * 1.02 + 1e-4
    assert 0.5 * err <= js.spot_error <= 2.0 * err
    assert gain > 3.0
