"""The surrogate's HIP layers and training on the GPU against the float64 oracle (oracle/surrogate_oracle.py, pinned
in test_oracle_surrogate.py), and the judged surrogate scan on a small generated workload.

Weights are bfloat16 in the kernels; the oracle gets the same bfloat16-rounded values, so the comparisons measure the
kernels' float32 arithmetic (relative 1e-5 of the largest value).
"""
import math

import numpy as np
import pytest

from oracle import surrogate_oracle as S

import simulatokit_amd as sk
from simulatokit_amd import surrogate as SG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + sk._native.last_error())


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def bf16_bits(a):
    return S.bf16_round(a).view(np.uint32) >> 16


def test_conv_layers_equal_the_oracle():
    import torch

    L = SG._bind()
    rng = np.random.default_rng(5)
    B, ci, co, h, w = 2, 3, 10, 16, 32
    x = rng.normal(size=(B, ci, h, w)).astype(np.float32)
    wt = S.bf16_round(rng.normal(size=(co, ci, 5, 5)) * 0.2)
    bias = rng.normal(size=co).astype(np.float32)
    dev = torch.device("cuda:0")
    X = torch.tensor(x, device=dev)
    W = torch.tensor(bf16_bits(wt).astype(np.int16), device=dev)
    Bv = torch.tensor(bias, device=dev)
    Y = torch.empty((B, co, h, w), device=dev)
    st = None
    assert L.sk_conv5_forward(X.data_ptr(), B, ci, h, w, W.data_ptr(), Bv.data_ptr(), co, 1, Y.data_ptr(), st) == 0
    want = S.conv5_forward(x, wt, bias, relu=True)
    assert rel(Y.cpu().numpy(), want) <= 1e-5
    dy = rng.normal(size=(B, co, h, w)).astype(np.float32)
    DY = torch.tensor(dy, device=dev)
    DX = torch.empty((B, ci, h, w), device=dev)
    scratch = torch.empty(co * ci * 25, dtype=torch.int16, device=dev)
    assert L.sk_conv5_backward_data(DY.data_ptr(), B, ci, h, w, W.data_ptr(), co, DX.data_ptr(), scratch.data_ptr(),
                                    st) == 0
    assert rel(DX.cpu().numpy(), S.conv5_backward_data(dy, wt)) <= 1e-5
    DW = torch.empty((co, ci, 5, 5), device=dev)
    DB = torch.empty(co, device=dev)
    assert L.sk_conv5_backward_weights(X.data_ptr(), DY.data_ptr(), B, ci, h, w, co, DW.data_ptr(), DB.data_ptr(),
                                       st) == 0
    dw, db = S.conv5_backward_weights(x, dy)
    assert rel(DW.cpu().numpy(), dw) <= 1e-5
    assert rel(DB.cpu().numpy(), db) <= 1e-5
    A = S.resample_matrix(24, 32)
    img = rng.normal(size=(3, 32, 32)).astype(np.float32)
    tmp = torch.empty((3, 32, 24), device=dev)
    out = torch.empty((3, 24, 24), device=dev)
    At = torch.tensor(A.astype(np.float32), device=dev)
    assert L.sk_separable_map(torch.tensor(img, device=dev).data_ptr(), 3, 32, At.data_ptr(), 24, tmp.data_ptr(),
                              out.data_ptr(), st) == 0
    assert rel(out.cpu().numpy(), S.separable_map(img, A)) <= 1e-5


def test_network_gradients_equal_the_oracle():
    model = SG.HaadfSurrogate(seed=3, channels=[2, 8, 8, 1], input_n=32, output_n=24)
    rng = np.random.default_rng(6)
    import torch

    x = rng.normal(size=(2, 2, 32, 32)).astype(np.float32)
    for k in range(3):  # give the biases values so every ReLU sees both signs
        model.b[k].copy_(torch.tensor(rng.normal(size=model.b[k].numel()).astype(np.float32) * 0.1))
    out, acts = model.forward(torch.tensor(x, device=model.dev), keep=True)
    dout = rng.normal(size=tuple(out.shape)).astype(np.float32)
    grads = model.backward(acts, torch.tensor(dout, device=model.dev))
    layers = [(S.bf16_round(w.cpu().numpy()), b.cpu().numpy()) for w, b in zip(model.w32, model.b)]
    oacts = S.network_forward(x, layers)
    assert model.resample == "fourier"
    A = S.fourier_resample_matrix(24, 32)
    assert rel(out.cpu().numpy(), S.separable_map(oacts[-1][:, 0], A)) <= 1e-5
    g_last = S.separable_map(dout, A.T)[:, None]
    ograds = S.network_backward(oacts, layers, g_last)
    for k, ((dw, db), (odw, odb)) in enumerate(zip(grads, ograds)):
        assert rel(dw.cpu().numpy(), odw) <= 1e-4, k
        assert rel(db.cpu().numpy(), odb) <= 1e-4, k


def test_judged_surrogate_scan_on_a_small_workload():
    # a reduced workload so that data generation and training take seconds: 256^2 waves (annular 50-100 mrad fits),
    # 48 x 48 scan, 64^2 input, thin cells; 24 training cells
    work = SG.SurrogateWorkload(grid=256, scan_n=48, input_n=64, outer=100.0, t_min=8.0, t_max=20.0, n_min=60,
                                n_max=160)
    data = SG.make_dataset(work, list(range(100, 124)))
    model = SG.HaadfSurrogate(seed=1, channels=[2, 16, 16, 16, 1], input_n=64, output_n=48)
    hist = model.train(data, epochs=40, batch=8, lr=3e-3)
    print("mse first/last", hist[0], hist[-1])
    assert hist[-1] < 0.25 * hist[0]
    st = work.cell(999)
    js = SG.judged_scan(model, work, st, fraction=0.05, tolerance=0.02)
    truth = work.haadf(st)
    err = np.linalg.norm(js.surrogate - truth) / np.linalg.norm(truth)
    print("held-out network error", err, "spot", js.spot_error, "seconds", js.seconds)
    assert len(js.checked) == round(0.05 * 48 * 48)
    # the judge's estimate tracks the network's true error, and decides at the tolerance
    assert 0.5 * err <= js.spot_error <= 2.0 * err
    assert js.fallback == (js.spot_error > 0.02)
    if js.fallback:
        assert np.array_equal(js.image, truth)
    else:
        assert js.image is js.surrogate
    # just above the spot error it accepts the network's image, just below it falls back to multislice
    above = SG.judged_scan(model, work, st, fraction=0.05, tolerance=js.spot_error * 1.01)
    assert not above.fallback and above.image is above.surrogate
    below = SG.judged_scan(model, work, st, fraction=0.05, tolerance=js.spot_error * 0.99)
    assert below.fallback
    assert np.array_equal(below.image, truth)
