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


def test_matrix_core_convolution_equals_the_oracle():
    # the 32 -> 32 layers on v_mfma_f32_32x32x16_bf16: two 8-row and two 32-column tiles per image, so that the
    # circular halo crosses tile edges both ways; forward with bias and ReLU, and the backward pass over the data as
    # the forward convolution of dy with the flipped tap-major weights
    import torch

    L = SG._bind()
    rng = np.random.default_rng(15)
    B, c, h, w = 2, 32, 16, 64
    x = rng.normal(size=(B, c, h, w)).astype(np.float32)
    wt = S.bf16_round(rng.normal(size=(c, c, 5, 5)) * 0.05)
    bias = rng.normal(size=c).astype(np.float32)
    dev = torch.device("cuda:0")
    W = torch.tensor(bf16_bits(wt).astype(np.int16), device=dev)
    fwd, bwd = (torch.empty(c * c * 25, dtype=torch.int16, device=dev) for _ in range(2))
    assert L.sk_conv5_weights_tap_major(W.data_ptr(), c, c, 0, fwd.data_ptr(), None) == 0
    assert L.sk_conv5_weights_tap_major(W.data_ptr(), c, c, 1, bwd.data_ptr(), None) == 0
    X, Bv = torch.tensor(x, device=dev), torch.tensor(bias, device=dev)
    Y = torch.empty((B, c, h, w), device=dev)
    assert L.sk_conv5_forward_tap_major(X.data_ptr(), B, c, h, w, fwd.data_ptr(), Bv.data_ptr(), c, 1, Y.data_ptr(),
                                        None) == 0
    assert rel(Y.cpu().numpy(), S.conv5_forward(x, wt, bias, relu=True)) <= 1e-5
    dy = rng.normal(size=(B, c, h, w)).astype(np.float32)
    DX = torch.empty((B, c, h, w), device=dev)
    assert L.sk_conv5_forward_tap_major(torch.tensor(dy, device=dev).data_ptr(), B, c, h, w, bwd.data_ptr(), None, c,
                                        0, DX.data_ptr(), None) == 0
    assert rel(DX.cpu().numpy(), S.conv5_backward_data(dy, wt)) <= 1e-5
    # the gradient of the weights: both operands split, three products (3.6e-6 of the largest value in float64 for
    # these inputs), float32 sums per block
    ws = torch.empty(SG.WGRAD_WORKSPACE_FLOATS, dtype=torch.float32, device=dev)
    DW, DB = torch.empty((c, c, 5, 5), device=dev), torch.empty(c, device=dev)
    assert L.sk_conv5_backward_weights_mfma(X.data_ptr(), torch.tensor(dy, device=dev).data_ptr(), B, c, h, w, c,
                                            DW.data_ptr(), DB.data_ptr(), ws.data_ptr(), None) == 0
    dw, db = S.conv5_backward_weights(x, dy)
    assert rel(DW.cpu().numpy(), dw) <= 1e-5
    assert rel(DB.cpu().numpy(), db) <= 1e-6
    # shapes outside the kernels' tiling are refused, not computed
    assert L.sk_conv5_forward_tap_major(X.data_ptr(), B, c, h, 48, fwd.data_ptr(), None, c, 0, Y.data_ptr(),
                                        None) != 0
    assert L.sk_conv5_forward_tap_major(X.data_ptr(), B, 33, h, w, fwd.data_ptr(), None, c, 0, Y.data_ptr(),
                                        None) != 0
    assert L.sk_conv5_backward_weights_mfma(X.data_ptr(), X.data_ptr(), B, c, h, 48, c, DW.data_ptr(), None,
                                            ws.data_ptr(), None) != 0


def test_matrix_core_layers_with_fewer_channels_equal_the_oracle():
    # the first layer (2 -> 32), the last (32 -> 1) and its backward pass over the data (1 -> 32): channels padded
    # with zeros to 16 or 32 inside the kernels, only the real ones stored
    import torch

    L = SG._bind()
    rng = np.random.default_rng(18)
    dev = torch.device("cuda:0")
    B, h, w = 3, 16, 32
    ws = torch.empty(SG.WGRAD_WORKSPACE_FLOATS, dtype=torch.float32, device=dev)
    for ci, co in ((2, 32), (32, 1), (5, 17)):
        x = rng.normal(size=(B, ci, h, w)).astype(np.float32)
        wt = S.bf16_round(rng.normal(size=(co, ci, 5, 5)) * 0.1)
        bias = rng.normal(size=co).astype(np.float32)
        W = torch.tensor(bf16_bits(wt).astype(np.int16), device=dev)
        fwd, bwd = (torch.empty(25 * 32 * 32, dtype=torch.int16, device=dev) for _ in range(2))
        assert L.sk_conv5_weights_tap_major(W.data_ptr(), ci, co, 0, fwd.data_ptr(), None) == 0
        assert L.sk_conv5_weights_tap_major(W.data_ptr(), ci, co, 1, bwd.data_ptr(), None) == 0
        X = torch.tensor(x, device=dev)
        Y = torch.empty((B, co, h, w), device=dev)
        assert L.sk_conv5_forward_tap_major(X.data_ptr(), B, ci, h, w, fwd.data_ptr(),
                                            torch.tensor(bias, device=dev).data_ptr(), co, 1, Y.data_ptr(), None) == 0
        assert rel(Y.cpu().numpy(), S.conv5_forward(x, wt, bias, relu=True)) <= 1e-5, (ci, co)
        dy = rng.normal(size=(B, co, h, w)).astype(np.float32)
        DY = torch.tensor(dy, device=dev)
        DX = torch.empty((B, ci, h, w), device=dev)
        assert L.sk_conv5_forward_tap_major(DY.data_ptr(), B, co, h, w, bwd.data_ptr(), None, ci, 0, DX.data_ptr(),
                                            None) == 0
        assert rel(DX.cpu().numpy(), S.conv5_backward_data(dy, wt)) <= 1e-5, (ci, co)
        DW, DB = torch.empty((co, ci, 5, 5), device=dev), torch.empty(co, device=dev)
        assert L.sk_conv5_backward_weights_mfma(X.data_ptr(), DY.data_ptr(), B, ci, h, w, co, DW.data_ptr(),
                                                DB.data_ptr(), ws.data_ptr(), None) == 0
        dw, db = S.conv5_backward_weights(x, dy)
        assert rel(DW.cpu().numpy(), dw) <= 1e-5, (ci, co)
        assert rel(DB.cpu().numpy(), db) <= 1e-6, (ci, co)


def test_network_gradients_on_the_matrix_cores_equal_the_oracle():
    import torch

    model = SG.HaadfSurrogate(seed=4, channels=[2, 32, 32, 32, 1], input_n=32, output_n=24)
    assert sorted(model.tap_major) == [0, 1, 2, 3]  # every layer takes the matrix-core path
    rng = np.random.default_rng(16)
    x = rng.normal(size=(2, 2, 32, 32)).astype(np.float32)
    for k in range(4):
        model.b[k].copy_(torch.tensor(rng.normal(size=model.b[k].numel()).astype(np.float32) * 0.1))
    out, acts = model.forward(torch.tensor(x, device=model.dev), keep=True)
    dout = rng.normal(size=tuple(out.shape)).astype(np.float32)
    grads = model.backward(acts, torch.tensor(dout, device=model.dev))
    layers = [(S.bf16_round(w.cpu().numpy()), b.cpu().numpy()) for w, b in zip(model.w32, model.b)]
    oacts = S.network_forward(x, layers)
    A = S.fourier_resample_matrix(24, 32)
    assert rel(out.cpu().numpy(), S.separable_map(oacts[-1][:, 0], A)) <= 1e-5
    ograds = S.network_backward(oacts, layers, S.separable_map(dout, A.T)[:, None])
    for k, ((dw, db), (odw, odb)) in enumerate(zip(grads, ograds)):
        assert rel(dw.cpu().numpy(), odw) <= 1e-4, k
        assert rel(db.cpu().numpy(), odb) <= 1e-4, k
    # an update refreshes the tap-major copies: the vector-unit network with the same weights agrees
    model.w32[1].mul_(1.5)
    model._publish()
    plain = SG.HaadfSurrogate(seed=4, channels=[2, 32, 32, 32, 1], input_n=32, output_n=24, matrix_cores=False)
    assert not plain.tap_major
    for k in range(4):
        plain.w32[k].copy_(model.w32[k])
        plain.b[k].copy_(model.b[k])
    plain._publish()
    xs = torch.tensor(x, device=model.dev)
    a, b = model.forward(xs).cpu().numpy(), plain.forward(xs).cpu().numpy()
    assert rel(a, b.astype(np.float64)) <= 1e-5


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
