"""Pins of the surrogate's float64 oracle (the reference has no learned part): the direct convolution against
circular convolution by FFT, the backward passes against the adjoint identities and finite differences, the
resampling matrix against interpolation of linear functions."""
import numpy as np

from oracle import surrogate_oracle as S


def test_direct_convolution_equals_circular_convolution_by_fft():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 16, 20))
    wt = rng.normal(size=(4, 3, 5, 5))
    b = rng.normal(size=4)
    y = S.conv5_forward(x, wt, b)
    H, W = x.shape[2:]
    want = np.zeros_like(y)
    for o in range(4):
        for c in range(3):
            g = np.zeros((H, W))
            for ky in range(5):
                for kx in range(5):
                    g[(ky - 2) % H, (kx - 2) % W] = wt[o, c, ky, kx]
            want[:, o] += np.real(np.fft.ifft2(np.fft.fft2(x[:, c]) * np.conj(np.fft.fft2(g))))
        want[:, o] += b[o]
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()


def test_backward_passes_are_the_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 3, 16, 16))
    wt = rng.normal(size=(5, 3, 5, 5))
    dy = rng.normal(size=(2, 5, 16, 16))
    y = S.conv5_forward(x, wt)
    lhs = np.sum(dy * y)
    assert abs(np.sum(S.conv5_backward_data(dy, wt) * x) - lhs) <= 1e-11 * abs(lhs)
    dw, db = S.conv5_backward_weights(x, dy)
    assert abs(np.sum(dw * wt) - lhs) <= 1e-11 * abs(lhs)
    assert np.allclose(db, dy.sum(axis=(0, 2, 3)))


def test_network_gradient_equals_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 2, 16, 16))
    layers = [(rng.normal(size=(4, 2, 5, 5)) * 0.3, rng.normal(size=4) * 0.1),
              (rng.normal(size=(4, 4, 5, 5)) * 0.3, rng.normal(size=4) * 0.1),
              (rng.normal(size=(1, 4, 5, 5)) * 0.3, rng.normal(size=1) * 0.1)]
    target = rng.normal(size=(1, 1, 16, 16))

    def loss(ls):
        out = S.network_forward(x, ls)[-1]
        return 0.5 * np.sum((out - target) ** 2)

    acts = S.network_forward(x, layers)
    grads = S.network_backward(acts, layers, acts[-1] - target)
    for k, idx in ((0, (1, 0, 2, 3)), (1, (2, 3, 0, 4)), (2, (0, 1, 2, 2))):
        h = 1e-6
        plus = [(w.copy(), b.copy()) for w, b in layers]
        minus = [(w.copy(), b.copy()) for w, b in layers]
        plus[k][0][idx] += h
        minus[k][0][idx] -= h
        fd = (loss(plus) - loss(minus)) / (2 * h)
        assert abs(fd - grads[k][0][idx]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grads[k][0][idx])
        plus = [(w.copy(), b.copy()) for w, b in layers]
        minus = [(w.copy(), b.copy()) for w, b in layers]
        plus[k][1][0] += h
        minus[k][1][0] -= h
        fd = (loss(plus) - loss(minus)) / (2 * h)
        assert abs(fd - grads[k][1][0]) <= 1e-5 * max(1.0, abs(fd))


def test_resampling_matrix_interpolates_linear_functions_and_keeps_constants():
    A = S.resample_matrix(96, 128)
    assert np.allclose(A.sum(axis=1), 1.0)
    t = np.arange(128.0)
    f = 3.0 + 0.25 * t
    got = A @ f
    u = np.arange(96) * 128 / 96
    inside = u < 127  # away from the periodic seam
    assert np.allclose(got[inside], 3.0 + 0.25 * u[inside], atol=1e-12)
    x = np.random.default_rng(4).normal(size=(2, 128, 128))
    assert np.allclose(S.separable_map(x, A)[0], A @ x[0] @ A.T)


def test_fourier_resampling_reproduces_band_limited_signals():
    m, n = 96, 128
    A = S.fourier_resample_matrix(m, n)
    t, u = np.arange(n) / n, np.arange(m) / m
    rng = np.random.default_rng(7)
    for k in (0, 1, 5, 31, 47):  # every frequency below the 96-point Nyquist limit comes through exactly
        ph = rng.uniform(0, 2 * np.pi)
        assert np.abs(A @ np.cos(2 * np.pi * k * t + ph) - np.cos(2 * np.pi * k * u + ph)).max() <= 1e-12
    # the 96-point Nyquist cosine keeps half its amplitude (mean of the +48 and -48 bins); frequencies above vanish
    assert np.abs(A @ np.cos(2 * np.pi * 48 * t) - 0.5 * np.cos(2 * np.pi * 48 * u)).max() <= 1e-12
    for k in (49, 63, 64):
        assert np.abs(A @ np.cos(2 * np.pi * k * t) - 0.0).max() <= 1e-12 or k % m == 0
    # odd m: no Nyquist bin
    B = S.fourier_resample_matrix(45, 60)
    assert np.abs(B @ np.sin(2 * np.pi * 22 * np.arange(60) / 60) - np.sin(2 * np.pi * 22 * np.arange(45) / 45)).max() \
        <= 1e-12


def test_product_fourier_resampling_equals_the_oracle():
    from simulatokit_amd.surrogate import fourier_resample_matrix

    for m, n in ((96, 128), (48, 64), (45, 60), (24, 32)):
        assert np.abs(fourier_resample_matrix(m, n) - S.fourier_resample_matrix(m, n)).max() <= 1e-13


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.0 ** -133, 3.140625], dtype=np.float32)
    r = S.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway 1 + 2^-8 rounds to even (1.0)
    assert r[2] == 1.0078125 and r[3] == 1.0078125
    assert r[5] == 3.140625
