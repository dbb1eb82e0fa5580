"""Float64 restatement of the HAADF surrogate's layers (test infrastructure only; the reference has no learned part,
so these are pinned by the mathematics in tests/test_oracle_surrogate.py: direct loops against circular convolution
by FFT, backward passes against finite differences, the linear resampling matrix against interpolation of linear
functions, the band-limited one against exact reproduction of trigonometric polynomials).

Shapes: x [batch, cin, h, w], weights [cout, cin, 5, 5], circular padding of 2 on both axes.
"""
from __future__ import annotations

import numpy as np


def conv5_forward(x, wt, bias=None, relu=False):
    """y[b, o] = bias[o] + sum_{c, ky, kx} wt[o, c, ky, kx] x[b, c, (p + (ky - 2, kx - 2)) mod (h, w)]."""
    x = np.asarray(x, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    B, C, H, W = x.shape
    O = wt.shape[0]
    y = np.zeros((B, O, H, W))
    for ky in range(5):
        for kx in range(5):
            shifted = np.roll(x, shift=(-(ky - 2), -(kx - 2)), axis=(2, 3))  # shifted[p] = x[p + (ky - 2, kx - 2)]
            y += np.einsum("oc,bchw->bohw", wt[:, :, ky, kx], shifted)
    if bias is not None:
        y += np.asarray(bias, dtype=np.float64)[None, :, None, None]
    return np.maximum(y, 0.0) if relu else y


def conv5_backward_data(dy, wt):
    """dx = the transpose of conv5_forward applied to dy."""
    dy = np.asarray(dy, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    B, O, H, W = dy.shape
    C = wt.shape[1]
    dx = np.zeros((B, C, H, W))
    for ky in range(5):
        for kx in range(5):
            contrib = np.einsum("oc,bohw->bchw", wt[:, :, ky, kx], dy)
            dx += np.roll(contrib, shift=(ky - 2, kx - 2), axis=(2, 3))
    return dx


def conv5_backward_weights(x, dy):
    """(dw [cout, cin, 5, 5], db [cout]) of sum(dy * conv5_forward(x, w, b))."""
    x = np.asarray(x, dtype=np.float64)
    dy = np.asarray(dy, dtype=np.float64)
    dw = np.zeros((dy.shape[1], x.shape[1], 5, 5))
    for ky in range(5):
        for kx in range(5):
            shifted = np.roll(x, shift=(-(ky - 2), -(kx - 2)), axis=(2, 3))
            dw[:, :, ky, kx] = np.einsum("bohw,bchw->oc", dy, shifted)
    return dw, dy.sum(axis=(0, 2, 3))


def resample_matrix(m, n):
    """The m x n matrix of periodic linear interpolation from n samples to m samples over the same period: output u
    sits at input coordinate u n / m."""
    A = np.zeros((m, n))
    for u in range(m):
        t = u * n / m
        i = int(np.floor(t))
        f = t - i
        A[u, i % n] += 1 - f
        A[u, (i + 1) % n] += f
    return A


def fourier_resample_matrix(m, n):
    """The m x n matrix of band-limited periodic interpolation from n samples to m <= n samples over one period, built
    column by column through the DFT: the n-point spectrum of each unit input, cropped to the frequencies |k| < m / 2
    with the m-point Nyquist bin (even m) taking the mean of the +m/2 and -m/2 bins, back through the m-point inverse
    DFT and scaled by m / n."""
    A = np.zeros((m, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        X = np.fft.fft(e)
        Y = np.zeros(m, dtype=complex)
        half = (m + 1) // 2
        Y[:half] = X[:half]
        if m > 1:
            Y[m - (m - 1) // 2:] = X[n - (m - 1) // 2:]
        if m % 2 == 0:
            Y[m // 2] = 0.5 * (X[m // 2] + X[n - m // 2])
        A[:, i] = np.real(np.fft.ifft(Y)) * m / n
    return A


def separable_map(x, A):
    """out[b, u, v] = sum_{i, j} A[u, i] A[v, j] x[b, i, j]."""
    return np.einsum("ui,bij,vj->buv", A, np.asarray(x, dtype=np.float64), A)


def bf16_round(a):
    """Round to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def network_forward(x, layers):
    """The surrogate's stack: every layer but the last with ReLU; layers = [(wt, b), ...]. Returns every activation."""
    acts = [np.asarray(x, dtype=np.float64)]
    for k, (wt, b) in enumerate(layers):
        acts.append(conv5_forward(acts[-1], wt, b, relu=k < len(layers) - 1))
    return acts


def network_backward(acts, layers, dout):
    """Gradients [(dw, db), ...] of sum(dout * network output), from the activations of network_forward."""
    grads = [None] * len(layers)
    g = np.asarray(dout, dtype=np.float64)
    for k in range(len(layers) - 1, -1, -1):
        if k < len(layers) - 1:
            g = g * (acts[k + 1] > 0)
        grads[k] = conv5_backward_weights(acts[k], g)
        g = conv5_backward_data(g, layers[k][0])
    return grads
