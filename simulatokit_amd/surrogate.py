"""The HAADF surrogate (build-only; the reference has no learned part): a six-layer convolutional network that maps a
cell's summed projected potential (128 x 128) and its thickness to the HAADF image of a 96 x 96 scan over the cell,
trained on images the multislice kernels generate, with plain multislice re-simulating a random 5 % of the positions
as its error judge and standing in for the whole image when that error exceeds the tolerance.

Network: 5 x 5 convolutions with circular padding (the cells are periodic), 2 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1
channels, ReLU between layers, bfloat16 weights, float32 activations and accumulation, all in the library's own HIP
kernels (sk_conv5_forward / _backward_data / _backward_weights, sk_weighted_mse_grad, sk_adam_step); the 128 x 128
output is resampled to the 96 x 96 scan grid by a fixed separable map (sk_separable_map): band-limited (Fourier)
interpolation by default, which is exact for images without spatial frequencies at or above the 96-point grid's
Nyquist limit (a HAADF image is band-limited by the probe at 2 alpha / lambda, 47.8 cycles per 30 Angstrom at
20 mrad and 200 keV), or periodic linear interpolation. Torch only holds device memory.

Input: the cell's summed projected potential on a 128 x 128 grid and a thickness channel (the specified input), or,
as labelled variants, the projected potential summed over each of K equal depth ranges of [0, t_max)
(`depth_slabs` = K > 0), and/or split by element (`by_element`), each beside the thickness channel: the probe spreads
with depth, so the same projected atom scatters differently at the entrance and the exit of the cell, and atoms that
overlap in projection scatter by their atomic numbers, neither of which the summed projection can tell apart.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .core import AtomicStructure, ElectronBeam, ExecutionOptions
from .detector import STEM, AnnularDetector
from .multislice import Potential, _raise
from .probe import Aperture, GridScan, Probe
from .workloads import random_cell

CHANNELS = [2, 32, 32, 32, 32, 32, 1]
WGRAD_WORKSPACE_FLOATS = 512 * 25 * 32 * 32 + 2 * 32 * 64  # SK_CONV5_MFMA_WORKSPACE_FLOATS
INPUT_N = 128
SCAN_N = 96
ELEMENTS = (6, 8, 14, 22, 38, 79)


# ---- the cells and their images --------------------------------------------------------------------------------------
@dataclass(frozen=True)
class SurrogateWorkload:
    """Seeded random cells (extent x extent, thickness uniform in [t_min, t_max], atom count uniform in
    [n_min, n_max], Z from `elements`, 1.2 Angstrom minimum separation) scanned as the HAADF workload is: 200 keV,
    20 mrad soft aperture, annular 50-200 mrad, `grid` x `grid` waves, 2 Angstrom slices, a scan_n x scan_n scan."""

    extent: float = 30.0
    t_min: float = 20.0
    t_max: float = 100.0
    n_min: int = 200
    n_max: int = 800
    elements: Tuple[int, ...] = ELEMENTS
    min_separation: float = 1.2
    grid: int = 512
    scan_n: int = SCAN_N
    input_n: int = INPUT_N
    energy: float = 200e3
    semiangle: float = 20.0
    inner: float = 50.0
    outer: float = 200.0
    slice_thickness: float = 2.0

    def cell(self, seed: int) -> AtomicStructure:
        rng = np.random.default_rng([seed, 0x5EED])
        t = float(rng.uniform(self.t_min, self.t_max))
        n = int(rng.integers(self.n_min, self.n_max + 1))
        return random_cell(seed, self.extent, t, n, self.elements, min_separation=self.min_separation)

    def scan(self) -> GridScan:
        return GridScan((0.0, 0.0), (self.extent, self.extent), self.scan_n, self.scan_n)

    def probe(self, potential: Potential) -> Probe:
        return Probe(ElectronBeam(self.energy), potential.grid, Aperture(self.semiangle))

    def potential(self, st: AtomicStructure) -> Potential:
        return Potential(st, self.grid, self.grid, slice_thickness=self.slice_thickness)

    def haadf(self, st: AtomicStructure, positions: Optional[np.ndarray] = None,
              options: Optional[ExecutionOptions] = None) -> np.ndarray:
        """Plain multislice HAADF at `positions` (default: the whole scan, as a scan_n x scan_n image), with the
        library's default execution plan unless `options` names another."""
        pot = self.potential(st)
        probe = self.probe(pot)
        prepared = pot.prepare(probe.beam)
        scan = self.scan() if positions is None else positions
        out = STEM.scan(probe, [prepared], scan, [AnnularDetector(self.inner, self.outer)],
                        options or ExecutionOptions())[0].values
        prepared.release()
        return out.reshape(self.scan_n, self.scan_n) if positions is None else out

    def features(self, st: AtomicStructure, depth_slabs: int = 0, by_element: bool = False) -> np.ndarray:
        """The network's input: the summed projected potential on an input_n^2 grid and a thickness channel; with
        depth_slabs = K > 0 the projected potential summed over each of K equal depth ranges of [0, t_max) (a slice
        counts in the range that holds its centre) instead of the one sum; with by_element, those channels once per
        element of `elements` (the potential of that element's atoms alone)."""
        groups = [st.numbers == z for z in self.elements] if by_element else [np.ones(len(st.numbers), bool)]
        chans = []
        for sel in groups:
            if not sel.any():
                chans += [np.zeros((self.input_n, self.input_n))] * max(1, depth_slabs)
                continue
            sub = AtomicStructure(st.box, st.numbers[sel], st.positions[sel])
            pot = Potential(sub, self.input_n, self.input_n, slice_thickness=self.slice_thickness)
            prepared = pot.prepare(ElectronBeam(self.energy))
            V = prepared.values()
            prepared.release()
            if depth_slabs <= 0:
                chans.append(V.sum(axis=0))
                continue
            edges = np.cumsum(pot.slice_thicknesses)
            centres = edges - 0.5 * np.asarray(pot.slice_thicknesses)
            slab = np.minimum((centres / (self.t_max / depth_slabs)).astype(int), depth_slabs - 1)
            out = np.zeros((depth_slabs,) + V.shape[1:])
            for j, k in enumerate(slab):
                out[k] += V[j]
            chans += list(out)
        chans.append(np.full((self.input_n, self.input_n), st.box[2]))
        return np.stack(chans).astype(np.float32)


@dataclass
class Dataset:
    seeds: List[int]
    x: np.ndarray   # [n, 2, input_n, input_n] float32 (summed V in V*Angstrom, thickness in Angstrom)
    y: np.ndarray   # [n, scan_n, scan_n] float32 HAADF (fraction of the incident intensity)


def input_channels(depth_slabs: int = 0, by_element: bool = False, elements: Sequence[int] = ELEMENTS) -> int:
    """Channels of the network's input: potential channels (max(1, K) per element group) and the thickness."""
    return (len(elements) if by_element else 1) * max(1, depth_slabs) + 1


def make_dataset(work: SurrogateWorkload, seeds: Sequence[int], log=None, depth_slabs: int = 0,
                 labels: Optional[np.ndarray] = None, by_element: bool = False) -> Dataset:
    """The cells of `seeds`: their features and their multislice HAADF images (or the given `labels`, images made
    earlier from the same seeds)."""
    xs, ys = [], []
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        st = work.cell(int(s))
        xs.append(work.features(st, depth_slabs, by_element))
        ys.append(work.haadf(st).astype(np.float32) if labels is None else np.asarray(labels[k], np.float32))
        if log and (k + 1) % max(1, len(seeds) // 20) == 0:
            log(f"cells {k + 1}/{len(seeds)} in {time.perf_counter() - t0:.1f} s")
    return Dataset(list(seeds), np.stack(xs), np.stack(ys))


# ---- the network ----------------------------------------------------------------------------------------------------
def _check(rc, what):
    _raise(rc, what)


def _bind():
    L = N.lib()
    if getattr(L, "_surrogate_bound", False):
        return L
    vp, ll = C.c_void_p, C.c_longlong
    L.sk_conv5_forward.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp]
    L.sk_conv5_backward_data.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, vp, vp]
    L.sk_conv5_backward_weights.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp]
    L.sk_relu_backward.argtypes = [vp, vp, ll, vp]
    L.sk_separable_map.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, vp, vp, vp]
    L.sk_mse_grad.argtypes = [vp, vp, ll, vp, vp, vp]
    L.sk_weighted_mse_grad.argtypes = [vp, vp, vp, ll, ll, vp, vp, vp]
    L.sk_adam_step.argtypes = [vp, vp, vp, vp, vp, ll, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, vp]
    L.sk_to_bf16.argtypes = [vp, vp, ll, vp]
    L.sk_conv5_weights_tap_major.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp]
    L.sk_conv5_forward_tap_major.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp]
    L.sk_conv5_backward_weights_mfma.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
    L._surrogate_bound = True
    return L


def fourier_resample_matrix(m: int, n: int) -> np.ndarray:
    """Band-limited periodic interpolation from n samples to m <= n samples over one period (output u at input
    coordinate u n / m): the n-point DFT cropped to the m-point grid's frequencies |k| < m / 2, with the m-point
    Nyquist frequency (even m) kept as a cosine, then the inverse m-point DFT. Real; exact for periodic signals with
    no frequency at or above m / 2 cycles per period."""
    if m > n:
        raise ValueError("fourier_resample_matrix: m must not exceed n")
    d = np.arange(m)[:, None] / m - np.arange(n)[None, :] / n  # output position minus input position, in periods
    A = np.ones((m, n))
    for k in range(1, (m + 1) // 2):
        A += 2.0 * np.cos(2.0 * np.pi * k * d)
    if m % 2 == 0:
        A += np.cos(np.pi * m * d)
    return A / n


def resample_matrix(m: int, n: int) -> np.ndarray:
    """Periodic linear interpolation from n to m samples over one period (output u at input coordinate u n / m)."""
    A = np.zeros((m, n), dtype=np.float64)
    for u in range(m):
        t = u * n / m
        i = int(math.floor(t))
        f = t - i
        A[u, i % n] += 1 - f
        A[u, (i + 1) % n] += f
    return A


class HaadfSurrogate:
    """The network on the device. Weights: float32 masters (trained) and their bfloat16 rounding (used).

    With matrix_cores true, an input size that is a multiple of 32 and at most 32 channels a layer, every layer runs
    on the matrix cores: the forward pass and the backward pass over the data through sk_conv5_forward_tap_major (from
    tap-major copies of the bfloat16 weights, refreshed after every update), the gradient of the weights through
    sk_conv5_backward_weights_mfma. Otherwise, or with matrix_cores=False, on the vector units (sk_conv5_forward,
    sk_conv5_backward_data, sk_conv5_backward_weights)."""

    def __init__(self, seed: int = 0, channels: Sequence[int] = CHANNELS, input_n: int = INPUT_N,
                 output_n: int = SCAN_N, resample: str = "fourier", depth_slabs: int = 0, by_element: bool = False,
                 matrix_cores: bool = True):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.channels = list(channels)
        self.input_n, self.output_n = input_n, output_n
        if resample not in ("fourier", "linear"):
            raise ValueError("resample must be 'fourier' or 'linear'")
        self.resample, self.depth_slabs, self.by_element = resample, int(depth_slabs), bool(by_element)
        want = input_channels(self.depth_slabs, self.by_element)
        if self.channels[0] != want:
            raise ValueError(f"the first layer takes {want} input channels with depth_slabs={self.depth_slabs}, "
                             f"by_element={self.by_element}")
        rng = np.random.default_rng(seed)
        self.w32, self.wbf, self.b = [], [], []
        for k, (ci, co) in enumerate(zip(self.channels[:-1], self.channels[1:])):
            std = math.sqrt(2.0 / (ci * 25)) * (0.1 if k == len(self.channels) - 2 else 1.0)
            w = rng.normal(0.0, std, size=(co, ci, 5, 5)).astype(np.float32)
            self.w32.append(torch.tensor(w, device=self.dev))
            self.wbf.append(torch.empty(w.size, dtype=torch.int16, device=self.dev))
            self.b.append(torch.zeros(co, dtype=torch.float32, device=self.dev))
        A = fourier_resample_matrix(output_n, input_n) if resample == "fourier" else resample_matrix(output_n, input_n)
        self.A = torch.tensor(A, dtype=torch.float32, device=self.dev)
        self.At = self.A.t().contiguous()
        self.norm = {"v": 1.0, "t": 100.0, "y": 1.0}
        # tap-major bfloat16 copies of every layer: [forward, flipped for the backward pass over the data]
        self.tap_major, self._wgrad_ws = {}, None
        if matrix_cores and input_n % 32 == 0 and max(self.channels) <= 32:
            for k in range(len(self.w32)):
                self.tap_major[k] = [torch.empty(25 * 32 * 32, dtype=torch.int16, device=self.dev) for _ in range(2)]
        self._publish()

    # -- plumbing
    @property
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def _publish(self):
        L = _bind()
        for w, wb in zip(self.w32, self.wbf):
            _check(L.sk_to_bf16(w.data_ptr(), wb.data_ptr(), w.numel(), self._stream), "to_bf16")
        self._refresh_tap_major()

    def _refresh_tap_major(self):
        L = _bind()
        for k, (fwd, bwd) in self.tap_major.items():
            ci, co = self.channels[k], self.channels[k + 1]
            for flip, out in ((0, fwd), (1, bwd)):
                _check(L.sk_conv5_weights_tap_major(self.wbf[k].data_ptr(), ci, co, flip, out.data_ptr(),
                                                    self._stream), "conv5_weights_tap_major")

    def parameters(self) -> List:
        return [t for pair in zip(self.w32, self.b) for t in pair]

    # -- normalisation
    def fit_normalisation(self, data: Dataset):
        self.norm = {"v": float(np.std(data.x[:, :-1])) or 1.0, "t": 100.0, "y": float(np.mean(data.y)) or 1.0}

    def _inputs(self, x: np.ndarray):
        x = np.array(x, dtype=np.float32, copy=True)
        x[:, :-1] /= self.norm["v"]  # the potential channels
        x[:, -1] /= self.norm["t"]   # the thickness channel
        return self.torch.tensor(x, device=self.dev)

    # -- forward / backward
    def forward(self, x, keep: bool = False):
        """x [B, 2, n, n] on the device -> the output_n^2 image [B, output_n, output_n] (normalised units);
        keep=True also returns every activation (for backward)."""
        torch, L = self.torch, _bind()
        B = x.shape[0]
        acts = [x]
        for k in range(len(self.w32)):
            ci, co = self.channels[k], self.channels[k + 1]
            y = torch.empty((B, co, self.input_n, self.input_n), dtype=torch.float32, device=self.dev)
            relu = 1 if k < len(self.w32) - 1 else 0
            if k in self.tap_major:
                _check(L.sk_conv5_forward_tap_major(acts[-1].data_ptr(), B, ci, self.input_n, self.input_n,
                                                    self.tap_major[k][0].data_ptr(), self.b[k].data_ptr(), co, relu,
                                                    y.data_ptr(), self._stream), "conv5_forward_tap_major")
            else:
                _check(L.sk_conv5_forward(acts[-1].data_ptr(), B, ci, self.input_n, self.input_n,
                                          self.wbf[k].data_ptr(), self.b[k].data_ptr(), co, relu, y.data_ptr(),
                                          self._stream), "conv5_forward")
            acts.append(y)
        tmp = torch.empty((B, self.input_n, self.output_n), dtype=torch.float32, device=self.dev)
        out = torch.empty((B, self.output_n, self.output_n), dtype=torch.float32, device=self.dev)
        _check(L.sk_separable_map(acts[-1].data_ptr(), B, self.input_n, self.A.data_ptr(), self.output_n,
                                  tmp.data_ptr(), out.data_ptr(), self._stream), "separable_map")
        return (out, acts) if keep else out

    def backward(self, acts, dout):
        """Gradients [(dw, db), ...] of sum(dout * output) from the kept activations."""
        torch, L = self.torch, _bind()
        B = dout.shape[0]
        tmp = torch.empty((B, self.output_n, self.input_n), dtype=torch.float32, device=self.dev)
        g = torch.empty((B, 1, self.input_n, self.input_n), dtype=torch.float32, device=self.dev)
        _check(L.sk_separable_map(dout.data_ptr(), B, self.output_n, self.At.data_ptr(), self.input_n,
                                  tmp.data_ptr(), g.data_ptr(), self._stream), "separable_map")
        grads = [None] * len(self.w32)
        for k in range(len(self.w32) - 1, -1, -1):
            ci, co = self.channels[k], self.channels[k + 1]
            if k < len(self.w32) - 1:
                _check(L.sk_relu_backward(g.data_ptr(), acts[k + 1].data_ptr(), g.numel(), self._stream), "relu")
            dw = torch.empty((co, ci, 5, 5), dtype=torch.float32, device=self.dev)
            db = torch.empty(co, dtype=torch.float32, device=self.dev)
            if k in self.tap_major:
                if self._wgrad_ws is None:
                    self._wgrad_ws = torch.empty(WGRAD_WORKSPACE_FLOATS, dtype=torch.float32, device=self.dev)
                _check(L.sk_conv5_backward_weights_mfma(acts[k].data_ptr(), g.data_ptr(), B, ci, self.input_n,
                                                        self.input_n, co, dw.data_ptr(), db.data_ptr(),
                                                        self._wgrad_ws.data_ptr(), self._stream),
                       "backward_weights_mfma")
            else:
                _check(L.sk_conv5_backward_weights(acts[k].data_ptr(), g.data_ptr(), B, ci, self.input_n,
                                                   self.input_n, co, dw.data_ptr(), db.data_ptr(), self._stream),
                       "backward_weights")
            grads[k] = (dw, db)
            if k > 0:
                gx = torch.empty((B, ci, self.input_n, self.input_n), dtype=torch.float32, device=self.dev)
                if k in self.tap_major:  # the forward convolution of dy with the flipped, transposed kernel
                    _check(L.sk_conv5_forward_tap_major(g.data_ptr(), B, co, self.input_n, self.input_n,
                                                        self.tap_major[k][1].data_ptr(), None, ci, 0, gx.data_ptr(),
                                                        self._stream), "conv5_forward_tap_major")
                else:
                    scratch = torch.empty(co * ci * 25, dtype=torch.int16, device=self.dev)
                    _check(L.sk_conv5_backward_data(g.data_ptr(), B, ci, self.input_n, self.input_n,
                                                    self.wbf[k].data_ptr(), co, gx.data_ptr(), scratch.data_ptr(),
                                                    self._stream), "backward_data")
                g = gx
        return grads

    # -- augmentation: the exact symmetries of a periodic cell scanned on a grid that starts at the origin
    def _augment(self, xb, yb, rng):
        """A random one of the 8 rotations/reflections and a random circular shift by a whole number of `q` input
        pixels (= `p` output pixels, the shifts that map both grids onto themselves), applied to inputs and labels."""
        torch = self.torch
        n_in, n_out = self.input_n, self.output_n
        g = math.gcd(n_in, n_out)
        step_in, step_out = n_in // g, n_out // g
        k = int(rng.integers(8))
        sx, sy = (int(v) for v in rng.integers(g, size=2))

        def transform(a, n, step):
            i = torch.arange(n, device=self.dev)
            neg = (-i) % n
            if k & 1:  # reflection x -> -x
                a = a.index_select(-2, neg)
            for _ in range(k >> 1):  # rotation by 90 degrees about the origin: out[i][j] = in[-j][i]
                a = a.index_select(-1, neg).transpose(-1, -2)
            return torch.roll(a, shifts=(sx * step, sy * step), dims=(-2, -1))

        return transform(xb, n_in, step_in).contiguous(), transform(yb, n_out, step_out).contiguous()

    # -- training
    def train(self, data: Dataset, epochs: int = 60, batch: int = 16, lr: float = 2e-3, seed: int = 0,
              log=None, augment: bool = True, loss: str = "relative") -> List[float]:
        """Adam on the squared error of the normalised images; loss="relative" weights each image by the inverse of
        its squared norm (scaled to a mean weight of one), so that thin cells, whose HAADF images are faint, count as
        much as thick ones by their relative error; loss="mse" weights every pixel alike."""
        torch, L = self.torch, _bind()
        if loss not in ("relative", "mse"):
            raise ValueError("loss must be 'relative' or 'mse'")
        self.fit_normalisation(data)
        X = self._inputs(data.x)
        Y = torch.tensor(data.y / self.norm["y"], dtype=torch.float32, device=self.dev)
        norms2 = np.sum(np.asarray(data.y, np.float64).reshape(len(data.y), -1) ** 2, axis=1)
        weights = np.mean(norms2) / np.maximum(norms2, 1e-300) if loss == "relative" else np.ones(len(norms2))
        Wimg = torch.tensor(weights, dtype=torch.float32, device=self.dev)
        m = [torch.zeros_like(p) for p in self.parameters()]
        v = [torch.zeros_like(p) for p in self.parameters()]
        rng = np.random.default_rng(seed)
        step, history = 0, []
        n = X.shape[0]
        for ep in range(epochs):
            order = rng.permutation(n)
            total, count = 0.0, 0
            rate = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))  # cosine schedule
            for s0 in range(0, n - batch + 1, batch):
                idx = torch.tensor(order[s0:s0 + batch], device=self.dev)
                xb, yb = X.index_select(0, idx).contiguous(), Y.index_select(0, idx).contiguous()
                wb = Wimg.index_select(0, idx).contiguous()
                if augment:
                    xb, yb = self._augment(xb, yb, rng)
                out, acts = self.forward(xb, keep=True)
                g = torch.empty_like(out)
                nb = out.numel()
                partial = torch.empty((nb + 255) // 256, dtype=torch.float64, device=self.dev)
                _check(L.sk_weighted_mse_grad(out.data_ptr(), yb.data_ptr(), wb.data_ptr(), nb // len(idx), nb,
                                              g.data_ptr(), partial.data_ptr(), self._stream), "mse_grad")
                grads = self.backward(acts, g)
                step += 1
                for k, (dw, db) in enumerate(grads):
                    for p, gp, mm, vv, pub in ((self.w32[k], dw, m[2 * k], v[2 * k], self.wbf[k]),
                                               (self.b[k], db, m[2 * k + 1], v[2 * k + 1], None)):
                        _check(L.sk_adam_step(p.data_ptr(), gp.data_ptr(), mm.data_ptr(), vv.data_ptr(),
                                              pub.data_ptr() if pub is not None else None, p.numel(), rate, 0.9,
                                              0.999, 1e-8, step, self._stream), "adam")
                self._refresh_tap_major()
                total += float(partial.sum().item())
                count += nb
            history.append(total / max(1, count))
            if log:
                log(f"epoch {ep + 1}/{epochs}: mse {history[-1]:.4e}")
        torch.cuda.synchronize()
        return history

    # -- inference
    def predict(self, x: np.ndarray) -> np.ndarray:
        """HAADF images [B, output_n, output_n] (fraction of the incident intensity) for features x [B, 2, n, n]."""
        out = self.forward(self._inputs(x))
        self.torch.cuda.synchronize()
        return out.cpu().numpy() * self.norm["y"]

    # -- persistence
    def state(self, compact: bool = False) -> Dict[str, np.ndarray]:
        """The weights (float32 masters, or with compact=True only the bfloat16 bits the network uses), biases,
        normalisation, channel list, resampling and depth slabs."""
        if compact:
            d = {f"wbf{k}": w.cpu().numpy().view(np.uint16) for k, w in enumerate(self.wbf)}
        else:
            d = {f"w{k}": w.cpu().numpy() for k, w in enumerate(self.w32)}
        d.update({f"b{k}": b.cpu().numpy() for k, b in enumerate(self.b)})
        d["norm"] = np.array([self.norm["v"], self.norm["t"], self.norm["y"]])
        d["channels"] = np.array(self.channels)
        d["resample"] = np.array(self.resample)
        d["depth_slabs"] = np.array(self.depth_slabs)
        d["by_element"] = np.array(int(self.by_element))
        d["io"] = np.array([self.input_n, self.output_n])
        return d

    def save(self, path: str, compact: bool = False):
        np.savez(path, **self.state(compact))

    @classmethod
    def from_file(cls, path: str) -> "HaadfSurrogate":
        """A network with the shape, resampling and weights saved in `path`."""
        d = np.load(path, allow_pickle=False)
        io = d["io"] if "io" in d.files else np.array([INPUT_N, SCAN_N])
        model = cls(channels=[int(c) for c in d["channels"]], input_n=int(io[0]), output_n=int(io[1]),
                    resample=str(d["resample"]) if "resample" in d.files else "linear",
                    depth_slabs=int(d["depth_slabs"]) if "depth_slabs" in d.files else 0,
                    by_element=bool(int(d["by_element"])) if "by_element" in d.files else False)
        model.load(path)
        return model

    def load(self, path: str):
        d = np.load(path, allow_pickle=False)
        torch = self.torch
        for k in range(len(self.w32)):
            if f"w{k}" in d.files:
                self.w32[k].copy_(torch.tensor(d[f"w{k}"], device=self.dev))
            else:  # compact file: the masters become the bfloat16 values themselves
                bits = d[f"wbf{k}"].astype(np.uint32) << 16
                self.w32[k].copy_(torch.tensor(bits.view(np.float32).reshape(self.w32[k].shape), device=self.dev))
            self.b[k].copy_(torch.tensor(d[f"b{k}"], device=self.dev))
        nv, nt, ny = d["norm"]
        self.norm = {"v": float(nv), "t": float(nt), "y": float(ny)}
        self._publish()


# ---- the surrogate scan with its multislice judge ---------------------------------------------------------------------
@dataclass
class JudgedScan:
    image: np.ndarray          # the delivered HAADF image (surrogate, or multislice after a fallback)
    surrogate: np.ndarray      # the network's image
    checked: np.ndarray        # flat indices of the re-simulated positions
    spot_error: float          # relative L2 error of the network on the re-simulated positions
    fallback: bool
    seconds: Dict[str, float] = field(default_factory=dict)


def judged_scan(model: HaadfSurrogate, work: SurrogateWorkload, st: AtomicStructure, fraction: float = 0.05,
                tolerance: float = 0.02, seed: int = 0) -> JudgedScan:
    """The surrogate's HAADF image of cell `st`, judged by plain multislice on a random `fraction` of the positions;
    when the network's relative L2 error there exceeds `tolerance`, the whole image is re-simulated."""
    t0 = time.perf_counter()
    img = model.predict(work.features(st, model.depth_slabs, model.by_element)[None])[0]
    t1 = time.perf_counter()
    npos = work.scan_n * work.scan_n
    k = max(1, int(round(fraction * npos)))
    checked = np.sort(np.random.default_rng(seed).choice(npos, size=k, replace=False))
    positions = work.scan().positions[checked]
    truth = work.haadf(st, positions)
    t2 = time.perf_counter()
    pred = img.reshape(-1)[checked]
    err = float(np.linalg.norm(pred - truth) / max(np.linalg.norm(truth), 1e-300))
    fallback = err > tolerance
    final = img
    if fallback:
        final = work.haadf(st)
    t3 = time.perf_counter()
    return JudgedScan(final, img, checked, err, fallback,
                      {"surrogate": t1 - t0, "judge": t2 - t1, "fallback": t3 - t2, "total": t3 - t0})


__all__ = ["SurrogateWorkload", "Dataset", "make_dataset", "HaadfSurrogate", "JudgedScan", "judged_scan",
           "resample_matrix", "fourier_resample_matrix", "generate_labels", "evaluate", "run_study", "input_channels",
           "CHANNELS"]


# ---- the study: train on generated cells, judge on held-out cells -----------------------------------------------------
def generate_labels(work: SurrogateWorkload, seeds: Sequence[int], log=None) -> Tuple[np.ndarray, np.ndarray]:
    """The multislice HAADF images [len(seeds), scan_n, scan_n] of the seeded cells (the training labels) and the
    cells' thicknesses."""
    out = np.empty((len(seeds), work.scan_n, work.scan_n), dtype=np.float32)
    thick = np.empty(len(seeds))
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        st = work.cell(int(s))
        thick[k] = st.box[2]
        out[k] = work.haadf(st)
        if log and (k + 1) % max(1, len(seeds) // 40) == 0:
            log(f"labels {k + 1}/{len(seeds)} in {time.perf_counter() - t0:.1f} s")
    return out, thick


def evaluate(model: HaadfSurrogate, work: SurrogateWorkload, seeds: Sequence[int], tolerance: float = 0.02,
             fraction: float = 0.05, log=print) -> Dict:
    """Scans each held-out cell both ways, plain multislice with the library's default plan (the truth and the
    baseline time) and the judged surrogate, and returns the errors, fallbacks and throughputs."""
    import torch

    rows = []
    for s in seeds:
        st = work.cell(int(s))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        truth = work.haadf(st)
        plain_s = time.perf_counter() - t0
        js = judged_scan(model, work, st, fraction=fraction, tolerance=tolerance, seed=int(s))
        err_net = float(np.linalg.norm(js.surrogate - truth) / np.linalg.norm(truth))
        err_out = float(np.linalg.norm(js.image - truth) / np.linalg.norm(truth))
        slices = math.ceil(st.box[2] / work.slice_thickness)
        rows.append({"seed": int(s), "thickness": float(st.box[2]), "atoms": len(st.numbers), "slices": slices,
                     "plain_s": plain_s, "judged_s": js.seconds["total"],
                     "surrogate_and_judge_s": js.seconds["surrogate"] + js.seconds["judge"],
                     "surrogate_err": err_net, "delivered_err": err_out, "spot_err": js.spot_error,
                     "fallback": bool(js.fallback)})
        if log:
            log(f"held-out {s} ({st.box[2]:.0f} A): network {err_net:.4f}, spot {js.spot_error:.4f}, fallback "
                f"{js.fallback}, plain {plain_s * 1e3:.0f} ms, judged {js.seconds['total'] * 1e3:.0f} ms")
    npos = work.scan_n ** 2
    work_ps = sum(npos * r["slices"] for r in rows)
    plain_rate = work_ps / sum(r["plain_s"] for r in rows)
    judged_rate = work_ps / sum(r["judged_s"] for r in rows)
    accepted_rate = work_ps / sum(r["surrogate_and_judge_s"] for r in rows)
    errs = np.array([r["surrogate_err"] for r in rows])
    return {"held_out_cells": len(rows), "tolerance": tolerance, "fraction": fraction,
            "network_err_mean": float(errs.mean()), "network_err_median": float(np.median(errs)),
            "network_err_min": float(errs.min()), "network_err_max": float(errs.max()),
            "accepted": int(sum(not r["fallback"] for r in rows)),
            "delivered_err_max": float(np.max([r["delivered_err"] for r in rows])),
            "delivered_err_mean": float(np.mean([r["delivered_err"] for r in rows])),
            "fallbacks": int(sum(r["fallback"] for r in rows)),
            "plain_pos_slices_per_s": plain_rate, "judged_pos_slices_per_s": judged_rate,
            "speedup": judged_rate / plain_rate,
            "accepted_pos_slices_per_s": accepted_rate, "speedup_when_accepted": accepted_rate / plain_rate,
            "rows": rows}


def run_study(train: int = 256, held_out: int = 32, epochs: int = 60, batch: int = 16, seed: int = 20261202,
              work: SurrogateWorkload = SurrogateWorkload(), tolerance: float = 0.02, fraction: float = 0.05,
              weights: Optional[str] = None, log=print, lr: float = 2e-3, depth_slabs: int = 0) -> Dict:
    """Generates `train` + `held_out` seeded cells with the multislice kernels, trains the network on the first, and
    scans the others both ways: plain multislice (the truth and the baseline time) and the judged surrogate. Returns
    the errors, fallback count and the speed-up."""
    seeds = [seed + i for i in range(train + held_out)]
    t0 = time.perf_counter()
    data = make_dataset(work, seeds[:train], log=log, depth_slabs=depth_slabs)
    gen_s = time.perf_counter() - t0
    channels = [input_channels(depth_slabs)] + CHANNELS[1:]
    model = HaadfSurrogate(seed=seed, channels=channels, depth_slabs=depth_slabs)
    t0 = time.perf_counter()
    history = model.train(data, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    train_s = time.perf_counter() - t0
    if weights:
        model.save(weights)
    res = evaluate(model, work, seeds[train:], tolerance=tolerance, fraction=fraction, log=log)
    res.update({"train_cells": train, "epochs": epochs, "generate_s": gen_s, "train_s": train_s,
                "final_loss": history[-1], "first_loss": history[0], "depth_slabs": depth_slabs})
    return res


def _main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="the HAADF surrogate: generate labels, train, evaluate (or all three)")
    sub = ap.add_subparsers(dest="cmd")
    g = sub.add_parser("generate", help="multislice HAADF labels of seeded cells")
    g.add_argument("--first", type=int, default=20261202, help="seed of the first cell")
    g.add_argument("--count", type=int, required=True)
    g.add_argument("--out", required=True, help=".npz: seeds, thickness, labels")
    g.add_argument("--dtype", default="float32", choices=["float32", "float16"])
    t = sub.add_parser("train", help="train on label files written by `generate`")
    t.add_argument("--labels", nargs="+", required=True)
    t.add_argument("--count", type=int, default=0, help="use the first COUNT cells (0: all)")
    t.add_argument("--epochs", type=int, default=150)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--loss", default="relative", choices=["relative", "mse"])
    t.add_argument("--resample", default="fourier", choices=["fourier", "linear"])
    t.add_argument("--depth-slabs", type=int, default=0)
    t.add_argument("--by-element", action="store_true", help="potential channels per element (a labelled variant)")
    t.add_argument("--seed", type=int, default=20261202)
    t.add_argument("--weights", required=True, help="output .npz")
    t.add_argument("--compact", action="store_true", help="save only the bfloat16 weights the network uses")
    e = sub.add_parser("evaluate", help="judge held-out cells with saved weights")
    e.add_argument("--weights", required=True)
    e.add_argument("--first", type=int, default=20261202 + 2000)
    e.add_argument("--count", type=int, default=256)
    e.add_argument("--tolerance", type=float, default=0.02)
    e.add_argument("--fraction", type=float, default=0.05)
    e.add_argument("--out", default=None, help="write the summary JSON here")
    st = sub.add_parser("study", help="generate, train and evaluate in one run")
    st.add_argument("--train", type=int, default=256)
    st.add_argument("--held-out", type=int, default=32)
    st.add_argument("--epochs", type=int, default=60)
    st.add_argument("--batch", type=int, default=16)
    st.add_argument("--seed", type=int, default=20261202)
    st.add_argument("--lr", type=float, default=2e-3)
    st.add_argument("--depth-slabs", type=int, default=0)
    st.add_argument("--weights", default=None)
    st.add_argument("--out", default=None, help="write the summary JSON here")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    work = SurrogateWorkload()
    if a.cmd == "generate":
        seeds = list(range(a.first, a.first + a.count))
        t0 = time.perf_counter()
        y, thick = generate_labels(work, seeds, log=log)
        np.savez(a.out, seeds=np.array(seeds), thickness=thick, labels=y.astype(a.dtype))
        log(json.dumps({"cells": len(seeds), "seconds": time.perf_counter() - t0, "out": a.out}))
    elif a.cmd == "train":
        seeds, labels = [], []
        for f in a.labels:
            d = np.load(f, allow_pickle=False)
            seeds += [int(v) for v in d["seeds"]]
            labels.append(d["labels"].astype(np.float32))
        labels = np.concatenate(labels)
        if a.count:
            seeds, labels = seeds[:a.count], labels[:a.count]
        t0 = time.perf_counter()
        data = make_dataset(work, seeds, log=log, depth_slabs=a.depth_slabs, labels=labels, by_element=a.by_element)
        log(f"features of {len(seeds)} cells in {time.perf_counter() - t0:.1f} s")
        channels = [input_channels(a.depth_slabs, a.by_element)] + CHANNELS[1:]
        model = HaadfSurrogate(seed=a.seed, channels=channels, resample=a.resample, depth_slabs=a.depth_slabs,
                               by_element=a.by_element)
        t0 = time.perf_counter()
        hist = model.train(data, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log, loss=a.loss)
        model.save(a.weights, compact=a.compact)
        log(json.dumps({"cells": len(seeds), "epochs": a.epochs, "train_s": time.perf_counter() - t0,
                        "first_loss": hist[0], "final_loss": hist[-1], "weights": a.weights}))
    elif a.cmd == "evaluate":
        model = HaadfSurrogate.from_file(a.weights)
        res = evaluate(model, work, list(range(a.first, a.first + a.count)), tolerance=a.tolerance,
                       fraction=a.fraction, log=log)
        log(json.dumps({k: v for k, v in res.items() if k != "rows"}))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
    else:
        if a.cmd is None:
            a = ap.parse_args(["study"] + list(argv or []))
        res = run_study(a.train, a.held_out, a.epochs, a.batch, a.seed, weights=a.weights, log=log, lr=a.lr,
                        depth_slabs=a.depth_slabs)
        log(json.dumps({k: v for k, v in res.items() if k != "rows"}))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)


if __name__ == "__main__":
    _main()
