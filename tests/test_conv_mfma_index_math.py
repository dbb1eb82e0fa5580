"""The operand layouts of the surrogate's matrix-core convolutions (conv5_forward_mfma_kernel and
conv5_wgrad_mfma_kernel, sk_conv.hip), emulated on the CPU with the kernels' index arithmetic.

v_mfma_f32_32x32x16_bf16 takes A[m = lane % 32][k = 8 (lane / 32) + j] and B[k = 8 (lane / 32) + j][n = lane % 32]
(j < 8) and leaves D[8 (q / 4) + 4 (lane / 32) + q % 4][lane % 32] in register q. The forward kernel feeds A from
the tap-major weights wtm[tap][co][ci] and B from the halo tile [row][column][channel], one k-step per (tap, 16
channels); the weight-gradient kernel feeds A from dy and B from shifted windows of x, one k-step per 16 pixels.
Built that way, the convolution, the backward pass over the data (flipped tap-major weights) and the gradient of the
weights equal the oracle's (float64, no bfloat16 rounding), also with channel counts padded to 16 or 32."""
import numpy as np

from oracle import surrogate_oracle as S

ROWS, COLS = 8, 32  # output rows (4 wavefronts x 2) and columns per block


def mfma(A, B, D):
    """One v_mfma_f32_32x32x16 from per-lane fragments: A[lane][j], B[lane][j]; D[lane][q]."""
    Am = np.zeros((32, 16))
    Bm = np.zeros((16, 32))
    for lane in range(64):
        for j in range(8):
            Am[lane % 32, 8 * (lane // 32) + j] = A[lane][j]
            Bm[8 * (lane // 32) + j, lane % 32] = B[lane][j]
    Dm = Am @ Bm
    for lane in range(64):
        for q in range(16):
            D[lane][q] += Dm[8 * (q // 4) + 4 * (lane // 32) + q % 4, lane % 32]


def tap_major(wt, flip):
    cout, cin = wt.shape[:2]
    w25 = wt.reshape(cout, cin, 25)
    if flip:
        out = np.zeros((25, cin, cout))
        for tap in range(25):
            out[24 - tap] = w25[:, :, tap].T
        return out
    return np.transpose(w25, (2, 0, 1)).copy()


def tap_major_padded(wt, flip):
    """weights_tap_major_kernel's index arithmetic: out[i], i = (tap * 32 + row) * kp + col, kp = 16 for at most 16
    input channels of the convolution served (cin with flip = 0, cout with flip = 1), else 32."""
    cout, cin = wt.shape[:2]
    kp = 16 if (cout if flip else cin) <= 16 else 32
    flat = wt.reshape(-1)
    out = np.zeros(25 * 32 * kp)
    for i in range(out.size):
        col, row, tap = i % kp, (i // kp) % 32, i // (32 * kp)
        co, ci, src_tap = (col, row, 24 - tap) if flip else (row, col, tap)
        if co < cout and ci < cin:
            out[i] = flat[(co * cin + ci) * 25 + src_tap]
    return out.reshape(25, 32, kp)


def emulate(x, wtm, cout=32):
    """conv5_forward_mfma_kernel: the input channels padded with zeros to the weights' kp, 32 output channels
    computed, the first cout stored."""
    B_, C, H, W = x.shape
    kp = wtm.shape[2]
    if C < kp:
        x = np.concatenate([x, np.zeros((B_, kp - C, H, W))], axis=1)
        C = kp
    y = np.zeros((B_, 32, H, W))
    for b in range(B_):
        for y0 in range(0, H, ROWS):
            for x0 in range(0, W, COLS):
                halo = np.zeros((ROWS + 4, COLS + 4, C))
                for rr in range(ROWS + 4):
                    for cc in range(COLS + 4):
                        halo[rr, cc] = x[b, :, (y0 + rr - 2) % H, (x0 + cc - 2) % W]
                for row in range(ROWS):  # wavefront row // 2, tile row % 2
                    D = np.zeros((64, 16))
                    for tap in range(25):
                        ky, kx = divmod(tap, 5)
                        for c16 in range(0, C, 16):
                            A = [[wtm[tap, lane % 32, c16 + 8 * (lane // 32) + j] for j in range(8)] for lane in range(64)]
                            Bf = [[halo[row + ky, lane % 32 + kx, c16 + 8 * (lane // 32) + j] for j in range(8)]
                                  for lane in range(64)]
                            mfma(A, Bf, D)
                    for lane in range(64):
                        for q in range(16):
                            co = 8 * (q // 4) + 4 * (lane // 32) + q % 4
                            y[b, co, y0 + row, x0 + lane % 32] = D[lane][q]
    return y[:, :cout]


def test_matrix_core_convolution_layout_equals_the_oracle():
    rng = np.random.default_rng(9)
    x = rng.normal(size=(1, 32, 8, 32))
    wt = rng.normal(size=(32, 32, 5, 5))
    got = emulate(x, tap_major(wt, flip=False))
    want = S.conv5_forward(x, wt)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_padded_tap_major_layers_equal_the_oracle():
    # the first layer (2 -> 32), the last (32 -> 1) and an odd one (5 -> 17), forward and backward over the data
    rng = np.random.default_rng(13)
    for ci, co in ((2, 32), (32, 1), (5, 17)):
        x = rng.normal(size=(1, ci, 8, 32))
        wt = rng.normal(size=(co, ci, 5, 5))
        fwd = tap_major_padded(wt, flip=False)
        assert fwd.shape[2] == (16 if ci <= 16 else 32)
        want = S.conv5_forward(x, wt)
        assert np.abs(emulate(x, fwd, cout=co) - want).max() <= 1e-12 * np.abs(want).max()
        dy = rng.normal(size=(1, co, 8, 32))
        want = S.conv5_backward_data(dy, wt)
        got = emulate(dy, tap_major_padded(wt, flip=True), cout=ci)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # for 32 x 32 weights the padded layout is the plain tap-major one
    wt = rng.normal(size=(32, 32, 5, 5))
    assert np.array_equal(tap_major_padded(wt, False), tap_major(wt, False))
    assert np.array_equal(tap_major_padded(wt, True), tap_major(wt, True))


def test_matrix_core_backward_data_layout_equals_the_oracle():
    rng = np.random.default_rng(10)
    dy = rng.normal(size=(1, 32, 8, 32))
    wt = rng.normal(size=(32, 32, 5, 5))
    got = emulate(dy, tap_major(wt, flip=True))
    want = S.conv5_backward_data(dy, wt)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


# ---- the gradient of the weights on the matrix cores (conv5_wgrad_mfma_kernel) ----------------------------------------
WG_ROWS, WG_COLS, WG_XSTRIDE, WG_DSTRIDE = 4, 32, 56, 40


def wgrad_emulate(x, dy, blocks):
    """dw[co][ci][tap] as the kernel computes it: strips of 4 rows x 32 columns of one image, taken round-robin by
    `blocks` blocks; x staged as 8 rows x 32 channels x 48 columns (groups of 8 columns from x0 - 8, circular) in
    LDS [row][channel][column] with a 56-column stride, dy as 4 rows x 32 channels x 32 columns with a 40-column
    stride (channels past cin / cout zero); wavefront ky, output row r, k-step t: A[co][k] = dy row r, columns
    16 t + k; B[k][ci] = x staged row r + ky, staged columns 16 t + 8 (k / 8) + (k % 8) + 6 + kx, read as the window
    W = staged columns 16 t + 8 h .. + 23 of lane half h = k / 8 and its elements 6 + kx .. 13 + kx."""
    B_, cin, h, w = x.shape
    cout = dy.shape[1]
    sx, sy = w // WG_COLS, h // WG_ROWS
    strips = B_ * sx * sy
    partial = np.zeros((blocks, 25, 32, 32))
    for g in range(blocks):
        for s in range(g, strips, blocks):
            b, r_ = divmod(s, sx * sy)
            y0, x0 = (r_ // sx) * WG_ROWS, (r_ % sx) * WG_COLS
            xs = np.full((WG_ROWS + 4) * 32 * WG_XSTRIDE, np.nan)
            ds = np.full(WG_ROWS * 32 * WG_DSTRIDE, np.nan)
            for e in range((WG_ROWS + 4) * 32 * 6):
                grp, rc = e % 6, e // 6
                ci, rr = rc % 32, rc // 32
                col0 = (x0 - 8 + 8 * grp) % w
                vals = x[b, ci, (y0 + rr - 2) % h, col0:col0 + 8] if ci < cin else np.zeros(8)
                assert len(vals) == 8
                base = (rr * 32 + ci) * WG_XSTRIDE + 8 * grp
                assert base % 8 == 0 and base + 8 <= xs.size
                xs[base:base + 8] = vals
            for e in range(WG_ROWS * 32 * 4):
                grp, rc = e % 4, e // 4
                co, rr = rc % 32, rc // 32
                vals = dy[b, co, y0 + rr, x0 + 8 * grp:x0 + 8 * grp + 8] if co < cout else np.zeros(8)
                base = (rr * 32 + co) * WG_DSTRIDE + 8 * grp
                assert base % 8 == 0 and base + 8 <= ds.size
                ds[base:base + 8] = vals
            for ky in range(5):
                for r in range(WG_ROWS):
                    for t in range(2):
                        A = np.zeros((32, 16))
                        Bm = [np.zeros((16, 32)) for _ in range(5)]
                        for lane in range(64):
                            n, half = lane % 32, lane // 32
                            da = (r * 32 + n) * WG_DSTRIDE + 16 * t + 8 * half
                            A[n, 8 * half:8 * half + 8] = ds[da:da + 8]
                            xa = ((r + ky) * 32 + n) * WG_XSTRIDE + 16 * t + 8 * half
                            assert xa % 8 == 0 and xa + 24 <= ((r + ky) * 32 + n + 1) * WG_XSTRIDE
                            win = xs[xa:xa + 24]
                            for kx in range(5):
                                Bm[kx][8 * half:8 * half + 8, n] = win[6 + kx:14 + kx]
                        for kx in range(5):
                            assert not np.isnan(A).any() and not np.isnan(Bm[kx]).any()
                            partial[g, ky * 5 + kx] += A @ Bm[kx]
    total = partial.sum(axis=0)  # [tap][co][ci]
    return np.transpose(total, (1, 2, 0))[:cout, :cin].reshape(cout, cin, 5, 5)


def test_matrix_core_weight_gradient_layout_equals_the_oracle():
    rng = np.random.default_rng(11)
    x = rng.normal(size=(2, 32, 8, 64))
    dy = rng.normal(size=(2, 32, 8, 64))
    got = wgrad_emulate(x, dy, blocks=3)  # 8 strips over 3 blocks: uneven round-robin
    want, _ = S.conv5_backward_weights(x, dy)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_matrix_core_weight_gradient_pads_fewer_channels():
    rng = np.random.default_rng(12)
    x = rng.normal(size=(1, 2, 4, 32))   # the first layer: 2 input channels
    dy = rng.normal(size=(1, 32, 4, 32))
    want, _ = S.conv5_backward_weights(x, dy)
    assert np.abs(wgrad_emulate(x, dy, blocks=1) - want).max() <= 1e-12 * np.abs(want).max()
    x = rng.normal(size=(1, 32, 4, 32))  # the last layer: 1 output channel
    dy = rng.normal(size=(1, 1, 4, 32))
    want, _ = S.conv5_backward_weights(x, dy)
    assert np.abs(wgrad_emulate(x, dy, blocks=2) - want).max() <= 1e-12 * np.abs(want).max()


def test_xs_and_dy_strides_are_free_of_bank_conflicts():
    # 16-byte reads by lanes n = 0..31 of one half (16 lanes per LDS cycle): the channel stride in bytes must put
    # 16 consecutive lanes on 16 distinct 16-byte bank groups of the 256-byte bank width
    for stride in (WG_XSTRIDE, WG_DSTRIDE):
        for first in (0, 16):
            groups = {((first + n) * stride * 2 // 16) % 16 for n in range(16)}
            assert len(groups) == 16, stride
