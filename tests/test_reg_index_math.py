"""The index arithmetic of the register-resident passes (simulatokit_amd/csrc/sk_reg.hip), emulated on the CPU:

- the three-stage Stockham transforms with V values per lane (384 = 8*8*6 with 24 values, 1024 = 16*8*8 with 16),
  built from the kernel's formulas (butterfly bi = i + T q reads x[bi + (M / R) j], twiddle w_{LR}^{jk} with
  k = bi mod L, writes y[(bi - k) R + k + L j]; first stage from the registers, last into them), equal numpy's FFT;
- Bluestein's transform of n points on 1024 (chirp, forward transform, kernel product, inverse, chirp) equals it;
- the LDS slot map e -> e ^ (((e >> 3) ^ (e >> 7)) & 31) is a permutation of each line, and a bank model (64 banks of
  4 bytes, a float2 access served in groups of 32 lanes) finds every exchange of both lengths free of conflicts;
- the column pass with its first and last stages in the block phases, run through the tile's slot map, equals
  IFFT(FFT(g) P) per column, and the XCD-ordered block numbering runs every (tile, position group) unit once.
"""
import numpy as np
import pytest

PLANS = {1024: (16, (16, 8, 8), 1024), 384: (24, (8, 8, 6), 400)}  # M: V, radices, line stride (slots)


def rslot(e):
    return e ^ (((e >> 3) ^ (e >> 7)) & 31)


def dft(u, sign):
    k = np.arange(len(u))
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / len(u)) @ u


def vfft(x, M, V, radices, sign):
    T = M // V
    roots = np.exp(-2j * np.pi * np.arange(M) / M)
    v = np.array([[x[i + T * r] for r in range(V)] for i in range(T)], dtype=complex)  # lane i, register r
    buf = np.zeros(M, complex)
    L = 1
    for s, R in enumerate(radices):
        NB, STRIDE = V // R, M // R
        first, last = s == 0, s == len(radices) - 1
        out, nv = np.zeros(M, complex), v.copy()
        for i in range(T):
            for q in range(NB):
                bi = i + T * q
                u = np.array([v[i, q + NB * j] if first else buf[bi + STRIDE * j] for j in range(R)])
                k = 0 if L == 1 else bi % L
                for j in range(1, R):
                    if L > 1:
                        w = roots[j * k * (M // (L * R))]
                        u[j] *= w if sign < 0 else np.conj(w)
                u = dft(u, sign)
                for j in range(R):
                    if last:
                        nv[i, q + NB * j] = u[j]
                    else:
                        out[(bi - k) * R + k + j * L] = u[j]
        buf, v, L = out, nv, L * R
    return np.array([v[e % T, e // T] for e in range(M)])


@pytest.mark.parametrize("M", [1024, 384])
@pytest.mark.parametrize("sign", [-1, 1])
def test_register_stages_equal_numpy(M, sign):
    V, radices, _ = PLANS[M]
    x = np.random.default_rng(M).normal(size=(M, 2)) @ np.array([1, 1j])
    got = vfft(x, M, V, radices, sign)
    want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * M
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("n", [257, 500, 509])
def test_bluestein_on_1024_points_equals_numpy(n):
    M = 1024
    k = np.arange(n)
    chirp = np.exp(-1j * np.pi * ((k * k) % (2 * n)) / n)
    b = np.zeros(M, complex)
    b[0] = np.conj(chirp[0])
    for q in range(1, n):
        b[q] = b[M - q] = np.conj(chirp[q])
    khat = np.fft.fft(b) / M
    x = np.random.default_rng(n).normal(size=(n, 2)) @ np.array([1, 1j])
    for sign in (-1, 1):
        c, kh = (chirp, khat) if sign < 0 else (np.conj(chirp), np.conj(khat))
        u = np.zeros(M, complex)
        u[:n] = x * c
        w = vfft(vfft(u, M, 16, (16, 8, 8), -1) * kh, M, 16, (16, 8, 8), +1)
        want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        assert np.abs(w[:n] * c - want).max() <= 1e-13 * np.abs(want).max()


def worst_conflict(slots):
    return max(np.bincount(np.asarray(slots[g:g + 32]) % 32).max() for g in range(0, len(slots), 32))


@pytest.mark.parametrize("M", [1024, 384])
def test_slot_map_is_a_permutation_and_every_exchange_is_free_of_bank_conflicts(M):
    V, radices, LS = PLANS[M]
    assert sorted(rslot(e) for e in range(M)) == list(range(M))
    T = M // V
    L = 1
    for s, R in enumerate(radices):
        NB, STRIDE = V // R, M // R
        for q in range(NB):
            for j in range(R):
                reads, writes = [], []
                for lane in range(64):
                    line, i = divmod(lane, T)
                    bi = i + T * q
                    k = 0 if L == 1 else bi % L
                    reads.append(line * LS + rslot(bi + STRIDE * j))
                    writes.append(line * LS + rslot((bi - k) * R + k + j * L))
                if s > 0:
                    assert worst_conflict(reads) == 1, (M, s, q, j)
                if s < len(radices) - 1:
                    assert worst_conflict(writes) == 1, (M, s, q, j)
        L *= R


@pytest.mark.parametrize("M,TC,CS", [(384, 16, 400), (384, 8, 400), (1024, 4, 1032), (1024, 8, 1028)])
def test_column_tiles_are_free_of_bank_conflicts(M, TC, CS):
    """The column pass's tile: column c at slots [c CS, c CS + M), element e at rslot(e), XORed with c (16-column
    tiles) or 4 c (8-column tiles) at 384 points.
    The block phases (thread t -> column t mod TC, rows t / TC + (threads / TC) r) and, at 384 points, the exchanges
    of the four columns a wavefront holds fall on distinct banks."""
    V, radices, _ = PLANS[M]
    T = M // V
    threads = TC * T

    def elem(e, c):
        return rslot(e) ^ ((4 * c if TC == 8 else c) & 31) if M == 384 else rslot(e)

    assert all(sorted(elem(e, c) for e in range(M)) == list(range(M)) for c in range(TC))
    step = threads // TC
    for r in range(M // step):
        for w in range(threads // 64):
            slots = [(t % TC) * CS + elem(t // TC + step * r, t % TC) for t in range(64 * w, 64 * w + 64)]
            assert worst_conflict(slots) == 1, (r, w)
    if M == 384:
        L = 1
        for s, R in enumerate(radices):
            NB, STRIDE = V // R, M // R
            for w in range(threads // 64):
                for q in range(NB):
                    for j in range(R):
                        reads, writes = [], []
                        for lane in range(64):
                            c, i = w * (64 // T) + lane // T, lane % T
                            bi = i + T * q
                            k = 0 if L == 1 else bi % L
                            reads.append(c * CS + elem(bi + STRIDE * j, c))
                            writes.append(c * CS + elem((bi - k) * R + k + j * L, c))
                        if s > 0:
                            assert worst_conflict(reads) == 1
                        if s < len(radices) - 1:
                            assert worst_conflict(writes) == 1
            L *= R


def _dft16_half(u, sign):
    """dft16_half (sk_dft.h): the 4 x 4 DFT16 of inputs whose upper half is zero, in the kernel's order: first-layer
    DFT4s of (a0, a1, 0, 0), the w16 twiddles, the second-layer DFT4s."""
    si = 1j * sign

    def dft4(a0, a1, a2, a3):
        t0, t1, t2, t3 = a0 + a2, a0 - a2, a1 + a3, si * (a1 - a3)
        return t0 + t2, t1 + t3, t0 - t2, t1 - t3

    y = np.zeros((4, 4), complex)
    for n2 in range(4):
        a0, a1 = u[n2], u[4 + n2]
        y[n2] = [a0 + a1, a0 + si * a1, a0 - a1, a0 - si * a1]
    w16 = np.exp(sign * 2j * np.pi / 16)
    for n2 in range(4):
        for k1 in range(4):
            y[n2][k1] *= w16 ** (n2 * k1)
    out = np.zeros(16, complex)
    for k1 in range(4):
        b = dft4(y[0][k1], y[1][k1], y[2][k1], y[3][k1])
        for k2 in range(4):
            out[k1 + 4 * k2] = b[k2]
    return out


@pytest.mark.parametrize("sign", [-1, 1])
def test_dft16_of_a_zero_upper_half_equals_numpy(sign):
    rng = np.random.default_rng(21)
    u = np.zeros(16, complex)
    u[:8] = rng.normal(size=8) + 1j * rng.normal(size=8)
    want = np.fft.fft(u) if sign < 0 else np.fft.ifft(u) * 16
    assert np.abs(_dft16_half(u, sign) - want).max() <= 1e-12 * np.abs(want).max()


def _in_band(M, r):
    """in_band<M> (sk_reg.hip): r T > floor(M / 3) and r T + T - 1 < M - floor(M / 3)."""
    T = M // PLANS[M][0]
    K = M // 3
    return r * T > K and r * T + T - 1 < M - K


@pytest.mark.parametrize("M", [384, 1024])
def test_the_folded_band_holds_no_kept_frequency_on_any_grid(M):
    # the antialias aperture keeps |k| <= 0.6666666 / max(sx, sy) / 2 (Antialias.swift:18-39, radial, so its rows and
    # the kept columns both lie in that range on each axis); the elements the register passes treat as known zeros
    # (i + T r for r in the band, every lane i) must hold no kept frequency, for square and unequal samplings
    V = PLANS[M][0]
    T = M // V
    band = [i + T * r for r in range(V) if _in_band(M, r) for i in range(T)]
    assert band, M
    for ex, ey in ((30.0, 30.0), (31.2, 31.2), (30.0, 45.0), (45.0, 30.0), (12.0, 80.0), (100.0, 7.5)):
        for n, extent in ((M, ex), (M, ey)):
            s = extent / n
            kc = 0.6666666 / max(ex / M, ey / M) / 2
            k = np.fft.fftfreq(n, s)
            kept = set(np.nonzero(np.abs(k) <= kc)[0].tolist())
            assert kept.isdisjoint(band), (M, ex, ey)
    # and the band is as wide as that argument allows: the element just outside it can be kept
    r0 = min(r for r in range(V) if _in_band(M, r))
    assert not _in_band(M, r0 - 1) and (r0 - 1) * T + T - 1 >= M // 3


def _dft4_z(a, Z, sign):
    """dft4_z (sk_dft.h): known-zero inputs (bit j of Z) are not read."""
    z = [(Z >> j) & 1 for j in range(4)]
    a = [0j if z[j] else a[j] for j in range(4)]
    t0, t1, t2, t3 = a[0] + a[2], a[0] - a[2], a[1] + a[3], 1j * sign * (a[1] - a[3])
    return [t0 + t2, t1 + t3, t0 - t2, t1 - t3]


def _dft8_z(u, Z, sign):
    ZE = (Z & 1) | ((Z >> 1) & 2) | ((Z >> 2) & 4) | ((Z >> 3) & 8)
    ZO = ((Z >> 1) & 1) | ((Z >> 2) & 2) | ((Z >> 3) & 4) | ((Z >> 4) & 8)
    e = _dft4_z([u[0], u[2], u[4], u[6]], ZE, sign)
    o = _dft4_z([u[1], u[3], u[5], u[7]], ZO, sign)
    w = [o[k] * np.exp(sign * 1j * np.pi * k / 4) for k in range(4)]
    return [e[k] + w[k] for k in range(4)] + [e[k] - w[k] for k in range(4)]


def _dft16_z(u, Z, sign):
    masks = [sum(((Z >> (n2 + 4 * j)) & 1) << j for j in range(4)) for n2 in range(4)]
    y = [_dft4_z([u[n2], u[4 + n2], u[8 + n2], u[12 + n2]], masks[n2], sign) for n2 in range(4)]
    w16 = np.exp(sign * 2j * np.pi / 16)
    y = [[y[n2][k1] * w16 ** (n2 * k1) for k1 in range(4)] for n2 in range(4)]
    out = np.zeros(16, complex)
    for k1 in range(4):
        b = _dft4_z([y[0][k1], y[1][k1], y[2][k1], y[3][k1]], 0, sign)
        for k2 in range(4):
            out[k1 + 4 * k2] = b[k2]
    return out


@pytest.mark.parametrize("R,Z", [(8, 0x38), (8, 0x18), (8, 0), (8, 0xff), (16, 0x3c0), (16, 0xff00), (16, 0)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_dfts_with_known_zero_inputs_equal_numpy(R, Z, sign):
    # the first-stage masks of the band (384: inputs 3..5 of butterfly 0, 3..4 of the others; 1024: inputs 6..9),
    # no zeros and all zeros; the inputs flagged zero hold garbage, which the masked transforms must not read
    rng = np.random.default_rng(R + Z)
    u = rng.normal(size=R) + 1j * rng.normal(size=R)
    clean = np.array([0j if (Z >> j) & 1 else u[j] for j in range(R)])
    want = np.fft.fft(clean) if sign < 0 else np.fft.ifft(clean) * R
    got = np.array(_dft8_z(u, Z, sign) if R == 8 else _dft16_z(u, Z, sign))
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_band_masks_match_the_band():
    for M, R in ((384, 8), (1024, 16)):
        V = PLANS[M][0]
        NB = V // R
        for q in range(NB):
            z = sum(1 << j for j in range(R) if _in_band(M, q + NB * j))
            assert z == {(384, 0): 0x38, (384, 1): 0x18, (384, 2): 0x18, (1024, 0): 0x3c0}[(M, q)]


def _xcd_unit(bx, by, gx, gy):
    """The column pass's (tile, position group) of block (bx, by) in a gx x gy grid (xcd_unit in sk_reg.hip)."""
    units = gx * gy
    L = bx + gx * by
    u = (L % 8) * (units // 8) + L // 8 if units % 8 == 0 else L
    return u % gx, u // gx


@pytest.mark.parametrize("gx,gy", [(34, 512), (86, 512), (88, 512), (43, 128), (22, 1), (17, 3), (5, 8), (1, 1)])
def test_column_units_cover_every_tile_and_position_group_once(gx, gy):
    seen = {}
    for by in range(gy):
        for bx in range(gx):
            seen.setdefault(_xcd_unit(bx, by, gx, gy), []).append((bx, by))
    assert set(seen) == {(x, y) for y in range(gy) for x in range(gx)}
    assert all(len(v) == 1 for v in seen.values())
    if (gx * gy) % 8 == 0:
        # the blocks that one XCD receives (dispatch order L = 8 k + xcd) run consecutive units, tile fastest
        for xcd in range(8):
            order = [_xcd_unit(L % gx, L // gx, gx, gy) for L in range(xcd, gx * gy, 8)]
            flat = [x + gx * y for x, y in order]
            assert flat == list(range(flat[0], flat[0] + len(flat)))


def _stage(M, V, R, L, sign, src, dst, lanes):
    """One Stockham stage (radix R, sub-length L) of the kernel's vstage for the lanes `lanes`: (i, load, store)
    with load(e) -> the element e of the line's input (registers or LDS slot) and store(e, value) -> its output;
    returns nothing, the lanes' reads all happen before their writes (the wavefront barrier between them)."""
    T, NB, STRIDE = M // V, V // R, M // R
    roots = np.exp(-2j * np.pi * np.arange(M) / M)
    outs = []
    for i, load, store in lanes:
        for q in range(NB):
            bi = i + T * q
            k = 0 if L == 1 else bi % L
            u = np.array([load(i, q + NB * j, bi + STRIDE * j) for j in range(R)], dtype=complex)
            for j in range(1, R):
                if L > 1:
                    w = roots[j * k * (M // (L * R))]
                    u[j] *= w if sign < 0 else np.conj(w)
            u = dft(u, sign)
            outs += [(store, i, q + NB * j, (bi - k) * R + k + j * L, bi + STRIDE * j, u[j]) for j in range(R)]
    for store, i, reg, slot_e, nat_e, val in outs:
        store(i, reg, slot_e, nat_e, val)


@pytest.mark.parametrize("M,TC,CS", [(384, 8, 400), (384, 16, 400), (1024, 8, 1028)])
def test_fused_column_pass_through_the_tile_equals_numpy(M, TC, CS):
    """The register column pass with the first and last stages in the block phases (col_reg_kernel, not
    Bluestein's): block-phase thread t = (column t mod TC, lane t / TC) runs the forward transform's first stage from
    its loaded rows into the tile; each column's lanes run the second and last stages, the product with P (zero in
    the antialias band), the inverse's first (band inputs zero) and second stages; the block phase runs the inverse's
    last stage from the tile (last_stage_store) and stores rows bi + (M / R3) j. Through the tile's slot map (rslot,
    XORed per column at 384 points) the result is IFFT(FFT(g) P) for every column."""
    V, radices, _ = PLANS[M]
    R1, R2, R3 = radices
    T = M // V
    rng = np.random.default_rng(M + TC)
    g = rng.normal(size=(M, TC)) + 1j * rng.normal(size=(M, TC))
    band = np.array([_in_band(M, r) for r in range(V)])
    P = rng.normal(size=(M, TC)) + 1j * rng.normal(size=(M, TC))
    for e in range(M):
        if band[e // T]:
            P[e] = 0

    def elem(e, c):
        return rslot(e) ^ ((4 * c if TC == 8 else c) & 31) if M == 384 else rslot(e)

    tile = np.full(TC * CS, np.nan, dtype=complex)
    block = [(t % TC, t // TC) for t in range(TC * T)]
    # block phase: the forward transform's first stage, thread (cb, xi) as lane xi of column cb
    for cb in range(TC):
        lanes = [(xi, lambda i, r, e, cb=cb: g[i + T * r, cb],
                  lambda i, r, s, n, val, cb=cb: tile.__setitem__(cb * CS + elem(s, cb), val))
                 for (c, xi) in block if c == cb]
        _stage(M, V, R1, 1, -1, None, None, lanes)
    out = np.zeros((M, TC), complex)
    for wc in range(TC):
        v = np.zeros((T, V), complex)
        rd = lambda i, r, e, wc=wc: tile[wc * CS + elem(e, wc)]
        wr = lambda i, r, s, n, val, wc=wc: tile.__setitem__(wc * CS + elem(s, wc), val)
        to_reg = lambda i, r, s, n, val: v.__setitem__((i, r), val)
        _stage(M, V, R2, R1, -1, None, None, [(i, rd, wr) for i in range(T)])
        _stage(M, V, R3, R1 * R2, -1, None, None, [(i, rd, to_reg) for i in range(T)])
        for i in range(T):
            for r in range(V):
                v[i, r] = 0 if band[r] else v[i, r] * P[i + T * r, wc]
        from_reg = lambda i, r, e: 0 if band[r] else v[i, r]
        _stage(M, V, R1, 1, +1, None, None, [(i, from_reg, wr) for i in range(T)])
        _stage(M, V, R2, R1, +1, None, None, [(i, rd, wr) for i in range(T)])
    # block phase: the inverse's last stage, outputs to their natural rows
    for cb in range(TC):
        rd = lambda i, r, e, cb=cb: tile[cb * CS + elem(e, cb)]
        st = lambda i, r, s, n, val, cb=cb: out.__setitem__((n, cb), val)
        _stage(M, V, R3, R1 * R2, +1, None, None, [(xi, rd, st) for (c, xi) in block if c == cb])
    want = np.fft.ifft(np.fft.fft(g, axis=0) * P, axis=0) * M
    assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()


def _bluestein(x, n, sign, chirp_in=True, chirp_out=True):
    """line_fft's Bluestein transform of n points on 1024 (sk_reg.hip), with or without its opening and closing chirp
    products; elements past n are zero on the way in and dropped on the way out."""
    M = 1024
    k = np.arange(n)
    chirp = np.exp(-1j * np.pi * ((k * k) % (2 * n)) / n)
    b = np.zeros(M, complex)
    b[0] = np.conj(chirp[0])
    for q in range(1, n):
        b[q] = b[M - q] = np.conj(chirp[q])
    khat = np.fft.fft(b) / M
    c, kh = (chirp, khat) if sign < 0 else (np.conj(chirp), np.conj(khat))
    u = np.zeros(M, complex)
    u[:n] = x[:n] * c if chirp_in else x[:n]
    w = vfft(vfft(u, M, 16, (16, 8, 8), -1) * kh, M, 16, (16, 8, 8), +1)[:n]
    return w * c if chirp_out else w


@pytest.mark.parametrize("n", [257, 509])
def test_adjacent_bluestein_chirps_cancel(n):
    """The row pass's inverse transform without its closing chirp, the product with t, and the forward transform
    without its opening chirp give FFT(t IFFT(x)) n; the column pass's forward without its closing chirp, the product
    with P and the inverse without its opening chirp give IFFT(FFT(x) P) n; the forward without its closing chirp
    gives the same intensities |FFT(x)|^2."""
    rng = np.random.default_rng(n)
    x, t, P = (rng.normal(size=(n, 2)) @ np.array([1, 1j]) for _ in range(3))
    row = _bluestein(_bluestein(x, n, +1, chirp_out=False) * t, n, -1, chirp_in=False)
    want = np.fft.fft(t * np.fft.ifft(x) * n)
    assert np.abs(row - want).max() <= 1e-12 * np.abs(want).max()
    col = _bluestein(_bluestein(x, n, -1, chirp_out=False) * P, n, +1, chirp_in=False)
    want = np.fft.ifft(np.fft.fft(x) * P) * n
    assert np.abs(col - want).max() <= 1e-12 * np.abs(want).max()
    inten = np.abs(_bluestein(x, n, -1, chirp_out=False)) ** 2
    assert np.abs(inten - np.abs(np.fft.fft(x)) ** 2).max() <= 1e-12 * (np.abs(np.fft.fft(x)) ** 2).max()


@pytest.mark.parametrize("n", [256, 512])
def test_band_row_pass_writes_every_padding_column(n):
    """The band-pruned row pass (row_fast_kernel, sk_fast.hip) writes whole rows: a lane i whose element 2 or 5
    (i + 2 T, i + 5 T; T = n / 8) is not kept stores a zero into padding column nkept + i mod npad. On the HAADF
    grids every padding column gets one."""
    T = n // 8
    ky = np.fft.fftfreq(n, 1.0 / n)  # integer frequencies; the kept set is |k| <= floor(n / 3), as on square grids
    kept = np.abs(ky) <= n // 3
    nkept = int(kept.sum())
    nc = -(-nkept // 16) * 16
    npad = nc - nkept
    assert npad > 0
    written = set()
    for i in range(T):
        for r in (2, 5):
            if not kept[i + r * T]:
                written.add(nkept + i % npad)
    assert written == set(range(nkept, nc))


def test_bluestein_block_phase_rows():
    """Bluestein's column pass (col_reg_kernel, BLUE): a block-phase thread xi moves rows xi + 64 r for r < 8; a row
    past n stores the lane's last row below n, xlast = xi + 64 floor((n - 1 - xi) / 64), again. For every n the pass
    takes (257..512) and every lane that row exists, lies below n, is the lane's own and is its last; every row below
    n is stored by exactly one (lane, r) as itself."""
    for n in range(257, 513):
        stored = []
        for xi in range(64):
            xlast = xi + 64 * ((n - 1 - xi) // 64)
            assert xlast < n and xlast % 64 == xi and xlast + 64 >= n
            for r in range(8):
                x0 = xi + 64 * r
                if x0 < n:
                    stored.append(x0)
                else:
                    assert xlast in stored  # a repeat of a row the lane already stores
        assert sorted(stored) == list(range(n))
