"""The index arithmetic of the generic LDS passes (simulatokit_amd/csrc/sk_passes.hip), checked on the CPU.

`Div` divides by a divisor known only at run time from its float reciprocal: q = trunc(float(n) * (1 / d)), then one
correction step either way. The kernels use it for 0 <= n < 65536 and 0 < d <= 16384. float32 multiplication rounds
to nearest and the conversion truncates, here as on the device; the reciprocal is checked as the correctly rounded
1 / d and one ulp either side of it, so the result does not rest on the device's division rounding. Every (n, d) of
the range is checked.

The padded LDS layout keeps entry q at q + q // 16: it must be injective and stay inside the buffer the host sizes,
q + q // 16 < S + S // 16 for q < S.
"""
import numpy as np


def div_by_reciprocal(n, d, inv):
    """Div::quot of sk_passes.hip for an int32 array n, divisor d and float32 reciprocal inv."""
    q = (n.astype(np.float32) * np.float32(inv)).astype(np.int32)
    r = n - q * d
    low = r < 0
    q = np.where(low, q - 1, q)
    r = np.where(low, r + d, r)
    high = r >= d
    q = np.where(high, q + 1, q)
    r = np.where(high, r - d, r)
    return q, r


def test_reciprocal_division_is_exact_over_the_kernels_range():
    n = np.arange(65536, dtype=np.int32)
    for d in range(1, 16385):
        inv = np.float32(1.0) / np.float32(d)
        want_q, want_r = n // d, n - (n // d) * d
        for v in (np.nextafter(inv, np.float32(0)), inv, np.nextafter(inv, np.float32(1))):
            q, r = div_by_reciprocal(n, d, v)
            if not (np.array_equal(q, want_q) and np.array_equal(r, want_r)):
                bad = np.flatnonzero((q != want_q) | (r != want_r))[0]
                raise AssertionError(f"n={bad} d={d} inv={v!r}: got ({q[bad]}, {r[bad]})")


def test_padded_layout_is_injective_and_inside_the_buffer():
    for S in (1, 15, 16, 17, 1920, 2048, 4096, 8192):
        q = np.arange(S)
        slot = q + (q >> 4)
        assert len(np.unique(slot)) == S
        assert slot.max() < S + (S >> 4)
