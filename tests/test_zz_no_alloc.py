"""sk_hip_scan_noalloc (the scan for callers that cannot allocate device memory) against sk_hip_scan, both called
through the C ABI on the same prepared potentials and positions.

The entry point runs the same kernels as sk_hip_scan with the positions per launch its static scratch array holds;
sk_hip_scan is asked for that many positions per launch on one stream, so both reduce every position through the same
launches. Two calls in a row, the second smaller (a 256 x 256 grid, one potential, a few positions, device outputs),
then a scan too large for the scratch, which must be refused before any work.
"""
import ctypes as C

import numpy as np
import pytest

import simulatokit_amd as sk
from simulatokit_amd import _native as N
from simulatokit_amd.workloads import SRTIO3_A, srtio3

pytestmark = pytest.mark.gpu
TOL = 1e-12  # same launches, same kernels: equal up to nothing but the order of identical float64 sums


@pytest.fixture(scope="module")
def lib():
    if not sk.hip_is_available():
        pytest.fail("no HIP device: " + N.last_error())
    L = N.lib()
    L.sk_hip_scan_noalloc.argtypes = L.sk_hip_scan.argtypes
    return L


def _call(L, fn, prepared, probe, pos, dets, opt, outputs=None):
    handles = (C.c_void_p * len(prepared))(*[p.handle.value for p in prepared])
    descs = (N.SkDetector * len(dets))(*[d.descriptor() for d in dets])
    host = [np.full(len(pos), np.nan) for _ in dets]
    ptrs = [int(o) for o in outputs] if outputs is not None else [h.ctypes.data for h in host]
    parr = (C.c_void_p * len(dets))(*ptrs)
    rc = fn(handles, len(prepared), C.byref(probe.descriptor()), len(pos), N.dptr(pos), descs, len(dets), parr,
            1 if outputs is not None else 0, C.byref(opt))
    assert rc == 0, N.last_error()
    t = N.SkTiming()
    assert L.sk_hip_last_timing(C.byref(t)) == 0
    return host, int(t.batch)


def _opt(batch=0, streams=0):
    o = N.SkOptions()
    o.fast_path = 1
    o.batch = batch
    o.streams = streams
    return o


def _compare(name, got, want):
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"{name}: max rel diff {err:.3e}, bit-equal {bool(np.array_equal(got, want))}", flush=True)
    assert np.all(np.isfinite(got))
    assert err <= TOL


def test_noalloc_scan_equals_the_scan_twice_in_a_row(lib):
    L = lib
    # call 1: 512 x 512, two frozen-phonon-like configurations of different slice counts, BF + HAADF, host outputs
    beam = sk.ElectronBeam(200e3)
    pots = [sk.Potential(srtio3(8, 8, k), 512, 512, slice_thickness=2.0) for k in (6, 4)]
    prepared = [p.prepare(beam) for p in pots]
    probe = sk.Probe(beam, pots[0].grid, sk.Aperture(20.0))
    pos = np.ascontiguousarray(sk.GridScan((0.0, 0.0), (SRTIO3_A, SRTIO3_A), 10, 10).positions)
    dets = [sk.AnnularDetector(0, 20), sk.AnnularDetector(50, 200)]
    print("call 1: 512 x 512, 2 potentials", [p.slice_count for p in prepared], "100 positions", flush=True)
    got, T = _call(L, L.sk_hip_scan_noalloc, prepared, probe, pos, dets, _opt())
    assert 1 <= T < 100  # the scratch holds fewer than 100 waves at 512 x 512
    want, Tw = _call(L, L.sk_hip_scan, prepared, probe, pos, dets, _opt(batch=T, streams=1))
    assert Tw == T
    for k in range(len(dets)):
        _compare(f"call 1 detector {k} (T = {T})", got[k], want[k])

    # call 2: smaller, 256 x 256, one potential, 7 positions along a line, device outputs
    import torch

    pot2 = sk.Potential(srtio3(4, 4, 3), 256, 256, slice_thickness=2.0)
    prep2 = [pot2.prepare(beam)]
    probe2 = sk.Probe(beam, pot2.grid, sk.Aperture(20.0))
    pos2 = np.ascontiguousarray(np.stack([np.linspace(0.1, 3.0, 7), np.linspace(0.2, 1.5, 7)], axis=1))
    dets2 = [sk.AnnularDetector(50, 200)]
    dev = torch.full((len(pos2),), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    print("call 2: 256 x 256, 1 potential, 7 positions, device output", flush=True)
    _, T2 = _call(L, L.sk_hip_scan_noalloc, prep2, probe2, pos2, dets2, _opt(), outputs=[dev.data_ptr()])
    got2 = dev.cpu().numpy()
    want2, _ = _call(L, L.sk_hip_scan, prep2, probe2, pos2, dets2, _opt(batch=T2, streams=1))
    assert T2 == 7
    _compare("call 2", got2, want2[0])
    for p in prepared + prep2:
        p.release()


@pytest.mark.parametrize("n", [384, 509, 1024])
def test_noalloc_scan_on_the_register_passes(lib, n):
    # the register passes (384 and 1024 points, 509 by Bluestein on 1024): the detecting pass's partial sums per
    # column tile live in the scratch array as they do in the library's own workspace
    L = lib
    beam = sk.ElectronBeam(200e3)
    pot = sk.Potential(srtio3(4, 4, 2), n, n, slice_thickness=2.0)
    prepared = [pot.prepare(beam)]
    probe = sk.Probe(beam, pot.grid, sk.Aperture(20.0))
    pos = np.ascontiguousarray(sk.GridScan((0.0, 0.0), (SRTIO3_A, SRTIO3_A), 5, 5).positions)
    dets = [sk.AnnularDetector(0, 20), sk.AnnularDetector(40, 100)]
    print(f"{n} x {n}, 1 potential, 25 positions", flush=True)
    got, T = _call(L, L.sk_hip_scan_noalloc, prepared, probe, pos, dets, _opt())
    assert 1 <= T <= 25
    want, Tw = _call(L, L.sk_hip_scan, prepared, probe, pos, dets, _opt(batch=T, streams=1))
    assert Tw == T
    for k in range(len(dets)):
        _compare(f"{n} x {n} detector {k} (T = {T})", got[k], want[k])
    prepared[0].release()


def test_noalloc_scan_refuses_what_its_scratch_cannot_hold(lib):
    L = lib
    beam = sk.ElectronBeam(200e3)
    pot = sk.Potential(srtio3(2, 2, 1), 128, 128, slice_thickness=2.0)
    prepared = [pot.prepare(beam)]
    probe = sk.Probe(beam, pot.grid, sk.Aperture(20.0))
    # 3 M positions: their coordinates and sums alone exceed the 64 MiB array
    pos = np.zeros((3 << 20, 2))
    handles = (C.c_void_p * 1)(prepared[0].handle.value)
    descs = (N.SkDetector * 1)(sk.AnnularDetector(50, 200).descriptor())
    out = np.empty(len(pos))
    parr = (C.c_void_p * 1)(out.ctypes.data)
    rc = L.sk_hip_scan_noalloc(handles, 1, C.byref(probe.descriptor()), len(pos), N.dptr(pos), descs, 1, parr, 0,
                               C.byref(_opt()))
    msg = N.last_error()
    print("refused:", rc, msg, flush=True)
    assert rc == N.SK_ERR_MEMORY
    assert "scratch" in msg and "sk_hip_scan" in msg
    # a pixelated detector is not one the entry point reduces
    descs = (N.SkDetector * 1)(sk.PixelatedDetector().descriptor())
    rc = L.sk_hip_scan_noalloc(handles, 1, C.byref(probe.descriptor()), 4, N.dptr(pos[:4]), descs, 1, parr, 0,
                               C.byref(_opt()))
    assert rc == N.SK_ERR_INVALID and "annular" in N.last_error()
    prepared[0].release()
