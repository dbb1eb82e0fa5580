"""ctypes binding of the HIP library's C ABI (include/simulato_hip.h).

Loads the in-tree build ``simulatokit_amd/_lib/libsimulato_hip.so``; there is no other implementation behind this
package, so a missing library is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libsimulato_hip.so")

SK_OK, SK_ERR_UNAVAILABLE, SK_ERR_INVALID, SK_ERR_DEVICE, SK_ERR_MEMORY = 0, -1, -2, -3, -4


class SkGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("extent_x", C.c_double), ("extent_y", C.c_double)]


class SkProbe(C.Structure):
    _fields_ = [("energy", C.c_double), ("semiangle_mrad", C.c_double), ("soft", C.c_int),
                ("aberrations", C.c_double * 25), ("tilt", C.c_double * 2)]


class SkDetector(C.Structure):
    _fields_ = [("kind", C.c_int), ("inner", C.c_double), ("outer", C.c_double), ("max_angle", C.c_int),
                ("units", C.c_int), ("crop_nx", C.c_int), ("crop_ny", C.c_int), ("out_f32", C.c_int),
                ("bins_radial", C.c_int), ("bins_azimuthal", C.c_int), ("rotation", C.c_double)]


class SkOptions(C.Structure):
    _fields_ = [("batch", C.c_int), ("use_graph", C.c_int), ("fast_path", C.c_int), ("pipeline", C.c_int),
                ("lane_waves", C.c_int), ("col_tile", C.c_int),
                ("precision", C.c_int), ("streams", C.c_int), ("tile", C.c_int)]


class SkTiming(C.Structure):
    _fields_ = [("row_ms", C.c_double), ("col_ms", C.c_double), ("row_launches", C.c_longlong),
                ("col_launches", C.c_longlong), ("row_bytes", C.c_double), ("col_bytes", C.c_double),
                ("total_ms", C.c_double), ("batch", C.c_longlong), ("group", C.c_longlong)]


class SkCtf(C.Structure):
    _fields_ = [("aberrations", C.c_double * 25), ("semiangle_mrad", C.c_double), ("soft", C.c_int),
                ("focal_spread", C.c_double), ("angular_spread", C.c_double)]


SK_TUNE_MAX = 32


class SkTuneReport(C.Structure):
    _fields_ = [("best", SkOptions), ("best_rate", C.c_double), ("ncandidates", C.c_int),
                ("candidates", SkOptions * SK_TUNE_MAX), ("rates", C.c_double * SK_TUNE_MAX),
                ("device", C.c_char * 64)]


_lib = None
_dp = C.POINTER(C.c_double)
# sk_deliver_fn: (user, first position, count, count x nx x ny float64 intensities) -> 0 to go on
DELIVER_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double))
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def lib():
    """The loaded library (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"the HIP library is not built ({LIB_PATH}); run `python -m simulatokit_amd.build`")
        # torch, when installed, is loaded first: its wheel bundles its own ROCm runtime libraries, and loaded after
        # this library has initialised the system's runtime it finds no devices; loaded first, both share its HSA
        # runtime (the device memory the surrogate and the benchmarks use is torch's)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        # the binary must have been built from the sources beside it: a stale build fails loudly here
        from .build import source_hash
        L.sk_hip_build_id.restype = C.c_char_p
        built, want = L.sk_hip_build_id().decode(), source_hash()
        if built != want:
            raise RuntimeError(f"{LIB_PATH} was built from other sources (build id {built[:16]}, sources "
                               f"{want[:16]}); rebuild it with `python -m simulatokit_amd.build`")
        L.sk_hip_available.restype = C.c_int
        L.sk_hip_last_error.restype = C.c_char_p
        L.sk_hip_potential_from_values.argtypes = [C.POINTER(SkGrid), C.c_int, _dp, _dp, C.c_double, _dp,
                                                   C.POINTER(_vp)]
        L.sk_hip_potential_from_atoms.argtypes = [C.POINTER(SkGrid), C.c_int, _dp, C.c_int, _ip, _dp, _ip, C.c_int,
                                                  C.c_double, _dp, C.POINTER(_vp)]
        L.sk_hip_potential_from_atoms_finite.argtypes = [C.POINTER(SkGrid), C.c_int, _dp, C.c_int, _ip, _dp,
                                                         C.c_int, C.c_double, _dp, C.POINTER(_vp)]
        L.sk_hip_potential_release.argtypes = [_vp]
        L.sk_hip_potential_release.restype = None
        L.sk_hip_potential_slice_count.argtypes = [_vp]
        L.sk_hip_potential_values.argtypes = [_vp, _dp]
        L.sk_hip_potential_transmissions.argtypes = [_vp, _dp, _dp]
        L.sk_hip_run.argtypes = [_vp, _dp, _dp, _ip, C.c_int, _dp, _dp, C.POINTER(SkOptions)]
        L.sk_hip_run_batch.argtypes = [_vp, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(SkOptions)]
        L.sk_hip_probes.argtypes = [C.POINTER(SkGrid), C.POINTER(SkProbe), C.c_int, _dp, _dp, _dp]
        L.sk_hip_scan.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(SkProbe), C.c_int, _dp, C.POINTER(SkDetector),
                                  C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(SkOptions)]
        L.sk_hip_scan_noalloc.argtypes = L.sk_hip_scan.argtypes
        L.sk_hip_scan_intensities.argtypes = [_vp, C.POINTER(SkProbe), C.c_int, _dp, DELIVER_FN, _vp,
                                              C.POINTER(SkOptions)]
        L.sk_hip_scan_waves.argtypes = [_vp, C.POINTER(SkProbe), C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                        C.POINTER(SkOptions)]
        L.sk_hip_ctf_kernel.argtypes = [C.POINTER(SkGrid), C.c_double, C.POINTER(SkCtf), _dp, _dp]
        L.sk_hip_apply_ctf.argtypes = [C.POINTER(SkGrid), C.c_double, C.POINTER(SkCtf), C.c_int, _dp, _dp, _dp, _dp]
        L.sk_hip_detector_shape.argtypes = [C.POINTER(SkGrid), C.c_double, C.POINTER(SkDetector), _ip]
        L.sk_hip_last_timing.argtypes = [C.POINTER(SkTiming)]
        L.sk_hip_set_timing.argtypes = [C.c_int]
        L.sk_hip_set_timing.restype = None
        L.sk_hip_potential_grid.argtypes = [_vp, C.POINTER(SkGrid)]
        L.sk_hip_device_name.argtypes = [C.c_char_p, C.c_int]
        L.sk_hip_autotune.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(SkProbe), C.POINTER(SkDetector), C.c_int,
                                      C.c_int, C.POINTER(SkTuneReport)]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().sk_hip_last_error().decode()


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


EXPORTED = ("sk_hip_available", "sk_hip_last_error", "sk_hip_build_id", "sk_hip_potential_from_values", "sk_hip_potential_from_atoms",
            "sk_hip_potential_release", "sk_hip_potential_slice_count", "sk_hip_potential_values",
            "sk_hip_potential_transmissions", "sk_hip_run", "sk_hip_run_batch", "sk_hip_probes", "sk_hip_scan",
            "sk_hip_detector_shape", "sk_hip_last_timing", "sk_hip_set_timing", "sk_hip_potential_grid",
            "sk_hip_autotune", "sk_hip_device_name", "sk_hip_potential_from_atoms_finite", "sk_hip_scan_waves", "sk_hip_scan_intensities",
            "sk_hip_ctf_kernel", "sk_hip_apply_ctf", "sk_hip_scan_noalloc",
            # include/simulato_surrogate.h
            "sk_conv5_forward", "sk_conv5_backward_data", "sk_conv5_backward_weights", "sk_relu_backward",
            "sk_separable_map", "sk_mse_grad", "sk_weighted_mse_grad", "sk_adam_step", "sk_to_bf16",
            "sk_conv5_weights_tap_major", "sk_conv5_forward_tap_major", "sk_conv5_backward_weights_mfma")
