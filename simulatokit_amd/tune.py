"""The execution-plan autotuner (build-only; it sets the knobs of ExecutionOptions, ExecutionOptions.swift:21-49).

`autotune` times the scan of the caller's own workload under each candidate plan on this device
(`sk_hip_autotune`: positions per launch, the 512-point column tile width, hipGraph replay, the lane pipeline when
there are several potentials) and returns the fastest as ExecutionOptions. A plan changes rounding only, never which
numbers are computed. Results persist in a JSON cache keyed by device, grid lengths, slice counts, the power-of-two
bucket of the position count and detector kinds, so a second process with the same workload reuses the measurement.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

from . import _native as N
from .core import ExecutionOptions, SimulatoError
from .detector import Detector
from .multislice import AnyPotential, _prepared, _raise
from .probe import Probe

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "simulatokit_amd", "tune.json")


@dataclass
class TuneResult:
    options: ExecutionOptions
    rate: float                 # position-slices per second of `options`
    device: str
    candidates: List[Dict]      # every plan timed, with its rate
    key: str
    from_cache: bool = False


def _as_options(o: N.SkOptions) -> ExecutionOptions:
    return ExecutionOptions(batch=o.batch, use_graph=bool(o.use_graph), fast_path=bool(o.fast_path),
                            pipeline=o.pipeline, lane_waves=o.lane_waves, col_tile=o.col_tile, streams=o.streams,
                            tile=o.tile)


def _plan_dict(o: ExecutionOptions) -> Dict:
    return {"batch": o.batch, "use_graph": o.use_graph, "fast_path": o.fast_path, "pipeline": o.pipeline,
            "lane_waves": o.lane_waves, "col_tile": o.col_tile, "streams": o.streams, "tile": o.tile}


def positions_bucket(positions: int) -> int:
    """The power of two at or above the position count. The best plan depends strongly on how many positions a
    scan has (the lane pipeline wins by 3.4x for one position over 16 configurations and loses by 8 % for 64), so
    plans are cached per bucket."""
    n = max(1, int(positions))
    return 1 << (n - 1).bit_length()


def tune_key(device: str, grid, slice_counts: Sequence[int], detectors: Sequence[Detector], positions: int) -> str:
    """The cache key: the device, the grid lengths, the slice counts, the bucket of the position count and the
    detector kinds (the plan's cost does not depend on the atoms or where the positions lie)."""
    kinds = ",".join(type(d).__name__ for d in detectors)
    return (f"{device}|{grid.nx}x{grid.ny}|S={'+'.join(str(s) for s in slice_counts)}|"
            f"P<={positions_bucket(positions)}|{kinds}")


def load_cache(path: str) -> Dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(path: str, key: str, result: TuneResult) -> None:
    data = load_cache(path)
    data[key] = {"options": _plan_dict(result.options), "rate": result.rate, "device": result.device,
                 "candidates": result.candidates}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def device_name() -> str:
    """The device's architecture and compute-unit count, as the tuner reports them (e.g. "gfx950 256 CU")."""
    buf = C.create_string_buffer(64)
    _raise(N.lib().sk_hip_device_name(buf, len(buf)), "device_name")
    return buf.value.decode()


def autotune(probe: Probe, potentials: Sequence[AnyPotential], detectors: Sequence[Detector],
             positions: int = 4096, cache: Optional[str] = DEFAULT_CACHE, retune: bool = False) -> TuneResult:
    """The fastest execution plan for scanning `probe` through `potentials` into `detectors` on this device.
    `cache`: JSON file of earlier results (None = do not read or write one); `retune`: measure even when cached."""
    if not potentials:
        raise SimulatoError("invalidScan", "invalid scan: no potential")
    prepared = [_prepared(p, probe.beam, probe.tilt) for p in potentials]
    handles = (C.c_void_p * len(prepared))(*[p.handle.value for p in prepared])
    descs = (N.SkDetector * max(1, len(detectors)))(*[d.descriptor() for d in detectors])
    slice_counts = [p.slice_count for p in prepared]
    if cache and not retune:
        key = tune_key(device_name(), probe.grid, slice_counts, detectors, positions)
        entry = load_cache(cache).get(key)
        if entry is not None:
            return TuneResult(ExecutionOptions(**entry["options"]), float(entry["rate"]), entry["device"],
                              entry.get("candidates", []), key, from_cache=True)
    rep = N.SkTuneReport()
    _raise(N.lib().sk_hip_autotune(handles, len(prepared), C.byref(probe.descriptor()), descs, len(detectors),
                                   int(positions), C.byref(rep)), "autotune")
    cands = [dict(_plan_dict(_as_options(rep.candidates[k])), rate=rep.rates[k]) for k in range(rep.ncandidates)]
    device = rep.device.decode()
    result = TuneResult(_as_options(rep.best), rep.best_rate, device, cands,
                        tune_key(device, probe.grid, slice_counts, detectors, positions))
    if cache:
        store_cache(cache, result.key, result)
    return result


__all__ = ["TuneResult", "autotune", "device_name", "tune_key", "positions_bucket", "load_cache", "store_cache",
           "DEFAULT_CACHE"]
