"""Shared test setup: the `gpu` marker, the repo root on sys.path, the golden fixtures."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU parity cases")


class Golden:
    """One fixture set: ``meta`` (the JSON) and ``array(name)`` (the arrays, complex as complex128)."""

    def __init__(self, stem):
        with open(os.path.join(GOLDEN, stem + ".json")) as fh:
            self.meta = json.load(fh)
        self._npz = np.load(os.path.join(GOLDEN, stem + ".npz"), allow_pickle=False)

    def array(self, name):
        return self._npz[name.replace("/", "__")]

    def has(self, name):
        return name.replace("/", "__") in self._npz.files


_cache = {}


def golden(stem):
    if stem not in _cache:
        _cache[stem] = Golden(stem)
    return _cache[stem]


def rel_max(a, b):
    """max |a - b| / max |b| (Tests/SimulatoKitTestSupport/Comparison.swift:9-29)."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(a - b).max() if a.size else 0.0
    m = np.abs(b).max() if b.size else 0.0
    return d / m if m > 0 else d


def rel_diff(a, b):
    """Comparison.swift:31"""
    return abs(a) if b == 0 else abs(a - b) / abs(b)


def scaled_max(a, b, scale):
    """max |a - b| / scale (Tests/SimulatoKitParityTests/Detector/STEMPins.swift:110-115)."""
    return np.abs(np.asarray(a) - np.asarray(b)).max() / scale


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
