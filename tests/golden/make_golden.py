"""Write the golden fixtures the tests read, from the reference's committed abTEM fixtures.

The reference ships its parity data as ``*-pins.json`` (scalars, cases, an index into a sidecar) plus
``*-arrays.bin`` (float64 little-endian, complex as interleaved re, im), sha256-checked
(Tests/SimulatoKitParityTests/Fixtures/README.md:1-13, Tests/SimulatoKitParityTests/Multislice/Pins.swift:21-44).
This script copies the arrays the multislice/STEM parity tests need into compressed ``.npz`` files and the
case descriptions into ``.json`` files beside it. It only reads data files; it runs no reference code.

Run once, where the reference tree is mounted at ``/root/reference`` (the tests read only the committed fixtures):

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

REF = "/root/reference/Tests/SimulatoKitParityTests/Fixtures"
OUT = os.path.dirname(os.path.abspath(__file__))


def load(pins_name: str):
    with open(os.path.join(REF, pins_name)) as fh:
        js = json.load(fh)
    arrays = js["arrays"]
    with open(os.path.join(REF, arrays["file"]), "rb") as fh:
        blob = fh.read()
    digest = hashlib.sha256(blob).hexdigest()
    if digest != arrays["sha256"]:
        sys.exit(f"{arrays['file']}: sha256 mismatch")
    return js, blob


def array(js, blob, name):
    e = js["arrays"]["index"][name]
    count = e["count"] * (2 if e["kind"] == "c16" else 1)
    d = np.frombuffer(blob, dtype="<f8", count=count, offset=e["offset"]).copy()
    if e["kind"] == "c16":
        d = d[0::2] + 1j * d[1::2]
    return d.reshape(e["shape"])


def dump(stem, js, blob, names, meta):
    arrs = {n.replace("/", "__"): array(js, blob, n) for n in names}
    np.savez_compressed(os.path.join(OUT, stem + ".npz"), **arrs)
    with open(os.path.join(OUT, stem + ".json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print(stem, len(arrs), "arrays")


def main():
    # slice-1: potentials, transmission, multislice, probe (tools/pins/potential-multislice/pins.py)
    js, blob = load("slice-1-pins.json")
    names = sorted(js["arrays"]["index"].keys())
    meta = {k: js[k] for k in ("beam", "cases", "f", "f_k2", "kappa", "pi2_float32", "pi_float32",
                               "truth_far_field_Al", "versions")}
    dump("slice1", js, blob, names, meta)

    # slice-2: frozen phonons, grid scans, detectors (tools/pins/stem-detectors/pins.py)
    js, blob = load("slice-2-pins.json")
    names = sorted(js["arrays"]["index"].keys())
    meta = {k: js[k] for k in ("cases", "detectors", "versions")}
    dump("slice2", js, blob, names, meta)

    # slice-4: odd grid 90 x 75 and beam tilt (tools/pins/sampling-tilt/pins.py)
    js, blob = load("slice-4-pins.json")
    names = sorted(js["arrays"]["index"].keys())
    meta = {k: v for k, v in js.items() if k != "arrays"}
    dump("slice4", js, blob, names, meta)

    # slice-4-scan: pixelated cutoff crop and centre of mass over a scan (tools/pins/scan-patterns/pins.py)
    js, blob = load("slice-4-scan-pins.json")
    names = ["dp", "com_mrad", "com_inv"]
    meta = {k: v for k, v in js.items() if k != "arrays"}
    dump("slice4scan", js, blob, names, meta)

    # line scans: positions and STEM over a line with an annular detector (tools/pins/linescan/pins.py)
    js, blob = load("linescan-pins.json")
    names = sorted(n for n in js["arrays"]["index"] if n.startswith("line.") or n.endswith((".adf", ".positions")))
    meta = {k: js[k] for k in ("versions", "box", "gpts", "sliceThickness", "energy", "semiangle", "atoms",
                               "lineScans")}
    meta["scans"] = [{k: v for k, v in s.items() if k in ("label", "start", "end", "gpts", "sampling", "endpoint")}
                     for s in js["scans"]]
    dump("linescan", js, blob, names, meta)

    # exit waves over a scan, on the simulation grid and Fourier-resampled (tools/pins/waves/pins.py)
    js, blob = load("waves-pins.json")
    names = sorted(js["arrays"]["index"].keys())
    meta = {k: v for k, v in js.items() if k != "arrays"}
    dump("waves", js, blob, names, meta)

    # CTF kernels and HRTEM images of a SrTiO3 exit wave (tools/pins/ctf/pins.py)
    js, blob = load("slice-3-pins.json")
    names = sorted(js["arrays"]["index"].keys())
    meta = {k: v for k, v in js.items() if k != "arrays"}
    dump("slice3", js, blob, names, meta)


if __name__ == "__main__":
    main()
