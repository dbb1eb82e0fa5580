"""Build the HIP library in-tree: simulatokit_amd/_lib/libsimulato_hip.so (gfx950).

    python -m simulatokit_amd.build

Each source compiles to its own object in parallel, then one link.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libsimulato_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["sk_passes.hip", "sk_fast.hip", "sk_reg.hip", "sk_prepare.hip", "sk_api.hip", "sk_tune.hip", "sk_conv.hip"]
ARCH = os.environ.get("SK_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions (v_pk_*_f32) are kept out of every device code object: the compiler would otherwise form
# them from plain float code, and some of their forms have given wrong results without any error.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def source_hash() -> str:
    """SHA-256 over the library's sources: every file of csrc/ and include/, by name and content, in name order.
    The build embeds it (sk_hip_build_id) and the loader compares it with the sources beside the binary."""
    h = hashlib.sha256()
    for d, tag in ((CSRC, "csrc"), (INCLUDE, "include")):
        for name in sorted(os.listdir(d)):
            path = os.path.join(d, name)
            if not os.path.isfile(path) or not name.endswith((".hip", ".h", ".inc")):
                continue
            with open(path, "rb") as fh:
                data = fh.read()
            h.update(f"{tag}/{name}:{len(data)}:".encode())
            h.update(data)
    return h.hexdigest()


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))


def build(verbose: bool = False, out: str = LIB_PATH, extra_flags=()) -> str:
    obj_dir = os.path.join(os.path.dirname(out), "obj")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    build_id = f'-DSK_BUILD_ID="{source_hash()}"'
    cmds = [["hipcc", *HIPCC_FLAGS, build_id, *extra_flags, "-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        list(pool.map(lambda c: _run(c, verbose), cmds))
    tmp = out + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
