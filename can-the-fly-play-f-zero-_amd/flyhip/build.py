"""Build recipe for the native library (hipcc, gfx950) -- run by ``__graft_entry__.build()``.

``libflyhip.so`` is built in-tree next to this file so that it travels with the source tree.
"""

from __future__ import annotations

import hashlib
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                       # can-the-fly-play-f-zero-_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_NAME = "libflyhip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
SOURCES = ["lif_engine.hip", "plastic.hip", "surrogate.hip", "eye.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLYHIP_ARCH", "gfx950")

# -ffp-contract=off: the LIF update is evaluated exactly as written (bit-exact vs the CPU oracle).
# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources() -> list[str]:
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) of every file the library is compiled from: the HIP
    sources, the internal header and the C-ABI header, in a fixed order."""
    h = hashlib.sha256()
    for path in sorted(sources() + [os.path.join(CSRC, "errors.h"), os.path.join(INCLUDE, "flyhip.h")]):
        if os.path.exists(path):
            h.update(os.path.basename(path).encode())
            h.update(open(path, "rb").read())
    return h.hexdigest()[:16]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ``libflyhip.so`` (always recompiles)."""
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *FLAGS, f'-DFZ_SRC_HASH="{source_hash()}"', f"-I{INCLUDE}", *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}) building {LIB_NAME}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
