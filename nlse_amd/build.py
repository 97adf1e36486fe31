"""Build libnlse_hip.so in-tree for gfx950 with hipcc.

    python -m nlse_amd.build        # compile (always recompiles)

Packed FP32 VALU instructions are disabled (``-packed-fp32-ops``): the
engine's FP32 arithmetic is issued as ordinary VOP3 instructions.
"""

from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "engine.hip")
OUT = os.path.join(HERE, "libnlse_hip.so")
ARCH = os.environ.get("NLSE_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, SRC]
    t0 = time.time()
    if verbose:
        print("[nlse_amd.build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature: harmless
    noise = ("packed-fp32-ops", "is not a recognized feature")
    err = "\n".join(l for l in res.stderr.splitlines() if not any(n in l for n in noise))
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    os.replace(tmp, OUT)
    if verbose:
        if err.strip():
            print(err, file=sys.stderr)
        print(f"[nlse_amd.build] built {OUT} in {time.time() - t0:.1f}s", flush=True)
    return OUT


if __name__ == "__main__":
    build()
