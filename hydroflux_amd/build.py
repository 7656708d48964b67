"""Build recipe for the in-tree gfx950 library (hipcc, no cmake).

    python -m hydroflux_amd.build

compiles hydroflux_amd/csrc/hf2d.hip (single grids) and hf2d_batch.hip (batched
scenarios) into hydroflux_amd/lib/libhf2d.so.
-ffp-contract=off keeps every a*b+c unfused (the reference's Rust arithmetic
never contracts), which is what makes the f64 step bit-identical to the CPU
restatement; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=ROOT)


def build_library() -> str:
    out_dir = os.path.join(HERE, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libhf2d.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(HERE, "csrc", "hf2d.hip"),
          os.path.join(HERE, "csrc", "hf2d_batch.hip")])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build_library()
    sys.exit(0)
