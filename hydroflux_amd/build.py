"""Build recipe for the in-tree gfx950 library (hipcc, no cmake).

    python -m hydroflux_amd.build

compiles hydroflux_amd/csrc/hf2d.hip (single grids), hf2d_batch.hip (batched
scenarios) and hf2d_fast.hip (the fast FP64 kind's stage kernels) into
hydroflux_amd/lib/libhf2d.so. -ffp-contract=off keeps every a*b+c unfused in
the first two (the reference's Rust arithmetic never contracts), which is what
makes the f64 step bit-identical to the CPU restatement; the fast FP64 kind
is built with contraction. Packed FP32 instructions are disabled.
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


# Translation units and their contraction flag: the bitwise kinds never fuse
# a*b+c; the fast FP64 kind's kernels (hf2d_fast.hip) are built with FMA
# contraction.
SOURCES = (("hf2d.hip", "-ffp-contract=off"), ("hf2d_batch.hip", "-ffp-contract=off"),
           ("hf2d_fast.hip", "-ffp-contract=fast"))


def build_library() -> str:
    out_dir = os.path.join(HERE, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libhf2d.so")
    tmp = out + ".tmp"
    objs = []
    for src, contract in SOURCES:
        obj = os.path.join(out_dir, src.replace(".hip", ".o"))
        flags = [f for f in HIP_FLAGS if f not in ("-shared", "-ffp-contract=off")]
        _run([HIPCC, *flags, contract, "-c", "-o", obj, os.path.join(HERE, "csrc", src)])
        objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", tmp, *objs])
    for obj in objs:
        os.remove(obj)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build_library()
    sys.exit(0)
