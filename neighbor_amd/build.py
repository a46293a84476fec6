"""Build the HIP library libneighbor_amd.so (gfx950) in-tree with hipcc.

Flags that matter for the results:
  -ffp-contract=off   no FMA contraction: every float op is the single IEEE op the
                      reference's HIP path performs (explicit fmaf stays fused)
  (no fast-math, no -fgpu-flush-denormals-to-zero: FP32 denormals are kept)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libneighbor_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "capi.hip")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
    "-I", os.path.join(HERE, "include"),
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + SOURCES + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed building libneighbor_amd.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


EXAMPLES = [os.path.join(HERE, "examples", "periodic_neighbors.hip"),
            # device probe of the rounding helpers and bounding volumes, run by the GPU tests
            os.path.join(ROOT, "tests", "device", "device_math.hip")]


def build_examples(verbose: bool = False) -> list:
    """Compile the C++ programs over the template API (the examples, used as a reference caller
    uses the API, and the tests' device probes) into <name>.bin beside their sources."""
    flags = [f for f in FLAGS if f not in ("-fPIC", "-shared")]
    out = []
    for src in EXAMPLES:
        exe = os.path.splitext(src)[0] + ".bin"
        cmd = [HIPCC] + flags + [src, "-o", exe + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"hipcc failed building {src}")
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_examples(verbose=True))
