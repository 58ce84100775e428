"""Build `libcuwave_amd.so` from the HIP sources for gfx950: `python -m cuwave_amd.build`.

Flags: no contraction (the kernels round as the reference's IEEE path does, an
fma only where written), denormals kept, correctly rounded division (hipcc's
default), and no packed-fp32 instruction selection.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = [HERE / "csrc" / "scalar.hip"]
OUTPUT = HERE / "libcuwave_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into the library next to this file; returns its path."""
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode:
        raise RuntimeError(f"hipcc failed:\n{done.stderr}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
