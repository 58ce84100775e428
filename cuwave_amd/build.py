"""Build `libcuwave_amd.so` from the HIP sources for gfx950: `python -m cuwave_amd.build`.

Every translation unit (the C ABI of the pressure family in `csrc/scalar.hip`
and of the elastic family in `csrc/elastic.hip`, of the vector Maxwell
family in `csrc/maxwell.hip`, of the cell-assembled elastic family in
`csrc/anisotropic.hip`, the U-Net's matrix-core
kernels in `csrc/unet.hip`, the step kernels of one
precision and radius in each `csrc/steps_*.hip` and `csrc/elastic_*.hip`)
compiles in its own process, all at once, and the objects link into one
shared library.

Flags: no contraction (the kernels round as the reference's IEEE path does, an
fma only where written), denormals kept, correctly rounded division (hipcc's
default), and no packed-fp32 instruction selection.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUTPUT = HERE / "libcuwave_amd.so"
BUILD = HERE / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[Path]:
    return [CSRC / "scalar.hip", *sorted(CSRC.glob("steps_*.hip")), CSRC / "elastic.hip",
            *sorted(CSRC.glob("elastic_*.hip")), CSRC / "maxwell.hip", CSRC / "anisotropic.hip",
            CSRC / "unet.hip"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode:
        raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{done.stderr}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source into the library next to this file; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda so: _run([HIPCC, *FLAGS, "-c", str(so[0]), "-o", str(so[1])], verbose),
                      zip(srcs, objs)))
    tmp = OUTPUT.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], verbose)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
