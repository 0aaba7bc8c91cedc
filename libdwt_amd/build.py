"""Build the HIP library in-tree: libdwt_amd/libdwt_gpu.so for gfx950.

Every source is compiled with
  * -ffp-contract=off       : lifting updates stay add, mul, add (no FMA), the
                              arithmetic the bit-exact contract is stated in;
  * -packed-fp32-ops off    : no packed FP32 instructions are generated.
Usage: python -m libdwt_amd.build   (or build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdwt_gpu.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["line_kernels.hip", "tile_kernels.hip", "block_kernels.hip", "inplace.hip", "mid_kernels.hip", "pyr_kernels.hip", "batch.hip", "capi.hip",
           "batch_host.hip", "server.hip", "lines.hip", "util.hip", "verify.hip", "sparse.hip"]
ARCH = os.environ.get("DWT_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src: str, objdir: str = OBJDIR, defines=()) -> str:
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: str = OUT, defines=()) -> str:
    """Compile and link the library to `out` (extra -D defines optional)."""
    objdir = OBJDIR if not defines else os.path.join(
        tempfile.gettempdir(), "libdwt_amd_build_" + "_".join(d.replace("=", "") for d in defines))
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir, defines), SOURCES))
    tmp = out + ".tmp"
    # --no-undefined: a symbol missing from every object fails the build here
    # rather than the first dlopen
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-Wl,--no-undefined", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
