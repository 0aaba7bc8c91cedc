"""Guards on the generated gfx950 device code (CPU only: hipcc cross-compiles).

* No flat memory instructions.  LDS must be addressed by ds_* instructions:
  a flat pointer to LDS lets the compiler fold x[i-1] into (x - 4) plus an
  instruction offset, and for a buffer at LDS offset 0 that base lies below
  the LDS aperture, which the hardware faults on.
* No scalar-cache stores or atomics, and no packed FP32 arithmetic
  (v_pk_add/mul/fma_f32), which the build disables.
* No register spills (private segment) in the level kernels.
"""
import os
import re
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from libdwt_amd import build as hipbuild

SOURCES = hipbuild.SOURCES


@pytest.fixture(scope="module")
def asm():
    """Device assembly of every source, compiled in parallel."""
    with tempfile.TemporaryDirectory() as d:
        def compile_one(src):
            s = os.path.join(d, src + ".s")
            cmd = [hipbuild.HIPCC, *hipbuild.FLAGS, "--cuda-device-only", "-S",
                   os.path.join(hipbuild.CSRC, src), "-o", s]
            r = subprocess.run(cmd, capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            with open(s) as fh:
                return src, fh.read()

        workers = max(1, min(len(SOURCES), os.cpu_count() or 1))
        with ThreadPoolExecutor(max_workers=workers) as ex:
            return dict(ex.map(compile_one, SOURCES))


def test_no_flat_memory_instructions(asm):
    for src, text in asm.items():
        hits = re.findall(r"^\s+(flat_(?:load|store|atomic)\w*)", text, flags=re.M)
        assert not hits, f"{src}: {sorted(set(hits))}"


def test_no_scalar_stores(asm):
    for src, text in asm.items():
        hits = re.findall(r"^\s+(s_(?:store|buffer_store|atomic|dcache_wb)\w*)", text, flags=re.M)
        assert not hits, f"{src}: {sorted(set(hits))}"


def test_no_packed_fp32(asm):
    for src, text in asm.items():
        hits = re.findall(r"^\s+(v_pk_(?:add|mul|fma)_f32)", text, flags=re.M)
        assert not hits, f"{src}: {sorted(set(hits))}"


def test_level_kernels_do_not_spill(asm):
    text = asm["tile_kernels.hip"]
    sizes = re.findall(r"\.set (_ZN3dwt1\d(?:fwd|inv)_level_kernel\w+)\.private_seg_size, (\d+)", text)
    assert sizes
    spilled = [k for k, v in sizes if int(v) != 0]
    assert not spilled, spilled
