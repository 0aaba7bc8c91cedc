"""GPU: a C program built with plain gcc against include/libdwt.h and
libdwt_gpu.so uses the drop-in transforms and dwt_util_* helpers on host
images (staged through the device), and the device-resident performance
harness, as a libdwt program would."""
import os
import subprocess
import sys

import numpy as np
import pytest

import libdwt_amd

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_c_program_roundtrips_and_perf(tmp_path):
    inc = os.path.dirname(libdwt_amd.HEADER)
    libdir = os.path.dirname(libdwt_amd.LIB_PATH)
    exe = tmp_path / "dropin_program"
    r = subprocess.run(["gcc", "-std=gnu99", "-O2", "-Wall", "-I", inc,
                        os.path.join(HERE, "c", "dropin_program.c"), "-L", libdir, "-ldwt_gpu",
                        f"-Wl,-rpath,{libdir}", "-lm", "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    print("running the C program", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    sys.stdout.write(r.stderr[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert r.stdout.strip().endswith("perf ok")
    assert r.stderr.count("roundtrip ok") == 3
    lines = (tmp_path / "perf_fwd.txt").read_text().split("\n")
    sizes = [int(l.split("\t")[0]) for l in lines if l.strip()]
    # the reference's sweep: x = 16, then (int)(x * 1.28f + 1 - FLT_EPSILON) in
    # single precision while x <= 100
    f = np.float32
    want, x = [], 16
    while x <= 100:
        want.append(x * x)
        x = int(f(f(f(x) * f(1.28)) + f(1.0)) - f(np.finfo(np.float32).eps))
    assert sizes == want
