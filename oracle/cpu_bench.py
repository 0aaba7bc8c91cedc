"""ORACLE -- test infrastructure only: the CPU baseline leg of bench.py.

Times the reference's 2D drivers (oracle/_ref, the reference compiled from its
own sources; the clean-room restatement liboracle.so when _ref is absent) on
this host, the way the reference's own harness does
(src/libdwt.c:8440-8492): forward and inverse timed separately, the minimum
over repeated runs, each accel mode tried (9 = SSE, 0 = scalar; they give
bit-identical results) and the faster kept.  Run as a child process so that
OMP_NUM_THREADS / OMP_PROC_BIND take effect before the OpenMP runtime starts:

    OMP_PROC_BIND=close OMP_NUM_THREADS=16 python oracle/cpu_bench.py \\
        --width 4096 --height 4096 --wavelet cdf97 --levels 8 --threads 16 --seconds 6

Prints one JSON line.  Only bench.py's cpu_baseline leg runs this script.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, this directory comes first on sys.path and would shadow
# the `oracle` package with its own oracle.py: resolve from the repository root
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as orc  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=4096)
    ap.add_argument("--wavelet", default="cdf97")
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=6.0)
    ap.add_argument("--seed", type=int, default=3)
    a = ap.parse_args()
    from libdwt_amd import workloads  # the benchmark's image generator (numpy only)

    img = workloads.sinusoid_image(a.width, a.height, seed=a.seed)
    H, W = img.shape
    if orc.have_ref():
        lib, kind, modes = orc.Reference(), "reference", (9, 0)
    else:
        lib, kind, modes = orc.oracle(), "port", (None,)
    lib.set_threads(a.threads)
    buf = np.empty_like(img)
    best = {"fwd": float("inf"), "inv": float("inf"), "step": float("inf")}
    best_mode, runs = None, 0
    per_mode = a.seconds / len(modes)
    for mode in modes:
        if mode is not None:
            lib.set_accel(mode)
        t_end = time.perf_counter() + per_mode
        n = 0
        while n < 2 or (time.perf_counter() < t_end and n < 200):
            buf[...] = img
            t0 = time.perf_counter()
            j = lib.fwd2_raw(buf, a.wavelet, 4 * W, 4, W, H, W, H, a.levels)
            t1 = time.perf_counter()
            lib.inv2_raw(buf, a.wavelet, 4 * W, 4, W, H, W, H, j)
            t2 = time.perf_counter()
            n += 1
            if t2 - t0 < best["step"]:
                best["step"], best_mode = t2 - t0, mode
            best["fwd"] = min(best["fwd"], t1 - t0)
            best["inv"] = min(best["inv"], t2 - t1)
        runs += n
    out = {
        "kind": kind, "threads": a.threads, "runs": runs, "accel": best_mode,
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
        "fwd_s": best["fwd"], "inv_s": best["inv"], "step_s": best["step"],
        "Mpix_s": W * H / best["step"] / 1e6,
        "Mpix_s_fwd_plus_inv_mins": W * H / (best["fwd"] + best["inv"]) / 1e6,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
