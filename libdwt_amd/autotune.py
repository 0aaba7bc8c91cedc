"""Tile-shape autotuner for the packed level kernels.

A level kernel's shape is (S, U): a wavefront walks S rows of its 128-column
strip and keeps U row pairs of loads in flight (U = 0: it requests every row
of its segment at once).  Tall segments amortise the halo rows, short ones
give a level more wavefronts; deeper prefetch keeps more bytes in flight.  The best shape depends on the level's size, so the tuner times every
level of a workload for each candidate shape (HIP events around each launch,
recorded by the library on the launch stream), keeps the fastest shape per
(direction, height class = ceil(log2 H)) and installs it with
dwt_gpu_tile_set.  The table can be saved as JSON and re-applied later.
"""
from __future__ import annotations

import json
import math

import libdwt_amd

# (S, U): U = 4 / 8 rolling prefetch with any S; U = 0 single shot (the whole
# segment requested at once), S in 2, 4, 8, 12, 16, 20, 24, 32
CANDIDATES = ([(S, 4) for S in (8, 16, 32, 64)] + [(S, 8) for S in (16, 32)]
              + [(S, 0) for S in (2, 4, 8, 12, 16, 20, 24, 32)])


def height_class(h: int) -> int:
    return 0 if h <= 1 else math.ceil(math.log2(h))


def apply(table: dict) -> None:
    """Install a table {"fwd": {class: [S, U]}, "inv": {...}}."""
    lib = libdwt_amd.lib()
    for d, fwd in (("fwd", 1), ("inv", 0)):
        for cls, (S, U) in table.get(d, {}).items():
            rc = lib.dwt_gpu_tile_set(fwd, int(cls), int(S), int(U))
            if rc != 0:
                raise ValueError(f"bad tile entry {d} {cls}: {S},{U}")


def reset() -> None:
    lib = libdwt_amd.lib()
    for fwd in (0, 1):
        for cls in range(32):
            lib.dwt_gpu_tile_set(fwd, cls, 0, 0)


def load(path: str) -> dict:
    with open(path) as fh:
        table = json.load(fh)
    apply(table)
    return table


def tune(width: int, height: int, levels: int, wavelet: str = "cdf97", reps: int = 5,
         candidates=CANDIDATES) -> dict:
    """Time every level of a width x height, `levels`-level transform for each
    candidate shape and install the fastest per level.  Needs a GPU."""
    import torch

    from libdwt_amd import workloads

    img = workloads.sinusoid_image(width, height, seed=5)
    t = torch.from_numpy(img).cuda()
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    pitch = 4 * width
    lib = libdwt_amd.lib()
    sizes = workloads.level_sizes(width, height, levels)
    best = {"fwd": {}, "inv": {}}
    timing = {"fwd": {}, "inv": {}}

    def step():
        j = libdwt_amd.transform_2d_async(wavelet, True, t.data_ptr(), pitch, 4, width, height,
                                          j=levels, stream=sp)
        libdwt_amd.transform_2d_async(wavelet, False, t.data_ptr(), pitch, 4, width, height, j=j,
                                      stream=sp)

    for (S, U) in candidates:
        for fwd in (0, 1):
            for cls in range(32):
                lib.dwt_gpu_tile_set(fwd, cls, S, U)
        step()
        torch.cuda.synchronize()
        libdwt_amd.profile_reset()
        libdwt_amd.profile_enable(True)
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        libdwt_amd.profile_enable(False)
        for kind, d in ((0, "fwd"), (1, "inv")):
            for lev, (w, h) in enumerate(sizes):
                n, ms = libdwt_amd.profile_read(kind, lev)
                if not n:
                    continue
                us = 1000.0 * ms / n
                cls = str(height_class(h))
                timing[d].setdefault(cls, {})[f"{S},{U}"] = round(us, 3)
                if cls not in best[d] or us < best[d][cls][2]:
                    best[d][cls] = (S, U, us)
        libdwt_amd.profile_reset()
    reset()
    table = {d: {c: [v[0], v[1]] for c, v in best[d].items()} for d in best}
    apply(table)
    table["timing_us"] = timing
    table["workload"] = f"{width}x{height} {wavelet} J={levels}"
    return table


def save(table: dict, path: str) -> None:
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
