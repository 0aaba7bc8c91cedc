"""Tile-shape autotuner for the packed level kernels and the batched path.

A level kernel's shape is (S, U): a wavefront walks S rows of its 128-column
strip and keeps U row pairs of loads in flight (U = 0: it requests every row
of its segment at once).  Tall segments amortise the halo rows, short ones
give a level more wavefronts; deeper prefetch keeps more bytes in flight.  The best shape depends on the level's size, so the tuner times every
level of a workload for each candidate shape (HIP events around each launch,
recorded by the library on the launch stream), keeps the fastest shape per
(direction, height class = ceil(log2 H)) and installs it with
dwt_gpu_tile_set.  The table can be saved as JSON and re-applied later.

Batches (tune_batch): the segment rows of each level's batched launch
(dwt_gpu_batch_rows_set: 8, 16 or 32) and the region size from which every
image's remaining levels run as one launch in LDS (dwt_gpu_batch_tail_set),
timed on the batch to be run: rows per level by the per-launch events of the
batched level launches (forward + inverse), then the tail threshold by the
whole step.
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


BATCH_ROWS = (8, 16, 32)
BATCH_TAILS = (0, 16384, 32768, 65536, 131072)


def apply_batch(table: dict) -> None:
    """Install a batch table {"rows": {level: rows}, "tail_bytes": n}."""
    lib = libdwt_amd.lib()
    for lev, rows in table.get("rows", {}).items():
        if lib.dwt_gpu_batch_rows_set(int(lev), int(rows)) != 0:
            raise ValueError(f"bad batch rows entry {lev}: {rows}")
    if "tail_bytes" in table and lib.dwt_gpu_batch_tail_set(int(table["tail_bytes"])) != 0:
        raise ValueError(f"bad batch tail threshold {table['tail_bytes']}")


def reset_batch() -> None:
    lib = libdwt_amd.lib()
    for lev in range(32):
        lib.dwt_gpu_batch_rows_set(lev, 0)
    lib.dwt_gpu_batch_tail_set(65536)


def tune_batch(images, wavelet: str = "cdf97", storage: str = "f32", reps: int = 5,
               rows=BATCH_ROWS, tails=BATCH_TAILS, levels=-1) -> dict:
    """Tune the batched path for a list of 2D float32 CUDA tensors (the batch
    to be run, with its levels: one count or one per image) and install the
    result.  Needs a GPU."""
    import torch

    lib = libdwt_amd.lib()

    def run(n):
        for _ in range(n):
            c, s, done = libdwt_amd.batch_forward(images, wavelet, storage, levels)
            libdwt_amd.batch_inverse(c, done, wavelet, storage, s)

    def timed_levels():
        run(1)  # builds the plans (untimed)
        torch.cuda.synchronize()
        libdwt_amd.profile_reset()
        libdwt_amd.profile_enable(True)
        run(reps)
        torch.cuda.synchronize()
        libdwt_amd.profile_enable(False)
        out = {}
        for kind in (8, 9):
            for lev in range(32):
                n, ms = libdwt_amd.profile_read(kind, lev)
                if n:
                    out[lev] = out.get(lev, 0.0) + 1000.0 * ms / n
        libdwt_amd.profile_reset()
        return out

    reset_batch()
    lib.dwt_gpu_batch_tail_set(0)  # every level as a batched launch while rows are tuned
    timing = {"rows_us": {}, "tail_step_us": {}}
    best_rows = {}
    for r in rows:
        for lev in range(32):
            lib.dwt_gpu_batch_rows_set(lev, r)
        for lev, us in timed_levels().items():
            timing["rows_us"].setdefault(str(lev), {})[str(r)] = round(us, 3)
            if lev not in best_rows or us < best_rows[lev][1]:
                best_rows[lev] = (r, us)
    for lev in range(32):
        lib.dwt_gpu_batch_rows_set(lev, best_rows[lev][0] if lev in best_rows else 0)
    best_tail = None
    for tb in tails:
        lib.dwt_gpu_batch_tail_set(tb)
        run(1)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run(reps)
        b.record()
        torch.cuda.synchronize()
        us = 1000.0 * a.elapsed_time(b) / reps
        timing["tail_step_us"][str(tb)] = round(us, 3)
        if best_tail is None or us < best_tail[1]:
            best_tail = (tb, us)
    table = {"rows": {str(lev): v[0] for lev, v in sorted(best_rows.items())},
             "tail_bytes": best_tail[0]}
    reset_batch()
    apply_batch(table)
    table["timing_us"] = timing
    table["workload"] = f"{len(images)} images {wavelet} {storage}"
    return table
