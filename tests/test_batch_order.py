"""CPU check of the batched level kernels' workgroup order, through the
library's own table builder and mapping (dwt_gpu_batch_order_check).

A batched level launch deals its workgroups to images in two ways: large
images (at least 1024 workgroups at a level) fill whole 64-workgroup rounds
and are banded per image across the XCDs; everything else goes through the
launch-wide order, with padding workgroups in between.  The check replays
every workgroup of every level launch and requires each tile (segment, strip)
of each image to be computed by exactly one wave.  Batches mix large images
(2048^2 and up, which take the per-image bands at level 0 and 1) with small
and odd-sized ones, in both directions and for both wavelets.
"""
import ctypes as C

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import BatchImage, workloads

WAVELETS = {"cdf97": 0, "cdf53": 1}


def descriptors(sizes, levels=-1, direction=0):
    arr = (BatchImage * len(sizes))()
    for k, (w, h) in enumerate(sizes):
        # distinct fake buffers: the check never touches them
        arr[k].src = 0x100000000 + k * 0x10000000
        arr[k].dst = 0x800000000 + k * 0x10000000
        arr[k].src_pitch = 4 * w
        arr[k].dst_pitch = 4 * w
        arr[k].width = w
        arr[k].height = h
        if direction == 0:
            arr[k].levels = levels
        else:  # the inverse needs explicit levels
            arr[k].levels = min(8, int(np.floor(np.log2(min(w, h)))))
    return arr


def check(sizes, wavelet="cdf97", direction=0):
    L = libdwt_amd.lib()
    L.dwt_gpu_batch_order_check.argtypes = [C.c_int, C.c_int, C.POINTER(BatchImage), C.c_int]
    arr = descriptors(sizes, direction=direction)
    rc = L.dwt_gpu_batch_order_check(WAVELETS[wavelet], direction, arr, len(sizes))
    assert rc == 0, L.dwt_gpu_last_error().decode()


@pytest.mark.parametrize("wavelet", ["cdf97", "cdf53"])
@pytest.mark.parametrize("direction", [0, 1])
def test_large_and_small_images(wavelet, direction):
    sizes = [(4096, 4096), (129, 300), (2048, 2048), (3001, 2047), (64, 64), (2500, 3100), (1023, 1021)]
    check(sizes, wavelet, direction)


@pytest.mark.parametrize("seed", range(6))
def test_random_batches(seed):
    rng = np.random.default_rng(seed)
    sizes = workloads.uneven_sizes(int(rng.integers(1, 40)), 64, 2048, seed=seed)
    # put a few large images at random places
    for _ in range(int(rng.integers(0, 4))):
        w, h = int(rng.integers(1900, 4100)), int(rng.integers(1900, 4100))
        sizes.insert(int(rng.integers(0, len(sizes) + 1)), (w, h))
    check(sizes, "cdf97", int(seed % 2))


def test_benchmark_batch():
    check(workloads.uneven_sizes(256, 64, 2048, seed=4))
    check(workloads.uneven_sizes(256, 64, 2048, seed=4), direction=1)
