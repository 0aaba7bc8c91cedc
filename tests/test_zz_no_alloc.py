"""The allocation-free entry point dwt_gpu_2d_noalloc against the existing
dwt_gpu_2d_async, through the C ABI: bit-identical coefficients and
reconstructions on two calls in a row (the second with a smaller image, so
the shared static scratch holds the first call's contents), both wavelets,
a padded row pitch, an image the single-workgroup kernel takes whole, and an
image too large for the static scratch (rejected, data untouched)."""
import ctypes as C

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

WAV = {"cdf97": 0, "cdf53": 1}
FWD, INV = 0, 1


def lib():
    L = libdwt_amd.lib()
    assert L.dwt_gpu_init() == 0, L.dwt_gpu_last_error()
    return L


def frame(W, H, pitch, seed):
    """A W x H sinusoid image in device memory with row pitch `pitch` floats
    (padding filled with a marker that must survive)."""
    buf = torch.full((H, pitch), -7.25, dtype=torch.float32)
    buf[:, :W] = torch.from_numpy(workloads.sinusoid_image(W, H, seed=seed))
    return buf.cuda()


def via_async(L, t, W, H, wav, direction, j, st):
    jj = C.c_int(j)
    rc = L.dwt_gpu_2d_async(WAV[wav], direction, 0, C.c_void_p(t.data_ptr()), 4 * t.shape[1], 4, W, H, W, H,
                            C.byref(jj), 0, 0, C.c_void_p(st))
    assert rc == 0, L.dwt_gpu_last_error()
    return jj.value


def via_noalloc(L, t, W, H, wav, direction, j, st):
    jj = C.c_int(j)
    rc = L.dwt_gpu_2d_noalloc(WAV[wav], direction, C.c_void_p(t.data_ptr()), 4 * t.shape[1], W, H, C.byref(jj),
                              C.c_void_p(st))
    assert rc == 0, L.dwt_gpu_last_error()
    return jj.value


@pytest.mark.parametrize("wav", ["cdf97", "cdf53"])
def test_noalloc_matches_async_two_calls(wav):
    L = lib()
    s = torch.cuda.Stream()
    st = s.cuda_stream
    # second image smaller than the first: the static scratch still holds the
    # first call's data when the second runs
    for (W, H, pitch, J, seed) in [(1500, 1100, 1536, -1, 1), (700, 333, 700, 4, 2), (3001, 2047, 3001, 8, 3)]:
        print(f"noalloc {wav} {W}x{H} pitch {pitch} J={J}", flush=True)
        a = frame(W, H, pitch, seed)
        b = a.clone()
        torch.cuda.synchronize()
        ja = via_async(L, a, W, H, wav, FWD, J, st)
        jb = via_noalloc(L, b, W, H, wav, FWD, J, st)
        s.synchronize()
        assert ja == jb
        assert torch.equal(a, b), f"forward differs: {(a - b).abs().max().item()}"
        via_async(L, a, W, H, wav, INV, ja, st)
        via_noalloc(L, b, W, H, wav, INV, jb, st)
        s.synchronize()
        assert torch.equal(a, b), f"inverse differs: {(a - b).abs().max().item()}"
        img = torch.from_numpy(workloads.sinusoid_image(W, H, seed=seed))
        assert float((b[:, :W].cpu() - img).abs().max()) < 1e-4
        assert bool((b[:, W:] == -7.25).all()), "row padding was written"


def test_noalloc_whole_image_block_and_repeat():
    """An image the single-workgroup kernel takes whole needs no scratch;
    several calls in a row on one stream stay equal to the async path."""
    L = lib()
    s = torch.cuda.Stream()
    st = s.cuda_stream
    for (W, H) in [(128, 128), (64, 40), (90, 91)]:
        a = frame(W, H, W, 5)
        b = a.clone()
        for _ in range(3):
            j = via_async(L, a, W, H, "cdf97", FWD, -1, st)
            via_noalloc(L, b, W, H, "cdf97", FWD, -1, st)
            via_async(L, a, W, H, "cdf97", INV, j, st)
            via_noalloc(L, b, W, H, "cdf97", INV, j, st)
        s.synchronize()
        assert torch.equal(a, b)


def test_noalloc_rejects_what_it_cannot_hold():
    L = lib()
    s = torch.cuda.Stream()
    st = s.cuda_stream
    W = H = 4096  # needs about 88 MB of scratch
    a = frame(W, H, W, 7)
    before = a.clone()
    jj = C.c_int(8)
    rc = L.dwt_gpu_2d_noalloc(0, FWD, C.c_void_p(a.data_ptr()), 4 * W, W, H, C.byref(jj), C.c_void_p(st))
    s.synchronize()
    assert rc == -1
    assert b"scratch" in L.dwt_gpu_last_error()
    assert torch.equal(a, before)
    # an element stride the packed path cannot take, and host memory
    rc = L.dwt_gpu_2d_noalloc(0, FWD, C.c_void_p(a.data_ptr()), 4 * W + 2, 100, 100, C.byref(jj), C.c_void_p(st))
    assert rc == -1
    host = np.zeros((64, 64), np.float32)
    rc = L.dwt_gpu_2d_noalloc(0, FWD, C.c_void_p(host.ctypes.data), 256, 64, 64, C.byref(jj), C.c_void_p(st))
    assert rc == -2
