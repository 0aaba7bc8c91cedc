"""GPU: the in-place level 0 of the packed 2D path against the oracle and
against the scratch route (dwt_gpu_level0_set), through the C ABI.

Bar: bit-exact to the oracle (the reference's _s / _d arithmetic) forward and
inverse, the same with level 0 in place and through scratch, and no
dependency wait past its limit (dwt_gpu_device_check).  The in-place route is
off by default; these tests switch it on.
"""
import sys

import numpy as np
import pytest

import cases
import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHAPES = [(4096, 4096, 8), (3001, 2047, 8), (1000, 700, -1), (517, 389, 2), (64, 1000, -1), (1000, 64, 3),
          (130, 71, -1), (2049, 1025, 5), (241, 2, 1), (2, 241, -1)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0
    assert libdwt_amd.lib().dwt_gpu_level0_set(1, 8) == 0
    yield
    libdwt_amd.lib().dwt_gpu_level0_set(0, 8)  # the library's default


def check_clean():
    lib = libdwt_amd.lib()
    assert lib.dwt_gpu_device_check(None) == 0, lib.dwt_gpu_last_error()


def round_trip(img, w, J):
    t = torch.from_numpy(img).cuda()
    j = libdwt_amd.fwd2(t, w, J)
    torch.cuda.synchronize()
    c = t.cpu().numpy()
    libdwt_amd.inv2(t, w, j)
    torch.cuda.synchronize()
    return j, c, t.cpu().numpy()


@pytest.mark.parametrize("W,H,J", SHAPES)
def test_inplace_level0_matches_oracle(W, H, J):
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    img = workloads.sinusoid_image(W, H, seed=W + H)
    for w in cases.WAVELETS:
        say(f"level 0 in place {W}x{H} {w} J={J}")
        ref, jr = o.fwd2(img, w, J)
        rec = o.inv2(ref, w, jr)
        assert lib.dwt_gpu_level0_set(1, 8) == 0
        j, c, r = round_trip(img, w, J)
        assert j == jr
        np.testing.assert_array_equal(c, ref, err_msg="forward, in place")
        np.testing.assert_array_equal(r, rec, err_msg="inverse, in place")
        check_clean()
        assert lib.dwt_gpu_level0_set(0, 0) == 0
        j2, c2, r2 = round_trip(img, w, J)
        assert lib.dwt_gpu_level0_set(1, 0) == 0
        np.testing.assert_array_equal(c2, c, err_msg="forward, scratch route")
        np.testing.assert_array_equal(r2, r, err_msg="inverse, scratch route")


def test_inplace_level0_fp64():
    from oracle.oracle import oracle
    o = oracle()
    for (W, H, J) in ((1031, 523, -1), (640, 480, 4)):
        img = np.random.default_rng(3).standard_normal((H, W))
        for w in cases.WAVELETS:
            say(f"level 0 in place fp64 {W}x{H} {w}")
            ref, jr = o.fwd2(img, w, J)
            j, c, r = round_trip(img, w, J)
            np.testing.assert_array_equal(c, ref)
            np.testing.assert_array_equal(r, o.inv2(ref, w, jr))
    check_clean()


@pytest.mark.parametrize("band", [1, 3, 64])
def test_inplace_level0_other_bands(band):
    """Other XCD band widths (the order and the fallback parts change)."""
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    img = workloads.sinusoid_image(1537, 1283, seed=4)
    try:
        assert lib.dwt_gpu_level0_set(1, band) == 0
        for w in cases.WAVELETS:
            say(f"level 0 band {band} {w}")
            ref, jr = o.fwd2(img, w, -1)
            j, c, r = round_trip(img, w, -1)
            np.testing.assert_array_equal(c, ref)
            np.testing.assert_array_equal(r, o.inv2(ref, w, jr))
        check_clean()
    finally:
        lib.dwt_gpu_level0_set(1, 8)


def test_inplace_level0_four_row_tiles():
    """Level-0 tiles of 4 rows (the tile table's single shot S = 4); a
    rolling shape sends level 0 through scratch instead."""
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    img = workloads.sinusoid_image(1200, 900, seed=8)
    cls = 10  # height class of 900 rows
    try:
        for (S, U) in ((4, 0), (16, 4)):
            for fwd in (0, 1):
                assert lib.dwt_gpu_tile_set(fwd, cls, S, U) == 0
            for w in cases.WAVELETS:
                say(f"level 0 tiles S={S} U={U} {w}")
                ref, jr = o.fwd2(img, w, -1)
                j, c, r = round_trip(img, w, -1)
                np.testing.assert_array_equal(c, ref)
                np.testing.assert_array_equal(r, o.inv2(ref, w, jr))
        check_clean()
    finally:
        for fwd in (0, 1):
            lib.dwt_gpu_tile_set(fwd, cls, 0, 0)


def test_inplace_level0_many_launches_in_a_row():
    """Forty transforms back to back on one stream (a new epoch each), then
    the result against the oracle's forward of the same image."""
    from oracle.oracle import oracle
    o = oracle()
    img = workloads.sinusoid_image(2048, 1536, seed=6)
    t = torch.from_numpy(img).cuda()
    say("level 0: 20 forward+inverse round trips")
    for _ in range(20):
        j = libdwt_amd.fwd2(t, "cdf97", 6)
        libdwt_amd.inv2(t, "cdf97", j)
    libdwt_amd.fwd2(t, "cdf97", 6)
    torch.cuda.synchronize()
    check_clean()
    # the round trips are not the identity in FP32: compare with the oracle
    # doing the same 20 round trips
    x = img
    for _ in range(20):
        c, jj = o.fwd2(x, "cdf97", 6)
        x = o.inv2(c, "cdf97", jj)
    ref, _ = o.fwd2(x, "cdf97", 6)
    np.testing.assert_array_equal(t.cpu().numpy(), ref)
