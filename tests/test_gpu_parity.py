"""GPU parity: the HIP library against the reference's golden digests and the
oracle, through the C ABI.

Bar: bit-exact (NaN-free inputs) to the reference's _s and _d outputs for
every case; FP32 results within 1e-5 max-abs of the FP64 oracle on the
benchmark images.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FP32_TOL = 1e-5  # max-abs vs the FP64 oracle on the sinusoid+noise images
TAIL_DEFAULT = 64 * 1024  # the library's default single-workgroup region
TAIL_MAX = 150 * 1024     # the largest region dwt_gpu_tail_set accepts


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "golden_hashes.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="module")
def arrays():
    return np.load(os.path.join(GOLD, "golden_arrays.npz"))


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    libdwt_amd.lib()
    assert libdwt_amd.lib().dwt_gpu_init() == 0, libdwt_amd.lib().dwt_gpu_last_error()
    return torch.device("cuda:0")


def dropin(mode, timeout=140):
    """Run the drop-in entry points over the fixture cases in a child process:
    they abort on error like the reference, and a child keeps that abort (and
    the library's message, and the case it names) contained and visible."""
    worker = os.path.join(os.path.dirname(__file__), "gpu_dropin_worker.py")
    r = subprocess.run([sys.executable, "-u", worker, mode], capture_output=True, text=True,
                       timeout=timeout)
    tail = "\n".join(r.stderr.strip().splitlines()[-6:])
    assert r.returncode == 0, f"worker {mode} exited {r.returncode}:\n{tail}"
    res = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"  {mode}: {res['cases']} cases, {len(res['mismatches'])} mismatches")
    return res


@pytest.mark.parametrize("mode", ["dense", "sparse", "host", "lines", "slice", "pinned", "example"])
def test_dropin_entry_points(mode):
    """Dense frames, sparse frames / zero padding / strided layouts, staged
    host memory, 1D lines N = 0..40 (strided, aliased), the 1D full-depth
    slice, pinned host memory and the reference example's 2053-byte pitch,
    all through the drop-in C ABI, against the reference's digests."""
    say(f"drop-in sweep: {mode}")
    res = dropin(mode)
    assert res["cases"] > 0
    assert not res["mismatches"], res["mismatches"][:8]


@pytest.mark.parametrize("cfg", cases.CONFIG_IMAGES, ids=lambda c: c["name"])
@pytest.mark.parametrize("wavelet", cases.WAVELETS)
def test_config_images(gold, cfg, wavelet):
    from oracle.oracle import oracle
    say(f"config image {cfg['name']} {wavelet}")
    img = workloads.sinusoid_image(cfg["W"], cfg["H"], cfg["seed"])
    rec = gold["config"][cfg["name"]]
    assert cases.digest(img) == rec["input"]
    t = torch.from_numpy(img).cuda()
    j = libdwt_amd.fwd2(t, wavelet, cfg["J"])
    torch.cuda.synchronize()
    coef = t.cpu().numpy()
    assert j == rec[wavelet]["j"]
    assert cases.digest(coef) == rec[wavelet]["fwd"], "forward differs from the reference _s"
    # FP32 vs FP64 oracle
    o = oracle()
    cd, _ = o.fwd2(img.astype(np.float64), wavelet, cfg["J"])
    err = float(np.abs(coef - cd).max())
    say(f"  max-abs vs FP64 oracle: {err:.3e}")
    assert err <= FP32_TOL
    libdwt_amd.inv2(t, wavelet, j)
    torch.cuda.synchronize()
    rec_img = t.cpu().numpy()
    assert cases.digest(rec_img) == rec[wavelet]["inv"], "inverse differs from the reference _s"
    assert float(np.abs(rec_img - img).max()) <= FP32_TOL


def test_every_tile_shape_is_bit_exact(gold):
    """Every level-kernel instantiation the tile table can select (rolling
    prefetch and single shot, all segment heights), on the non-power-of-two
    benchmark image, both wavelets."""
    from libdwt_amd import autotune
    cfg = cases.CONFIG_IMAGES[1]
    img = workloads.sinusoid_image(cfg["W"], cfg["H"], cfg["seed"])
    rec = gold["config"][cfg["name"]]
    lib = libdwt_amd.lib()
    try:
        for (S, U) in autotune.CANDIDATES:
            say(f"tile shape S={S} U={U}")
            for fwd in (0, 1):
                for cls in range(32):
                    assert lib.dwt_gpu_tile_set(fwd, cls, S, U) == 0
            for w in cases.WAVELETS:
                t = torch.from_numpy(img).cuda()
                j = libdwt_amd.fwd2(t, w, cfg["J"])
                torch.cuda.synchronize()
                assert cases.digest(t.cpu().numpy()) == rec[w]["fwd"], (S, U, w)
                libdwt_amd.inv2(t, w, j)
                torch.cuda.synchronize()
                assert cases.digest(t.cpu().numpy()) == rec[w]["inv"], (S, U, w)
    finally:
        autotune.reset()


def test_fp64_config_image_matches_oracle():
    from oracle.oracle import oracle
    say("fp64 512x512 image")
    img = workloads.sinusoid_image(512, 512, 1).astype(np.float64)
    o = oracle()
    for w in cases.WAVELETS:
        t = torch.from_numpy(img).cuda()
        j = libdwt_amd.fwd2(t, w, 5)
        torch.cuda.synchronize()
        ref, _ = o.fwd2(img, w, 5)
        np.testing.assert_array_equal(t.cpu().numpy(), ref)
        libdwt_amd.inv2(t, w, j)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(t.cpu().numpy(), o.inv2(ref, w, j))


def test_async_api_on_torch_stream(gold):
    """Three forward+inverse round trips and a final forward on a side stream,
    enqueued without synchronisation; equal to the oracle doing the same."""
    from oracle.oracle import oracle
    say("async API on a torch stream")
    cfg = cases.CONFIG_IMAGES[1]
    img = workloads.sinusoid_image(cfg["W"], cfg["H"], cfg["seed"])
    o = oracle()
    want = img
    for _ in range(3):
        c, j = o.fwd2(want, "cdf97", cfg["J"])
        want = o.inv2(c, "cdf97", j)
    want, _ = o.fwd2(want, "cdf97", cfg["J"])
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        t = torch.from_numpy(img).cuda()
        for _ in range(3):
            j = libdwt_amd.transform_2d_async("cdf97", True, t.data_ptr(), 4 * cfg["W"], 4,
                                              cfg["W"], cfg["H"], j=cfg["J"], stream=s.cuda_stream)
            libdwt_amd.transform_2d_async("cdf97", False, t.data_ptr(), 4 * cfg["W"], 4,
                                          cfg["W"], cfg["H"], j=j, stream=s.cuda_stream)
        libdwt_amd.transform_2d_async("cdf97", True, t.data_ptr(), 4 * cfg["W"], 4, cfg["W"],
                                      cfg["H"], j=cfg["J"], stream=s.cuda_stream)
    s.synchronize()
    np.testing.assert_array_equal(t.cpu().numpy(), want)


BLOCK_SHAPES = [(128, 128), (127, 129), (255, 64), (100, 163), (2, 2), (3, 17), (33, 9),
                (1500, 5), (7, 2000), (90, 91), (190, 190), (376, 256), (301, 511)]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_block_kernel_whole_images(prec):
    """Images the single-workgroup LDS kernel transforms completely (and at
    the larger threshold images it takes over after level 0 or 1), bit-exact
    to the oracle, forward and inverse, both wavelets, full depth and J=2."""
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    dt = np.float32 if prec == "s" else np.float64
    rng = np.random.default_rng(5)
    try:
        for tb in (TAIL_MAX, 4 * 1024):
            assert lib.dwt_gpu_tail_set(tb) == 0
            for (W, H) in BLOCK_SHAPES:
                img = rng.standard_normal((H, W)).astype(dt)
                for w in cases.WAVELETS:
                    for J in (-1, 2):
                        say(f"block {W}x{H} {w} {prec} J={J} threshold {tb}")
                        ref, jr = o.fwd2(img, w, J)
                        t = torch.from_numpy(img).cuda()
                        j = libdwt_amd.fwd2(t, w, J)
                        assert j == jr
                        np.testing.assert_array_equal(t.cpu().numpy(), ref, err_msg="forward")
                        libdwt_amd.inv2(t, w, j)
                        np.testing.assert_array_equal(t.cpu().numpy(), o.inv2(ref, w, j),
                                                      err_msg="inverse")
    finally:
        lib.dwt_gpu_tail_set(TAIL_DEFAULT)


MID_SHAPES = [(1000, 700), (513, 511), (777, 1500), (257, 4000), (4000, 130), (96, 96), (300, 257)]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_mid_kernel_shapes(prec):
    """Three middle levels fused into one launch (each workgroup computes
    overlapping windows of the three levels in LDS), switched on from level 1,
    with and without the single-workgroup tail, full depth and J=4 (where the
    fused launch does the last levels): bit-exact to the oracle, forward and
    inverse, both wavelets."""
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    dt = np.float32 if prec == "s" else np.float64
    rng = np.random.default_rng(9)
    try:
        assert lib.dwt_gpu_mid_set(1 << 40) == 0
        for tb in (TAIL_DEFAULT, 0):
            assert lib.dwt_gpu_tail_set(tb) == 0
            for (W, H) in MID_SHAPES:
                img = rng.standard_normal((H, W)).astype(dt)
                for w in cases.WAVELETS:
                    for J in (-1, 4):
                        say(f"mid {W}x{H} {w} {prec} J={J} tail {tb}")
                        ref, jr = o.fwd2(img, w, J)
                        t = torch.from_numpy(img).cuda()
                        j = libdwt_amd.fwd2(t, w, J)
                        assert j == jr
                        np.testing.assert_array_equal(t.cpu().numpy(), ref, err_msg="forward")
                        libdwt_amd.inv2(t, w, j)
                        np.testing.assert_array_equal(t.cpu().numpy(), o.inv2(ref, w, j),
                                                      err_msg="inverse")
    finally:
        lib.dwt_gpu_mid_set(0)
        lib.dwt_gpu_tail_set(TAIL_DEFAULT)


SPARSE_FRAMES = [((512, 512), (300, 257)), ((1024, 1024), (1000, 700)), ((64, 256), (33, 200)),
                 ((256, 64), (255, 2))]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_sparse_frames_match_oracle(prec):
    """Sparse frames (inner image smaller than the outer frame, the rest of the
    frame random, as the garbage the reference transforms along): rows on the
    line kernel, columns on the coalesced column kernel through scratch; with
    and without zero padding and decompose_one, full depth and J=3, forward
    then inverse, bit-exact to the oracle."""
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    rng = np.random.default_rng(17)
    cases_ = [(f, zp, d1, J) for f in SPARSE_FRAMES for zp in (0, 1) for d1 in (0, 1) for J in (-1, 3)]
    cases_.append((((4096, 2048), (3001, 2047)), 0, 0, -1))
    for ((Wo, Ho), (Wi, Hi)), zp, d1, J in cases_:
        buf = rng.standard_normal((Ho, Wo)).astype(dt)
        for w in cases.WAVELETS:
            say(f"sparse {Wo}x{Ho} inner {Wi}x{Hi} {w} {prec} zp={zp} d1={d1} J={J}")
            ref, jr = o.fwd2(buf, w, J, d1=d1, zp=zp, inner=(Wi, Hi))
            t = torch.from_numpy(buf).cuda()
            j = libdwt_amd.transform_2d_async(w, True, t.data_ptr(), Wo * es, es, Wo, Ho, Wi, Hi, j=J,
                                              decompose_one=d1, zero_padding=zp, f64=prec == "d")
            assert j == jr
            np.testing.assert_array_equal(t.cpu().numpy(), ref, err_msg="forward")
            libdwt_amd.transform_2d_async(w, False, t.data_ptr(), Wo * es, es, Wo, Ho, Wi, Hi, j=j,
                                          decompose_one=d1, zero_padding=zp, f64=prec == "d")
            want = o.inv2(ref, w, j, d1=d1, zp=zp, inner=(Wi, Hi))
            np.testing.assert_array_equal(t.cpu().numpy(), want, err_msg="inverse")


RESTAGE_CASES = [  # (Wo, Ho, Wi, Hi, layout)
    (300, 257, 300, 257, "chan"), (1000, 700, 1000, 700, "chan"), (513, 129, 513, 129, "trans"),
    (512, 512, 300, 257, "chan"), (64, 256, 33, 200, "trans"), (2, 1000, 2, 1000, "chan")]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_restaged_layouts_match_oracle(prec):
    """Interleaved 3-channel (each channel through the drivers' strides, as the
    reference's OpenCV wrapper calls them) and transposed frames: copied into
    packed scratch, transformed there, copied back; bit-exact to the oracle on
    the same frame, with and without zero padding and decompose_one, and the
    other channels untouched."""
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    rng = np.random.default_rng(23)
    for (Wo, Ho, Wi, Hi, kind) in RESTAGE_CASES:
        for zp, d1 in ((0, 0), (1, 1)):
            vals = rng.standard_normal((3, Ho, Wo)).astype(dt)
            if kind == "chan":  # H x W x 3, channel 1 transformed
                host = np.ascontiguousarray(np.moveaxis(vals, 0, -1))
                sx, sy, off = 3 * Wo * es, 3 * es, es
                plane = lambda a: a[..., 1]
            else:  # column-major plane: element (y, x) at x * Ho + y
                host = np.ascontiguousarray(vals[0].T)
                sx, sy, off = es, Ho * es, 0
                plane = lambda a: a.T
            for w in cases.WAVELETS:
                say(f"restage {kind} {Wo}x{Ho} inner {Wi}x{Hi} {w} {prec} zp={zp} d1={d1}")
                ref, jr = o.fwd2(np.ascontiguousarray(plane(host)), w, -1, d1=d1, zp=zp, inner=(Wi, Hi))
                t = torch.from_numpy(host.copy()).cuda()
                j = libdwt_amd.transform_2d_async(w, True, t.data_ptr() + off, sx, sy, Wo, Ho, Wi, Hi,
                                                  j=-1, decompose_one=d1, zero_padding=zp,
                                                  f64=prec == "d")
                assert j == jr
                got = t.cpu().numpy()
                np.testing.assert_array_equal(plane(got), ref, err_msg="forward")
                if kind == "chan":  # channels 0 and 2 untouched
                    np.testing.assert_array_equal(got[..., 0], host[..., 0])
                    np.testing.assert_array_equal(got[..., 2], host[..., 2])
                libdwt_amd.transform_2d_async(w, False, t.data_ptr() + off, sx, sy, Wo, Ho, Wi, Hi,
                                              j=j, decompose_one=d1, zero_padding=zp, f64=prec == "d")
                want = o.inv2(ref, w, j, d1=d1, zp=zp, inner=(Wi, Hi))
                np.testing.assert_array_equal(plane(t.cpu().numpy()), want, err_msg="inverse")


PYR_SHAPES = [(2, 2), (3, 5), (17, 33), (130, 70), (300, 257), (1000, 40), (9, 250), (513, 511),
              (2048, 3)]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_pyramid_schedule_shapes(prec):
    """Level 0 plus one pyramid (dataflow) launch for every later level, with
    the single-workgroup block kernel switched off so that small images take
    this path too: bit-exact to the oracle and equal to the per-level
    schedule, forward and inverse, both wavelets."""
    from oracle.oracle import oracle
    o = oracle()
    lib = libdwt_amd.lib()
    dt = np.float32 if prec == "s" else np.float64
    rng = np.random.default_rng(11)
    st = torch.cuda.current_stream().cuda_stream
    try:
        assert lib.dwt_gpu_tail_set(0) == 0
        for (W, H) in PYR_SHAPES:
            for w in cases.WAVELETS:
                img = rng.standard_normal((H, W)).astype(dt)
                ref, J = o.fwd2(img, w, -1)
                want = o.inv2(ref, w, J)
                for sched in (1, 0):
                    say(f"pyramid {W}x{H} {w} {prec} schedule {sched}")
                    libdwt_amd.schedule_set(sched)
                    t = torch.from_numpy(img).cuda()
                    j = libdwt_amd.fwd2(t, w, -1)
                    libdwt_amd.device_check(st)
                    assert j == J
                    np.testing.assert_array_equal(t.cpu().numpy(), ref, err_msg=f"forward {sched}")
                    libdwt_amd.inv2(t, w, j)
                    libdwt_amd.device_check(st)
                    np.testing.assert_array_equal(t.cpu().numpy(), want, err_msg=f"inverse {sched}")
    finally:
        lib.dwt_gpu_tail_set(TAIL_DEFAULT)
        libdwt_amd.schedule_set(0)


def test_pyramid_repeated_launches_reuse_counters():
    """Forty back-to-back forward+inverse pairs on one stream reuse the
    pyramid's counters (each launch leaves them zero for the next); the
    result equals the oracle doing the same forty round trips."""
    from oracle.oracle import oracle
    say("pyramid repeated launches")
    o = oracle()
    img = workloads.sinusoid_image(640, 480, 9)
    want = img
    for _ in range(40):
        c, j = o.fwd2(want, "cdf53", 6)
        want = o.inv2(c, "cdf53", j)
    s = torch.cuda.Stream()
    t = torch.from_numpy(img).cuda()
    torch.cuda.synchronize()
    libdwt_amd.schedule_set(1)
    try:
        for _ in range(40):
            j = libdwt_amd.transform_2d_async("cdf53", True, t.data_ptr(), 4 * 640, 4, 640, 480,
                                              j=6, stream=s.cuda_stream)
            libdwt_amd.transform_2d_async("cdf53", False, t.data_ptr(), 4 * 640, 4, 640, 480,
                                          j=j, stream=s.cuda_stream)
        libdwt_amd.device_check(s.cuda_stream)
    finally:
        libdwt_amd.schedule_set(0)
    np.testing.assert_array_equal(t.cpu().numpy(), want)


def test_sparse_fused_odd_pitch_and_generic_agree():
    """Sparse frames at an odd row pitch and an odd element offset from the
    allocation: the fused sparse level launches (default) and the generic
    per-line passes give the oracle's result bit for bit, forward and
    inverse, with and without zero padding."""
    from oracle.oracle import oracle
    o = oracle()
    L = libdwt_amd.lib()
    rng = np.random.default_rng(29)
    for (Wo, Ho, Wi, Hi, pitch, off) in ((700, 300, 513, 257, 703, 1), (1030, 517, 1030, 260, 1031, 3)):
        for zp in (0, 1):
            base = rng.standard_normal(pitch * Ho + off + 8).astype(np.float32)
            frame = base[off:off + pitch * Ho].reshape(Ho, pitch)
            want = frame.copy()
            jj = o.fwd2_raw(want, "cdf97", pitch * 4, 4, Wo, Ho, Wi, Hi, -1, 0, zp)
            back = want.copy()
            o.inv2_raw(back, "cdf97", pitch * 4, 4, Wo, Ho, Wi, Hi, jj, 0, zp)
            for fused in (1, 0):
                assert L.dwt_gpu_sparse_set(fused) == 0
                say(f"sparse pitch {pitch} off {off} {Wo}x{Ho} in {Wi}x{Hi} zp={zp} fused={fused}")
                t = torch.from_numpy(base.copy()).cuda()
                ptr = t.data_ptr() + 4 * off
                j = libdwt_amd.transform_2d_async("cdf97", True, ptr, pitch * 4, 4, Wo, Ho, Wi, Hi, j=-1,
                                                  zero_padding=zp)
                assert j == jj
                got = t.cpu().numpy()[off:off + pitch * Ho].reshape(Ho, pitch)
                np.testing.assert_array_equal(got, want, err_msg=f"forward fused={fused}")
                libdwt_amd.transform_2d_async("cdf97", False, ptr, pitch * 4, 4, Wo, Ho, Wi, Hi, j=j,
                                              zero_padding=zp)
                got = t.cpu().numpy()[off:off + pitch * Ho].reshape(Ho, pitch)
                np.testing.assert_array_equal(got, back, err_msg=f"inverse fused={fused}")
    assert L.dwt_gpu_sparse_set(1) == 0


def test_sparse_random_frames_match_oracle():
    """60 random sparse frames (outer sides 2..700, inner sides 1..outer, any
    levels, zero padding and decompose_one, both wavelets, FP32 and FP64,
    random garbage outside the inner image): forward then inverse through the
    drop-in driver's arguments, bit-exact to the oracle.  Covers the fused
    sparse levels, the sparse block tail and their hand-over to the generic
    passes where an inner side reaches 1."""
    from oracle.oracle import oracle
    o = oracle()
    rng = np.random.default_rng(123)
    for case in range(60):
        Wo, Ho = int(rng.integers(2, 701)), int(rng.integers(2, 701))
        Wi, Hi = int(rng.integers(1, Wo + 1)), int(rng.integers(1, Ho + 1))
        zp, d1 = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        J = int(rng.integers(-1, 10))
        prec = "d" if case % 5 == 4 else "s"
        w = cases.WAVELETS[case % 2]
        dt = np.float32 if prec == "s" else np.float64
        es = np.dtype(dt).itemsize
        buf = rng.standard_normal((Ho, Wo)).astype(dt)
        say(f"random sparse {case}: {Wo}x{Ho} inner {Wi}x{Hi} {w} {prec} zp={zp} d1={d1} J={J}")
        ref, jr = o.fwd2(buf, w, J, d1=d1, zp=zp, inner=(Wi, Hi))
        t = torch.from_numpy(buf).cuda()
        j = libdwt_amd.transform_2d_async(w, True, t.data_ptr(), Wo * es, es, Wo, Ho, Wi, Hi, j=J,
                                          decompose_one=d1, zero_padding=zp, f64=prec == "d")
        assert j == jr
        np.testing.assert_array_equal(t.cpu().numpy(), ref, err_msg=f"forward case {case}")
        libdwt_amd.transform_2d_async(w, False, t.data_ptr(), Wo * es, es, Wo, Ho, Wi, Hi, j=j,
                                      decompose_one=d1, zero_padding=zp, f64=prec == "d")
        want = o.inv2(ref, w, j, d1=d1, zp=zp, inner=(Wi, Hi))
        np.testing.assert_array_equal(t.cpu().numpy(), want, err_msg=f"inverse case {case}")
