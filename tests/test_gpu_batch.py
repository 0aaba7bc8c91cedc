"""GPU: batched transforms of images of uneven sizes (one launch per level for
the whole batch), through the C ABI.

Bars:
  * FP32 storage: every image's coefficients and reconstruction bit-identical
    to the oracle's single-image transform (the reference's _s arithmetic);
  * FP16 / BF16 storage: every stored coefficient equals the FP32 one scaled
    by its subband's power of two and rounded to nearest (so within half an
    ulp of the 16-bit format, relative 2^-11 / 2^-8, or the FP16 subnormal
    spacing), the FP16 scales are powers of two with |c * s| <= 2^14, the BF16
    scales are 1, and the reconstruction from 16-bit coefficients stays within
    the tolerance below.
"""
import sys

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (65, 33), (2, 2), (130, 71), (300, 257), (1031, 64), (64, 1031), (513, 511),
         (2, 9), (127, 3)]
# reconstruction error bound relative to max |x| for 16-bit storage (measured
# 2e-4 for FP16 and 2e-3 for BF16 on these images, J up to 9)
RECON_TOL = {"f16": 2e-3, "bf16": 2e-2}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0
    return torch.device("cuda:0")


def images(seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for k, (W, H) in enumerate(SIZES):
        if k % 2:
            out.append(workloads.sinusoid_image(W, H, seed + k))
        else:
            out.append(rng.standard_normal((H, W)).astype(np.float32))
    return out


@pytest.mark.parametrize("wavelet", ["cdf97", "cdf53"])
def test_batch_f32_matches_single_image_oracle(wavelet):
    from oracle.oracle import oracle
    o = oracle()
    imgs = images()
    say(f"batch f32 {wavelet}: {len(imgs)} images")
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    levels = [-1, 3, -1, 5, -1, 2, -1, 9, 1, 1]
    coefs, _, done = libdwt_amd.batch_forward(ts, wavelet, "f32", levels)
    torch.cuda.synchronize()
    for x, c, lv, j in zip(imgs, coefs, levels, done):
        ref, jr = o.fwd2(x, wavelet, lv)
        assert j == jr
        np.testing.assert_array_equal(c.cpu().numpy(), ref, err_msg=f"forward {x.shape} J={j}")
    refs = [torch.from_numpy(o.fwd2(x, wavelet, lv)[0]).cuda() for x, lv in zip(imgs, levels)]
    outs = libdwt_amd.batch_inverse(refs, done, wavelet, "f32")
    torch.cuda.synchronize()
    for x, r, out, j in zip(imgs, refs, outs, done):
        want = o.inv2(r.cpu().numpy(), wavelet, j)
        np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=f"inverse {x.shape} J={j}")


def check_16bit(imgs, wavelet, storage):
    """Forward in FP32 and in 16-bit storage: every scale a power of two,
    every stored coefficient within 2^14 after scaling and the FP32 one
    scaled and rounded to nearest; the inverse within RECON_TOL of max |x|."""
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    c32, _, done = libdwt_amd.batch_forward(ts, wavelet, "f32", -1)
    c16, scales, done16 = libdwt_amd.batch_forward(ts, wavelet, storage, -1)
    torch.cuda.synchronize()
    assert done16 == done
    mant = 11 if storage == "f16" else 8
    for x, a, b, sc, j in zip(imgs, c32, c16, scales, done):
        H, W = x.shape
        s = sc.cpu().numpy().astype(np.float64)
        assert s.shape == (3 * j + 1,)
        assert np.all(np.isfinite(s)) and np.all(s > 0)
        assert np.all(np.frexp(s)[0] == 0.5), "scales are powers of two"
        if storage == "bf16":
            assert np.all(s == 1.0), "BF16 keeps FP32's exponent range: scales 1"
        smap = np.full((H, W), np.nan)
        for lev in range(j):
            w, h = -(-W // 2 ** lev), -(-H // 2 ** lev)
            wd, hd = -(-w // 2), -(-h // 2)
            smap[:hd, wd:w] = s[3 * lev + 0]
            smap[hd:h, :wd] = s[3 * lev + 1]
            smap[hd:h, wd:w] = s[3 * lev + 2]
        wl, hl = -(-W // 2 ** j), -(-H // 2 ** j)
        smap[:hl, :wl] = s[3 * j]
        assert not np.isnan(smap).any()
        c = a.cpu().numpy().astype(np.float64)
        q = b.float().cpu().numpy().astype(np.float64)
        if storage == "f16":
            assert np.abs(c * smap).max() <= 2.0 ** 14 * (1 + 2.0 ** -mant), x.shape
        # round to nearest: error of the scaled value <= half an ulp (or half
        # the FP16 subnormal spacing)
        err = np.abs(q - c * smap)
        ulp_half = np.maximum(np.abs(c * smap) * 2.0 ** -mant, 2.0 ** -25 if storage == "f16" else 0)
        assert np.all(err <= ulp_half * 1.0000001 + 1e-45), (x.shape, float((err / ulp_half).max()))
    outs = libdwt_amd.batch_inverse(c16, done, wavelet, storage, scales)
    torch.cuda.synchronize()
    for x, out in zip(imgs, outs):
        e = float(np.abs(out.cpu().numpy() - x).max()) / float(np.abs(x).max())
        say(f"  {x.shape} relative reconstruction error {e:.2e}")
        assert e <= RECON_TOL[storage]


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("wavelet", ["cdf97", "cdf53"])
def test_batch_16bit_storage(wavelet, storage):
    say(f"batch {storage} {wavelet}")
    check_16bit(images(9), wavelet, storage)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_batch_16bit_rows_longer_than_one_pass(storage):
    """Rows longer than one max |x| pass (4096 pixels), with the largest value
    past the first pass: the scales must still see it.  Widths with and
    without whole 16-byte groups."""
    rng = np.random.default_rng(13)
    imgs = []
    for W, H in ((5000, 6), (4100, 3), (4097, 2), (8200, 2), (4096, 2)):
        x = rng.standard_normal((H, W)).astype(np.float32)
        x[H // 2, W - 3] = 3.0e4
        imgs.append(x)
    say(f"batch {storage} long rows")
    check_16bit(imgs, "cdf97", storage)


def test_batch_plan_replays_and_graph_capture():
    """A plan executed repeatedly, and captured in a HIP graph and replayed
    three times (eager work and syncs in between), gives the eager result
    every time."""
    import ctypes
    imgs = images(3)[:6]
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    outs = [torch.empty_like(t) for t in ts]
    arr = libdwt_amd.batch_descriptors(ts, outs, -1)
    L = libdwt_amd.lib()
    s = torch.cuda.Stream()
    plan = ctypes.c_void_p()
    assert L.dwt_gpu_batch_plan_create(0, 0, 0, arr, len(ts), s.cuda_stream, ctypes.byref(plan)) == 0
    try:
        assert L.dwt_gpu_batch_plan_execute(plan, s.cuda_stream) == 0
        s.synchronize()
        eager = [o.clone() for o in outs]
        for o in outs:
            o.zero_()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            assert L.dwt_gpu_batch_plan_execute(plan, s.cuda_stream) == 0
        for rep in range(3):
            for o in outs:
                o.fill_(float(rep + 1))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(outs, eager):
                assert torch.equal(a, b), f"replay {rep}"
    finally:
        L.dwt_gpu_batch_plan_destroy(plan)


def test_batch_rejects_bad_descriptors():
    t = torch.zeros((8, 8), device="cuda")
    L = libdwt_amd.lib()
    arr = libdwt_amd.batch_descriptors([t], [t], -1)  # in place
    assert L.dwt_gpu_batch_2d(0, 0, 0, arr, 1, None) == -1
    u = torch.zeros((8, 8), device="cuda")
    arr = libdwt_amd.batch_descriptors([t], [u], -1)
    assert L.dwt_gpu_batch_2d(0, 0, 1, arr, 1, None) == -1  # 16-bit without scales
    one = torch.zeros((1, 8), device="cuda")
    arr = libdwt_amd.batch_descriptors([one], [u[:1]], -1)
    assert L.dwt_gpu_batch_2d(0, 0, 0, arr, 1, None) == -1  # height 1


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_batch_tail_threshold_changes_nothing(storage):
    """The batched tail (every image's small levels in one launch, in LDS) off,
    at its default and at its largest region give bit-identical coefficients,
    scales and reconstructions: its arithmetic and 16-bit conversion are the
    per-level launches'."""
    L = libdwt_amd.lib()
    imgs = images(21) + [np.random.default_rng(4).standard_normal((181, 181)).astype(np.float32)]
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    results = []
    try:
        for tail in (0, 65536, 131072):
            assert L.dwt_gpu_batch_tail_set(tail) == 0
            say(f"batch {storage} tail {tail}")
            c, s, done = libdwt_amd.batch_forward(ts, "cdf97", storage, -1)
            r = libdwt_amd.batch_inverse(c, done, "cdf97", storage, s)
            torch.cuda.synchronize()
            results.append(([x.cpu() for x in c], None if s is None else [x.cpu() for x in s],
                            [x.cpu() for x in r]))
    finally:
        assert L.dwt_gpu_batch_tail_set(65536) == 0
    base = results[0]
    for other in results[1:]:
        for a, b in zip(base[0], other[0]):
            assert torch.equal(a, b)
        if base[1] is not None:
            for a, b in zip(base[1], other[1]):
                assert torch.equal(a, b)
        for a, b in zip(base[2], other[2]):
            assert torch.equal(a, b)
    assert L.dwt_gpu_batch_tail_set(131073) == -1


def test_batch_autotune_installs_a_valid_table():
    """The batch autotuner times rows per level and tail thresholds on a small
    batch, installs its choice, and the batch still matches the oracle bit
    for bit under it."""
    from libdwt_amd import autotune
    from oracle.oracle import oracle
    imgs = images(17)
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    say("batch autotune: 10 images")
    try:
        table = autotune.tune_batch(ts, "cdf97", "f32", reps=1)
        assert table["tail_bytes"] in autotune.BATCH_TAILS
        assert all(r in autotune.BATCH_ROWS for r in table["rows"].values())
        coefs, _, done = libdwt_amd.batch_forward(ts, "cdf97", "f32", -1)
        torch.cuda.synchronize()
        o = oracle()
        for x, c, j in zip(imgs, coefs, done):
            ref, _ = o.fwd2(x, "cdf97", j)
            np.testing.assert_array_equal(c.cpu().numpy(), ref, err_msg=f"{x.shape} J={j}")
    finally:
        autotune.reset_batch()


LARGE_SIZES = [(129, 300), (4096, 4096), (64, 64), (2500, 2100), (1023, 1021), (3001, 2047), (2, 9)]


@pytest.mark.parametrize("wavelet", ["cdf97", "cdf53"])
def test_batch_large_images_match_single_image(wavelet):
    """Images large enough that their level-0 and level-1 launches take the
    per-image XCD bands (at least 1024 workgroups: 4096^2, 2500x2100,
    3001x2047), between small ones: FP32 coefficients and reconstructions
    bit-identical to the single-image transform of each image (itself
    bit-identical to the reference, tests/test_gpu_parity.py)."""
    rng = np.random.default_rng(31)
    imgs = [rng.standard_normal((H, W)).astype(np.float32) for W, H in LARGE_SIZES]
    say(f"batch large {wavelet}: {LARGE_SIZES}")
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    coefs, _, done = libdwt_amd.batch_forward(ts, wavelet, "f32", -1)
    outs = libdwt_amd.batch_inverse(coefs, done, wavelet, "f32")
    torch.cuda.synchronize()
    for x, c, out, j in zip(ts, coefs, outs, done):
        H, W = x.shape
        t = x.clone()
        jj = libdwt_amd.transform_2d_async(wavelet, True, t.data_ptr(), 4 * W, 4, W, H, j=j)
        torch.cuda.synchronize()
        assert jj == j
        assert torch.equal(c, t), f"forward {W}x{H} J={j}"
        libdwt_amd.transform_2d_async(wavelet, False, t.data_ptr(), 4 * W, 4, W, H, j=j)
        torch.cuda.synchronize()
        assert torch.equal(out, t), f"inverse {W}x{H} J={j}"


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_batch_large_images_16bit(storage):
    """16-bit storage on the same large-image batch: the rounding, scale and
    reconstruction bars of check_16bit."""
    rng = np.random.default_rng(37)
    imgs = [workloads.sinusoid_image(W, H, 40 + k) if k % 2 else rng.standard_normal((H, W)).astype(np.float32)
            for k, (W, H) in enumerate(LARGE_SIZES)]
    say(f"batch large {storage}")
    check_16bit(imgs, "cdf97", storage)
