"""GPU: the batched and serving workloads at their benchmark scale.

  * 256 images, sides log-uniform in [64, 2048] (workloads.uneven_sizes seed
    4), J = min(8, floor(log2(min side))), one batched launch per level:
    FP32 storage bit-exact to the oracle for every image, forward and
    inverse; FP16 / BF16 storage within half an ulp of the 16-bit format at
    each subband's scale (plus the FP32 transform's own error) of the FP64
    transform, and the reconstruction's PSNR above the bound stated below.
  * The serving stream: 64 resident images with sides in [128, 4096] (seed
    7), requests 50/50 CDF 9/7 / 5/3 and forward / inverse, J uniform in
    [1, min(8, j_limit)], submitted at Poisson times (30000 requests/s) by
    the native replay loop: every result equal to the same request run
    alone, and to the oracle.
"""
import math
import sys

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import serving, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

# reconstruction PSNR (peak-to-peak of the image) from 16-bit coefficients:
# the worst of the 256 images measured 88.5 dB (FP16) and 70.5 dB (BF16) on
# MI355X; the bounds leave 6 dB
PSNR_MIN = {"f16": 82.0, "bf16": 64.0}
MANT = {"f16": 11, "bf16": 8}  # significand bits: half an ulp is 2^-MANT relative


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0


@pytest.fixture(scope="module")
def batch():
    sizes = workloads.uneven_sizes(256, 64, 2048, seed=4)
    levels = [min(8, int(math.floor(math.log2(min(w, h))))) for w, h in sizes]
    imgs = [workloads.sinusoid_image(w, h, seed=1000 + i) for i, (w, h) in enumerate(sizes)]
    return sizes, levels, imgs


def subbands(W, H, J):
    """(row slice, column slice) of every subband of a J-level Mallat layout."""
    out = []
    for lev in range(J):
        w, h = -(-W // 2 ** lev), -(-H // 2 ** lev)
        wd, hd = -(-w // 2), -(-h // 2)
        out += [(slice(0, hd), slice(wd, w)), (slice(hd, h), slice(0, wd)), (slice(hd, h), slice(wd, w))]
    out.append((slice(0, -(-H // 2 ** J)), slice(0, -(-W // 2 ** J))))
    return out


def test_batch_256_uneven_fp32_bit_exact(batch):
    from oracle.oracle import oracle
    o = oracle()
    sizes, levels, imgs = batch
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    say(f"batch of {len(ts)} images, {sum(w * h for w, h in sizes) / 1e6:.1f} Mpixel")
    coefs, _, done = libdwt_amd.batch_forward(ts, "cdf97", "f32", levels)
    outs = libdwt_amd.batch_inverse(coefs, done, "cdf97", "f32")
    torch.cuda.synchronize()
    assert done == levels
    for k, (x, c, r, j) in enumerate(zip(imgs, coefs, outs, done)):
        if k % 64 == 0:
            say(f"  image {k}: {x.shape[1]}x{x.shape[0]} J={j}")
        ref, jr = o.fwd2(x, "cdf97", j)
        assert jr == j
        np.testing.assert_array_equal(c.cpu().numpy(), ref, err_msg=f"forward image {k}")
        np.testing.assert_array_equal(r.cpu().numpy(), o.inv2(ref, "cdf97", j), err_msg=f"inverse image {k}")


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_batch_256_uneven_16bit_error(batch, storage):
    from oracle.oracle import oracle
    o = oracle()
    sizes, levels, imgs = batch
    ts = [torch.from_numpy(x).cuda() for x in imgs]
    c32, _, done = libdwt_amd.batch_forward(ts, "cdf97", "f32", levels)
    c16, scales, done16 = libdwt_amd.batch_forward(ts, "cdf97", storage, levels)
    outs = libdwt_amd.batch_inverse(c16, done16, "cdf97", storage, scales)
    torch.cuda.synchronize()
    assert done16 == done
    worst_coef, worst_psnr = 0.0, float("inf")
    for k, (x, a, b, sc, out, j) in enumerate(zip(imgs, c32, c16, scales, outs, done)):
        H, W = x.shape
        # FP64 coefficients: the library's FP64 transform, itself pinned to the
        # oracle's FP64 arithmetic on the first images
        d = torch.from_numpy(x.astype(np.float64)).cuda()
        libdwt_amd.fwd2(d, "cdf97", j)
        torch.cuda.synchronize()
        c64 = d.cpu().numpy()
        if k < 4:
            np.testing.assert_array_equal(c64, o.fwd2(x.astype(np.float64), "cdf97", j)[0])
        s = sc.cpu().numpy().astype(np.float64)
        q = b.float().cpu().numpy().astype(np.float64)
        f = a.cpu().numpy().astype(np.float64)
        for i, (rs, cs) in enumerate(subbands(W, H, j)):
            e16 = np.abs(q[rs, cs] / s[i] - c64[rs, cs]).max()
            e32 = np.abs(f[rs, cs] - c64[rs, cs]).max()
            top = np.abs(f[rs, cs]).max()
            # half an ulp of the 16-bit value at this subband's scale, the FP16
            # subnormal spacing, and the FP32 transform's own error
            bound = top * 2.0 ** -MANT[storage] + (2.0 ** -25 / s[i] if storage == "f16" else 0) + e32
            assert e16 <= bound * (1 + 1e-6) + 1e-30, (k, i, e16, bound)
            worst_coef = max(worst_coef, e16 / max(top, 1e-30))
        e = out.cpu().numpy().astype(np.float64) - x
        peak = float(x.max() - x.min())
        mse = float((e * e).mean())
        psnr = 10 * math.log10(peak * peak / mse) if mse > 0 else float("inf")
        worst_psnr = min(worst_psnr, psnr)
        assert psnr >= PSNR_MIN[storage], (k, x.shape, psnr)
    say(f"{storage}: worst subband-relative coefficient error {worst_coef:.2e}, worst PSNR {worst_psnr:.1f} dB")


def test_serving_poisson_stream_matches_alone_and_oracle():
    from oracle.oracle import oracle
    o = oracle()
    rng = np.random.default_rng(7)
    sizes = workloads.uneven_sizes(64, 128, 4096, seed=7)
    pool_np = [workloads.sinusoid_image(w, h, seed=2000 + i) for i, (w, h) in enumerate(sizes)]
    pool = [torch.from_numpy(x).cuda() for x in pool_np]
    n = 160
    reqs = []
    for _ in range(n):
        k = int(rng.integers(len(pool)))
        w, h = sizes[k]
        jmax = min(8, workloads.ceil_log2(min(w, h)))
        reqs.append((k, "cdf97" if rng.random() < 0.5 else "cdf53", bool(rng.random() < 0.5),
                     int(rng.integers(1, jmax + 1))))
    at = np.cumsum(rng.exponential(1.0 / 30000.0, n)).tolist()
    torch.cuda.synchronize()
    say(f"serving: {n} requests, {sum(pool[k].numel() for k, *_ in reqs) / 1e6:.0f} Mpixel, Poisson 30000/s")
    mixed = [torch.empty_like(pool[k]) for k, *_ in reqs]
    with serving.Server(16 << 30, max_batch=256, streams=4) as s:
        rl = [s.request(wv, fwd, pool[k], out, j) for (k, wv, fwd, j), out in zip(reqs, mixed)]
        ids = s.replay(rl, at)
        s.drain()
        got = {}
        while True:
            r = s.poll()
            if not r:
                break
            got.update({i: st for i, _, st in r})
        batches, count = s.stats()
    assert sorted(got) == sorted(ids) and all(v == 0 for v in got.values())
    assert batches < count, "requests were batched"
    say(f"  {count} requests in {batches} batches")
    alone = [torch.empty_like(pool[k]) for k, *_ in reqs]
    with serving.Server(16 << 30, max_batch=1, streams=1) as s:
        for (k, wv, fwd, j), out in zip(reqs, alone):
            s.submit(wv, fwd, pool[k], out, j)
            s.drain()
        assert all(st == 0 for _, _, st in s.poll(1 << 12))
    torch.cuda.synchronize()
    cache = {}
    for i, ((k, wv, fwd, j), m, a) in enumerate(zip(reqs, mixed, alone)):
        mm = m.cpu().numpy()
        np.testing.assert_array_equal(mm, a.cpu().numpy(), err_msg=f"request {i}: mixed vs alone")
        key = (k, wv, fwd, j)
        if key not in cache:
            x = pool_np[k]
            cache[key] = o.fwd2(x, wv, j)[0] if fwd else o.inv2(x, wv, j)
        np.testing.assert_array_equal(mm, cache[key], err_msg=f"request {i}: vs oracle {key}")
