"""GPU: device-side compare (dwt_gpu_compare) and FP64 verification
(dwt_gpu_verify_2d).  Every statistic is checked against numpy on the same
arrays; the differ verdict against the host dwt_util_compare_s / _d (the
reference's rule, src/libdwt.c:1223-1277) on host copies."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0


def ulp_hist(a, b):
    """Histogram of the distance in ulps of a's dtype between a and b rounded
    to it (bin 0: equal, bin k: [2^(k-1), 2^k), last bin: the rest)."""
    bt = b.astype(a.dtype)
    if a.dtype == np.float32:
        ia, ib, lo = a.view(np.int32).astype(np.int64), bt.view(np.int32).astype(np.int64), -2 ** 31
    else:
        ia, ib, lo = a.view(np.int64), bt.view(np.int64), np.iinfo(np.int64).min
    oa = np.where(ia >= 0, ia, lo - ia)
    ob = np.where(ib >= 0, ib, lo - ib)
    d = [abs(int(x) - int(y)) for x, y in zip(oa.ravel(), ob.ravel())]
    h = [0] * libdwt_amd.ULP_BINS
    for v in d:
        h[min(v.bit_length(), libdwt_amd.ULP_BINS - 1)] += 1
    return h


def host_compare(a, b):
    """dwt_util_compare_s / _d on host copies (1 = differ)."""
    L = libdwt_amd.lib()
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    f = L.dwt_util_compare_s if a.dtype == np.float32 else L.dwt_util_compare_d
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    f.restype = C.c_int
    H, W = a.shape
    return f(a.ctypes.data, b.ctypes.data, a.strides[0], a.strides[1], W, H)


def check_stats(s, a, b, peak=None):
    fin = np.isfinite(a) & np.isfinite(b)
    d = np.abs(a[fin].astype(np.float64) - b[fin].astype(np.float64))
    assert s["count"] == a.size
    assert s["nonfinite"] == int((~fin).sum())
    assert s["max_abs"] == (float(d.max()) if d.size else 0.0)
    ss = float(np.sum(d * d))
    assert s["sum_sq"] == pytest.approx(ss, rel=1e-12, abs=1e-300)
    assert s["peak_b"] == float(np.abs(b[fin]).max())
    mse = ss / fin.sum()
    assert s["mse"] == pytest.approx(mse, rel=1e-12, abs=1e-300)
    pk = peak if peak else float(np.abs(b[fin]).max())
    if mse > 0:
        assert s["psnr"] == pytest.approx(10 * math.log10(pk * pk / mse), rel=1e-9)
    else:
        assert math.isinf(s["psnr"])
    assert s["ulp"] == ulp_hist(a[fin], b[fin])


def test_compare_f32_statistics_and_verdict():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((301, 517)).astype(np.float32)
    b = a.copy()
    idx = rng.choice(a.size, 400, replace=False)
    b.flat[idx[:200]] = np.nextafter(a.flat[idx[:200]], np.float32(np.inf))  # 1 ulp
    b.flat[idx[200:300]] += np.float32(0.01)                                # over eps
    b.flat[idx[300:350]] *= np.float32(1 + 2 ** -12)                         # a few ulps
    for case in ("finite", "nan", "inf"):
        bb = b.copy()
        if case == "nan":
            bb.flat[idx[399]] = np.nan
        if case == "inf":
            bb.flat[idx[398]] = -np.inf
        say(f"compare f32 {case}")
        s = libdwt_amd.compare(torch.from_numpy(a).cuda(), torch.from_numpy(bb).cuda())
        check_stats(s, a, bb)
        assert s["over_eps"] == int(np.sum(np.abs(a - bb)[np.isfinite(bb)] > np.float32(1e-3)))
        assert s["differ"] == host_compare(a, bb) == 1
    same = libdwt_amd.compare(torch.from_numpy(a).cuda(), torch.from_numpy(a).cuda())
    assert same["differ"] == host_compare(a, a) == 0
    assert same["max_abs"] == 0 and same["ulp"][0] == a.size and math.isinf(same["psnr"])


def test_compare_mixed_precision_strided_views():
    """FP32 view inside a larger tensor against an FP64 transposed view."""
    rng = np.random.default_rng(4)
    big = torch.from_numpy(rng.standard_normal((400, 700)).astype(np.float32)).cuda()
    a = big[13:313, 21:521]                      # row pitch 700 floats
    noise = torch.from_numpy(rng.standard_normal((300, 500))).cuda() * 1e-4
    bt = (a.double() + noise).t().contiguous().t()  # FP64, column-major strides
    s = libdwt_amd.compare(a, bt, peak=4.0)
    check_stats(s, a.cpu().numpy(), bt.cpu().numpy(), peak=4.0)
    assert s["differ"] == 0 and s["over_eps"] == 0
    s = libdwt_amd.compare(bt, a, eps=1e-5)
    assert s["over_eps"] == int(np.sum(np.abs(bt.cpu().numpy() - a.cpu().numpy().astype(np.float64)) > 1e-5))
    assert s["differ"] == (1 if s["over_eps"] else 0)


def test_compare_f64_matches_util_compare_d():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((64, 90))
    b = a + rng.choice([0.0, 5e-7, 2e-6], size=a.shape)
    s = libdwt_amd.compare(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    check_stats(s, a, b)
    assert s["over_eps"] == int(np.sum(np.abs(a - b) > 1e-6))
    assert s["differ"] == host_compare(a, b) == 1


@pytest.mark.parametrize("wavelet", ["cdf97", "cdf53"])
def test_verify_2d_against_oracle_fp64(wavelet):
    """The FP64 reference computed on the device is the oracle's _d transform
    bit for bit, so the statistics equal numpy's |FP32 result - oracle _d|."""
    from oracle.oracle import oracle
    o = oracle()
    img = workloads.sinusoid_image(512, 512, 1)
    t = torch.from_numpy(img).cuda()
    res = t.clone()
    j = libdwt_amd.fwd2(res, wavelet, 5)
    s = libdwt_amd.verify_2d(t, res, wavelet, True, j)
    ref, _ = o.fwd2(img.astype(np.float64), wavelet, 5)
    got = res.cpu().numpy()
    assert s["max_abs"] == float(np.abs(got.astype(np.float64) - ref).max())
    check_stats(s, got, ref)
    assert s["differ"] == 0
    say(f"verify fwd {wavelet}: max_abs {s['max_abs']:.3e} psnr {s['psnr']:.1f} dB")
    back = res.clone()
    libdwt_amd.inv2(back, wavelet, j)
    s = libdwt_amd.verify_2d(res, back, wavelet, False, j)
    ref_inv = o.inv2(got.astype(np.float64), wavelet, j)
    check_stats(s, back.cpu().numpy(), ref_inv)
