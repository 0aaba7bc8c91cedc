"""GPU: dwt_gpu_2d_channels, every channel of a multi-channel image in one
call.  Each channel must come out exactly as the oracle's transform of that
channel alone (the result the reference's OpenCV wrapper gets from one driver
call per channel, src/cvdwt.cpp:67-147), the channels not named must stay
untouched, and the result must equal the per-channel dwt_gpu_2d_async calls."""
import sys

import numpy as np
import pytest

import libdwt_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0


# (W, H, channels stored, channels transformed, layout, J, decompose_one)
CASES = [(640, 480, 3, 3, "interleaved", -1, 0), (300, 257, 4, 3, "interleaved", 4, 0),
         (513, 129, 2, 2, "planar", -1, 1), (64, 1000, 3, 3, "interleaved", -1, 1),
         (2, 3, 3, 3, "interleaved", -1, 0)]


def layout_args(kind, W, H, Cs, es):
    """(stride_x, stride_y, channel_stride) in bytes and a view from the
    stored array to channel c."""
    if kind == "interleaved":  # H x W x Cs
        return Cs * W * es, Cs * es, es, lambda a, c: a[..., c]
    return W * es, es, W * H * es, lambda a, c: a[c]  # Cs x H x W


@pytest.mark.parametrize("prec", ["s", "d"])
def test_channels_match_oracle_and_per_channel_calls(prec):
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    rng = np.random.default_rng(41)
    for W, H, Cs, Ct, kind, J, d1 in CASES:
        shape = (H, W, Cs) if kind == "interleaved" else (Cs, H, W)
        host = rng.standard_normal(shape).astype(dt)
        sx, sy, sc, chan = layout_args(kind, W, H, Cs, es)
        for w in ("cdf97", "cdf53"):
            say(f"channels {kind} {W}x{H} {Ct}/{Cs} {w} {prec} J={J} d1={d1}")
            t = torch.from_numpy(host.copy()).cuda()
            j = libdwt_amd.transform_channels(w, True, t.data_ptr(), sx, sy, sc, Ct, W, H, j=J,
                                              decompose_one=d1, f64=prec == "d")
            torch.cuda.synchronize()
            got = t.cpu().numpy()
            # the same frames through one dwt_gpu_2d_async call per channel
            u = torch.from_numpy(host.copy()).cuda()
            for c in range(Ct):
                jc = libdwt_amd.transform_2d_async(w, True, u.data_ptr() + c * sc, sx, sy, W, H, j=J,
                                                   decompose_one=d1, f64=prec == "d")
                assert jc == j
            torch.cuda.synchronize()
            np.testing.assert_array_equal(got, u.cpu().numpy(), err_msg="vs per-channel calls")
            refs = []
            for c in range(Ct):
                ref, jr = o.fwd2(np.ascontiguousarray(chan(host, c)), w, J, d1=d1)
                assert j == jr
                np.testing.assert_array_equal(chan(got, c), ref, err_msg=f"forward channel {c}")
                refs.append(ref)
            for c in range(Ct, Cs):  # not transformed
                np.testing.assert_array_equal(chan(got, c), chan(host, c))
            libdwt_amd.transform_channels(w, False, t.data_ptr(), sx, sy, sc, Ct, W, H, j=j,
                                          decompose_one=d1, f64=prec == "d")
            torch.cuda.synchronize()
            back = t.cpu().numpy()
            for c in range(Ct):
                want = o.inv2(refs[c], w, j, d1=d1)
                np.testing.assert_array_equal(chan(back, c), want, err_msg=f"inverse channel {c}")
            for c in range(Ct, Cs):
                np.testing.assert_array_equal(chan(back, c), chan(host, c))


def test_channels_reject_partial_elements():
    t = torch.zeros((8, 8, 3), device="cuda")
    with pytest.raises(RuntimeError):
        libdwt_amd.transform_channels("cdf97", True, t.data_ptr(), 8 * 3 * 4, 3 * 4, 2, 3, 8, 8)
    with pytest.raises(RuntimeError):
        libdwt_amd.transform_channels("cdf97", True, t.data_ptr(), 8 * 3 * 4, 3 * 4, 4, 0, 8, 8)
    torch.cuda.synchronize()
    assert not t.any()
