"""GPU: the serving loop (continuous batching).  Each request's result is the
same alone and in the mix (bit for bit, FP32 storage, and equal to the
oracle's single-image transform); admission by a small HBM budget still
completes every request with the same results."""
import sys

import numpy as np
import pytest

import libdwt_amd
from libdwt_amd import serving, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0


def make_requests(n=40, seed=21):
    rng = np.random.default_rng(seed)
    sizes = workloads.uneven_sizes(n, 128, 600, seed=seed)
    reqs = []
    for k, (W, H) in enumerate(sizes):
        wavelet = "cdf97" if rng.random() < 0.5 else "cdf53"
        fwd = bool(rng.random() < 0.5)
        jl = workloads.ceil_log2(min(W, H))
        J = int(rng.integers(1, min(8, jl) + 1))
        x = workloads.sinusoid_image(W, H, seed + k)
        reqs.append(dict(wavelet=wavelet, fwd=fwd, J=J, x=x))
    return reqs


def run(server, reqs, one_at_a_time):
    outs = []
    ids = []
    for r in reqs:
        src = torch.from_numpy(r["x"]).cuda()
        dst = torch.empty_like(src)
        torch.cuda.synchronize()  # inputs ready before the server's streams read them
        ids.append(server.submit(r["wavelet"], r["fwd"], src, dst, r["J"]))
        outs.append((src, dst))
        if one_at_a_time:
            server.drain()
    server.drain()
    done = {}
    while True:
        got = server.poll()
        if not got:
            break
        for i, lat, st in got:
            done[i] = (lat, st)
    assert sorted(done) == sorted(ids)
    assert all(st == 0 for _, st in done.values())
    assert all(lat > 0 for lat, _ in done.values())
    return [d.cpu().numpy() for _, d in outs]


def test_results_same_alone_and_in_the_mix():
    from oracle.oracle import oracle
    o = oracle()
    reqs = make_requests()
    with serving.Server(2 << 30, max_batch=64, streams=3) as s:
        say("serving: mixed")
        mixed = run(s, reqs, one_at_a_time=False)
        b_mixed, n_mixed = s.stats()
        say(f"  {n_mixed} requests in {b_mixed} batches")
        say("serving: one at a time")
        alone = run(s, reqs, one_at_a_time=True)
    assert b_mixed < n_mixed, "requests were batched"
    for r, a, m in zip(reqs, alone, mixed):
        np.testing.assert_array_equal(m, a)
        if r["fwd"]:
            want, j = o.fwd2(r["x"], r["wavelet"], r["J"])
            assert j == r["J"]
        else:
            want = o.inv2(r["x"], r["wavelet"], r["J"])
        np.testing.assert_array_equal(m, want)


def test_small_budget_admission():
    """A budget that holds only a few requests' scratch at a time."""
    reqs = make_requests(16, seed=5)
    with serving.Server(8 << 20, max_batch=64, streams=2) as s:
        small = run(s, reqs, one_at_a_time=False)
    with serving.Server(2 << 30, max_batch=64, streams=2) as s:
        big = run(s, reqs, one_at_a_time=False)
    for a, b in zip(small, big):
        np.testing.assert_array_equal(a, b)


def test_rejects_request_larger_than_budget():
    x = torch.zeros((512, 512), device="cuda")
    y = torch.empty_like(x)
    with serving.Server(64 << 10, max_batch=4, streams=1) as s:
        with pytest.raises(libdwt_amd.DwtError):
            s.submit("cdf97", True, x, y, 3)
