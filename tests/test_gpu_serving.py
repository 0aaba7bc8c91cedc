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
    """Every buffer is made first.  One at a time: submit, then wait for it.
    Mixed: all requests in one burst from the native replay loop, so that
    they arrive while earlier ones are on the GPU (a Python loop that
    uploads each input between submissions can be slower than the server,
    which then never has more than one request waiting)."""
    outs = []
    for r in reqs:
        src = torch.from_numpy(r["x"]).cuda()
        outs.append((src, torch.empty_like(src)))
    torch.cuda.synchronize()  # inputs ready before the server's streams read them
    if one_at_a_time:
        ids = []
        for r, (src, dst) in zip(reqs, outs):
            ids.append(server.submit(r["wavelet"], r["fwd"], src, dst, r["J"]))
            server.drain()
    else:
        ids = server.replay([server.request(r["wavelet"], r["fwd"], src, dst, r["J"])
                             for r, (src, dst) in zip(reqs, outs)], [0.0] * len(reqs))
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


def test_16bit_storage_requests_same_alone_in_the_mix_and_as_a_batch():
    """FP16 and BF16 requests (forward: image -> 16-bit coefficients and
    per-subband scales; inverse: back to FP32) mixed with FP32 ones: the
    coefficients, scales and reconstructions are bit-identical alone and in
    the mix, and to the batched path's (dwt_gpu_batch_2d) for the same
    storage, which tests/test_gpu_batch.py holds to its rounding bars."""
    rng = np.random.default_rng(41)
    sizes = workloads.uneven_sizes(18, 128, 700, seed=41)
    storages = ["f16", "bf16", "f32"]
    reqs = []
    for k, (W, H) in enumerate(sizes):
        J = int(rng.integers(1, min(8, workloads.ceil_log2(min(W, H))) + 1))
        reqs.append(dict(wavelet="cdf97" if k % 4 < 2 else "cdf53", storage=storages[k % 3], J=J,
                         x=torch.from_numpy(workloads.sinusoid_image(W, H, 300 + k)).cuda()))
    dtypes = {"f16": torch.float16, "bf16": torch.bfloat16, "f32": torch.float32}

    def submit_all(server, args, one_at_a_time):
        # buffers made beforehand; mixed: one burst from the native replay loop
        torch.cuda.synchronize()
        if one_at_a_time:
            for a in args:
                server.submit(*a)
                server.drain()
        else:
            server.replay([server.request(*a) for a in args], [0.0] * len(args))
        server.drain()
        while server.poll():
            pass

    def forward(server, one_at_a_time):
        outs = [(torch.empty(r["x"].shape, dtype=dtypes[r["storage"]], device="cuda"),
                 torch.zeros(3 * r["J"] + 1, device="cuda") if r["storage"] != "f32" else None) for r in reqs]
        submit_all(server, [(r["wavelet"], True, r["x"], c, r["J"], r["storage"], s)
                            for r, (c, s) in zip(reqs, outs)], one_at_a_time)
        return outs

    def inverse(server, coefs, one_at_a_time):
        outs = [torch.empty_like(r["x"]) for r in reqs]
        submit_all(server, [(r["wavelet"], False, c, y, r["J"], r["storage"], s)
                            for r, (c, s), y in zip(reqs, coefs, outs)], one_at_a_time)
        return outs

    with serving.Server(1 << 30, max_batch=64, streams=3) as srv:
        say("serving 16-bit: forward mixed / alone")
        mixed = forward(srv, False)
        alone = forward(srv, True)
        say("serving 16-bit: inverse mixed / alone")
        rec_mixed = inverse(srv, mixed, False)
        rec_alone = inverse(srv, alone, True)
    for r, (cm, sm), (ca, sa), ym, ya in zip(reqs, mixed, alone, rec_mixed, rec_alone):
        tag = f"{tuple(r['x'].shape)} {r['wavelet']} {r['storage']} J={r['J']}"
        assert torch.equal(cm.view(torch.int16) if r["storage"] != "f32" else cm,
                           ca.view(torch.int16) if r["storage"] != "f32" else ca), tag
        if sm is not None:
            assert torch.equal(sm, sa), tag
        assert torch.equal(ym, ya), tag
        cb, sb, done = libdwt_amd.batch_forward([r["x"]], r["wavelet"], r["storage"], [r["J"]])
        yb = libdwt_amd.batch_inverse(cb, done, r["wavelet"], r["storage"], sb)
        torch.cuda.synchronize()
        assert done == [r["J"]], tag
        if r["storage"] != "f32":
            assert torch.equal(cm.view(torch.int16), cb[0].view(torch.int16)), tag
            assert torch.equal(sm, sb[0]), tag
        else:
            assert torch.equal(cm, cb[0]), tag
        assert torch.equal(ym, yb[0]), tag
