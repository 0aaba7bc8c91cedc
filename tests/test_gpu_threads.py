"""GPU: the drop-in drivers called from several host threads at once, as the
reference allows (its only global, the accel setting, is read-only during a
call; src/libdwt.c:522).  Scratch is kept per (thread, stream, device), so
concurrent calls must not disturb each other: every thread's results are
bit-exact to the oracle, for packed, sparse and interleaved-channel frames,
through the synchronous drivers and through the async entry point on each
thread's own stream."""
import ctypes as C
import sys
import threading

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


# (outer W, H, inner W, H, layout): packed, sparse (inner in a larger frame),
# chan (channel 1 of an interleaved 3-channel image)
JOBS = [(640, 480, 640, 480, "packed"), (1024, 1024, 1000, 700, "packed_sparse"),
        (300, 257, 300, 257, "chan"), (2048, 512, 2048, 512, "packed"),
        (513, 129, 513, 129, "packed"), (256, 256, 200, 100, "packed_sparse")]


def make_job(k, rng):
    Wo, Ho, Wi, Hi, kind = JOBS[k % len(JOBS)]
    vals = rng.standard_normal((Ho, Wo) if kind != "chan" else (Ho, Wo, 3)).astype(np.float32)
    return dict(Wo=Wo, Ho=Ho, Wi=Wi, Hi=Hi, kind=kind, vals=vals,
                w="cdf97" if k % 2 == 0 else "cdf53", J=-1 if k % 3 else 3)


def plane(job, a):
    return a[..., 1] if job["kind"] == "chan" else a


def expected(o, job):
    ref, j = o.fwd2(np.ascontiguousarray(plane(job, job["vals"])), job["w"], job["J"],
                    inner=(job["Wi"], job["Hi"]))
    back = o.inv2(ref, job["w"], j, inner=(job["Wi"], job["Hi"]))
    return ref, back, j


def run_dropin(L, job, t):
    """Forward then inverse through the synchronous drivers on device memory."""
    es = 4
    if job["kind"] == "chan":
        ptr, sx, sy = t.data_ptr() + es, 3 * job["Wo"] * es, 3 * es
    else:
        ptr, sx, sy = t.data_ptr(), job["Wo"] * es, es
    f = getattr(L, f"dwt_{job['w']}_2f_s")
    i = getattr(L, f"dwt_{job['w']}_2i_s")
    j = C.c_int(job["J"])
    f(ptr, sx, sy, job["Wo"], job["Ho"], job["Wi"], job["Hi"], C.byref(j), 0, 0)
    coef = t.cpu().numpy().copy()
    i(ptr, sx, sy, job["Wo"], job["Ho"], job["Wi"], job["Hi"], j.value, 0, 0)
    return coef, t.cpu().numpy(), j.value


@pytest.mark.parametrize("api", ["dropin", "async"])
def test_concurrent_host_threads(api):
    from oracle.oracle import oracle
    o = oracle()
    L = libdwt_amd.lib()
    nthreads, per_thread = 4, 6
    rng = np.random.default_rng(31)
    jobs = [[make_job(t * per_thread + k, rng) for k in range(per_thread)] for t in range(nthreads)]
    want = [[expected(o, job) for job in jl] for jl in jobs]
    results = [[None] * per_thread for _ in range(nthreads)]
    errors = []

    def worker(tid):
        try:
            stream = torch.cuda.Stream() if api == "async" else None
            for k, job in enumerate(jobs[tid]):
                t = torch.from_numpy(job["vals"].copy()).cuda()
                torch.cuda.synchronize()
                if api == "dropin":
                    results[tid][k] = run_dropin(L, job, t)
                    continue
                es = 4
                if job["kind"] == "chan":
                    ptr, sx, sy = t.data_ptr() + es, 3 * job["Wo"] * es, 3 * es
                else:
                    ptr, sx, sy = t.data_ptr(), job["Wo"] * es, es
                sp = stream.cuda_stream
                j = libdwt_amd.transform_2d_async(job["w"], True, ptr, sx, sy, job["Wo"], job["Ho"],
                                                  job["Wi"], job["Hi"], j=job["J"], stream=sp)
                stream.synchronize()
                coef = t.cpu().numpy().copy()
                libdwt_amd.transform_2d_async(job["w"], False, ptr, sx, sy, job["Wo"], job["Ho"],
                                              job["Wi"], job["Hi"], j=j, stream=sp)
                stream.synchronize()
                results[tid][k] = (coef, t.cpu().numpy(), j)
        except Exception as e:  # reported by the main thread
            errors.append((tid, repr(e)))

    say(f"{nthreads} threads x {per_thread} transforms ({api})")
    threads = [threading.Thread(target=worker, args=(tid,)) for tid in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
    assert not errors, errors
    for tid in range(nthreads):
        for k, job in enumerate(jobs[tid]):
            coef, back, j = results[tid][k]
            ref, rec, jr = want[tid][k]
            assert j == jr
            np.testing.assert_array_equal(plane(job, coef), ref, err_msg=f"thread {tid} job {k} forward")
            np.testing.assert_array_equal(plane(job, back), rec, err_msg=f"thread {tid} job {k} inverse")
            if job["kind"] == "chan":  # the other channels untouched
                np.testing.assert_array_equal(back[..., 0], job["vals"][..., 0])
