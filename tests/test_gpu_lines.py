"""GPU: full-depth 1D transforms of many strided signals (dwt_gpu_1d_batch),
bit-exact to the oracle's caller loop over the single-level line functions,
for contiguous signals, signals as matrix columns and a generic strided
layout; and the reference-generated digests of the 256 x 8192 random-walk
slice (stored as columns)."""
import json
import os
import sys

import numpy as np
import pytest

import cases
import libdwt_amd
from libdwt_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    assert libdwt_amd.lib().dwt_gpu_init() == 0


def oracle_full(o, sig, wavelet, fwd):
    f = o.f_full_1d if fwd else o.i_full_1d
    return np.stack([f(r, wavelet) for r in sig])


@pytest.mark.parametrize("layout", ["rows", "cols", "strided"])
@pytest.mark.parametrize("prec", ["s", "d"])
def test_lines_match_oracle(layout, prec):
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    for N in (2, 3, 5, 513, 777, 1000, 4097, 8192, 16384):
        count = 37
        sig = workloads.random_walk(count, N, seed=N).astype(dt)
        for w in ("cdf97", "cdf53"):
            say(f"lines {layout} {prec} N={N} {w}")
            want_f = oracle_full(o, sig, w, True)
            want_i = oracle_full(o, want_f, w, False)
            if layout == "rows":  # signal k = row k, padded pitch
                buf = np.zeros((count, N + 3), dtype=dt)
                buf[:, :N] = sig
                t = torch.from_numpy(buf).cuda()
                args = (t.data_ptr(), (N + 3) * es, es)
                read = lambda: t.cpu().numpy()[:, :N]
            elif layout == "cols":  # signal k = column k of an N x cp matrix (cp odd for odd N)
                cp = count + 1 + N % 2
                buf = np.zeros((N, cp), dtype=dt)
                buf[:, :count] = sig.T
                t = torch.from_numpy(buf).cuda()
                args = (t.data_ptr(), es, cp * es)
                read = lambda: t.cpu().numpy()[:, :count].T
            else:  # samples 3 elements apart, each signal in its own region
                sg = 3 * N + 1
                buf = np.zeros((count * sg + 3 * N,), dtype=dt)
                idx = sg * np.arange(count)[:, None] + 3 * np.arange(N)[None, :]
                buf[idx] = sig
                t = torch.from_numpy(buf).cuda()
                args = (t.data_ptr(), sg * es, 3 * es)
                read = lambda: t.cpu().numpy()[idx]
            torch.cuda.synchronize()
            j = libdwt_amd.lines_1d(w, True, args[0], args[1], args[2], N, count, -1, prec == "d")
            torch.cuda.synchronize()
            assert j == workloads.ceil_log2(N)
            np.testing.assert_array_equal(read(), want_f, err_msg="forward")
            libdwt_amd.lines_1d(w, False, args[0], args[1], args[2], N, count, j, prec == "d")
            torch.cuda.synchronize()
            np.testing.assert_array_equal(read(), want_i, err_msg="inverse")


@pytest.mark.parametrize("layout", ["rows", "cols"])
@pytest.mark.parametrize("prec", ["s", "d"])
def test_partial_depth(prec, layout):
    """Contiguous signals and signals as matrix columns (staged in-place level
    0 from two levels on), 1..4 levels (the last level's L to global memory),
    against the oracle's caller loop stopped at that depth."""
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    for N in (6, 300, 4096, 5000):
        sig = workloads.random_walk(9, N, seed=N + 1).astype(dt)
        for w in ("cdf97", "cdf53"):
            for lv in (1, 2, 4):
                lv = min(lv, workloads.ceil_log2(N))  # the forward clamps to ceil(log2 N)
                say(f"lines {layout} {prec} N={N} {w} levels={lv}")
                want = sig.copy()
                for r in want:  # the caller loop, lv levels
                    n = N
                    for _ in range(lv):
                        lo, hi = o.f_line(r[:n].copy(), w)
                        r[:len(lo)] = lo
                        r[len(lo):n] = hi
                        n = len(lo)
                if layout == "rows":
                    t = torch.from_numpy(sig.copy()).cuda()
                    args = (N * es, es)
                    read = lambda: t.cpu().numpy()
                else:  # signal k = column k of an N x 10 matrix
                    t = torch.from_numpy(np.ascontiguousarray(
                        np.concatenate([sig.T, np.zeros((N, 1), dt)], axis=1))).cuda()
                    args = (es, 10 * es)
                    read = lambda: t.cpu().numpy()[:, :9].T
                torch.cuda.synchronize()
                j = libdwt_amd.lines_1d(w, True, t.data_ptr(), args[0], args[1], N, len(sig), lv,
                                        prec == "d")
                torch.cuda.synchronize()
                assert j == lv
                np.testing.assert_array_equal(read(), want, err_msg="forward")
                libdwt_amd.lines_1d(w, False, t.data_ptr(), args[0], args[1], N, len(sig), lv,
                                    prec == "d")
                torch.cuda.synchronize()
                back = want.copy()
                sizes = [N]
                for _ in range(lv - 1):
                    sizes.append((sizes[-1] + 1) // 2)
                for r in back:  # the caller loop of inverse levels, coarsest first
                    for n in reversed(sizes):
                        h = (n + 1) // 2
                        r[:n] = o.i_line(r[:h].copy(), r[h:n].copy(), w)
                np.testing.assert_array_equal(read(), back, err_msg="inverse")


@pytest.mark.parametrize("prec", ["s", "d"])
def test_contiguous_level0_in_place(prec):
    """Contiguous signals around the longest signal one workgroup holds whole
    (level 0 in place; 16384 FP32 / 8192 FP64 samples) and just past it
    (level 0 through scratch), many signals of 64 samples per workgroup, the
    shorter ones (through scratch), and signals whose start or pitch allow
    only narrower loads."""
    from oracle.oracle import oracle
    o = oracle()
    dt = np.float32 if prec == "s" else np.float64
    es = np.dtype(dt).itemsize
    lim = 65536 // es
    cases_ = [(64, 700, 0, 0), (65, 300, 1, 1), (lim, 5, 0, 0), (lim + 1, 5, 0, 0),
              (lim - 2, 4, 1, 2), (lim + 3000, 3, 0, 0), (2, 1000, 0, 0), (63, 500, 0, 1)]
    for N, count, off, extra in cases_:
        pitch = N + extra
        sig = workloads.random_walk(count, N, seed=N + count).astype(dt)
        for w in ("cdf97", "cdf53"):
            say(f"lines rows {prec} N={N} count={count} offset={off} pitch={pitch} {w}")
            want_f = oracle_full(o, sig, w, True)
            want_i = oracle_full(o, want_f, w, False)
            buf = np.zeros(off + count * pitch, dtype=dt)
            view = lambda a: a[off:].reshape(count, pitch)[:, :N]
            view(buf)[:] = sig
            t = torch.from_numpy(buf).cuda()
            ptr = t.data_ptr() + off * es
            torch.cuda.synchronize()
            j = libdwt_amd.lines_1d(w, True, ptr, pitch * es, es, N, count, -1, prec == "d")
            torch.cuda.synchronize()
            got = t.cpu().numpy()
            np.testing.assert_array_equal(view(got), want_f, err_msg="forward")
            assert not got[:off].any() and not got[off:].reshape(count, pitch)[:, N:].any()
            libdwt_amd.lines_1d(w, False, ptr, pitch * es, es, N, count, j, prec == "d")
            torch.cuda.synchronize()
            got = t.cpu().numpy()
            np.testing.assert_array_equal(view(got), want_i, err_msg="inverse")
            assert not got[:off].any() and not got[off:].reshape(count, pitch)[:, N:].any()


def test_degenerate_arguments():
    """Length 0 and 1 (no level) and no signals touch nothing; an inverse
    asked for more levels than the length allows, and negative sizes, are
    errors that leave the data as it was."""
    sig = torch.arange(64 * 5, dtype=torch.float32, device="cuda").reshape(5, 64)
    keep = sig.clone()
    for N, count in ((0, 5), (1, 5), (64, 0)):
        for layout in ("rows", "cols"):
            args = (64 * 4, 4) if layout == "rows" else (4, 64 * 4)
            say(f"degenerate {layout} N={N} count={count}")
            j = libdwt_amd.lines_1d("cdf97", True, sig.data_ptr(), args[0], args[1], N, count, -1)
            assert j == workloads.ceil_log2(max(N, 1))  # the depth, also with no signals
            libdwt_amd.lines_1d("cdf97", False, sig.data_ptr(), args[0], args[1], N, count, j)
    torch.cuda.synchronize()
    assert torch.equal(sig, keep)
    for bad in (dict(N=64, count=5, lv=7, fwd=False), dict(N=-1, count=5, lv=-1, fwd=True),
                dict(N=64, count=-2, lv=-1, fwd=True)):
        say(f"error case {bad}")
        with pytest.raises(RuntimeError):
            libdwt_amd.lines_1d("cdf97", bad["fwd"], sig.data_ptr(), 64 * 4, 4, bad["N"], bad["count"],
                                bad["lv"])
    torch.cuda.synchronize()
    assert torch.equal(sig, keep)


def test_slice_digests():
    """The 256 x 8192 random-walk slice as matrix columns through the batched
    entry point, against the digests the reference produced."""
    with open(os.path.join(GOLD, "golden_hashes.json")) as fh:
        rec = json.load(fh)["lines"]["slice_256x8192"]
    ls = cases.LINE_SLICE
    sig = workloads.random_walk(ls["count"], ls["length"], ls["seed"])
    n_sig, N = sig.shape
    for w in cases.WAVELETS:
        say(f"slice {w}")
        mat = torch.from_numpy(np.ascontiguousarray(sig.T)).cuda()
        torch.cuda.synchronize()
        j = libdwt_amd.lines_1d(w, True, mat.data_ptr(), 4, n_sig * 4, N, n_sig)
        torch.cuda.synchronize()
        assert cases.digest(np.ascontiguousarray(mat.cpu().numpy().T)) == rec[w]["fwd"]
        libdwt_amd.lines_1d(w, False, mat.data_ptr(), 4, n_sig * 4, N, n_sig, j)
        torch.cuda.synchronize()
        assert cases.digest(np.ascontiguousarray(mat.cpu().numpy().T)) == rec[w]["inv"]
        # the same signals as rows
        rows = torch.from_numpy(np.ascontiguousarray(sig)).cuda()
        torch.cuda.synchronize()
        libdwt_amd.lines_1d(w, True, rows.data_ptr(), N * 4, 4, N, n_sig)
        torch.cuda.synchronize()
        assert cases.digest(rows.cpu().numpy()) == rec[w]["fwd"]
