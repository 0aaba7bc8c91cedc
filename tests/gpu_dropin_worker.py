"""Runs the drop-in C entry points (which abort the process on error, as the
reference does) over the fixture cases, in a child process of the GPU tests.

Usage: python tests/gpu_dropin_worker.py MODE
MODE: dense | sparse | host | lines | slice | pinned | example
Prints "case <key>" to stderr before each case (an abort then names it) and
one JSON object on stdout at the end: {"mode", "cases", "mismatches": [...]}.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(HERE, "golden"))

import cases  # noqa: E402
import libdwt_amd  # noqa: E402
from libdwt_amd import workloads  # noqa: E402


def note(key):
    sys.stderr.write(f"case {key}\n")
    sys.stderr.flush()


def run_case(torch, c, host=False):
    dt = np.float32 if c["prec"] == "s" else np.float64
    buf, sx, sy = cases.make_buffer(c["layout"], c["Wo"], c["Ho"], dt, c["seed"])
    if host:
        ptr = buf.ctypes.data
        get = lambda: buf  # noqa: E731
    else:
        d = torch.from_numpy(buf).cuda()
        torch.cuda.synchronize()
        ptr = d.data_ptr()
        get = lambda: d.cpu().numpy()  # noqa: E731
    j = libdwt_amd.cdf_2f(c["wavelet"], c["prec"], ptr, sx, sy, c["Wo"], c["Ho"], c["Wi"],
                          c["Hi"], c["j"], c["d1"], c["zp"])
    hf = cases.digest(get())
    libdwt_amd.cdf_2i(c["wavelet"], c["prec"], ptr, sx, sy, c["Wo"], c["Ho"], c["Wi"], c["Hi"],
                      j, c["d1"], c["zp"])
    return [j, hf, cases.digest(get())]


def sweep(torch, gold, which, host=False, every=1):
    gen = cases.dense_cases() if which == "dense" else cases.sparse_cases()
    n, bad = 0, []
    for i, c in enumerate(gen):
        if i % every:
            continue
        key = cases.case_key(c)
        note(key)
        n += 1
        if run_case(torch, c, host) != gold[which][key]:
            bad.append(key)
    return n, bad


def lines(torch, arrays):
    lib = libdwt_amd.lib()
    n, bad = 0, []
    for w in cases.WAVELETS:
        for prec, dt in (("s", np.float32), ("d", np.float64)):
            fwd = getattr(lib, f"dwt_{w}_f_ex_stride_{prec}")
            inv = getattr(lib, f"dwt_{w}_i_ex_stride_{prec}")
            it = np.dtype(dt).itemsize
            for N in range(41):
                rng = np.random.default_rng(7000 + N)
                x = rng.standard_normal(N).astype(dt)
                ci = rng.standard_normal(N).astype(dt)
                k = f"line_{w}_{prec}_{N}"
                nl = (N + 1) // 2
                for host in (False, True):
                    note(f"{k} host={host}")
                    n += 1
                    # stride of 3 elements: interleaved with two other signals
                    src = np.zeros(3 * N + 3, dt)
                    src[0:3 * N:3] = x
                    out = np.zeros(3 * N + 3, dt)
                    if host:
                        s_ptr, o_ptr = src.ctypes.data, out.ctypes.data
                    else:
                        ds, do = torch.from_numpy(src).cuda(), torch.from_numpy(out).cuda()
                        torch.cuda.synchronize()
                        s_ptr, o_ptr = ds.data_ptr(), do.data_ptr()
                    fwd(s_ptr, o_ptr, o_ptr + nl * 3 * it, None, N, 3 * it)
                    res = out if host else do.cpu().numpy()
                    ok = (np.array_equal(res[0:3 * nl:3], arrays[k + "_L"])
                          and np.array_equal(res[3 * nl:3 * N:3], arrays[k + "_H"]))
                    # inverse in place over the coefficient buffer (src aliases dst)
                    cbuf = np.zeros(N + 1, dt)
                    cbuf[:N] = ci
                    if host:
                        p = cbuf.ctypes.data
                        inv(p, p + nl * it, p, None, N, it)
                        got = cbuf[:N]
                    else:
                        dc = torch.from_numpy(cbuf).cuda()
                        torch.cuda.synchronize()
                        p = dc.data_ptr()
                        inv(p, p + nl * it, p, None, N, it)
                        got = dc.cpu().numpy()[:N]
                    ok = ok and np.array_equal(got, arrays[k + "_inv"])
                    if not ok:
                        bad.append(f"{k} host={host}")
    return n, bad


def line_slice(torch, gold):
    """256 random walks x 8192, full depth by the caller loop over single
    levels, stored as matrix columns (element stride 256 x 4 bytes)."""
    ls = cases.LINE_SLICE
    sig = workloads.random_walk(ls["count"], ls["length"], ls["seed"])
    rec = gold["lines"]["slice_256x8192"]
    lib = libdwt_amd.lib()
    n_sig, N = sig.shape
    bad = []
    for w in cases.WAVELETS:
        note(f"slice {w}")
        f = getattr(lib, f"dwt_{w}_f_ex_stride_s")
        i = getattr(lib, f"dwt_{w}_i_ex_stride_s")
        mat = torch.from_numpy(np.ascontiguousarray(sig.T)).cuda()
        torch.cuda.synchronize()
        base, stride = mat.data_ptr(), n_sig * 4
        sizes = []
        n = N
        while n >= 2:
            sizes.append(n)
            n = (n + 1) // 2
        for s in range(n_sig):
            p = base + 4 * s
            for n in sizes:
                nl = (n + 1) // 2
                f(p, p, p + nl * stride, None, n, stride)
        if cases.digest(np.ascontiguousarray(mat.cpu().numpy().T)) != rec[w]["fwd"]:
            bad.append(f"slice {w} fwd")
        for s in range(n_sig):
            p = base + 4 * s
            for n in reversed(sizes):
                nl = (n + 1) // 2
                i(p, p + nl * stride, p, None, n, stride)
        if cases.digest(np.ascontiguousarray(mat.cpu().numpy().T)) != rec[w]["inv"]:
            bad.append(f"slice {w} inv")
    return 2, bad


def pinned(torch, gold):
    cfg = cases.CONFIG_IMAGES[0]
    img = workloads.sinusoid_image(cfg["W"], cfg["H"], cfg["seed"])
    t = torch.from_numpy(img).pin_memory()
    note("pinned 512x512 cdf53")
    rec = gold["config"][cfg["name"]]["cdf53"]
    j = libdwt_amd.cdf_2f("cdf53", "s", t.data_ptr(), 4 * cfg["W"], 4, cfg["W"], cfg["H"],
                          cfg["W"], cfg["H"], cfg["J"])
    bad = [] if cases.digest(t.numpy()) == rec["fwd"] else ["pinned fwd"]
    libdwt_amd.cdf_2i("cdf53", "s", t.data_ptr(), 4 * cfg["W"], 4, cfg["W"], cfg["H"], cfg["W"],
                      cfg["H"], j)
    if cases.digest(t.numpy()) != rec["inv"]:
        bad.append("pinned inv")
    return 1, bad


def example(torch, gold):
    """The reference example program's case: 512x512 type-0 image at a
    2053-byte (unaligned) row pitch, full depth."""
    ex = cases.EXAMPLE_SIMPLE
    img = cases.type0_image(ex["W"], ex["H"])
    buf = np.zeros(ex["pitch"] * ex["H"] + 16, dtype=np.uint8)
    offs = (np.arange(ex["H"])[:, None] * ex["pitch"] + np.arange(ex["W"])[None, :] * 4)[..., None] \
        + np.arange(4)
    buf[offs] = img.view(np.uint8).reshape(ex["H"], ex["W"], 4)
    rec = gold["example_simple"]
    bad = [] if cases.digest(buf) == rec["input"] else ["input"]
    for w in cases.WAVELETS:
        for host in (False, True):
            note(f"example {w} host={host}")
            if host:
                b = buf.copy()
                ptr, get = b.ctypes.data, (lambda b=b: b)
            else:
                d = torch.from_numpy(buf.copy()).cuda()
                torch.cuda.synchronize()
                ptr, get = d.data_ptr(), (lambda d=d: d.cpu().numpy())
            j = libdwt_amd.cdf_2f(w, "s", ptr, ex["pitch"], 4, ex["W"], ex["H"], ex["W"], ex["H"], -1)
            if j != rec[w]["j"] or cases.digest(get()) != rec[w]["fwd"]:
                bad.append(f"example {w} host={host} fwd")
            libdwt_amd.cdf_2i(w, "s", ptr, ex["pitch"], 4, ex["W"], ex["H"], ex["W"], ex["H"], j)
            if cases.digest(get()) != rec[w]["inv"]:
                bad.append(f"example {w} host={host} inv")
    return 4, bad


def main():
    mode = sys.argv[1]
    import torch
    assert torch.cuda.is_available()
    with open(os.path.join(HERE, "golden", "golden_hashes.json")) as fh:
        gold = json.load(fh)
    if mode in ("dense", "sparse"):
        n, bad = sweep(torch, gold, mode)
    elif mode == "host":
        n1, b1 = sweep(torch, gold, "sparse", host=True, every=7)
        n2, b2 = sweep(torch, gold, "dense", host=True, every=11)
        n, bad = n1 + n2, b1 + b2
    elif mode == "lines":
        n, bad = lines(torch, np.load(os.path.join(HERE, "golden", "golden_arrays.npz")))
    elif mode == "slice":
        n, bad = line_slice(torch, gold)
    elif mode == "pinned":
        n, bad = pinned(torch, gold)
    elif mode == "example":
        n, bad = example(torch, gold)
    else:
        raise SystemExit(f"unknown mode {mode}")
    print(json.dumps({"mode": mode, "cases": n, "mismatches": bad}), flush=True)


if __name__ == "__main__":
    main()
