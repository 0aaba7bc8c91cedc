"""CPU: the dwt_util_* helpers behave like the reference's (compared call for
call with the reference built from its own sources in oracle/_ref), and the
reference's example programs compile and link unchanged against
include/libdwt.h and libdwt_gpu.so."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import libdwt_amd
from oracle import oracle as orc

REF_EXAMPLES = "/root/reference/examples"

need_ref = pytest.mark.skipif(not orc.have_ref(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def libs():
    ours = libdwt_amd.lib()
    ref = orc.Reference().lib
    for L in (ours, ref):
        L.dwt_util_next_prime.argtypes = [C.c_int]
        L.dwt_util_is_prime.argtypes = [C.c_int]
        L.dwt_util_get_opt_stride.argtypes = [C.c_int]
        L.dwt_util_test_image_fill_s.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dwt_util_test_image_fill_d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dwt_util_compare_s.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dwt_util_conv_show_s.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dwt_util_save_to_pgm_s.argtypes = [C.c_char_p, C.c_float, C.c_void_p, C.c_int, C.c_int,
                                             C.c_int, C.c_int]
        L.dwt_util_subband_s.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.dwt_util_addr_coeff_s.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dwt_util_addr_coeff_s.restype = C.c_void_p
    return ours, ref


@need_ref
def test_primes_and_strides(libs):
    ours, ref = libs
    for n in list(range(-3, 4000)) + [65535, 65536, 1 << 20]:
        assert ours.dwt_util_is_prime(n) == ref.dwt_util_is_prime(n), n
    for n in list(range(1, 3000, 7)) + [2048, 4096 * 4, 3001 * 4]:
        assert ours.dwt_util_next_prime(n) == ref.dwt_util_next_prime(n), n
        assert ours.dwt_util_get_opt_stride(n) == ref.dwt_util_get_opt_stride(n), n
    assert ours.dwt_util_get_opt_stride(2048) == 2053  # the example program's pitch
    assert ours.dwt_util_is_prime(341) == 1  # base-2 Fermat pseudoprime, as the reference


@need_ref
@pytest.mark.parametrize("rand", [0, 2])
def test_test_image_and_compare(libs, rand):
    ours, ref = libs
    W, H, pitch = 67, 45, 4 * 71
    a = np.zeros(pitch * H, np.uint8)
    b = np.zeros(pitch * H, np.uint8)
    ours.dwt_util_test_image_fill_s(a.ctypes.data, pitch, 4, W, H, rand)
    ref.dwt_util_test_image_fill_s(b.ctypes.data, pitch, 4, W, H, rand)
    assert a.tobytes() == b.tobytes()
    ad = np.zeros(8 * W * H, np.uint8)
    bd = np.zeros(8 * W * H, np.uint8)
    ours.dwt_util_test_image_fill_d(ad.ctypes.data, 8 * W, 8, W, H, rand)
    ref.dwt_util_test_image_fill_d(bd.ctypes.data, 8 * W, 8, W, H, rand)
    assert ad.tobytes() == bd.tobytes()
    c = b.copy()
    f = c.view(np.float32)
    for delta, val in ((0.0, 0), (5e-4, 0), (2e-3, 1)):
        c[:] = b
        f[3] += delta
        for L in (ours, ref):
            assert L.dwt_util_compare_s(b.ctypes.data, c.ctypes.data, pitch, 4, W, H) == val
    f[5] = np.nan
    assert ours.dwt_util_compare_s(b.ctypes.data, c.ctypes.data, pitch, 4, W, H) == 1


@need_ref
def test_conv_show_pgm_subband_addr(libs):
    ours, ref = libs
    W, H = 37, 29
    rng = np.random.default_rng(3)
    src = (rng.standard_normal((H, W)) * 3).astype(np.float32)
    outs = []
    for L in (ours, ref):
        dst = np.zeros_like(src)
        L.dwt_util_conv_show_s(src.ctypes.data, dst.ctypes.data, 4 * W, 4, W, H)
        outs.append(dst)
    np.testing.assert_array_equal(outs[0], outs[1])
    with tempfile.TemporaryDirectory() as d:
        img = np.abs(src) / 4
        files = []
        for k, L in enumerate((ours, ref)):
            fn = os.path.join(d, f"{k}.pgm").encode()
            assert L.dwt_util_save_to_pgm_s(fn, 1.0, img.ctypes.data, 4 * W, 4, W, H) == 0
            files.append(open(fn, "rb").read())
        assert files[0] == files[1]
    buf = np.zeros(4 * 100 * 90, np.uint8)
    for (Wo, Ho, Wi, Hi, j) in ((100, 90, 100, 90, 3), (100, 90, 77, 51, 2), (64, 64, 64, 64, 6)):
        for band in range(4):
            got = []
            for L in (ours, ref):
                p, sx, sy = C.c_void_p(), C.c_int(), C.c_int()
                L.dwt_util_subband_s(buf.ctypes.data, 4 * Wo, 4, Wo, Ho, Wi, Hi, j, band,
                                     C.byref(p), C.byref(sx), C.byref(sy))
                got.append((p.value, sx.value, sy.value))
            assert got[0] == got[1], (Wo, Ho, Wi, Hi, j, band)
    for L in (ours, ref):
        assert L.dwt_util_addr_coeff_s(buf.ctypes.data, 3, 5, 400, 4) == buf.ctypes.data + 3 * 400 + 20


@need_ref
def test_more_utilities_match_reference(libs):
    """ceil_log2, worker count and the float-vector helpers, call for call;
    clock availability agrees except that this library also implements
    CLOCK_MONOTONIC_RAW, which the reference build leaves out."""
    ours, ref = libs
    for x in list(range(1, 5000)) + [1 << 20, (1 << 20) + 1, 1 << 30]:
        assert ours.dwt_util_ceil_log2(x) == ref.dwt_util_ceil_log2(x), x
    assert ours.dwt_util_get_max_workers() == ref.dwt_util_get_max_workers() == 1
    for t in range(16):
        got, want = ours.dwt_util_clock_available(t), ref.dwt_util_clock_available(t)
        assert got == want or (t == 3 and got == 0), t
    for L in (ours, ref):
        L.dwt_util_allocate_vec_s.restype = C.c_void_p
        for f in ("generate", "zero"):
            getattr(L, f"dwt_util_{f}_vec_s").argtypes = [C.c_void_p, C.c_int]
        for f in ("copy", "cmp"):
            getattr(L, f"dwt_util_{f}_vec_s").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    res = []
    for L in (ours, ref):
        a = np.zeros(40, np.float32)
        b = np.zeros(40, np.float32)
        out = [L.dwt_util_generate_vec_s(a.ctypes.data, 40), L.dwt_util_zero_vec_s(b.ctypes.data, 40),
               L.dwt_util_cmp_vec_s(a.ctypes.data, b.ctypes.data, 40), L.dwt_util_copy_vec_s(a.ctypes.data, b.ctypes.data, 40),
               L.dwt_util_cmp_vec_s(a.ctypes.data, b.ctypes.data, 40)]
        b[9] = np.float32(1e-40)  # denormal: "not normal nor zero"
        out.append(L.dwt_util_cmp_vec_s(a.ctypes.data, b.ctypes.data, 40))
        b[9] = a[9] + np.float32(5e-5)
        out.append(L.dwt_util_cmp_vec_s(a.ctypes.data, b.ctypes.data, 40))
        out += [a.tobytes(), b.tobytes()]
        res.append(out)
    assert res[0] == res[1]


@need_ref
def test_c_program_using_util_helpers_links_and_matches_reference():
    """A C program calling dwt_util_ceil_log2, dwt_util_printf and the other
    helpers links unchanged against libdwt_gpu.so and prints exactly what it
    prints linked against the reference."""
    src = os.path.join(os.path.dirname(__file__), "c", "util_helpers.c")
    inc = os.path.dirname(libdwt_amd.HEADER)
    outs = []
    with tempfile.TemporaryDirectory() as d:
        for name, so in (("ours", libdwt_amd.LIB_PATH), ("ref", orc.REF_SO)):
            exe = os.path.join(d, name)
            r = subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror=implicit-function-declaration", "-I", inc,
                                src, so, f"-Wl,-rpath,{os.path.dirname(so)}", "-lm", "-o", exe],
                               capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
            assert r.returncode == 0, r.stderr[-2000:]
            outs.append(r.stdout)
    assert outs[0] == outs[1]
    assert "ceil_log2(70)=7" in outs[0] and "vprintf 42" in outs[0]


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples not present")
@pytest.mark.parametrize("example", ["start", "simple", "simple-double", "simple-perf", "subbands",
                                     "perf-plot"])
def test_reference_examples_link_unchanged(example):
    d0 = os.path.join(REF_EXAMPLES, example)
    src = os.path.join(d0, sorted(f for f in os.listdir(d0) if f.endswith(".c"))[0])
    inc = os.path.dirname(libdwt_amd.HEADER)
    libdir = os.path.dirname(libdwt_amd.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror=implicit-function-declaration",
                            "-I", inc, src, "-L", libdir, "-ldwt_gpu",
                            f"-Wl,-rpath,{libdir}", "-lm", "-o", os.path.join(d, example)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
