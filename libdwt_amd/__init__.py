"""Python mirror of libdwt's C interface over the MI355X HIP library.

The product is ``libdwt_gpu.so`` (C ABI, see include/libdwt.h).  This module
binds it with ctypes under the reference's own names and argument order
(src/libdwt.h), and adds helpers that run the transforms on torch device
tensors.  Loading fails loudly when the library has not been built: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DWT_GPU_LIB") or os.path.join(HERE, "libdwt_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "libdwt.h")

CDF97, CDF53 = 0, 1
FORWARD, INVERSE = 0, 1
F32, F64 = 0, 1
WAVELET_IDS = {"cdf97": CDF97, "cdf53": CDF53}

_lib = None


class BatchImage(C.Structure):
    """struct dwt_gpu_image."""
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("src_pitch", C.c_longlong),
                ("dst_pitch", C.c_longlong), ("width", C.c_int), ("height", C.c_int),
                ("levels", C.c_int), ("reserved", C.c_int), ("scales", C.c_void_p)]


STORAGE = {"f32": 0, "f16": 1, "bf16": 2}

ULP_BINS = 32


class CompareStats(C.Structure):
    """struct dwt_gpu_compare_stats."""
    _fields_ = [("count", C.c_longlong), ("nonfinite", C.c_longlong), ("over_eps", C.c_longlong),
                ("max_abs", C.c_double), ("sum_sq", C.c_double), ("mse", C.c_double),
                ("peak_b", C.c_double), ("psnr", C.c_double), ("ulp", C.c_longlong * ULP_BINS),
                ("differ", C.c_int), ("reserved", C.c_int)]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("ulp", "reserved")}
        d["ulp"] = list(self.ulp)
        return d


class LibraryMissing(RuntimeError):
    pass


def header_functions() -> list[str]:
    """Names of every function declared in include/libdwt.h."""
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?(?:void|int|long long|char|float|double|dwt_clock_t)"
                      r"\s*\*?\s*(\w+)\s*\(", text, flags=re.M)


def lib() -> C.CDLL:
    """The loaded HIP library (raises LibraryMissing when not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -m libdwt_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i = C.c_void_p, C.c_int
        for w in ("cdf97", "cdf53"):
            for s in ("s", "d"):
                f = getattr(L, f"dwt_{w}_2f_{s}")
                f.argtypes = [vp, i, i, i, i, i, i, C.POINTER(i), i, i]
                f.restype = None
                f = getattr(L, f"dwt_{w}_2i_{s}")
                f.argtypes = [vp, i, i, i, i, i, i, i, i, i]
                f.restype = None
                for d in ("f", "i"):
                    f = getattr(L, f"dwt_{w}_{d}_ex_stride_{s}")
                    f.argtypes = [vp, vp, vp, vp, i, i]
                    f.restype = None
                    f = getattr(L, f"dwt_{w}_{d}_ex_{s}")
                    f.argtypes = [vp, vp, vp, vp, i]
                    f.restype = None
                    f = getattr(L, f"dwt_{w}_{d}_{s}")
                    f.argtypes = [vp, vp, vp, i]
                    f.restype = None
        for s in ("s", "d"):
            getattr(L, f"dwt_zero_padding_f_stride_{s}").argtypes = [vp, vp, i, i, i, i]
            getattr(L, f"dwt_zero_padding_f_{s}").argtypes = [vp, vp, i, i, i]
            getattr(L, f"dwt_zero_padding_i_stride_{s}").argtypes = [vp, i, i, i]
            getattr(L, f"dwt_zero_padding_i_{s}").argtypes = [vp, i, i]
        L.dwt_gpu_init.restype = i
        L.dwt_gpu_last_error.restype = C.c_char_p
        L.dwt_gpu_version.restype = C.c_char_p
        L.dwt_gpu_2d_async.argtypes = [i, i, i, vp, i, i, i, i, i, i, C.POINTER(i), i, i, vp]
        L.dwt_gpu_2d_async.restype = i
        L.dwt_gpu_2d_noalloc.argtypes = [i, i, vp, i, i, i, C.POINTER(i), vp]
        L.dwt_gpu_2d_noalloc.restype = i
        L.dwt_gpu_2d_channels.argtypes = [i, i, i, vp, i, i, i, i, i, i, C.POINTER(i), i, i, vp]
        L.dwt_gpu_2d_channels.restype = i
        L.dwt_gpu_profile_enable.argtypes = [i]
        L.dwt_gpu_profile_enable.restype = i
        L.dwt_gpu_profile_read.argtypes = [i, i, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        L.dwt_gpu_profile_read.restype = i
        L.dwt_gpu_profile_reset.restype = i
        L.dwt_gpu_tile_set.argtypes = [i, i, i, i]
        L.dwt_gpu_tile_set.restype = i
        L.dwt_gpu_tile_get.argtypes = [i, i, i, i, C.POINTER(i), C.POINTER(i)]
        L.dwt_gpu_tile_get.restype = i
        L.dwt_gpu_tail_set.argtypes = [C.c_longlong]
        L.dwt_gpu_tail_set.restype = i
        L.dwt_gpu_level0_set.argtypes = [i, i]
        L.dwt_gpu_level0_set.restype = i
        ip = C.POINTER(C.c_int)
        L.dwt_gpu_level0_plan.argtypes = [i, i, i, i, i, ip, ip, ip, ip, ip, ip, ip, ip]
        L.dwt_gpu_level0_plan.restype = i
        L.dwt_gpu_mid_set.argtypes = [C.c_longlong]
        L.dwt_gpu_mid_set.restype = i
        L.dwt_gpu_scratch_bytes.argtypes = [vp]
        L.dwt_gpu_scratch_bytes.restype = C.c_longlong
        L.dwt_gpu_release_scratch.argtypes = [vp]
        L.dwt_gpu_release_scratch.restype = i
        L.dwt_gpu_schedule_set.argtypes = [i]
        L.dwt_gpu_schedule_set.restype = i
        L.dwt_gpu_device_check.argtypes = [vp]
        L.dwt_gpu_device_check.restype = i
        L.dwt_gpu_pyr_trace_enable.argtypes = [i]
        L.dwt_gpu_pyr_trace_enable.restype = i
        L.dwt_gpu_pyr_trace_read.argtypes = [vp, vp, C.c_longlong]
        L.dwt_gpu_pyr_trace_read.restype = C.c_longlong
        L.dwt_gpu_clock_rate_khz.restype = i
        L.dwt_gpu_pyr_grid_cap_set.argtypes = [i]
        L.dwt_gpu_pyr_grid_cap_set.restype = i
        L.dwt_gpu_batch_2d.argtypes = [i, i, i, C.POINTER(BatchImage), i, vp]
        L.dwt_gpu_batch_2d.restype = i
        L.dwt_gpu_batch_plan_create.argtypes = [i, i, i, C.POINTER(BatchImage), i, vp,
                                                C.POINTER(vp)]
        L.dwt_gpu_batch_plan_create.restype = i
        L.dwt_gpu_batch_plan_execute.argtypes = [vp, vp]
        L.dwt_gpu_batch_plan_execute.restype = i
        L.dwt_gpu_batch_plan_destroy.argtypes = [vp]
        L.dwt_gpu_batch_plan_destroy.restype = i
        L.dwt_gpu_batch_rows_set.argtypes = [i, i]
        L.dwt_gpu_batch_rows_set.restype = i
        L.dwt_gpu_batch_tail_set.argtypes = [C.c_longlong]
        L.dwt_gpu_batch_tail_set.restype = i
        L.dwt_gpu_sparse_set.argtypes = [i]
        L.dwt_gpu_sparse_set.restype = i
        L.dwt_gpu_1d_batch.argtypes = [i, i, i, vp, C.c_longlong, C.c_longlong, i, i,
                                       C.POINTER(i), vp]
        L.dwt_gpu_1d_batch.restype = i
        L.dwt_gpu_compare.argtypes = [i, vp, i, i, i, vp, i, i, i, i, C.c_double, C.c_double,
                                      C.POINTER(CompareStats), vp]
        L.dwt_gpu_compare.restype = i
        L.dwt_gpu_verify_2d.argtypes = [i, i, vp, vp, i, i, i, i, i, C.POINTER(CompareStats), vp]
        L.dwt_gpu_verify_2d.restype = i
        _lib = L
    return _lib


class DwtError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        raise DwtError(f"libdwt_gpu error {rc}: {lib().dwt_gpu_last_error().decode()}")


# -- reference-named calls (addresses are plain integers) ----------------------

def cdf_2f(wavelet: str, prec: str, ptr: int, stride_x: int, stride_y: int, size_o_x: int,
           size_o_y: int, size_i_x: int, size_i_y: int, j_max: int, decompose_one: int = 0,
           zero_padding: int = 0) -> int:
    """dwt_{wavelet}_2f_{prec}; returns the achieved number of levels."""
    j = C.c_int(j_max)
    getattr(lib(), f"dwt_{wavelet}_2f_{prec}")(ptr, stride_x, stride_y, size_o_x, size_o_y,
                                               size_i_x, size_i_y, C.byref(j), decompose_one,
                                               zero_padding)
    return j.value


def cdf_2i(wavelet: str, prec: str, ptr: int, stride_x: int, stride_y: int, size_o_x: int,
           size_o_y: int, size_i_x: int, size_i_y: int, j_max: int, decompose_one: int = 0,
           zero_padding: int = 0) -> None:
    getattr(lib(), f"dwt_{wavelet}_2i_{prec}")(ptr, stride_x, stride_y, size_o_x, size_o_y,
                                               size_i_x, size_i_y, j_max, decompose_one,
                                               zero_padding)


def transform_2d_async(wavelet: str, forward: bool, ptr: int, stride_x: int, stride_y: int,
                       size_o_x: int, size_o_y: int, size_i_x: int | None = None,
                       size_i_y: int | None = None, j: int = -1, decompose_one: int = 0,
                       zero_padding: int = 0, f64: bool = False, stream: int = 0) -> int:
    """dwt_gpu_2d_async on device memory; returns the (clamped) level count."""
    jj = C.c_int(j)
    rc = lib().dwt_gpu_2d_async(WAVELET_IDS[wavelet], FORWARD if forward else INVERSE,
                                F64 if f64 else F32, ptr, stride_x, stride_y, size_o_x, size_o_y,
                                size_o_x if size_i_x is None else size_i_x,
                                size_o_y if size_i_y is None else size_i_y, C.byref(jj),
                                decompose_one, zero_padding, stream)
    _check(rc)
    return jj.value


def transform_channels(wavelet: str, forward: bool, ptr: int, stride_x: int, stride_y: int,
                       channel_stride: int, channels: int, size_x: int, size_y: int, j: int = -1,
                       decompose_one: int = 0, zero_padding: int = 0, f64: bool = False,
                       stream: int = 0) -> int:
    """dwt_gpu_2d_channels: every channel of a multi-channel image on device
    memory (channel c at ptr + c * channel_stride, driver strides in bytes);
    returns the (clamped) level count."""
    jj = C.c_int(j)
    rc = lib().dwt_gpu_2d_channels(WAVELET_IDS[wavelet], FORWARD if forward else INVERSE,
                                   F64 if f64 else F32, ptr, stride_x, stride_y, channel_stride, channels,
                                   size_x, size_y, C.byref(jj), decompose_one, zero_padding, stream)
    _check(rc)
    return jj.value


def schedule_set(mode: int) -> None:
    """1: levels after level 0 as one pyramid (dataflow) launch; 0: one launch
    per level."""
    _check(lib().dwt_gpu_schedule_set(mode))


def device_check(stream: int = 0) -> None:
    """Raise if a pyramid launch on `stream` (this thread) hit its wait limit."""
    _check(lib().dwt_gpu_device_check(stream))


def pyr_trace(stream: int = 0, max_items: int = 1 << 20):
    """Per-item timeline of the last pyramid launch on `stream`: an int64
    array of rows (start, dependencies met, done, workgroup), clock ticks in
    microseconds from the first start."""
    import numpy as np
    buf = np.zeros((max_items, 4), dtype=np.uint64)
    n = lib().dwt_gpu_pyr_trace_read(stream, buf.ctypes.data, max_items)
    if n < 0:
        _check(int(n))
    rows = buf[:n].astype(np.int64)
    khz = lib().dwt_gpu_clock_rate_khz()
    t0 = rows[:, 0].min() if n else 0
    out = rows.astype(np.float64)
    out[:, :3] = (rows[:, :3] - t0) * (1000.0 / khz)
    return out


def lines_1d(wavelet: str, forward: bool, ptr: int, signal_stride: int, sample_stride: int,
             length: int, count: int, levels: int = -1, f64: bool = False,
             stream: int = 0) -> int:
    """dwt_gpu_1d_batch: full-depth 1D transform of `count` strided signals in
    place (strides in bytes); returns the levels done."""
    lv = C.c_int(levels)
    _check(lib().dwt_gpu_1d_batch(WAVELET_IDS[wavelet], FORWARD if forward else INVERSE,
                                  F64 if f64 else F32, ptr, signal_stride, sample_stride, length,
                                  count, C.byref(lv), stream))
    return lv.value


def _image_args(t):
    """(precision, pointer, stride_x, stride_y, width, height) of a 2D CUDA tensor."""
    import torch
    if t.dim() != 2 or t.dtype not in (torch.float32, torch.float64):
        raise ValueError("expected a 2D float32 or float64 tensor")
    es = t.element_size()
    return (F64 if t.dtype == torch.float64 else F32, t.data_ptr(), t.stride(0) * es, t.stride(1) * es,
            t.shape[1], t.shape[0])


def compare(a, b, eps: float = -1.0, peak: float = 0.0, stream: int = 0) -> dict:
    """dwt_gpu_compare of two same-shape 2D CUDA tensors (any strides, FP32 or
    FP64 each), on the device; returns the statistics as a dict."""
    if a.shape != b.shape:
        raise ValueError("shapes differ")
    pa, ptr_a, sxa, sya, W, H = _image_args(a)
    pb, ptr_b, sxb, syb, _, _ = _image_args(b)
    st = CompareStats()
    _check(lib().dwt_gpu_compare(pa, ptr_a, sxa, sya, pb, ptr_b, sxb, syb, W, H, eps, peak,
                                 C.byref(st), stream))
    return st.as_dict()


def verify_2d(inp, result, wavelet: str = "cdf97", forward: bool = True, j: int = -1,
              stream: int = 0) -> dict:
    """dwt_gpu_verify_2d: FP32 `result` of the transform of FP32 `inp` against
    the FP64 transform of `inp` computed on the device."""
    if inp.shape != result.shape or inp.stride() != result.stride():
        raise ValueError("input and result must have the same shape and strides")
    _, ptr_i, sx, sy, W, H = _image_args(inp)
    st = CompareStats()
    _check(lib().dwt_gpu_verify_2d(WAVELET_IDS[wavelet], FORWARD if forward else INVERSE, ptr_i,
                                   result.data_ptr(), sx, sy, W, H, j, C.byref(st), stream))
    return st.as_dict()


def _storage_dtype(storage: str):
    import torch
    return {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[storage]


def batch_descriptors(srcs, dsts, levels, scales=None):
    """dwt_gpu_image array for lists of 2D CUDA tensors (packed rows)."""
    n = len(srcs)
    arr = (BatchImage * n)()
    for k, (a, b) in enumerate(zip(srcs, dsts)):
        H, W = a.shape
        arr[k].src = a.data_ptr()
        arr[k].dst = b.data_ptr()
        arr[k].src_pitch = a.stride(0) * a.element_size()
        arr[k].dst_pitch = b.stride(0) * b.element_size()
        arr[k].width = W
        arr[k].height = H
        arr[k].levels = levels[k] if isinstance(levels, (list, tuple)) else levels
        arr[k].scales = scales[k].data_ptr() if scales is not None else None
    return arr


def batch_forward(images, wavelet: str = "cdf97", storage: str = "f32", levels=-1,
                  stream: int | None = None):
    """Forward transform of a list of 2D float32 CUDA tensors in one batch on
    torch's current stream (or `stream`).  Returns (coefficients, scales,
    levels): coefficient tensors in the storage dtype, per-image scale tensors
    (3 * levels + 1 floats; None for f32) and the levels done per image."""
    import torch
    dt = _storage_dtype(storage)
    coefs = [torch.empty(t.shape, dtype=dt, device=t.device) for t in images]
    scales = None
    if storage != "f32":
        lv = levels if isinstance(levels, (list, tuple)) else [levels] * len(images)
        scales = [torch.zeros(3 * (_max_levels(t) if l < 0 else min(l, _max_levels(t))) + 1,
                              dtype=torch.float32, device=t.device) for t, l in zip(images, lv)]
    arr = batch_descriptors(images, coefs, levels, scales)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    _check(lib().dwt_gpu_batch_2d(WAVELET_IDS[wavelet], FORWARD, STORAGE[storage], arr,
                                  len(images), st))
    return coefs, scales, [arr[k].levels for k in range(len(images))]


def batch_inverse(coefs, levels, wavelet: str = "cdf97", storage: str = "f32", scales=None,
                  stream: int | None = None):
    """Inverse of batch_forward: a list of float32 image tensors."""
    import torch
    outs = [torch.empty(c.shape, dtype=torch.float32, device=c.device) for c in coefs]
    arr = batch_descriptors(coefs, outs, list(levels), scales)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    _check(lib().dwt_gpu_batch_2d(WAVELET_IDS[wavelet], INVERSE, STORAGE[storage], arr, len(coefs),
                                  st))
    return outs


def _max_levels(t) -> int:
    H, W = t.shape
    m = min(H, W)
    r = 0
    while (1 << r) < m:
        r += 1
    return r


def profile_enable(on: bool, repeat_level0: int = 1) -> None:
    """Per-launch HIP-event timing; repeat_level0 > 1 runs every level-0
    forward kernel that many times back to back inside one event pair."""
    lib().dwt_gpu_profile_enable((repeat_level0 if repeat_level0 > 1 else 1) if on else 0)


def profile_reset() -> None:
    lib().dwt_gpu_profile_reset()


def profile_read(kind: int, level: int = -1) -> tuple[int, float]:
    """(launches, total milliseconds) of the recorded launches of one kind
    (0 forward level, 1 inverse level, 2 detail copy, 4 forward block kernel,
    5 inverse block kernel, 6 forward pyramid, 7 inverse pyramid) at one level
    (-1 all)."""
    n = C.c_longlong(0)
    ms = C.c_double(0)
    _check(lib().dwt_gpu_profile_read(kind, level, C.byref(n), C.byref(ms)))
    return n.value, ms.value


# -- torch helpers ----------------------------------------------------------------

def fwd2(t, wavelet: str = "cdf97", j: int = -1, decompose_one: int = 0) -> int:
    """In-place forward 2D transform of a contiguous (H, W) float32/float64 CUDA tensor."""
    return _run2(t, wavelet, True, j, decompose_one)


def inv2(t, wavelet: str = "cdf97", j: int = -1, decompose_one: int = 0) -> None:
    _run2(t, wavelet, False, j, decompose_one)


def _run2(t, wavelet, forward, j, d1):
    import torch
    if not t.is_cuda or t.dim() != 2 or t.dtype not in (torch.float32, torch.float64):
        raise ValueError("expects a 2D float32/float64 CUDA tensor")
    if t.stride(1) != 1:
        raise ValueError("expects unit element stride along rows")
    es = t.element_size()
    H, W = t.shape
    stream = torch.cuda.current_stream(t.device).cuda_stream
    jj = transform_2d_async(wavelet, forward, t.data_ptr(), t.stride(0) * es, es, W, H, W, H, j,
                            d1, 0, t.dtype == torch.float64, stream)
    return jj
