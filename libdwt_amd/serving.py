"""Python handle on the library's serving loop (dwt_gpu_server_*).

Requests are images already resident in device memory (torch CUDA tensors
here); the native scheduler thread batches whatever is waiting whenever one
of its streams is free, within an HBM budget for scratch.  Completed requests
are reported with their latency from submission to observed completion.
"""
from __future__ import annotations

import ctypes as C

import libdwt_amd
from libdwt_amd import BatchImage


class Request(C.Structure):
    """struct dwt_gpu_request."""
    _fields_ = [("wavelet", C.c_int), ("direction", C.c_int), ("storage", C.c_int),
                ("reserved", C.c_int), ("image", BatchImage)]


def _bind(L):
    if getattr(L, "_server_bound", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.dwt_gpu_server_create.argtypes = [C.c_longlong, i, i, C.POINTER(vp)]
    L.dwt_gpu_server_create.restype = i
    L.dwt_gpu_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_longlong)]
    L.dwt_gpu_server_submit.restype = i
    L.dwt_gpu_server_replay.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_double), i,
                                        C.POINTER(C.c_longlong)]
    L.dwt_gpu_server_replay.restype = i
    L.dwt_gpu_server_poll.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_double),
                                      C.POINTER(i), i]
    L.dwt_gpu_server_poll.restype = i
    L.dwt_gpu_server_drain.argtypes = [vp]
    L.dwt_gpu_server_drain.restype = i
    L.dwt_gpu_server_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.dwt_gpu_server_stats.restype = i
    L.dwt_gpu_server_destroy.argtypes = [vp]
    L.dwt_gpu_server_destroy.restype = i
    L._server_bound = True
    return L


class Server:
    def __init__(self, budget_bytes: int, max_batch: int = 256, streams: int = 3):
        self.L = _bind(libdwt_amd.lib())
        self.h = C.c_void_p()
        libdwt_amd._check(self.L.dwt_gpu_server_create(budget_bytes, max_batch, streams,
                                                       C.byref(self.h)))
        self._id = C.c_longlong()

    def submit(self, wavelet: str, forward: bool, src, dst, levels: int = -1,
               storage: str = "f32", scales=None) -> int:
        """src / dst: 2D CUDA tensors (image float32 / coefficients in the
        storage dtype, by direction); returns the request id."""
        r = self.request(wavelet, forward, src, dst, levels, storage, scales)
        libdwt_amd._check(self.L.dwt_gpu_server_submit(self.h, C.byref(r), C.byref(self._id)))
        return self._id.value

    def request(self, wavelet: str, forward: bool, src, dst, levels: int = -1,
                storage: str = "f32", scales=None) -> Request:
        """A dwt_gpu_request for src / dst (as submit takes them)."""
        r = Request()
        r.wavelet = libdwt_amd.WAVELET_IDS[wavelet]
        r.direction = libdwt_amd.FORWARD if forward else libdwt_amd.INVERSE
        r.storage = libdwt_amd.STORAGE[storage]
        im = r.image
        H, W = src.shape
        im.src = src.data_ptr()
        im.dst = dst.data_ptr()
        im.src_pitch = src.stride(0) * src.element_size()
        im.dst_pitch = dst.stride(0) * dst.element_size()
        im.width = W
        im.height = H
        im.levels = levels
        im.scales = scales.data_ptr() if scales is not None else None
        return r

    def replay(self, requests, at_s):
        """Submit requests[i] at at_s[i] seconds after the call, from a native
        loop (dwt_gpu_server_replay); returns the request ids."""
        n = len(requests)
        arr = (Request * n)(*requests)
        times = (C.c_double * n)(*at_s)
        ids = (C.c_longlong * n)()
        rc = self.L.dwt_gpu_server_replay(self.h, arr, times, n, ids)
        if rc < 0:
            libdwt_amd._check(rc)
        return list(ids)

    def poll(self, max_n: int = 4096):
        """[(id, latency_s, status)] of requests completed since the last poll."""
        ids = (C.c_longlong * max_n)()
        lat = (C.c_double * max_n)()
        st = (C.c_int * max_n)()
        n = self.L.dwt_gpu_server_poll(self.h, ids, lat, st, max_n)
        return [(ids[k], lat[k], st[k]) for k in range(n)]

    def drain(self) -> None:
        libdwt_amd._check(self.L.dwt_gpu_server_drain(self.h))

    def stats(self):
        b, r = C.c_longlong(), C.c_longlong()
        self.L.dwt_gpu_server_stats(self.h, C.byref(b), C.byref(r))
        return b.value, r.value

    def close(self) -> None:
        if self.h:
            self.L.dwt_gpu_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
