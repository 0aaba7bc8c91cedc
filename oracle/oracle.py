"""ORACLE -- test infrastructure only.

ctypes front end for two CPU libraries with the same C interface:

* ``liboracle.so`` -- the clean-room restatement in ``dwt_oracle.c`` (symbols
  prefixed ``orc_``);
* ``_ref/libdwt_ref.so`` -- the reference libdwt compiled from its own sources
  by ``oracle/Makefile`` (symbols prefixed ``dwt_``), present only where it was
  built.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product library never loads
it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdwt_ref.so")

WAVELETS = ("cdf97", "cdf53")


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class DwtLib:
    """One CPU implementation of the libdwt C interface."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix
        vp, i = C.c_void_p, C.c_int
        for w in WAVELETS:
            for suf in ("s", "d"):
                f = self._fn(f"{w}_2f_{suf}")
                f.argtypes = [vp, i, i, i, i, i, i, C.POINTER(i), i, i]
                f.restype = None
                f = self._fn(f"{w}_2i_{suf}")
                f.argtypes = [vp, i, i, i, i, i, i, i, i, i]
                f.restype = None
                for d in ("f", "i"):
                    f = self._fn(f"{w}_{d}_ex_stride_{suf}")
                    f.argtypes = [vp, vp, vp, vp, i, i]
                    f.restype = None

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    # -- raw calls on byte buffers ------------------------------------------
    def fwd2_raw(self, buf: np.ndarray, wavelet: str, sx: int, sy: int, Wo: int,
                 Ho: int, Wi: int, Hi: int, j: int, d1: int = 0, zp: int = 0,
                 prec: str = "s", offset: int = 0) -> int:
        jj = C.c_int(j)
        self._fn(f"{wavelet}_2f_{prec}")(buf.ctypes.data + offset, sx, sy, Wo, Ho,
                                          Wi, Hi, C.byref(jj), d1, zp)
        return jj.value

    def inv2_raw(self, buf: np.ndarray, wavelet: str, sx: int, sy: int, Wo: int,
                 Ho: int, Wi: int, Hi: int, j: int, d1: int = 0, zp: int = 0,
                 prec: str = "s", offset: int = 0) -> None:
        self._fn(f"{wavelet}_2i_{prec}")(buf.ctypes.data + offset, sx, sy, Wo, Ho,
                                          Wi, Hi, j, d1, zp)

    # -- packed image helpers ------------------------------------------------
    def fwd2(self, img: np.ndarray, wavelet: str = "cdf97", j: int = -1,
             d1: int = 0, zp: int = 0, inner=None):
        """Forward 2D transform of a packed (H, W) array; returns (coeffs, j)."""
        a = np.ascontiguousarray(img).copy()
        H, W = a.shape
        Wi, Hi = inner if inner is not None else (W, H)
        prec = "s" if a.dtype == np.float32 else "d"
        jj = self.fwd2_raw(a, wavelet, a.strides[0], a.strides[1], W, H, Wi, Hi,
                           j, d1, zp, prec)
        return a, jj

    def inv2(self, coef: np.ndarray, wavelet: str = "cdf97", j: int = -1,
             d1: int = 0, zp: int = 0, inner=None) -> np.ndarray:
        a = np.ascontiguousarray(coef).copy()
        H, W = a.shape
        Wi, Hi = inner if inner is not None else (W, H)
        prec = "s" if a.dtype == np.float32 else "d"
        self.inv2_raw(a, wavelet, a.strides[0], a.strides[1], W, H, Wi, Hi, j, d1,
                      zp, prec)
        return a

    # -- 1D lines --------------------------------------------------------------
    def f_line(self, x: np.ndarray, wavelet: str = "cdf97"):
        """One forward level of a contiguous line; returns (L, H)."""
        x = np.ascontiguousarray(x)
        N = x.shape[0]
        prec = "s" if x.dtype == np.float32 else "d"
        out = np.zeros(N + 2, dtype=x.dtype)
        tmp = np.zeros(N + 8, dtype=x.dtype)
        nl = (N + 1) // 2
        src = x.copy() if N else np.zeros(1, x.dtype)
        self._fn(f"{wavelet}_f_ex_stride_{prec}")(
            src.ctypes.data, out.ctypes.data, out.ctypes.data + nl * x.itemsize,
            tmp.ctypes.data, N, x.itemsize)
        return out[:nl].copy(), out[nl:N].copy()

    def i_line(self, lo: np.ndarray, hi: np.ndarray, wavelet: str = "cdf97"):
        lo = np.ascontiguousarray(lo)
        hi = np.ascontiguousarray(hi)
        N = lo.shape[0] + hi.shape[0]
        prec = "s" if lo.dtype == np.float32 else "d"
        out = np.zeros(N + 1, dtype=lo.dtype)
        tmp = np.zeros(N + 8, dtype=lo.dtype)
        l2 = np.concatenate([lo, np.zeros(1, lo.dtype)])
        h2 = np.concatenate([hi, np.zeros(1, lo.dtype)])
        self._fn(f"{wavelet}_i_ex_stride_{prec}")(
            l2.ctypes.data, h2.ctypes.data, out.ctypes.data, tmp.ctypes.data, N,
            lo.itemsize)
        return out[:N].copy()

    def f_full_1d(self, x: np.ndarray, wavelet: str = "cdf97") -> np.ndarray:
        """Full-depth 1D decomposition by the caller loop over single levels:
        level L (length n) writes L to [0, ceil(n/2)) and H to
        [ceil(n/2), n) of the same buffer, until n < 2."""
        a = np.ascontiguousarray(x).copy()
        n = a.shape[0]
        while n >= 2:
            lo, hi = self.f_line(a[:n], wavelet)
            a[: lo.shape[0]] = lo
            a[lo.shape[0]: n] = hi
            n = lo.shape[0]
        return a

    def i_full_1d(self, c: np.ndarray, wavelet: str = "cdf97") -> np.ndarray:
        a = np.ascontiguousarray(c).copy()
        N = a.shape[0]
        sizes = []
        n = N
        while n >= 2:
            sizes.append(n)
            n = (n + 1) // 2
        for n in reversed(sizes):
            nl = (n + 1) // 2
            a[:n] = self.i_line(a[:nl], a[nl:n], wavelet)
        return a


class Oracle(DwtLib):
    def __init__(self):
        super().__init__(ORACLE_SO, "orc_")
        self.lib.orc_set_threads.argtypes = [C.c_int]
        self.lib.orc_get_threads.restype = C.c_int

    def set_threads(self, n: int) -> None:
        self.lib.orc_set_threads(n)

    def get_threads(self) -> int:
        return self.lib.orc_get_threads()

    def constants_s(self) -> np.ndarray:
        out = np.zeros(10, dtype=np.float32)
        self.lib.orc_constants_s(out.ctypes.data_as(C.c_void_p))
        return out


class Reference(DwtLib):
    """The reference compiled from its own sources (oracle/_ref)."""

    def __init__(self):
        super().__init__(REF_SO, "dwt_")
        self.lib.dwt_util_set_accel.argtypes = [C.c_int]
        self.lib.dwt_util_set_num_threads.argtypes = [C.c_int]

    def set_accel(self, a: int) -> None:
        self.lib.dwt_util_set_accel(a)

    def set_threads(self, n: int) -> None:
        self.lib.dwt_util_set_num_threads(n)


def have_ref() -> bool:
    return os.path.exists(REF_SO)


_oracle = None


def oracle() -> Oracle:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _oracle = Oracle()
    return _oracle
