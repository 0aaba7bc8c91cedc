"""Case definitions shared by the golden-fixture generator and the tests.

A case is fully described by plain parameters and a seed; ``make_buffer``
rebuilds the exact input bytes on any host, so fixtures only need to store
digests of the outputs (plus full arrays for the smallest cases).
"""
from __future__ import annotations

import hashlib
import itertools

import numpy as np

SMALL_SIDES = (1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 33)
ARRAY_SIDES = (1, 2, 3, 4, 5, 8, 9, 17)
WAVELETS = ("cdf97", "cdf53")


def digest(buf: np.ndarray) -> str:
    """First 96 bits of the SHA-256 of the array's bytes, in hex (fixture
    inputs are finite, so the outputs carry no NaN payloads)."""
    a = np.ascontiguousarray(buf)
    return hashlib.sha256(a.tobytes()).hexdigest()[:24]


def layout(kind: str, Wo: int, Ho: int, item: int):
    """(stride_x, stride_y, nbytes) of an outer frame Wo x Ho in one layout.

    packed: rows of Wo elements; pitch: rows padded to a prime byte pitch
    (unaligned for odd values, as dwt_util_get_opt_stride produces);
    trans: element stride = one column of Ho elements (stride_x = item);
    chan: element stride of 3 items (interleaved 3-channel image)."""
    if kind == "packed":
        return Wo * item, item, Wo * Ho * item
    if kind == "pitch":
        p = Wo * item + 3 * item
        while not _is_prime(p):
            p += 1
        return p, item, p * Ho
    if kind == "trans":
        return item, Ho * item, Wo * Ho * item
    if kind == "chan":
        return 3 * Wo * item, 3 * item, 3 * Wo * Ho * item
    raise ValueError(kind)


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    for d in range(2, int(n ** 0.5) + 1):
        if n % d == 0:
            return False
    return True


def make_buffer(kind: str, Wo: int, Ho: int, dtype, seed: int):
    """Zero-filled byte buffer with N(0,1) values at every frame position;
    returns (buffer, stride_x, stride_y)."""
    item = np.dtype(dtype).itemsize
    sx, sy, nbytes = layout(kind, Wo, Ho, item)
    buf = np.zeros(nbytes + 16, dtype=np.uint8)
    vals = np.random.default_rng(seed).standard_normal((Ho, Wo)).astype(dtype)
    offs = (np.arange(Ho)[:, None] * sx + np.arange(Wo)[None, :] * sy)[..., None] \
        + np.arange(item)[None, None, :]
    buf[offs] = vals.view(np.uint8).reshape(Ho, Wo, item)
    return buf, sx, sy


def read_frame(buf: np.ndarray, kind: str, Wo: int, Ho: int, dtype) -> np.ndarray:
    """The (Ho, Wo) frame values of a buffer made by make_buffer."""
    item = np.dtype(dtype).itemsize
    sx, sy, _ = layout(kind, Wo, Ho, item)
    offs = (np.arange(Ho)[:, None] * sx + np.arange(Wo)[None, :] * sy)[..., None] \
        + np.arange(item)[None, None, :]
    return np.ascontiguousarray(buf[offs]).view(dtype).reshape(Ho, Wo)


def dense_cases():
    """Packed frames, inner == outer."""
    seed = 1000
    for w, prec, W, H, d1, j in itertools.product(
            WAVELETS, ("s", "d"), SMALL_SIDES, SMALL_SIDES, (0, 1), (-1, 1, 2)):
        seed += 1
        yield dict(wavelet=w, prec=prec, Wo=W, Ho=H, Wi=W, Hi=H, d1=d1, zp=0, j=j,
                   layout="packed", seed=seed)


SPARSE_FRAMES = ((8, 8, 5, 7), (16, 9, 3, 9), (17, 33, 16, 20), (33, 17, 1, 5),
                 (12, 12, 12, 1), (9, 16, 2, 2), (64, 32, 37, 29), (5, 5, 1, 1),
                 (40, 24, 40, 24))


def sparse_cases():
    """Sparse frames (outer > inner), zero padding, decompose_one and the four
    memory layouts."""
    seed = 50000
    for w, prec, fr, d1, zp, j, lay in itertools.product(
            WAVELETS, ("s", "d"), SPARSE_FRAMES, (0, 1), (0, 1), (-1, 1, 3),
            ("packed", "pitch", "trans", "chan")):
        seed += 1
        Wo, Ho, Wi, Hi = fr
        yield dict(wavelet=w, prec=prec, Wo=Wo, Ho=Ho, Wi=Wi, Hi=Hi, d1=d1, zp=zp,
                   j=j, layout=lay, seed=seed)


def case_key(c) -> str:
    return "{wavelet}_{prec}_{layout}_{Wo}x{Ho}_{Wi}x{Hi}_d{d1}_z{zp}_j{j}".format(**c)


# Workload images of the benchmark configurations (sinusoid + noise).
CONFIG_IMAGES = (
    dict(name="512x512_J5", W=512, H=512, J=5, seed=1),
    dict(name="3001x2047_J8", W=3001, H=2047, J=8, seed=2),
    dict(name="4096x4096_J8", W=4096, H=4096, J=8, seed=3),
)


def type0_image(Wo: int, Ho: int, dtype=np.float32) -> np.ndarray:
    """The reference's type-0 test image, 2xy / (x^2 + y^2 + 1), with the
    numerator converted from int and the division done in the storage type
    (src/libdwt.c:1076-1090, 1102-1116)."""
    x = np.arange(Wo, dtype=np.int64)[None, :]
    y = np.arange(Ho, dtype=np.int64)[:, None]
    num = (2 * x * y).astype(dtype)
    den = (x * x + y * y + 1).astype(dtype)
    return (num / den).astype(dtype)


# The reference example program's case (examples/simple/simple.c:15-19,
# 45-70): 512 x 512 type-0 image, row pitch = next prime >= 2048 bytes.
EXAMPLE_SIMPLE = dict(W=512, H=512, pitch=2053)

# 1D full-depth slice of the strided-line configuration (random walks).
LINE_SLICE = dict(count=256, length=8192, seed=4)
