"""CPU model of the block kernel's line-parallel lifting, checked bit for bit
against the oracle's 1D lines.

The single-workgroup block kernel (libdwt_amd/csrc/block_kernels.hip) lifts a
line with lane l holding PPL consecutive sample pairs, the neighbour pair
across a lane boundary taken by a DPP wave shift (lane 63 receives zero), the
line ends by substituting the lane's own or its neighbour's value, and several
short lines packed into one wavefront.  This restates exactly that lane
arithmetic in float32 numpy (same operands, same order) and runs it for every
line length the kernel accepts (2..512 samples), both wavelets, both
directions, with 1..64/LPL lines per wavefront: a wrong boundary rule, a
neighbour leaking across packed lines or a dropped lifting term fails here.
"""
import numpy as np
import pytest

from oracle.oracle import oracle

f32 = np.float32
P1, U1, P2, U2 = f32(1.58613434342059), f32(-0.0529801185729), f32(-0.8829110755309), f32(0.4435068520439)
K = f32(1.1496043988602)
INVK = f32(1) / K
S1, S2 = f32(1.41421356237309504880), f32(0.70710678118654752440)


def coeffs(w, fwd):
    """Lifting constants in the kernel's Coeffs order (src/libdwt.c:5542, 5811)."""
    if w == "cdf97":
        return ([-P1, U1, -P2, U2], K, INVK) if fwd else ([-U2, P2, -U1, P1], INVK, K)
    return ([f32(-0.5), f32(0.25)], S1, S2) if fwd else ([f32(-0.25), f32(0.5)], S2, S1)


def shl(v):  # DPP wave_shl:1 with bound_ctrl: lane l <- lane l+1, lane 63 <- 0
    r = np.zeros_like(v)
    r[:-1] = v[1:]
    return r


def shr(v):  # wave_shr:1: lane l <- lane l-1, lane 0 <- 0
    r = np.zeros_like(v)
    r[1:] = v[:-1]
    return r


def wave_pass(lines, n, w, fwd, mirror_right=True):
    """One wavefront lifting len(lines) lines of n samples together
    (mirror_right=False drops the right-end mirror of the odd step: a planted
    mistake for the sensitivity check)."""
    nl, nh = (n + 1) // 2, n // 2
    ppl = 1 if nl <= 64 else 2 if nl <= 128 else 4
    lpl = 64
    if ppl == 1:
        lpl = 1
        while lpl < nl:
            lpl <<= 1
    lane = np.arange(64)
    sub = lane // lpl
    j = ((lane - sub * lpl) * ppl)[:, None] + np.arange(ppl)[None, :]
    je, jo = np.minimum(j, nl - 1), np.minimum(j, nh - 1)
    src = np.stack([lines[min(s, len(lines) - 1)] for s in sub])  # lanes past the last line re-read one
    rows = np.arange(64)[:, None]
    if fwd:
        e, o = src[rows, 2 * je], src[rows, 2 * jo + 1]
    else:
        e, o = src[rows, je], src[rows, nl + jo]
    e, o = e.astype(f32), o.astype(f32)
    c, se, so = coeffs(w, fwd)

    def odd(cs, e, o):
        en = np.empty_like(e)
        en[:, :-1] = e[:, 1:]
        en[:, -1] = shl(e[:, 0])
        nb = np.where(2 * j + 2 == n, e, en) if mirror_right else en
        return (o + (cs * (e + nb).astype(f32)).astype(f32)).astype(f32)

    def even(cs, e, o):
        op = np.empty_like(o)
        op[:, 0] = shr(o[:, -1])
        op[:, 1:] = o[:, :-1]
        left = np.where(j == 0, o, op)
        right = np.where(2 * j + 1 == n, left, o)
        return (e + (cs * (left + right).astype(f32)).astype(f32)).astype(f32)

    if fwd:
        for s in range(0, len(c), 2):
            o = odd(c[s], e, o)
            e = even(c[s + 1], e, o)
        e, o = (e * se).astype(f32), (o * so).astype(f32)
    else:
        e, o = (e * se).astype(f32), (o * so).astype(f32)
        for s in range(0, len(c), 2):
            e = even(c[s], e, o)
            o = odd(c[s + 1], e, o)
    out = []
    for li in range(len(lines)):
        y = np.full(n, np.nan, f32)
        m = sub == li
        ev, od = j[m] < nl, j[m] < nh
        if fwd:
            y[j[m][ev]] = e[m][ev]
            y[nl + j[m][od]] = o[m][od]
        else:
            y[2 * j[m][ev]] = e[m][ev]
            y[2 * j[m][od] + 1] = o[m][od]
        out.append(y)
    return out, 64 // lpl


LENGTHS = list(range(2, 134)) + [191, 200, 255, 256, 257, 300, 383, 384, 385, 511, 512]


@pytest.mark.parametrize("w", ["cdf97", "cdf53"])
def test_lane_lifting_matches_oracle_lines(w):
    o = oracle()
    rng = np.random.default_rng(7)
    for n in LENGTHS:
        _, lpw = wave_pass([np.zeros(n, f32)], n, w, True)
        for count in sorted({1, max(1, lpw - 1), lpw}):
            xs = [rng.standard_normal(n).astype(f32) for _ in range(count)]
            got, _ = wave_pass(xs, n, w, True)
            for x, g in zip(xs, got):
                lo, hi = o.f_line(x, w)
                np.testing.assert_array_equal(g, np.concatenate([lo, hi]), err_msg=f"fwd n={n} lines={count}")
            ys = [rng.standard_normal(n).astype(f32) for _ in range(count)]
            got, _ = wave_pass(ys, n, w, False)
            nl = (n + 1) // 2
            for y, g in zip(ys, got):
                np.testing.assert_array_equal(g, o.i_line(y[:nl], y[nl:], w), err_msg=f"inv n={n} lines={count}")


def test_lane_model_catches_a_wrong_boundary():
    """The comparison is sensitive: the same model with the right-end mirror
    dropped (the neighbour of the last odd sample read from the next lane)
    differs from the oracle on even-length lines and agrees on odd ones, where
    that mirror is never used."""
    o = oracle()
    rng = np.random.default_rng(3)
    for n in (6, 64, 128, 256):
        x = rng.standard_normal(n).astype(f32)
        lo, hi = o.f_line(x, "cdf97")
        got, _ = wave_pass([x], n, "cdf97", True, mirror_right=False)
        assert not np.array_equal(got[0], np.concatenate([lo, hi])), n
    for n in (7, 65, 129):
        x = rng.standard_normal(n).astype(f32)
        lo, hi = o.f_line(x, "cdf97")
        got, _ = wave_pass([x], n, "cdf97", True, mirror_right=False)
        np.testing.assert_array_equal(got[0], np.concatenate([lo, hi]))
