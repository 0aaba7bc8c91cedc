"""CPU: the work decomposition of the fused sparse-frame level kernels,
replayed with the oracle's 1D lines, equals the oracle's 2D sparse transform
bit for bit (garbage rows and columns included).

A sparse frame (outer Wo x Ho, inner Wi x Hi at a level; src/libdwt.c:
6331-6400 forward, 6683-6747 inverse) transforms rows y < Ho over the inner
length with H at the outer half, then columns x < Wo likewise.  The GPU level
kernel splits one level into independent parts that read the level's input
and write a separate output buffer Q, so that no part depends on another's
writes:
  forward  (A) inner tiles: rows y < Hi and their columns, the four subbands
               at the outer frame's Mallat positions;
           (C) rows in [Hl, Hod) u [Hod + Hh, Ho): their row transform alone;
           (B) garbage columns [Wl, Wod) u [Wod + Wh, Wo): their column
               transform of the input (rows the columns do not write keep
               the input: part D);
  inverse  (A) inner tiles: output [0, Hi) x [0, Wi);
           (B) columns [Wi, Wo): their column inverse of the input;
           (C) rows [Hi, Ho): their row inverse;
           (D) [Hi, Ho) x [Wi, Wo): the input unchanged;
  zero padding: parts B, C, D write zeros instead.
The replay below performs exactly these parts and checks the whole frame.
"""
import numpy as np
import pytest

from oracle.oracle import oracle


def cdp2(i, j):
    return (i + (1 << j) - 1) >> j


def fwd_level(o, B, Wo, Ho, Wi, Hi, wv, zp):
    """One forward level of the region [0, Ho) x [0, Wo) of B, by parts."""
    Q = np.full((Ho, Wo), np.nan, dtype=B.dtype)
    Wl, Wh, Hl, Hh = (Wi + 1) // 2, Wi // 2, (Hi + 1) // 2, Hi // 2
    Wod, Hod = (Wo + 1) // 2, (Ho + 1) // 2
    # (A) inner rows, then the columns of their L and H parts
    R = np.zeros((Hi, Wi), dtype=B.dtype)
    for y in range(Hi):
        L, H = o.f_line(B[y, :Wi], wv)
        R[y, :Wl], R[y, Wl:] = L, H
    for c in range(Wi):
        cl, ch = o.f_line(R[:, c], wv)
        x = c if c < Wl else Wod + (c - Wl)
        Q[:Hl, x], Q[Hod:Hod + Hh, x] = cl, ch
    rows_r = [y for y in range(Ho) if not (y < Hl or Hod <= y < Hod + Hh)]
    gx = [x for x in range(Wo) if not (x < Wl or Wod <= x < Wod + Wh)]
    # (C) the rows the column pass leaves: row transform only
    for y in rows_r:
        if zp:
            Q[y, :] = 0
            continue
        L, H = o.f_line(B[y, :Wi], wv)
        Q[y, :Wl], Q[y, Wod:Wod + Wh] = L, H
    # (B) garbage columns: column transform of the input; (D) rows left alone
    for x in gx:
        if zp:
            Q[:, x] = 0
            continue
        cl, ch = o.f_line(B[:Hi, x], wv)
        Q[:Hl, x], Q[Hod:Hod + Hh, x] = cl, ch
        Q[rows_r, x] = B[rows_r, x]
    assert not np.isnan(Q).any(), "every position of the level region is written by one part"
    return Q


def inv_level(o, B, Wo, Ho, Wi, Hi, wv, zp):
    """One inverse level (output region [0, Ho) x [0, Wo)) of B, by parts."""
    Q = np.full((Ho, Wo), np.nan, dtype=B.dtype)
    Wl, Hl = (Wi + 1) // 2, (Hi + 1) // 2
    Wod, Hod = (Wo + 1) // 2, (Ho + 1) // 2
    Wh, Hh = Wi // 2, Hi // 2
    # (A) inner: row inverses of the rows the column pass reads, then columns
    src_rows = list(range(Hl)) + list(range(Hod, Hod + Hh))
    R = {}
    for y in src_rows:
        R[y] = o.i_line(B[y, :Wl], B[y, Wod:Wod + Wh], wv)
    for x in range(Wi):
        col = o.i_line(np.array([R[y][x] for y in range(Hl)], dtype=B.dtype),
                       np.array([R[y][x] for y in range(Hod, Hod + Hh)], dtype=B.dtype), wv)
        Q[:Hi, x] = col
    # (B) columns past the inner width: column inverse of the input
    for x in range(Wi, Wo):
        Q[:Hi, x] = 0 if zp else o.i_line(B[:Hl, x], B[Hod:Hod + Hh, x], wv)
    # (C) rows past the inner height: row inverse; (D) the corner unchanged
    for y in range(Hi, Ho):
        if zp:
            Q[y, :] = 0
            continue
        Q[y, :Wi] = o.i_line(B[y, :Wl], B[y, Wod:Wod + Wh], wv)
        Q[y, Wi:] = B[y, Wi:Wo]
    assert not np.isnan(Q).any()
    return Q


CASES = [((16, 16), (11, 9)), ((32, 16), (17, 16)), ((64, 32), (37, 23)), ((16, 64), (16, 33)),
         ((32, 32), (3, 30)), ((40, 24), (40, 19)), ((64, 64), (47, 61)), ((128, 32), (94, 5))]


@pytest.mark.parametrize("zp", [0, 1])
@pytest.mark.parametrize("wv", ["cdf97", "cdf53"])
@pytest.mark.parametrize("outer,inner", CASES)
def test_sparse_parts_equal_oracle(outer, inner, wv, zp):
    o = oracle()
    (Wo, Ho), (Wi, Hi) = outer, inner
    rng = np.random.default_rng(Wo * 131 + Hi)
    img = rng.standard_normal((Ho, Wo)).astype(np.float32)  # garbage outside the inner frame too
    J = min(int(np.ceil(np.log2(min(Wo, Ho)))), 6)
    # levels while the inner sides stay >= 2 (the fused kernels' range)
    while J > 0 and (cdp2(Wi, J - 1) < 2 or cdp2(Hi, J - 1) < 2):
        J -= 1
    ref, jr = o.fwd2(img, wv, J, zp=zp, inner=(Wi, Hi))
    assert jr == J
    a = img.copy()
    for j in range(J):
        Wos, Hos, Wis, His = cdp2(Wo, j), cdp2(Ho, j), cdp2(Wi, j), cdp2(Hi, j)
        a[:Hos, :Wos] = fwd_level(o, a, Wos, Hos, Wis, His, wv, zp)
    np.testing.assert_array_equal(a, ref)
    back = o.inv2(ref, wv, J, zp=zp, inner=(Wi, Hi))
    b = ref.copy()
    for j in range(J, 0, -1):
        Wod, Hod, Wid, Hid = cdp2(Wo, j - 1), cdp2(Ho, j - 1), cdp2(Wi, j - 1), cdp2(Hi, j - 1)
        b[:Hod, :Wod] = inv_level(o, b, Wod, Hod, Wid, Hid, wv, zp)
    np.testing.assert_array_equal(b, back)
