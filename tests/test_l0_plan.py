"""CPU checks of the in-place level-0 plan (dwt_gpu_level0_plan), no GPU.

The plan orders the level-0 tiles so that a tile stores into the image only
after every other tile that reads those pixels has finished its reads.  Here
the pixels each tile reads and writes are enumerated lane by lane from the
level kernels' index arithmetic (the per-level tile bodies' loads and stores,
written out again in numpy, independent of the plan's interval code), and:
  * every (writer, reader) pair on an in-place pixel is a dependency of the
    writer, and every dependency belongs to an earlier segment of the order;
  * the parts that fall back to scratch are exactly the rows the copy carries;
  * the XCD lists partition the tiles, in the order, by bands of segments.
"""
import ctypes

import numpy as np
import pytest

import libdwt_amd

WAVE = 64
WAVES = 4


def plan(fwd, W, H, ns, S):
    L = libdwt_amd.lib()
    info = (ctypes.c_int * 8)()
    wid = libdwt_amd.WAVELET_IDS["cdf97" if ns == 4 else "cdf53"]
    assert L.dwt_gpu_level0_plan(int(fwd), W, H, wid, S, info, *([None] * 7)) == 0
    nseg, ng, nstrips, ntiles, ndeps, nrows, grid, band = list(info)
    arr = lambda n: (ctypes.c_int * max(n, 1))()
    order, segflags, list_off, lst = arr(nseg), arr(nseg), arr(9), arr(ntiles)
    dep_off, deps, rows = arr(ntiles + 1), arr(ndeps), arr(nrows)
    assert L.dwt_gpu_level0_plan(int(fwd), W, H, wid, S, info, order, segflags, list_off, lst, dep_off, deps,
                                 rows) == 0
    a = lambda x, n: np.array(x[:n], dtype=np.int64)
    return dict(nseg=nseg, ng=ng, nstrips=nstrips, ntiles=ntiles, grid=grid, band=band,
                order=a(order, nseg), segflags=a(segflags, nseg), list_off=a(list_off, 9),
                list=a(lst, ntiles), dep_off=a(dep_off, ntiles + 1), deps=a(deps, ndeps), rows=a(rows, nrows))


def mirror(i, n):
    p = 2 * (n - 1)
    i = np.mod(i, p)
    return np.where(i < n, i, p - i)


def tile_access(fwd, W, H, ns, S, s, g, nstrips, segflags):
    """(reads, writes) of tile (s, g) as flat pixel indices; writes are the
    in-place ones (parts that fall back to scratch are left out)."""
    core = 2 * WAVE - 2 * ns
    lanes = np.arange(WAVE)
    Wd, Wh, Hd, Hh = (W + 1) // 2, W // 2, (H + 1) // 2, H // 2
    reads, writes = [], []
    for q in range(WAVES * g, min(WAVES * g + WAVES, nstrips)):
        Xe = q * core - ns + 2 * lanes
        core_l = (lanes >= ns // 2) & (lanes < WAVE - ns // 2)
        pbeg = ((s * S) >> 1) - ns // 2
        Ys = mirror(2 * pbeg + np.arange(2 * (S // 2 + ns)), H)
        if fwd:
            cols = np.concatenate([mirror(Xe, W), mirror(Xe + 1, W)])
            reads.append((Ys[:, None] * W + cols[None, :]).ravel())
            k = Xe >> 1
            m0, m1 = (s * S) >> 1, min((s * S + S) >> 1, Hd)
            for m in range(m0, m1):
                if not (segflags[s] & 1):  # HL rows in place
                    c = Wd + k[core_l & (k < Wh)]
                    writes.append(m * W + c)
                if not (segflags[s] & 2) and m < Hh:  # LH and HH rows in place
                    c = np.concatenate([k[core_l & (k < Wd)], Wd + k[core_l & (k < Wh)]])
                    writes.append((Hd + m) * W + c)
        else:
            ke, ko = mirror(Xe, W) >> 1, mirror(Xe + 1, W) >> 1
            for Yr in Ys:
                m = Yr >> 1
                if Yr & 1:
                    reads.append((Hd + m) * W + np.concatenate([ke, Wd + ko]))
                else:
                    reads.append(m * W + Wd + ko)
            if not (segflags[s] & 1):
                for r in range(s * S, min(s * S + S, H)):
                    c = np.concatenate([Xe[core_l & (Xe < W)], Xe[core_l & (Xe + 1 < W)] + 1])
                    writes.append(r * W + c)
    cat = lambda v: np.unique(np.concatenate(v)) if v else np.zeros(0, np.int64)
    return cat(reads), cat(writes)


def check_plan(fwd, W, H, ns, S):
    P = plan(fwd, W, H, ns, S)
    nseg, ng, ntiles = P["nseg"], P["ng"], P["ntiles"]
    assert nseg == -(-H // S) and ntiles == nseg * ng
    rank = np.empty(nseg, np.int64)
    rank[P["order"]] = np.arange(nseg)
    assert sorted(P["order"].tolist()) == list(range(nseg))
    # XCD lists: a partition of the tiles, each in the order, bands of segments
    seen = []
    for x in range(8):
        lst = P["list"][P["list_off"][x]:P["list_off"][x + 1]]
        segs = lst // ng
        assert np.all((segs // P["band"]) % 8 == x)
        assert np.all(np.diff(rank[segs]) >= 0)
        seen.extend(lst.tolist())
    assert sorted(seen) == list(range(ntiles))
    assert P["grid"] == 8 * max(P["list_off"][x + 1] - P["list_off"][x] for x in range(8))
    # every dependency is an earlier segment's tile
    dep_of = [set(P["deps"][P["dep_off"][t]:P["dep_off"][t + 1]].tolist()) for t in range(ntiles)]
    for t in range(ntiles):
        for d in dep_of[t]:
            assert rank[d // ng] < rank[t // ng], (t, d)
    # pixel-level hazards: readers of every in-place write are dependencies
    rd_pix, rd_tile, wr_pix, wr_tile = [], [], [], []
    for t in range(ntiles):
        r, w = tile_access(fwd, W, H, ns, S, t // ng, t % ng, P["nstrips"], P["segflags"])
        rd_pix.append(r)
        rd_tile.append(np.full(r.size, t))
        wr_pix.append(w)
        wr_tile.append(np.full(w.size, t))
    rp, rt = np.concatenate(rd_pix), np.concatenate(rd_tile)
    wp, wt = np.concatenate(wr_pix), np.concatenate(wr_tile)
    if not fwd:
        # a band preserved in scratch is read from there, not from the image
        pres = np.flatnonzero(P["segflags"] & 1)
        keep = ~np.isin((rp // W) // S, pres)
        rp, rt = rp[keep], rt[keep]
    o = np.argsort(rp, kind="stable")
    rp, rt = rp[o], rt[o]
    lo, hi = np.searchsorted(rp, wp, "left"), np.searchsorted(rp, wp, "right")
    n = hi - lo
    wi = np.repeat(np.arange(wp.size), n)
    ri = np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]) if wp.size else np.zeros(0, np.int64)
    writer, reader = wt[wi], rt[ri]
    other = writer != reader
    pairs = np.unique(writer[other] * ntiles + reader[other])
    allowed = np.array(sorted(t * ntiles + d for t in range(ntiles) for d in dep_of[t]), dtype=np.int64)
    missing = pairs[~np.isin(pairs, allowed)]
    assert missing.size == 0, [(int(p // ntiles), int(p % ntiles)) for p in missing[:10]]
    # a tile's own segment never needs waiting on (that part falls back)
    for t in range(ntiles):
        for d in dep_of[t]:
            assert d // ng != t // ng
    # the copy carries exactly the rows that fall back (forward) or are
    # preserved (inverse)
    Hd, Hh = (H + 1) // 2, H // 2
    want = set()
    for s in range(nseg):
        f = int(P["segflags"][s]) & 0xff
        if fwd:
            m0, m1 = (s * S) >> 1, min((s * S + S) >> 1, Hd)
            if f & 1:
                want.update(range(m0, m1))
            if f & 2:
                want.update(Hd + m for m in range(m0, m1) if m < Hh)
        elif f & 1:
            want.update(range(s * S, min(s * S + S, H)))
    assert sorted(want) == P["rows"].tolist()
    if not fwd:
        # read-row mask: bit 8 + u set exactly when single-shot row u of the
        # segment reads its details from a preserved band
        pres = P["segflags"] & 1
        for s in range(nseg):
            pbeg = ((s * S) >> 1) - ns // 2
            for u in range(2 * (S // 2 + ns)):
                Yr = int(mirror(2 * pbeg + u, H))
                y = Hd + (Yr >> 1) if Yr & 1 else Yr >> 1
                assert bool((P["segflags"][s] >> (8 + u)) & 1) == bool(pres[y // S]), (s, u)
    return P


SHAPES = [(300, 260), (517, 389), (1000, 64), (64, 1000), (130, 71), (2, 2), (3, 17), (17, 3), (1031, 523),
          (961, 962)]


@pytest.mark.parametrize("fwd", [True, False])
@pytest.mark.parametrize("ns,S", [(4, 8), (4, 4), (2, 8), (2, 4)])
@pytest.mark.parametrize("W,H", SHAPES)
def test_plan_orders_every_hazard(fwd, ns, S, W, H):
    check_plan(fwd, W, H, ns, S)


@pytest.mark.parametrize("W,H", [(4096, 4096), (3001, 2047)])
def test_plan_scratch_share_at_workload_size(W, H):
    """Structure and the share of detail through scratch on the benchmark
    shapes (forward about 14%, inverse 21-24% of the detail)."""
    for fwd, bound in ((True, 0.16), (False, 0.26)):
        P = plan(fwd, W, H, 4, 8)
        rank = np.empty(P["nseg"], np.int64)
        rank[P["order"]] = np.arange(P["nseg"])
        ng = P["ng"]
        for t in range(P["ntiles"]):
            d = P["deps"][P["dep_off"][t]:P["dep_off"][t + 1]]
            assert np.all(rank[d // ng] < rank[t // ng])
        Wd, Hd = (W + 1) // 2, (H + 1) // 2
        rows = P["rows"]
        # bytes through scratch: right half of rows in the top half, whole rows below
        moved = np.where(rows < Hd, W - Wd, W).sum()
        share = moved / (W * H - Wd * Hd)
        assert share <= bound, (fwd, share)


def test_plan_rejects_bad_arguments():
    L = libdwt_amd.lib()
    info = (ctypes.c_int * 8)()
    assert L.dwt_gpu_level0_plan(1, 1, 64, 0, 8, info, *([None] * 7)) != 0
    assert L.dwt_gpu_level0_plan(1, 64, 64, 0, 6, info, *([None] * 7)) != 0
