"""CPU replay of the staged in-place level 0 for 1D signals stored as matrix
columns (dwt_gpu_1d_batch), through the library's own plan (col_stages) and
the column kernels' own row arithmetic (col_load_row), exported for tests.

Level 0 writes its H rows over input rows of the same matrix.  The plan runs
segment ranges in an order where every input row (forward) or H row
(inverse) is read before any launch overwrites it; the top segments' H rows
go through scratch.  The replay walks the launches in that order and fails
if a launch reads a row that an earlier launch, or the launch itself (its
workgroups run in any order), has overwritten: every n from 2 to 3000 and
the benchmark sizes, both lifting-step counts, both directions.  A planted
mistake in the plan (one stage boundary one segment too far) must fail it.
"""
import ctypes as C

import pytest

import libdwt_amd

K = 3  # staged ranges before the top segments


def lib():
    L = libdwt_amd.lib()
    i, p = C.c_int, C.POINTER(C.c_int)
    L.dwt_gpu_col_plan.argtypes = [i, i, i, p, p, p, p]
    L.dwt_gpu_col_segment_rows.argtypes = [i, i, i, i, p, i, p, p, i, p]
    return L


def plan(L, fwd, n, ns):
    st = (C.c_int * (K + 1))()
    hs, nseg, rows = C.c_int(), C.c_int(), C.c_int()
    staged = L.dwt_gpu_col_plan(int(fwd), n, ns, st, C.byref(hs), C.byref(nseg), C.byref(rows))
    assert staged in (0, 1)
    return bool(staged), list(st), hs.value, nseg.value, rows.value


def access(L, fwd, n, ns, seg):
    r, w = (C.c_int * 256)(), (C.c_int * 256)()
    nr, nw = C.c_int(), C.c_int()
    assert L.dwt_gpu_col_segment_rows(int(fwd), n, ns, seg, r, 256, C.byref(nr), w, 256, C.byref(nw)) == 0
    return set(r[: nr.value]), set(w[: nw.value])


def replay(L, fwd, n, ns, stages, hs, nseg, acc):
    """Returns None when no launch reads a row already overwritten, else the
    offending (launch, row).  Rows are the matrix rows of the signal:
    forward H row m lives at h0 + m; inverse reads H rows m at h0 + m (those
    at m >= hs from the scratch copy taken before anything is written)."""
    h0 = (n + 1) // 2
    if fwd:
        # the top segments first (their H goes to scratch), then the staged
        # ranges from the highest down, H straight into the matrix
        launches = [(range(stages[K], nseg), False)] + [(range(stages[k], stages[k + 1]), True)
                                                         for k in reversed(range(K))]
    else:
        # staged ranges from the lowest up, then the top segments (H from scratch)
        launches = [(range(stages[k], stages[k + 1]), True) for k in range(K)] + [(range(stages[K], nseg), False)]
    written = set()
    for li, (segs, in_place) in enumerate(launches):
        reads, writes = set(), set()
        for s in segs:
            r, w = acc(s)
            if fwd:
                reads |= r
                if in_place:
                    writes |= {h0 + m for m in w}
            else:
                reads |= {h0 + m for m in r if not (not in_place and m >= hs)}
                writes |= w
        bad = reads & (written | writes)
        if bad:
            return li, min(bad)
        written |= writes
    return None


SIZES = list(range(2, 3001)) + [4096, 8192, 16384]


@pytest.mark.parametrize("ns", [4, 2])
@pytest.mark.parametrize("fwd", [True, False])
def test_staged_level0_reads_before_overwrite(fwd, ns):
    L = lib()
    staged_count = 0
    for n in SIZES:
        staged, stages, hs, nseg, rows = plan(L, fwd, n, ns)
        if not staged:
            continue
        staged_count += 1
        assert stages[0] == 0 and all(stages[k] < stages[k + 1] for k in range(K)) and stages[K] < nseg
        cache = {}

        def acc(s):
            if s not in cache:
                cache[s] = access(L, fwd, n, ns, s)
            return cache[s]

        assert replay(L, fwd, n, ns, stages, hs, nseg, acc) is None, (n, stages, hs)
        if not fwd:
            # the top segments read no H row below hs from the matrix
            low = min(min(acc(s)[0]) for s in range(stages[K], nseg))
            assert low >= hs, (n, low, hs)
    assert staged_count > 2000  # the plan is in use for almost every size


@pytest.mark.parametrize("fwd", [True, False])
def test_replay_catches_a_stage_one_segment_too_far(fwd):
    """Sensitivity: widening one staged range by one segment (the inequality
    the plan relies on broken) must be reported."""
    L = lib()
    caught = 0
    for n in (512, 1000, 2048, 3000, 8192):
        staged, stages, hs, nseg, rows = plan(L, fwd, n, 4)
        assert staged
        bad = list(stages)
        k = 1 if fwd else 2
        bad[k] += 1  # forward: range [s1, s2) grows down into rows not yet read; inverse: [s1, s2) up
        cache = {}

        def acc(s):
            if s not in cache:
                cache[s] = access(L, fwd, n, 4, s)
            return cache[s]

        if replay(L, fwd, n, 4, bad, hs, nseg, acc) is not None:
            caught += 1
    assert caught == 5
