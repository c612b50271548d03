"""Pins the oracle's window extension, merge and split (oracle/p7f.c
ixo_extend_and_merge, ixo_split_windows; cm_pipeline.rs:722-837) with cases
whose answers are worked out by hand below each case, not recomputed by
the formula.

Extension: max_length 50, so a window starts at n - 50 (0.1 + prefix[k -
length + 1]) and ends at n + length + 50 (0.1 + suffix[k]), the sum taken
in double and truncated only at the end. With prefix = 0.4f (0.4000000060
as a float) the start offset is 25.0000003: a window at n = 100 starts at
74.9999997 -> 74, where truncating the width first would give 75. With
prefix = suffix = -0.1f (-0.1000000015) the offsets are -7.45e-8, so a raw
window (n, length) becomes exactly [n, n + length - 1]; the merge cases use
that to state their inputs as intervals."""
import ctypes as C

import numpy as np

from oracle_bind import Msv, RawWin, Window, lib

M = 40


def _msv(prefix, suffix):
    f = Msv()
    f.m, f.max_length = M, 50
    pre = np.ascontiguousarray(prefix, np.float32)
    suf = np.ascontiguousarray(suffix, np.float32)
    f.prefix = pre.ctypes.data_as(C.POINTER(C.c_float))
    f.suffix = suf.ctypes.data_as(C.POINTER(C.c_float))
    return f, (pre, suf)


def extend_merge(raws, prefix, suffix, target_len, pct=0.0):
    f, keep = _msv(prefix, suffix)
    L = lib()
    L.ixo_extend_and_merge.argtypes = [C.POINTER(Msv), C.POINTER(RawWin), C.c_int64, C.c_int64, C.c_float,
                                       C.POINTER(Window)]
    L.ixo_extend_and_merge.restype = C.c_int64
    rw = (RawWin * len(raws))(*[RawWin(n, ln, k, sc) for n, ln, k, sc in raws])
    out = (Window * len(raws))()
    n = L.ixo_extend_and_merge(C.byref(f), rw, len(raws), target_len, pct, out)
    return [(w.start, w.end, w.score) for w in out[:n]]


def split(wins, cmw):
    L = lib()
    L.ixo_split_windows.argtypes = [C.POINTER(Window), C.c_int64, C.c_int64, C.POINTER(Window), C.c_int64]
    L.ixo_split_windows.restype = C.c_int64
    w = (Window * len(wins))(*[Window(s, e, sc) for s, e, sc in wins])
    n = L.ixo_split_windows(w, len(wins), cmw, None, 0)
    out = (Window * max(n, 1))()
    L.ixo_split_windows(w, len(wins), cmw, out, n)
    return [(x.start, x.end) for x in out[:n]]


def _flat(v):
    return np.full(M + 1, v, np.float32)


def test_extension_double_then_truncate():
    pre = _flat(0.0)
    suf = _flat(0.0)
    pre[21] = 0.4   # k - length + 1 = 30 - 10 + 1
    suf[30] = 0.9   # 0.8999999762 as a float: 110 + 49.9999988 = 159.9999988 -> 159
    assert extend_merge([(100, 10, 30, 1.0)], pre, suf, 1000) == [(74, 159, 1.0)]
    # clamped to 1 and to the target's end: 5 - 25.0000003 < 1, and
    # 5 + 10 + 49.9999988 = 64.9999988 > 50
    assert extend_merge([(5, 10, 30, 1.0)], pre, suf, 50) == [(1, 50, 1.0)]
    # k - length + 1 below 0 reads prefix[0]; k above M reads suffix[M]
    pre2, suf2 = _flat(0.0), _flat(0.0)
    pre2[0], suf2[M] = 0.4, 0.9
    # n = 200, length 10, k = 3: start 200 - 25.0000003 -> 174; k = 3 reads
    # suffix[3] = 0: end 210 + 5 = 215 (0.1 * 50 = 5.000000000000001 -> 215)
    assert extend_merge([(200, 10, 3, 2.0)], pre2, suf2, 1000) == [(174, 215, 2.0)]
    # n = 300, length 5, k = 45: prefix[40] = 0: start 300 - 5.000000000000001,
    # which is 295.0 in double -> 295; k above M reads suffix[M] = 0.9f: end
    # 305 + 49.9999988 -> 354
    assert extend_merge([(300, 5, 45, 3.0)], pre2, suf2, 1000) == [(295, 354, 3.0)]


def test_merge_cases():
    z = _flat(-0.1)

    def merge(ivals, pct=0.0):
        raws = [(s, e - s + 1, 20, float(q)) for q, (s, e) in enumerate(ivals)]
        return [(s, e, int(sc)) for s, e, sc in extend_merge(raws, z, z, 10_000, pct)]

    # the identity extension
    assert merge([(10, 20)]) == [(10, 20, 0)]
    # one shared residue merges; touching end to start does not
    assert merge([(10, 20), (20, 30)]) == [(10, 30, 0)]
    assert merge([(10, 20), (21, 30)]) == [(10, 20, 0), (21, 30, 1)]
    # each window is compared with the last kept (already grown) one
    assert merge([(10, 20), (15, 25), (24, 40)]) == [(10, 40, 0)]
    # input order is kept: a later window may start earlier
    assert merge([(50, 60), (10, 55)]) == [(10, 60, 0)]
    # a window inside the last kept one disappears into it
    assert merge([(10, 20), (30, 40), (35, 36)]) == [(10, 20, 0), (30, 40, 1)]
    # overlap fraction over the shorter window, strictly above pct:
    # [1,10] + [6,20]: 5 / 10 = 0.5 -> kept apart; [1,10] + [5,20]: 0.6 -> merged
    assert merge([(1, 10), (6, 20)], 0.5) == [(1, 10, 0), (6, 20, 1)]
    assert merge([(1, 10), (5, 20)], 0.5) == [(1, 20, 0)]


def test_split_cases():
    # cmW = 10: windows longer than 20 become 20-long pieces stepping by 10,
    # the last clipped to the window's end
    assert split([(1, 20, 0.0)], 10) == [(1, 20)]
    assert split([(1, 21, 0.0)], 10) == [(1, 20), (11, 21)]
    assert split([(3, 24, 0.0)], 10) == [(3, 22), (13, 24)]
    assert split([(5, 45, 0.0)], 10) == [(5, 24), (15, 34), (25, 44), (35, 45)]
    assert split([(1, 20, 0.0), (100, 130, 0.0)], 10) == [(1, 20), (100, 119), (110, 129), (120, 130)]
