"""Pins the oracle's allow_truncation (oracle/trcm.c, following
cm_dpsearch_trunc.rs:2376-2422, cm_hit_AllowTruncation) against its truth
table, written out below as data: for every pass, local and global, every
mode, whether the hit starts at the envelope's i0 and stops at its j0, and
four kinds of entry state (its subtree spanning the whole consensus, only
its first column, only its last, or neither; spans from the guide tree,
synth_cm.state_spans)."""
import ctypes as C

import numpy as np

from infernox_amd import synth_cm as S
from oracle_cm_bind import OracleTrCM, lib

STD_ANY, P5_FORCE, P3_FORCE, P53_FORCE, P53_ANY = 1, 2, 3, 4, 5
T, R, L, J, UNKNOWN = 0, 1, 2, 3, 4

# For the forced passes (5' only, 3' only, 5' and 3'), one row per
# (local, mode): 16 characters, '1' = reported, over
#   (starts at i0, stops at j0) in (F F, F T, T F, T T)  [outer]
#   x entry span in (whole, first column only, last column only, neither)  [inner]
# A hit covering the whole envelope (T T) is always reported; otherwise a
# local hit in J is, one in L only when it reaches j0, one in R only when it
# starts at i0; globally the entry state must also reach the consensus ends
# that the truncation leaves in place.
FORCED = {
    (1, J): "1111" "1111" "1111" "1111",
    (1, L): "0000" "1111" "0000" "1111",
    (1, R): "0000" "0000" "1111" "1111",
    (1, T): "0000" "0000" "0000" "1111",
    (1, UNKNOWN): "0000" "0000" "0000" "1111",
    (0, J): "1000" "1000" "1000" "1111",
    (0, L): "0000" "1100" "0000" "1111",
    (0, R): "0000" "0000" "1010" "1111",
    (0, T): "0000" "0000" "0000" "1111",
    (0, UNKNOWN): "0000" "0000" "0000" "1111",
}


def test_allow_truncation_truth_table():
    # three hairpins off a root bifurcation chain: entry states of every kind
    cm = S.build_cm("<<..>>.<<..>>.<<...>>", 4)
    clen = int(cm["clen"])
    o = OracleTrCM(cm)
    lpos, rpos = o.emit_map()
    lp, rp = S.state_spans(cm)
    kinds = {}
    for v in range(cm["M"]):
        kind = (0 if lp[v] == 1 and rp[v] == clen else 1 if lp[v] == 1 else 2 if rp[v] == clen else 3)
        kinds.setdefault(kind, v)
    assert sorted(kinds) == [0, 1, 2, 3], kinds  # the model has every kind of entry state
    f = lib().ixo_tr_allow
    i32p = C.POINTER(C.c_int32)
    f.argtypes = [C.c_void_p, i32p, i32p] + [C.c_int] * 8
    f.restype = C.c_int
    i0, j0 = 10, 90
    for pass_idx in (STD_ANY, P5_FORCE, P3_FORCE, P53_FORCE, P53_ANY):
        for local in (0, 1):
            for mode in (T, R, L, J, UNKNOWN):
                for q, (at_i0, at_j0) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
                    start, stop = (i0 if at_i0 else i0 + 3), (j0 if at_j0 else j0 - 4)
                    for kind in range(4):
                        got = f(C.addressof(o.s), lpos.ctypes.data_as(i32p), rpos.ctypes.data_as(i32p), pass_idx,
                                local, start, stop, i0, j0, mode, kinds[kind])
                        want = 1 if pass_idx in (STD_ANY, P53_ANY) else int(FORCED[(local, mode)][4 * q + kind])
                        assert got == want, (pass_idx, local, mode, at_i0, at_j0, kind)
