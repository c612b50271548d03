"""Pins the CM part of the CPU oracle: the banded scanners with bands that
cover every (i, j) must equal a textbook unbanded CYK / Inside over
(state, i, j) written from the grammar's semantics (emit left/right/pair,
transition to a child, bifurcate, local end into EL, local begin at the
root); Inside >= CYK; narrowing bands never raises CYK; hit reporting and the
score tables obey their closed forms."""
import math

import numpy as np
import pytest

from infernox_amd import synth_cm as S
from oracle_cm_bind import OracleCM

NEG = np.float32(-1e36)


def _log2sum(a, b):
    if a < b:
        a, b = b, a
    if b <= -1e30:
        return a
    return a + math.log2(1.0 + 2.0 ** (b - a))


def textbook(cm, o, dsq, i0, j0, inside):
    """alpha[v][(i, j)] for every subsequence i..j of i0..j0 (j = i-1 empty)."""
    M = cm["M"]
    st, cf, cn = cm["sttype"], cm["cfirst"], cm["cnum"]
    f32 = np.float32
    comb = (lambda a, b: _log2sum(a, b)) if inside else (lambda a, b: b if b > a else a)
    A = [dict() for _ in range(M)]
    for v in range(M - 1, -1, -1):
        s = int(st[v])
        sd = S.state_delta(s)
        sl, sr = S.state_left_delta(s), S.state_right_delta(s)
        oe = o.oesc[v]
        for j in range(i0 - 1, j0 + 1):
            for i in range(j + 1, i0 - 1, -1):  # d ascending: IL self child (i+1, j) first
                d = j - i + 1
                if s == S.E_ST:
                    if d == 0:
                        A[v][(i, j)] = 0.0 if inside else f32(0.0)
                    continue
                if d < sd:
                    continue
                val = NEG if not inside else -1e36
                if o.endsc[v] > -9.999e35:
                    val = f32(f32(o.el_selfsc) * f32(d - sd)) + f32(o.endsc[v])
                    if inside:
                        val = float(val)
                if s == S.B_ST:
                    y, z = int(cf[v]), int(cn[v])
                    for k in range(i - 1, j + 1):
                        a, b = A[y].get((i, k)), A[z].get((k + 1, j))
                        if a is None or b is None:
                            continue
                        val = comb(val, (a + b) if inside else f32(a + b))
                    A[v][(i, j)] = val
                    continue
                for yo in range(int(cn[v])):
                    y = int(cf[v]) + yo
                    c = A[y].get((i + sl, j - sr))
                    if c is None:
                        continue
                    t = o.tsc[v, yo]
                    val = comb(val, (c + float(t)) if inside else f32(c + t))
                if s in (S.ML_ST, S.IL_ST):
                    e = oe[dsq[i]]
                elif s in (S.MR_ST, S.IR_ST):
                    e = oe[dsq[j]]
                elif s == S.MP_ST:
                    e = oe[dsq[i] * 18 + dsq[j]]
                else:
                    e = None
                if e is not None:
                    val = (val + float(e)) if inside else f32(val + e)
                val = max(val, -1e36) if inside else (val if val >= NEG else NEG)
                A[v][(i, j)] = val
    # local begins into the root by MAX (cp9.rs:5264-5304)
    for key in list(A[0].keys()):
        best = A[0][key]
        for y in range(1, M):
            if o.beginsc[y] > -9.999e35 and key in A[y]:
                c = (A[y][key] + float(o.beginsc[y])) if inside else f32(A[y][key] + o.beginsc[y])
                if c > best:
                    best = c
        A[0][key] = best
    return A


def _root_cells(b, root):
    out = {}
    off = 0
    for j in range(b["jmin"][0], b["jmax"][0] + 1):
        jp = j - b["jmin"][0]
        lo, hi = b["hdmin"][b["hd_off"][0] + jp], b["hdmax"][b["hd_off"][0] + jp]
        for d in range(lo, hi + 1):
            out[(j - d + 1, j)] = root[off + d - lo]
        off += max(0, hi - lo + 1)
    return out


STRUCTS = ["<<..>>.", "<.>.<.>", ".<<...>.>.", "<<.>.<..>>"]


@pytest.mark.parametrize("ss,seed", [(s, k) for k, s in enumerate(STRUCTS)])
def test_hb_cyk_inside_equal_textbook_with_full_bands(ss, seed):
    cm = S.build_cm(ss, seed)
    o = OracleCM(cm)
    r = np.random.default_rng(seed)
    L = 9
    dsq = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
    i0, j0 = 1, L
    b = S.wide_bands(cm, i0, j0)
    for inside in (0, 1):
        got = o.scan(b, dsq, i0, j0, inside)
        want = textbook(cm, o, dsq, i0, j0, inside)
        cells = _root_cells(b, got["root"])
        assert len(cells) == len(want[0])
        for key, val in cells.items():
            w = want[0][key]
            if inside:
                assert abs(val - w) < 0.02 or (val <= -1e35 and w <= -1e35), (key, val, w)
            else:
                assert np.float32(val).tobytes() == np.float32(w).tobytes(), (key, val, w)
        assert got["vsc"] == max(cells.values())


@pytest.mark.parametrize("clen,seed", [(40, 1), (72, 2)])
def test_inside_ge_cyk_and_bands_only_lower_cyk(clen, seed):
    ss = S.random_nested(clen, seed)
    cm = S.build_cm(ss, seed)
    o = OracleCM(cm)
    seq = S.emit(cm, seed)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    wide = S.wide_bands(cm, 1, L)
    narrow = S.bands(cm, 1, L, r=4, seed=seed)
    cw = o.scan(wide, dsq, 1, L, 0)
    cn = o.scan(narrow, dsq, 1, L, 0)
    iw = o.scan(wide, dsq, 1, L, 1)
    assert iw["vsc"] >= cw["vsc"] - 1e-3
    assert cn["vsc"] <= cw["vsc"] + 1e-6
    assert cw["vsc"] > 0  # the emitted homolog scores positive


def test_score_tables_closed_forms():
    cm = S.build_cm(S.cloverleaf(), 1)
    o = OracleCM(cm)
    M = cm["M"]
    # local ends: endsc = log2(pend / nexits) at exit states, IMPOSSIBLE elsewhere
    ends = o.endsc[o.endsc > -1e30]
    assert len(ends) > 0 and np.allclose(ends, ends[0])
    nexits = len(ends)
    assert abs(2 ** float(ends[0]) - 0.05 / nexits) < 1e-6
    # local begins: state nodemap[1] gets 1 - pbegin
    assert abs(2 ** float(o.beginsc[cm["nodemap"][1]]) - 0.95) < 1e-6
    # transitions out of a non-exit state stay normalized
    for v in range(1, M):
        if cm["sttype"][v] in (S.B_ST, S.E_ST) or o.endsc[v] > -1e30:
            continue
        n = cm["cnum"][v]
        assert abs(sum(2.0 ** float(x) for x in o.tsc[v, :n]) - 1.0) < 1e-4
    # singlet emissions: oesc = log2(e / 0.25) for canonical residues
    for v in range(M):
        if cm["sttype"][v] in (S.ML_ST, S.MR_ST):
            for x in range(4):
                assert abs(o.oesc[v, x] - math.log2(cm["e"][v, x] / 0.25)) < 1e-5
            break


def test_report_hits_greedy_rules():
    cm = S.build_cm(S.cloverleaf(), 3)
    o = OracleCM(cm)
    seq = S.emit(cm, 5)
    r = np.random.default_rng(1)
    full = np.concatenate([r.integers(0, 4, 30), seq, r.integers(0, 4, 30)]).astype(np.uint8)
    L = len(full)
    dsq = np.concatenate([[255], full, [255]]).astype(np.uint8)
    b = S.bands(cm, 31, 30 + len(seq), r=10, seed=2)
    # clip bands into [1, L]
    res = o.scan(b, dsq, 1, L, 1)
    hits = o.hits(b, dsq, 1, L, res["root"], 0.0, True)
    assert len(hits) >= 1
    for i, j, sc, bias in hits:
        assert 1 <= i <= j <= L and sc >= 0.0 and bias >= 0.0
