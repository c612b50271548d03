"""Pins the oracle's CP9 Forward and Backward (oracle/cp9.c, following
cp9.rs:1275 and :1995) in search (scan) mode and in alignment mode by
summing over every path of the CP9 explicitly, on models of 4-6 columns
and windows of 3-4 residues. No recurrence is run: each Forward cell is
the log2-sum over all path prefixes that reach it, each Backward cell the
log2-sum over all path suffixes that leave it, in double precision.

The state graph (rows 0..L, row r > 0 consuming residue r): the begin
state M0 goes to M_k of the next row (t[k, BM]), to I0 of the next row, or
to D1 of its own row; M_k, I_k and D_k go to M_k+1 and I_k of the next row
and to D_k+1 of their own row; M_k enters its local end EL_k in its own
row, EL_k stays in EL_k for one more residue (el_selfsc, no emission) or
goes to the match states it feeds on the next row; the end state is
reached from every M_k (ME), from I_M and D_M, and from the ELs that feed
the end. Search mode lets a path begin at M0 of any row (cp9.rs:1358) and
end at any row; alignment mode only at row 0 and row L. In search mode
the Backward's ends at a row before L from M_k or I_M are added after the
state's emission score was (cp9.rs:2100-2104, 2130-2132), so those ends
carry no emission; the sums below state that as it is. States no path can
occupy at row 0 (the emitting states and the ELs) have no Backward value
there (-inf, as the reference sets them)."""
import numpy as np
import pytest

from infernox_amd import synth_cm as S
from infernox_amd import synth_cp9 as P
from oracle_cp9_bind import OracleCP9

MM, IM, DM, BM, MI, II, DI, MD, ID, DD, ME, MEL = range(12)
NEG = -np.inf


def _bits(a):
    a = np.asarray(a, np.float64)
    return np.where(a <= -P.INFTY, NEG, a / 1000.0)


class Paths:
    def __init__(self, h, dsq, L, scan):
        self.M, self.L, self.scan = h["M"], L, scan
        self.t, self.ms, self.isc = _bits(h["otsc"]), _bits(h["msc"]), _bits(h["isc"])
        self.el = bool(h["el"])
        self.els = h["el_selfsc"] / 1000.0
        ct, off, idx = h["el_from_ct"], h["el_from_off"], h["el_from_idx"]
        self.feeds = {}  # EL state y -> the match columns k (M+1: the end) it feeds
        for k in range(self.M + 2):
            for y in idx[off[k]:off[k] + ct[k]]:
                self.feeds.setdefault(int(y), []).append(k)
        self.has = h["has_el"]
        self.x = dsq  # residue of row r is dsq[r] (window 1..L)

    def exists(self, s):
        kind, k = s
        if kind == "D":
            return k >= 1
        if kind == "E":
            return self.el and k >= 1 and bool(self.has[k])
        return True

    def emis(self, s, r):
        kind, k = s
        if kind == "M" and k > 0:
            return self.ms[k, self.x[r]]
        if kind == "I":
            return self.isc[k, self.x[r]]
        return 0.0

    def moves(self, s, r):
        """(next state or 'END', next row, transition score) from s at row r."""
        M, t = self.M, self.t
        kind, k = s
        out = []
        nxt = r + 1 <= self.L
        can_end = self.scan or r == self.L
        if kind == "M" and k == 0:
            if nxt:
                out += [(("M", kk), r + 1, t[kk, BM]) for kk in range(1, M + 1)]
                out.append((("I", 0), r + 1, t[0, MI]))
            out.append((("D", 1), r, t[0, MD]))
            return out
        if kind == "E":
            if nxt:
                out.append((("E", k), r + 1, self.els))
                out += [(("M", kk), r + 1, 0.0) for kk in self.feeds.get(k, []) if kk <= M]
            if can_end and (M + 1) in self.feeds.get(k, []):
                out.append(("END", r, 0.0))
            return out
        row = {"M": (MM, MI, MD), "I": (IM, II, ID), "D": (DM, DI, DD)}[kind]
        if nxt:
            if k < M:
                out.append((("M", k + 1), r + 1, t[k, row[0]]))
            out.append((("I", k), r + 1, t[k, row[1]]))
        if k < M:
            out.append((("D", k + 1), r, t[k, row[2]]))
        if kind == "M" and self.el and self.has[k]:
            out.append((("E", k), r, t[k, MEL]))
        if can_end:
            if kind == "M":
                out.append(("END", r, t[k, ME]))
            elif k == M:
                out.append(("END", r, t[M, IM] if kind == "I" else t[M, DM]))
        return out

    def forward(self):
        F = {}

        def add(key, w):
            F[key] = np.logaddexp2(F.get(key, NEG), w)

        def walk(s, r, w):
            add((r, s), w)
            for s2, r2, tw in self.moves(s, r):
                if tw == NEG:
                    continue
                if s2 == "END":
                    add((r, "END"), w + tw)
                else:
                    walk(s2, r2, w + tw + self.emis(s2, r2))

        for r0 in (range(self.L + 1) if self.scan else [0]):
            walk(("M", 0), r0, 0.0)
        return F

    def backward(self, s, r):
        if not self.exists(s):
            return NEG
        if r == 0 and (s[0] in ("I", "E") or s[0] == "M" and s[1] > 0):
            return NEG  # no path occupies an emitting state or an EL before the first residue
        total = [NEG]

        def walk(s, r, w):
            for s2, r2, tw in self.moves(s, r):
                if tw == NEG:
                    continue
                if s2 == "END":
                    drop = self.scan and r < self.L and (s[0] == "M" and s[1] > 0 or s == ("I", self.M))
                    total[0] = np.logaddexp2(total[0], w + tw - (self.emis(s, r) if drop else 0.0))
                else:
                    walk(s2, r2, w + tw + self.emis(s2, r2))

        walk(s, r, self.emis(s, r))
        return total[0]


def _case(ss, seed, L, el):
    cm = S.build_cm(ss, seed)
    h = P.build_cp9(cm, seed, el=el)
    seq = np.random.default_rng(seed).integers(0, 4, L)
    return h, np.concatenate([[255], seq, [255]]).astype(np.uint8)


@pytest.mark.parametrize("ss,seed,L,el", [("<.>.", 1, 4, True), ("<<..>>", 2, 3, True), ("<<..>>", 3, 3, False),
                                          (".<.>", 4, 4, True)])
@pytest.mark.parametrize("scan", [True, False])
def test_cp9_cells_equal_path_sums(ss, seed, L, el, scan):
    h, dsq = _case(ss, seed, L, el)
    M = h["M"]
    o = OracleCP9(h)
    f, b = o.forward(dsq, 1, L, scan, not scan), o.backward(dsq, 1, L, scan, not scan)
    paths = Paths(h, dsq, L, scan)
    F = paths.forward()
    tol = 5e-4 * (L + M + 8)  # the oracle's 1/1000-bit ILogsum table, per log-sum on a chain
    live = 0
    for r in range(L + 1):
        for kind, name in (("M", "mmx"), ("I", "imx"), ("D", "dmx"), ("E", "elmx")):
            for k in range(M + 1):
                s = (kind, k)
                # a state the model does not have (D0, an EL without a local
                # end) is -inf on both sides
                for side, mat, want in (("F", f, F.get((r, s), NEG)), ("B", b, paths.backward(s, r))):
                    got = mat[name][r, k]
                    if want == NEG or want < -1e4:
                        assert got <= -P.INFTY // 2, (side, r, s, got)
                    else:
                        assert abs(got / 1000.0 - want) < tol, (side, r, s, got / 1000.0, want)
                        live += 1
    assert live > 20
