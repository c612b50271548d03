"""Pins the truncated part of the CPU oracle (TrCYK / FTrInside over HMM
bands, the truncated reporting, and the model/band inputs):

- emit map: every state's consensus span equals its guide-tree node span;
- marginal emission scores: closed forms (log2 of row/column sums over the
  null, null-weighted means for degenerate codes);
- the J plane of every non-root state equals the standard banded scanner's
  deck (the truncated recursion restricted to J mode is the standard one);
- mirror symmetry: reversing the sequence and mirroring the model (ML<->MR,
  IL<->IR, pair emissions transposed, bifurcation children swapped, L<->R
  validity, 5'<->3' pass) maps the L plane onto the R plane exactly (CYK);
- closed forms on chain models: a 3'-truncated alignment of an ML chain and a
  5'-truncated one of an MR chain are sums of emissions and transitions from
  the entry state, maximized over entry states plus their penalty;
- FTrInside >= TrCYK cell by cell on every plane;
- reporting: hits are non-overlapping, sorted by score, above the cutoff.
"""
import numpy as np
import pytest

from infernox_amd import synth_cm as S
from oracle_cm_bind import OracleTrCM

NEG = -1e36


def _cm(seed=1, clen=80, nstems=3):
    return S.build_cm(S.multibranch(clen, nstems, seed), seed)


def test_emit_map_spans_equal_node_spans():
    for seed, clen, ns in ((1, 80, 3), (2, 120, 5), (3, 60, 1)):
        cm = _cm(seed, clen, ns)
        o = OracleTrCM(cm)
        lpos, rpos = o.emit_map()
        nt = cm["ndtype"][cm["ndidx"]]
        nd = cm["ndidx"]
        lp = np.where(np.isin(nt, (S.MATP_ND, S.MATL_ND)), lpos[nd], lpos[nd] + 1)
        rp = np.where(np.isin(nt, (S.MATP_ND, S.MATR_ND)), rpos[nd], rpos[nd] - 1)
        sl, sr = S.state_spans(cm)
        assert np.array_equal(lp, sl) and np.array_equal(rp, sr)


@pytest.mark.parametrize("pass_idx", [S.PASS_5P_ONLY, S.PASS_3P_ONLY, S.PASS_5P_AND_3P_FORCE, S.PASS_5P_AND_3P_ANY])
def test_validity_synth_matches_oracle(pass_idx):
    cm = _cm(4, 100, 4)
    o = OracleTrCM(cm)
    for sp1, sp2, ep1, ep2 in ((1, 30, 60, 90), (1, 1, 100, 100), (10, 50, 50, 95)):
        a = o.validity(pass_idx, sp1, sp2, ep1, ep2)
        b = S.tr_validity(cm, pass_idx, sp1, sp2, ep1, ep2)
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_marginals_closed_form():
    cm = _cm(5, 90, 3)
    o = OracleTrCM(cm)
    st, e, null = cm["sttype"], cm["e"].astype(np.float64), cm["null"].astype(np.float64)
    degen = np.array([[1, 0, 1, 0], [0, 1, 0, 1], [1, 1, 0, 0], [0, 0, 1, 1], [0, 1, 1, 0], [1, 0, 0, 1],
                      [1, 1, 0, 1], [0, 1, 1, 1], [1, 1, 1, 0], [1, 0, 1, 1], [1, 1, 1, 1]], np.float64)
    for v in range(cm["M"]):
        lm, rm = o.lmesc[v], o.rmesc[v]
        assert lm[4] == NEG and rm[4] == NEG and lm[16] == NEG and rm[17] == NEG
        if st[v] == S.MP_ST:
            p = e[v].reshape(4, 4)
            want_l, want_r = np.log2(p.sum(1) / null), np.log2(p.sum(0) / null)
        elif st[v] in (S.ML_ST, S.IL_ST):
            want_l, want_r = np.log2(e[v, :4] / null), np.zeros(4)
        elif st[v] in (S.MR_ST, S.IR_ST):
            want_l, want_r = np.zeros(4), np.log2(e[v, :4] / null)
        else:
            assert np.all(lm[:4] == 0) and np.all(rm[:4] == 0)
            continue
        np.testing.assert_allclose(lm[:4], want_l, atol=2e-5)
        np.testing.assert_allclose(rm[:4], want_r, atol=2e-5)
        np.testing.assert_allclose(lm[5:16], degen @ (want_l * null) / (degen @ null), atol=2e-5)
        np.testing.assert_allclose(rm[5:16], degen @ (want_r * null) / (degen @ null), atol=2e-5)


def _full_valid(M, seed=0, p=0.7):
    r = np.random.default_rng(seed)
    return {k: (r.random(M) < p).astype(np.uint8) for k in ("jvalid", "lvalid", "rvalid", "tvalid")}


@pytest.mark.parametrize("inside", [0, 1])
def test_j_plane_equals_standard_deck(inside):
    cm = _cm(6, 70, 3)
    o = OracleTrCM(cm)
    rng = np.random.default_rng(6)
    dsq = np.concatenate([[0], rng.integers(0, 4, 90), [0]]).astype(np.uint8)
    i0, j0 = 5, 84
    b = S.bands(cm, i0, j0, r=6, seed=6)  # standard bands: every d band starts at the emission count
    std = o.scan(b, dsq, i0, j0, inside, want_alpha=True)
    val = _full_valid(cm["M"], 6)
    val["jvalid"][:] = 1
    g, l = S.tr_penalties(cm)
    tr = o.trscan(b, val, l[0], dsq, i0, j0, S.PASS_5P_AND_3P_FORCE, 1, inside, want_planes=True)
    nroot = len(std["root"])
    J = tr["planes"][0]
    assert np.array_equal(J[nroot:], std["alpha"][nroot:])


def _mirror_bands(cm, i0, j0, r, seed):
    imin, imax, jmin, jmax = S.ij_bands(cm, i0, j0, r, seed)
    s = i0 + j0
    fwd = S.ij2d(cm, imin, imax, jmin, jmax, do_trunc=True)
    mir = S.ij2d(cm, s - jmax, s - jmin, s - imax, s - imin, do_trunc=True)
    return fwd, mir


def _cells(b):
    """(v, j, d) of every cell in deck order."""
    out = []
    M = len(b["jmin"])
    for v in range(M):
        for jp in range(max(0, b["jmax"][v] - b["jmin"][v] + 1)):
            h = b["hd_off"][v] + jp
            for d in range(b["hdmin"][h], b["hdmax"][h] + 1):
                out.append((v, b["jmin"][v] + jp, d))
    return out


def test_mirror_symmetry_cyk():
    cm = _cm(7, 76, 3)
    o = OracleTrCM(cm)
    M = cm["M"]
    st = cm["sttype"]
    # mirrored model: same state indices, left/right roles swapped
    mcm = dict(cm)
    mst = st.copy()
    for a, c in ((S.ML_ST, S.MR_ST), (S.IL_ST, S.IR_ST)):
        mst[st == a], mst[st == c] = c, a
    mcm["sttype"] = mst
    mcf, mcn = cm["cfirst"].copy(), cm["cnum"].copy()
    bif = st == S.B_ST
    mcf[bif], mcn[bif] = cm["cnum"][bif], cm["cfirst"][bif]
    mcm["cfirst"], mcm["cnum"] = mcf, mcn
    oesc = o.oesc.copy()
    for v in np.nonzero(st == S.MP_ST)[0]:
        oesc[v] = o.oesc[v].reshape(18, 18).T.reshape(-1)
    scores = dict(tsc=o.tsc, oesc=oesc, beginsc=o.beginsc, endsc=o.endsc, el_selfsc=o.el_selfsc)
    om = OracleTrCM(mcm, scores=scores)
    rng = np.random.default_rng(7)
    L = 90
    dsq = np.concatenate([[0], rng.integers(0, 4, L), [0]]).astype(np.uint8)
    i0, j0 = 4, 85
    s = i0 + j0
    mdsq = dsq.copy()
    mdsq[i0:j0 + 1] = dsq[i0:j0 + 1][::-1]
    fwd, mir = _mirror_bands(cm, i0, j0, 7, 7)
    val = _full_valid(M, 7, 0.8)
    mval = dict(jvalid=val["jvalid"], lvalid=val["rvalid"], rvalid=val["lvalid"], tvalid=val["tvalid"])
    g, l = S.tr_penalties(cm, 7)
    for pass_idx, mpass in ((S.PASS_5P_AND_3P_FORCE, S.PASS_5P_AND_3P_FORCE), (S.PASS_5P_ONLY, S.PASS_3P_ONLY),
                            (S.PASS_3P_ONLY, S.PASS_5P_ONLY)):
        a = o.trscan(fwd, val, l[0], dsq, i0, j0, pass_idx, 1, 0, want_planes=True)
        m = om.trscan(mir, mval, l[0], mdsq, i0, j0, mpass, 1, 0, want_planes=True,
                      lmesc=o.rmesc, rmesc=o.lmesc)
        assert a["vsc"] == m["vsc"]
        ia = {c: k for k, c in enumerate(_cells(fwd))}
        im = {c: k for k, c in enumerate(_cells(mir))}
        n_checked = 0
        for (v, j, d), k in ia.items():
            km = im[(v, s - (j - d + 1), d)]
            assert a["planes"][0][k] == m["planes"][0][km]
            assert a["planes"][1][k] == m["planes"][2][km]
            assert a["planes"][2][k] == m["planes"][1][km]
            if st[v] == S.B_ST or v == 0:
                assert a["planes"][3][k] == m["planes"][3][km]
            n_checked += 1
        assert n_checked == len(ia)


def _chain_cm(n, kind, seed):
    """ROOT S -> n states of one kind (ML or MR) -> E, with hand-made scores."""
    r = np.random.default_rng(seed)
    M = n + 2
    st = np.array([S.S_ST] + [kind] * n + [S.E_ST], np.int8)
    cfirst = np.array(list(range(1, n + 2)) + [-1], np.int32)
    cnum = np.array([1] * (n + 1) + [0], np.int32)
    nt = S.MATL_ND if kind == S.ML_ST else S.MATR_ND
    ndtype = np.array([S.ROOT_ND] + [nt] * n + [S.END_ND], np.int8)
    t = np.zeros((M, 6), np.float32)
    t[:M - 1, 0] = 1.0
    e = np.zeros((M, 16), np.float32)
    e[1:n + 1, :4] = 0.25
    cm = dict(M=M, nodes=M, clen=n, W=2 * n + 8, null=np.full(4, 0.25, np.float32), el_selfsc=np.float32(-1.0),
              pbegin=np.float32(0.05), pend=np.float32(0.05), n3_omega=0.03125, sttype=st, ndtype=ndtype,
              nodemap=np.arange(M, dtype=np.int32), ndidx=np.arange(M, dtype=np.int32), cfirst=cfirst, cnum=cnum,
              t=t, e=e)
    tsc = np.zeros((M, 6), np.float32)
    tsc[1:n + 1, 0] = r.uniform(-2, 0, n).astype(np.float32)
    oesc = np.full((M, 324), NEG, np.float32)
    em = r.uniform(-3, 2, (n, 4)).astype(np.float32)
    oesc[1:n + 1, :4] = em
    scores = dict(tsc=tsc, oesc=oesc, beginsc=np.full(M, NEG, np.float32), endsc=np.full(M, NEG, np.float32),
                  el_selfsc=-1.0)
    pty = np.full(M, NEG, np.float32)
    pty[1:n + 1] = r.uniform(-8, -2, n).astype(np.float32)
    marg = np.zeros((M, 18), np.float32)
    marg[1:n + 1, :4] = em
    return cm, scores, tsc[:, 0].astype(np.float64), em.astype(np.float64), pty, marg


@pytest.mark.parametrize("kind", [S.ML_ST, S.MR_ST])
def test_chain_closed_form(kind):
    n = 9
    cm, scores, t, em, pty, marg = _chain_cm(n, kind, 11 + kind)
    o = OracleTrCM(cm, scores=scores)
    M = cm["M"]
    rng = np.random.default_rng(kind)
    i0, j0 = 1, 14
    dsq = np.concatenate([[0], rng.integers(0, 4, j0), [0]]).astype(np.uint8)
    b = S.wide_tr_bands(cm, i0, j0)
    ones = np.ones(M, np.uint8)
    left = kind == S.ML_ST
    pass_idx = S.PASS_3P_ONLY if left else S.PASS_5P_ONLY
    val = dict(jvalid=ones, lvalid=ones if left else 0 * ones, rvalid=0 * ones if left else ones, tvalid=0 * ones)
    zero = np.zeros((M, 18), np.float32)
    out = o.trscan(b, val, pty, dsq, i0, j0, pass_idx, 1, 0, want_planes=True,
                   lmesc=marg if left else zero, rmesc=zero if left else marg)
    plane = out["planes"][1 if left else 2]
    J = out["planes"][0]
    k_cell = 0
    h0 = b["hd_off"][0]
    best = NEG
    for jp in range(b["jmax"][0] - b["jmin"][0] + 1):
        j = b["jmin"][0] + jp
        for d in range(b["hdmin"][h0 + jp], b["hdmax"][h0 + jp] + 1):
            i = j - d + 1
            want_tr, want_j = NEG, NEG
            for k in range(1, n + 1):  # entry state k (chain position)
                if d >= 1 and k + d - 1 <= n:
                    pos = [i + m for m in range(d)] if left else [j - m for m in range(d)]
                    s = pty[k] + sum(em[k - 1 + m, dsq[p]] for m, p in enumerate(pos)) + sum(t[k:k + d - 1])
                    want_tr = max(want_tr, s)
                if d == n - k + 1:  # a full (J) parse must reach E
                    pos = [i + m for m in range(d)] if left else [j - m for m in range(d)]
                    s = pty[k] + sum(em[k - 1 + m, dsq[p]] for m, p in enumerate(pos)) + sum(t[k:k + d])
                    want_j = max(want_j, s)
            got = plane[k_cell]
            if want_tr == NEG:
                assert got <= NEG
            else:
                assert abs(got - want_tr) < 1e-4, (j, d, got, want_tr)
            if want_j == NEG:
                assert J[k_cell] <= NEG
            else:
                assert abs(J[k_cell] - want_j) < 1e-4, (j, d, J[k_cell], want_j)
            best = max(best, want_tr, want_j)
            k_cell += 1
    assert abs(out["vsc"] - best) < 1e-4


@pytest.mark.parametrize("pass_idx", [S.PASS_5P_ONLY, S.PASS_3P_ONLY, S.PASS_5P_AND_3P_FORCE])
def test_inside_dominates_cyk_and_hits_are_greedy(pass_idx):
    cm = _cm(8, 90, 4)
    o = OracleTrCM(cm)
    rng = np.random.default_rng(8)
    seq = S.emit(cm, 8)
    i0, j0 = 3, 2 + len(seq)
    dsq = np.concatenate([[0, 1, 2], seq, [3, 0, 0]]).astype(np.uint8)
    b = S.tr_bands(cm, i0, j0, r=8, seed=8)
    clen = cm["clen"]
    val = S.tr_validity(cm, pass_idx, 1, clen // 3, 2 * clen // 3, clen - 5)
    g, l = S.tr_penalties(cm, 8)
    for local, pty in ((1, l[0 if pass_idx >= 4 else pass_idx - 1]), (0, g[0 if pass_idx >= 4 else pass_idx - 1])):
        c = o.trscan(b, val, pty, dsq, i0, j0, pass_idx, local, 0, cutoff=-5.0, want_planes=True)
        f = o.trscan(b, val, pty, dsq, i0, j0, pass_idx, local, 1, cutoff=-5.0, want_planes=True)
        assert np.all(f["planes"] >= c["planes"])
        assert f["vsc"] >= c["vsc"]
        for r in (c, f):
            h = r["hits"]
            sc = [x[2] for x in h]
            assert sc == sorted(sc, reverse=True)
            assert all(x[2] >= -5.0 for x in h)
            for a in range(len(h)):
                for k in range(a + 1, len(h)):
                    assert h[k][1] < h[a][0] or h[a][1] < h[k][0]
            assert all(x[4] in (S.TRMODE_J, S.TRMODE_L, S.TRMODE_R, S.TRMODE_T) for x in h)
