"""HMM -> CM band conversion (cp9.rs:2965 hmm_bands_enforce_valid_parse,
:3380 cp9_hmm2ijbands, :3815 ij2d_bands, :3854 cm_hb_mx_size_needed): the
product's host conversion (ixg_hmm2ij_bands / ixg_ij2d_bands, a pass over a
precomputed guide-tree walk) against the oracle's restatement on CP9
posterior bands of envelopes around planted homologs, in search and
alignment mode, with local begins/ends on and off (off exercises the gap
filling and unreachable-node repair), over several guide-tree shapes.

The oracle is pinned by what the bands must do: every state's bands lie in
the envelope (j in [i0-1, j0], i in [i0, j0+1]), an END state's j band is
its i band minus one, and a CYK scan restricted to the converted bands of a
wide CP9 band (tau 1e-7) never beats the non-banded CYK of the same window
and comes within a bit of it on the planted homolog (a conversion that
mislabels a band loses the homolog's parse). Runs on the CPU."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S, synth_cp9 as P
from oracle_cm_bind import OracleCM, OracleQdbCM
from oracle_cp9_bind import OracleCP9, OracleHb

STRUCTS = [("<<<..>>>", 1), ("<.>.<.>", 2), (S.cloverleaf(), 4), (S.multibranch(120, 3, 4), 5),
           (S.random_nested(90, 6), 6), (S.multibranch(300, 12, 8), 8)]


def _case(ss, seed, flank=20, pad=30):
    cm = S.build_cm(ss, seed)
    hom = S.emit(cm, seed + 100)
    r = np.random.default_rng(seed)
    seq = np.concatenate([r.integers(0, 4, pad + flank), hom, r.integers(0, 4, flank + pad)]).astype(np.uint8)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    i0, j0 = pad + 1, len(seq) - pad
    return cm, dsq, len(seq), i0, j0, pad + flank + 1, pad + flank + len(hom)


@pytest.mark.parametrize("ss,seed", STRUCTS)
def test_hmm2ij_matches_oracle(ss, seed):
    cm, dsq, L, i0, j0, h0, h1 = _case(ss, seed)
    for el, local in ((True, True), (False, True), (False, False)):
        cp9 = P.build_cp9_like(cm, seed, el=el)
        o9 = OracleCP9(cp9)
        hb = ixg.HbMap(cm, cp9, cm_local=local, cp9_local=local)
        ob = OracleHb(cm, cp9, cm_local=local, cp9_local=local)
        for search in (True, False):
            for tau in (1e-7, 5e-4, 0.05):
                pn = o9.bands(dsq, i0, j0, tau, scan=search)
                got = hb.hmm2ij(pn, i0, j0, search)
                want = ob.hmm2ij(pn, i0, j0, search)
                for g, w, name in zip(got, want, ("imin", "imax", "jmin", "jmax", "pn")):
                    assert np.array_equal(g, w), (ss, el, local, search, tau, name)
                imin, imax, jmin, jmax, _ = got
                b = hb.ij2d(imin, imax, jmin, jmax)
                wb = ob.ij2d(imin, imax, jmin, jmax)
                for f in ("hd_off", "hdmin", "hdmax"):
                    assert np.array_equal(b[f], wb[f]), (ss, f)
                on = jmin >= 0
                assert np.all(jmin[on] >= i0 - 1) and np.all(jmax[on] <= j0)
                assert np.all(imin[on] >= i0) and np.all(imax[on] <= j0 + 1)
                ends = np.asarray(cm["sttype"]) == S.E_ST
                assert np.array_equal(jmin[on & ends], imin[on & ends] - 1)
        hb.free()


@pytest.mark.parametrize("ss,seed", STRUCTS[2:])
def test_converted_bands_keep_the_homolog_parse(ss, seed):
    cm, dsq, L, i0, j0, h0, h1 = _case(ss, seed, flank=3, pad=0)
    cp9 = P.build_cp9_like(cm, seed, el=True)
    o9 = OracleCP9(cp9)
    ob = OracleHb(cm, cp9)
    pn = o9.bands(dsq, i0, j0, 1e-7, scan=True)
    imin, imax, jmin, jmax, _ = ob.hmm2ij(pn, i0, j0, True)
    b = ob.ij2d(imin, imax, jmin, jmax)
    hb_cyk = OracleCM(cm).scan(b, dsq, i0, j0, 0)["vsc"]
    _, root = OracleQdbCM(cm, local=True).scan(dsq, i0, j0, 0, want_root=True)
    best = float(root.max())
    assert hb_cyk <= best + 1e-4, (hb_cyk, best)
    assert hb_cyk >= best - 1.0, (hb_cyk, best)
    mb, cells = ob.mx_mb(b, j0 - i0 + 1)
    assert cells > 0 and mb > 0


def _trunc_case(ss, seed, cut, pad=30):
    """A window whose planted homolog is cut at its 5' end (cut < 0: the
    first -cut fraction removed) or its 3' end (cut > 0), touching the window
    edge there."""
    cm = S.build_cm(ss, seed)
    hom = S.emit(cm, seed + 100)
    r = np.random.default_rng(seed)
    n = int(len(hom) * abs(cut))
    part = hom[n:] if cut < 0 else hom[:len(hom) - n]
    seq = (np.concatenate([part, r.integers(0, 4, pad)]) if cut < 0 else
           np.concatenate([r.integers(0, 4, pad), part])).astype(np.uint8)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cm, dsq, 1, len(seq)


@pytest.mark.parametrize("ss,seed", STRUCTS[2:])
def test_hmm2ij_trunc_matches_oracle(ss, seed):
    # cp9_seq2bands_trunc's conversion for each truncated pass on windows
    # whose homolog is cut at one end: the product's host code against the
    # oracle's restatement (start / end prediction, marginal bands,
    # candidacy, the widened and special-cased hmm2ij, ij2d with do_trunc)
    for cut, passes in ((-0.4, (2, 4, 5)), (0.4, (3, 4, 5))):
        cm, dsq, i0, j0 = _trunc_case(ss, seed, cut)
        for q, cp9 in enumerate(P.build_trunc_cp9s(cm, seed)):
            o9 = OracleCP9(cp9)
            hb = ixg.HbMap(cm, cp9)
            ob = OracleHb(cm, cp9)
            for pass_idx in passes:
                for tau in (1e-7, 1e-4):
                    pn, occ = o9.trunc_pass_bands(dsq, i0, j0, tau, pass_idx)
                    got = hb.hmm2ij_trunc(pn, occ, i0, j0, pass_idx)
                    want = ob.hmm2ij_trunc(pn, occ, i0, j0, pass_idx)
                    for k, (g, w) in enumerate(zip(got, want)):
                        assert np.array_equal(np.asarray(g), np.asarray(w)), (ss, q, pass_idx, tau, k)
                    imin, imax, jmin, jmax = got[:4]
                    b = hb.ij2d(imin, imax, jmin, jmax, trunc=True)
                    wb = ob.ij2d(imin, imax, jmin, jmax, trunc=True)
                    for f in ("hd_off", "hdmin", "hdmax"):
                        assert np.array_equal(b[f], wb[f]), (ss, f)
                    sp1, sp2, ep1, ep2 = got[8][:4]
                    assert 1 <= sp1 <= sp2 <= cp9["M"] + 1 and 0 <= ep2 <= ep1 <= cp9["M"]
                    assert got[4][0] or got[5][0] or got[6][0] or got[7][0]  # some mode reaches the root
