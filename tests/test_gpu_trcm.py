"""GPU parity of the truncated HMM-banded scanners (TrCYK, FTrInside) against
the CPU oracle, through the C ABI: vsc_root, its mode, the envelope bounds
and the reported truncated hits (coordinates, score and null3 bias bits,
mode, entry state) must be identical for every truncated pass, in local and
global configuration, single and batched, on both strands."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S
from oracle_cm_bind import OracleTrCM

pytestmark = pytest.mark.gpu

PASSES = [S.PASS_5P_ONLY, S.PASS_3P_ONLY, S.PASS_5P_AND_3P_FORCE, S.PASS_5P_AND_3P_ANY]


def _say(*a):
    print(*a, flush=True)


def _pty_index(pass_idx):
    return {S.PASS_5P_ONLY: 1, S.PASS_3P_ONLY: 2}.get(pass_idx, 0)


def _truncated_case(ss, seed, cut=0.3, flank=20, skew=False):
    """A homolog whose 5' part is missing at the start of the sequence and a
    second whose 3' part is missing at the end. Returns the CM, dsq, L and
    the scan cases (i0, j0, consensus columns present c1..c2, validity
    parameters sp1, sp2, ep1, ep2)."""
    cm = S.build_cm(ss, seed)
    if skew:
        cm = S.skew_inserts(cm, seed)
    rng = np.random.default_rng(seed)
    h1, h2 = S.emit(cm, seed + 100), S.emit(cm, seed + 200)
    k1, k2 = int(len(h1) * cut), int(len(h2) * cut)
    mid = rng.integers(0, 4, 2 * flank)
    seq = np.concatenate([h1[k1:], mid, h2[:len(h2) - k2]]).astype(np.uint8)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    clen = cm["clen"]
    c1, c2 = max(1, int(cut * clen)), max(2, int((1 - cut) * clen))
    cases = [(1, len(h1) - k1, c1, clen, (1, c1 + 6, clen, clen - 2)),
             (L - (len(h2) - k2) + 1, L, 1, c2, (1, 3, max(1, c2 - 5), c2 + 5)),
             (1, L, 1, clen, (1, clen + 1, clen, 0))]
    return cm, dsq, L, cases


def _check(got, want, tag):
    vsc, vm, ei, ej, hits = got
    assert np.float32(vsc).tobytes() == np.float32(want["vsc"]).tobytes(), (tag, vsc, want["vsc"])
    assert vm == want["vmode"], tag
    assert (ei, ej) == (want["envi"], want["envj"]), tag
    wh = want["hits"]
    assert [(h[0], h[1], h[4], h[5]) for h in hits] == [(h[0], h[1], h[4], h[5]) for h in wh], tag
    assert np.array_equal(np.float32([h[2] for h in hits]), np.float32([h[2] for h in wh])), tag
    assert np.array_equal(np.float32([h[3] for h in hits]), np.float32([h[3] for h in wh])), tag


STRUCTS = [("<<<..>>>", 1), ("<.>.<.>", 2), (S.cloverleaf(), 4), (S.multibranch(120, 3, 4), 5),
           (S.random_nested(90, 6), 6)]


# the last case: insert emissions away from the background (non-zero IL / IR
# emission scores)
@pytest.mark.parametrize("ss,seed,skew", [c + (False,) for c in STRUCTS] + [(S.cloverleaf(), 25, True)])
def test_truncated_parity(ctx, ss, seed, skew):
    cm, dsq, L, cases = _truncated_case(ss, seed, skew=skew)
    o = OracleTrCM(cm)
    g = ixg.CM(ctx, cm)
    gp, lp = S.tr_penalties(cm, seed)
    g.set_trpenalties(gp, lp)
    nhits = 0
    for (i0, j0, c1, c2, vp) in cases:
        b = S.tr_bands(cm, i0, j0, r=8, seed=seed + i0, c1=c1, c2=c2)
        for pass_idx in PASSES:
            val = S.tr_validity(cm, pass_idx, *vp)
            for local in (1, 0):
                pty = (lp if local else gp)[_pty_index(pass_idx)]
                for inside in (0, 1):
                    tag = (len(ss), seed, i0, j0, pass_idx, local, inside)
                    _say("trcm", *tag)
                    want = o.trscan(b, val, pty, dsq, i0, j0, pass_idx, local, inside, cutoff=-2.0, do_null3=True,
                                    env_cutoff=-1.0)
                    got = g.trscan(b, val, dsq, L, i0, j0, pass_idx, local=bool(local), inside=bool(inside),
                                   cutoff=-2.0, null3=True, env_cutoff=-1.0)
                    _check(got, want, tag)
                    nhits += len(want["hits"])
    assert nhits > 0
    g.free()


def test_truncated_all_valid_wide_bands(ctx):
    """Every state valid in every mode, bands covering every (j, d)."""
    cm, dsq, L, _ = _truncated_case("<<.>.<..>>", 9, flank=3)
    o = OracleTrCM(cm)
    g = ixg.CM(ctx, cm)
    gp, lp = S.tr_penalties(cm, 9)
    g.set_trpenalties(gp, lp)
    M = cm["M"]
    ones = np.ones(M, np.uint8)
    val = dict(jvalid=ones, lvalid=ones, rvalid=ones, tvalid=(cm["sttype"] == S.B_ST).astype(np.uint8))
    val["tvalid"][0] = 1
    b = S.wide_tr_bands(cm, 1, L)
    for inside in (0, 1):
        _say("wide", inside)
        want = o.trscan(b, val, lp[0], dsq, 1, L, S.PASS_5P_AND_3P_FORCE, 1, inside, cutoff=-20.0)
        got = g.trscan(b, val, dsq, L, 1, L, S.PASS_5P_AND_3P_FORCE, inside=bool(inside), cutoff=-20.0)
        _check(got, want, ("wide", inside))
    g.free()


def test_truncated_batch_both_strands(ctx):
    cm = S.build_cm(S.multibranch(100, 3, 12), 12)
    o = OracleTrCM(cm)
    g = ixg.CM(ctx, cm)
    gp, lp = S.tr_penalties(cm, 12)
    g.set_trpenalties(gp, lp)
    rng = np.random.default_rng(12)
    L = 5000
    seq = rng.integers(0, 4, L).astype(np.uint8)
    for k in range(6):
        h = S.emit(cm, 300 + k)
        s = 150 + k * 800
        seq[s:s + len(h)] = h
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    tgt = ixg.Target(ctx, dsq, L)
    clen = cm["clen"]
    envs = []
    for k in range(12):
        strand = k % 2
        i0 = 100 + k * 390
        j0 = min(L, i0 + 160 + (k % 3) * 20)
        b = S.tr_bands(cm, i0, j0, r=7 + k % 3, seed=k)
        sp2, ep1 = max(1, clen // 4 + k), max(1, clen - clen // 4 - k)
        val = S.tr_validity(cm, S.PASS_5P_AND_3P_FORCE, 1, sp2, ep1, clen - 2)
        envs.append((i0, j0, strand, b, val))
    for inside in (0, 1):
        for local in (1, 0):
            _say("trbatch", inside, local)
            pty = (lp if local else gp)[0]
            vsc, vm, ei, ej, hits, cells = ixg.trhb_scan_batch(ctx, g, tgt, envs, S.PASS_5P_AND_3P_FORCE,
                                                               local=bool(local), inside=bool(inside),
                                                               env_cutoff=0.0, cutoff=-1.0)
            assert cells > 0
            for k, (i0, j0, strand, b, val) in enumerate(envs):
                d = rc if strand else dsq
                want = o.trscan(b, val, pty, d, i0, j0, S.PASS_5P_AND_3P_FORCE, local, inside, cutoff=-1.0,
                                env_cutoff=0.0)
                mine = [h[1:] for h in hits if h[0] == k]
                _check((vsc[k], vm[k], ei[k], ej[k], mine), want, ("batch", k, inside, local))
    tgt.free()
    g.free()


def test_truncated_large_model(ctx):
    """The 300-column, 12-stem model of the benchmark, one envelope per pass."""
    cm = S.build_cm(S.multibranch(300, 12, 3), 3)
    o = OracleTrCM(cm)
    g = ixg.CM(ctx, cm)
    gp, lp = S.tr_penalties(cm, 3)
    g.set_trpenalties(gp, lp)
    rng = np.random.default_rng(3)
    h = S.emit(cm, 3)
    seq = np.concatenate([h[60:], rng.integers(0, 4, 40)]).astype(np.uint8)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    clen = cm["clen"]
    j0 = len(h) - 60
    c1 = 1 + 60 * clen // len(h)
    b = S.tr_bands(cm, 1, j0, r=8, seed=3, c1=c1, c2=clen)
    nh = 0
    for pass_idx in (S.PASS_5P_ONLY, S.PASS_5P_AND_3P_FORCE):
        val = S.tr_validity(cm, pass_idx, 1, c1 + 6, clen, clen - 2)
        for inside in (0, 1):
            _say("large", pass_idx, inside)
            want = o.trscan(b, val, lp[_pty_index(pass_idx)], dsq, 1, j0, pass_idx, 1, inside, cutoff=0.0)
            got = g.trscan(b, val, dsq, L, 1, j0, pass_idx, inside=bool(inside), cutoff=0.0)
            _check(got, want, ("large", pass_idx, inside))
            nh += len(want["hits"])
    assert nh > 0
    g.free()
