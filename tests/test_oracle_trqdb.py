"""Pins the QDB / non-banded truncated scanners of the CPU oracle
(cm_dpsearch_trunc.rs:1800 generic_tr_scan, float TrCYK and integer
TrInside) with the reference's own relations:

- with HMM bands open over the whole window, the banded truncated CYK
  (pinned in test_oracle_trcm.py) reaches the same best root score, bit for
  bit, and the same marginal mode (cm_dpsearch_trunc.rs:2668-2692);
- against the banded float FTrInside the integer Inside is never higher
  (beyond its 1/1000-bit rounding): where the HB recursion log-sums a J and
  an L (or R) child term separately (:1239-1273, :1359-1362, :1573-1634),
  generic_tr_scan log-sums their maximum (:2081-2160), and max <= logsum.
  (The reference's own 2e-2 equality, :2695-2727, holds for its fixture
  only; on these models the gap reaches 0.15 bits.);
- query-dependent bands never raise the CYK score (:2568);
- the Inside score is never below the CYK score (:2587);
- reported hits are non-overlapping, sorted by score, above the cutoff, and
  legal under the truncation rules of their pass.
"""
import numpy as np
import pytest

from infernox_amd import synth_cm as S
from oracle_cm_bind import OracleTrCM, OracleTrQdbCM, PTY_INDEX

PASSES = [2, 3, 4, 5]  # 5' only, 3' only, 5'&3' force, 5'&3' any


def _trunc_case(ss, seed, cut5, cut3, flank=6):
    cm = S.build_cm(ss, seed)
    r = np.random.default_rng(seed)
    h = S.emit(cm, seed + 100)
    h = h[cut5:len(h) - cut3] if cut3 else h[cut5:]
    seq = np.concatenate([r.integers(0, 4, flank), h, r.integers(0, 4, flank)])
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cm, dsq, L


def _wide_tr_bands(cm, i0, j0):
    """cm_dpsearch_trunc.rs:2617-2660 wide_bands: j in i0..j0 for every state,
    d in 1..jp+1 for emitters, 0..jp+1 otherwise, E at d = 0; all modes valid."""
    M = cm["M"]
    b = S.ij2d(cm, np.full(M, i0, np.int32), np.full(M, j0 + 1, np.int32), np.full(M, i0, np.int32),
               np.full(M, j0, np.int32), do_trunc=True)
    val = {k: np.ones(M, np.uint8) for k in ("jvalid", "lvalid", "rvalid", "tvalid")}
    return b, val


STRUCTS = [("<<<..>>>.<<..>>", 1, 3, 0), (S.cloverleaf(), 3, 12, 0), (S.cloverleaf(), 5, 0, 10),
           (".<<.<<...>>.>>..", 7, 2, 3)]


@pytest.mark.parametrize("ss,seed,cut5,cut3", STRUCTS)
@pytest.mark.parametrize("local", [False, True])
def test_nonbanded_equals_open_hmm_bands(ss, seed, cut5, cut3, local):
    cm, dsq, L = _trunc_case(ss, seed, cut5, cut3)
    assert L <= int(cm["W"])  # both scanners then cover the same (j, d) cells
    q = OracleTrQdbCM(cm, local=local, seed=seed)
    hb = OracleTrCM(cm, scores=dict(tsc=q.tsc, oesc=q.oesc, beginsc=q.beginsc, endsc=q.endsc,
                                    el_selfsc=q.el_selfsc))
    b, val = _wide_tr_bands(cm, 1, L)
    for pas in PASSES:
        pty = q.pen["l" if local else "g"][PTY_INDEX[pas]]
        for inside in (0, 1):
            got = q.trscan(dsq, 1, L, inside, pas, local, cutoff=-1e30)
            want = hb.trscan(b, val, pty, dsq, 1, L, pas, local, inside, cutoff=-1e30, env_cutoff=-1e30)
            tag = (ss, pas, inside, local)
            if inside:
                assert got["vsc"] <= want["vsc"] + 0.02, (tag, got["vsc"], want["vsc"])
                assert got["vsc"] >= cyk - 0.05, tag  # Inside >= CYK up to the integer rounding
            else:
                assert got["vmode"] == want["vmode"], tag
                assert np.float32(got["vsc"]).tobytes() == np.float32(want["vsc"]).tobytes(), tag
                cyk = got["vsc"]


def test_qdb_bands_only_lower_cyk_and_inside_ge_cyk():
    cm, dsq, L = _trunc_case(S.random_nested(60, 4), 4, 15, 0, flank=10)
    q = OracleTrQdbCM(cm, local=True, seed=4)
    lp, rp = S.state_spans(cm)
    span = np.maximum(0, rp - lp + 1)
    for slack in (2, 8):
        dmax = (span + slack).astype(np.int32)
        for pas in PASSES:
            full = q.trscan(dsq, 1, L, 0, pas, True, cutoff=-1e30)
            band = q.trscan(dsq, 1, L, 0, pas, True, dmax=dmax, cutoff=-1e30)
            ins = q.trscan(dsq, 1, L, 1, pas, True, cutoff=-1e30)
            assert band["vsc"] <= full["vsc"] + 1e-4, (slack, pas)
            assert ins["vsc"] >= full["vsc"] - 0.05, pas  # Inside >= CYK up to the integer rounding
            assert full["vsc"] > 0.0


def test_reported_hits_legal_and_disjoint():
    cm, dsq, L = _trunc_case(S.cloverleaf(), 9, 14, 0, flank=25)
    q = OracleTrQdbCM(cm, local=True, seed=9)
    lp, rp = S.state_spans(cm)
    n = 0
    for pas in PASSES:
        for inside in (0, 1):
            hits = q.trscan(dsq, 1, L, inside, pas, True, cutoff=0.0)["hits"]
            sc = [h[2] for h in hits]
            assert sc == sorted(sc, reverse=True)
            for a in range(len(hits)):
                i, j, s, bias, mode, root = hits[a]
                assert s >= 0.0 and 1 <= i <= j <= L
                assert mode in (0, 1, 2, 3) and 1 <= root < cm["M"]
                if pas not in (5,) and not (i == 1 and j == L):  # local rules (:2376-2422)
                    assert mode != 0
                    if mode == 2:
                        assert j == L
                    if mode == 1:
                        assert i == 1
                for c in range(a + 1, len(hits)):
                    assert hits[c][1] < i or j < hits[c][0]
            n += len(hits)
    assert n > 0
