"""GPU parity of the query-dependent-banded and non-banded scanners
(cm_nohmm.rs generic_scan: float CYK and integer Inside) against the CPU
oracle, through the C ABI: reported hits (coordinates, score and null3 bias
bits) identical in local and global configuration, banded and non-banded,
single window and batched over both strands; cyk_seq_filter envelopes
identical."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S
from oracle_cm_bind import OracleQdbCM

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _qdb(cm, slack):
    lp, rp = S.state_spans(cm)
    span = np.maximum(0, rp - lp + 1)
    return np.maximum(0, span - slack).astype(np.int32), (span + slack).astype(np.int32)


def _seq(cm, seed, nhom=2, flank=30):
    r = np.random.default_rng(seed)
    parts = [r.integers(0, 4, flank)]
    for k in range(nhom):
        parts += [S.emit(cm, seed * 10 + k), r.integers(0, 4, flank)]
    seq = np.concatenate(parts).astype(np.uint8)
    return np.concatenate([[255], seq, [255]]).astype(np.uint8), len(seq)


def _same(got, want, tag):
    assert [(h[0], h[1]) for h in got] == [(h[0], h[1]) for h in want], tag
    assert np.array_equal(np.float32([h[2] for h in got]), np.float32([h[2] for h in want])), tag
    assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want])), tag


STRUCTS = [("<<<..>>>", 1), ("<.>.<.>", 2), (S.cloverleaf(), 4), (S.random_nested(90, 6), 6),
           (S.multibranch(120, 3, 4), 5)]


# the last case: insert emissions away from the background (non-zero IL / IR
# emission scores)
@pytest.mark.parametrize("ss,seed,skew", [c + (False,) for c in STRUCTS] + [(S.random_nested(90, 24), 24, True)])
def test_qdb_parity(ctx, ss, seed, skew):
    cm = S.build_cm(ss, seed)
    if skew:
        cm = S.skew_inserts(cm, seed)
    dsq, L = _seq(cm, seed)
    nh = 0
    for local in (True, False):
        o = OracleQdbCM(cm, local=local)
        g = ixg.CM(ctx, cm, local=local)
        for bands in ("none", "qdb"):
            dmin, dmax = _qdb(cm, 8) if bands == "qdb" else (None, None)
            for inside in (0, 1):
                for (i0, j0) in ((1, L), (5, min(L, 5 + int(cm["W"]) // 2))):
                    tag = (len(ss), seed, local, bands, inside, i0, j0)
                    _say("qdb", *tag)
                    want = o.scan(dsq, i0, j0, inside, dmin, dmax, cutoff=-2.0)
                    got = g.qdb_scan(dsq, L, i0, j0, inside=bool(inside), dmin=dmin, dmax=dmax, cutoff=-2.0)
                    _same(got, want, tag)
                    nh += len(want)
        g.free()
    assert nh > 0


def test_qdb_batch_both_strands(ctx):
    cm = S.build_cm(S.cloverleaf(), 11)
    rng = np.random.default_rng(11)
    L = 3000
    seq = rng.integers(0, 4, L).astype(np.uint8)
    for k in range(5):
        h = S.emit(cm, 400 + k)
        s = 100 + k * 560
        seq[s:s + len(h)] = h if k % 2 == 0 else np.where(h < 4, 3 - h, h)[::-1]
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    tgt = ixg.Target(ctx, dsq, L)
    wins = [(1 + k * 290, min(L, 1 + k * 290 + 320), k % 2) for k in range(10)]
    dmin, dmax = _qdb(cm, 10)
    for local in (True, False):
        o = OracleQdbCM(cm, local=local)
        g = ixg.CM(ctx, cm, local=local)
        for inside in (0, 1):
            _say("qdbbatch", local, inside)
            hits = ixg.qdb_scan_batch(ctx, g, tgt, wins, inside=bool(inside), dmin=dmin, dmax=dmax, cutoff=0.0)
            for k, (i0, j0, strand) in enumerate(wins):
                want = o.scan(rc if strand else dsq, i0, j0, inside, dmin, dmax, cutoff=0.0)
                _same([h[1:] for h in hits if h[0] == k], want, (local, inside, k))
        g.free()
    tgt.free()


def test_cyk_seq_filter_parity(ctx):
    cm = S.build_cm(S.cloverleaf(), 3)
    r = np.random.default_rng(3)
    parts = [r.integers(0, 4, 150)]
    for k in range(3):
        parts += [S.emit(cm, 30 + k), r.integers(0, 4, 120)]
    seq = np.concatenate(parts)
    n = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    o = OracleQdbCM(cm, local=True)
    g = ixg.CM(ctx, cm, local=True)
    dmin, dmax = _qdb(cm, 10)
    want = o.seq_filter(dsq, n, dmin, dmax, 5.0)
    got = g.cyk_seq_filter(dsq, n, dmin, dmax, 5.0)
    assert got == want and len(got) >= 1
    g.free()


def test_qdb_large_model(ctx):
    """The 300-column, 12-stem model, one banded window holding a homolog."""
    cm = S.build_cm(S.multibranch(300, 12, 3), 3)
    dsq, L = _seq(cm, 3, nhom=1, flank=25)
    dmin, dmax = _qdb(cm, 12)
    o = OracleQdbCM(cm, local=True)
    g = ixg.CM(ctx, cm, local=True)
    for inside in (0, 1):
        _say("qdblarge", inside)
        want = o.scan(dsq, 1, L, inside, dmin, dmax, cutoff=0.0)
        got = g.qdb_scan(dsq, L, 1, L, inside=bool(inside), dmin=dmin, dmax=dmax, cutoff=0.0)
        _same(got, want, ("large", inside))
        assert len(want) >= 1
    g.free()
