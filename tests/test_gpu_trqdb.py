"""GPU parity of the QDB / non-banded truncated scanners
(cm_dpsearch_trunc.rs:1800 generic_tr_scan: float TrCYK and integer
TrInside) against the CPU oracle, through the C ABI: the best root score
(bits), its mode, the full-window best, and the reported hits (coordinates,
score and null3 bias bits, mode, entry state) identical for every truncated
pass, in local and global configuration, banded and non-banded, single
window and batched over both strands."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S
from oracle_cm_bind import OracleTrQdbCM

pytestmark = pytest.mark.gpu

PASSES = [2, 3, 4, 5]


def _say(*a):
    print(*a, flush=True)


def _bits(x):
    return np.float32(x).tobytes()


def _same(got, want, tag):
    assert got["vmode"] == want["vmode"], tag
    assert _bits(got["vsc"]) == _bits(want["vsc"]), (tag, got["vsc"], want["vsc"])
    assert _bits(got["bsc_full"]) == _bits(want["bsc_full"]), tag
    gh, wh = got["hits"], want["hits"]
    assert [(h[0], h[1], h[4], h[5]) for h in gh] == [(h[0], h[1], h[4], h[5]) for h in wh], tag
    assert np.array_equal(np.float32([h[2] for h in gh]), np.float32([h[2] for h in wh])), tag
    assert np.array_equal(np.float32([h[3] for h in gh]), np.float32([h[3] for h in wh])), tag


def _case(ss, seed, cut5, cut3, flank=12, skew=False):
    cm = S.build_cm(ss, seed)
    if skew:
        cm = S.skew_inserts(cm, seed)
    r = np.random.default_rng(seed)
    h = S.emit(cm, seed + 100)
    h = h[cut5:len(h) - cut3] if cut3 else h[cut5:]
    seq = np.concatenate([r.integers(0, 4, flank), h, r.integers(0, 4, flank)])
    L = len(seq)
    return cm, np.concatenate([[255], seq, [255]]).astype(np.uint8), L


def _gpu_cm(ctx, cm, local, seed):
    g = ixg.CM(ctx, cm, local=local)
    gp, lp = S.tr_penalties(cm, seed)
    ig, il = S.tr_ipenalties(cm, seed)
    g.set_trpenalties(gp, lp)
    g.set_itrpenalties(ig, il)
    return g


CASES = [("<<<..>>>.<<..>>", 1, 3, 0), (S.cloverleaf(), 3, 12, 0), (S.cloverleaf(), 5, 0, 10),
         (S.random_nested(80, 6), 6, 20, 0)]


# the last case: insert emissions away from the background (non-zero IL / IR
# emission scores)
@pytest.mark.parametrize("ss,seed,cut5,cut3,skew", [c + (False,) for c in CASES] + [(S.cloverleaf(), 26, 8, 0, True)])
def test_trqdb_parity(ctx, ss, seed, cut5, cut3, skew):
    cm, dsq, L = _case(ss, seed, cut5, cut3, skew=skew)
    lp, rp = S.state_spans(cm)
    dmax = (np.maximum(0, rp - lp + 1) + 6).astype(np.int32)
    nh = 0
    for local in (True, False):
        o = OracleTrQdbCM(cm, local=local, seed=seed)
        g = _gpu_cm(ctx, cm, local, seed)
        for banded in (False, True):
            dm = dmax if banded else None
            for pas in PASSES:
                for inside in (0, 1):
                    tag = (len(ss), seed, local, banded, pas, inside)
                    _say("trqdb", *tag)
                    want = o.trscan(dsq, 1, L, inside, pas, local, dmax=dm, cutoff=0.0)
                    got = g.trqdb_scan(dsq, L, 1, L, pas, local=local, inside=bool(inside), dmax=dm, cutoff=0.0)
                    _same(got, want, tag)
                    nh += len(want["hits"])
        g.free()
    assert nh > 0


def test_trqdb_batch_both_strands(ctx):
    cm = S.build_cm(S.cloverleaf(), 11)
    rng = np.random.default_rng(11)
    L = 1600
    seq = rng.integers(0, 4, L).astype(np.uint8)
    for k in range(4):
        h = S.emit(cm, 300 + k)
        h = h[10:] if k % 2 == 0 else h[:-10]
        s = 50 + k * 390
        seq[s:s + len(h)] = h if k < 2 else np.where(h < 4, 3 - h, h)[::-1]
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    tgt = ixg.Target(ctx, dsq, L)
    wins = [(1 + k * 190, min(L, 1 + k * 190 + 150), k % 2) for k in range(8)]
    for local in (True, False):
        o = OracleTrQdbCM(cm, local=local, seed=11)
        g = _gpu_cm(ctx, cm, local, 11)
        for pas in (2, 4):
            for inside in (0, 1):
                _say("trqdb batch", local, pas, inside)
                vsc, vm, bf, hits = ixg.trqdb_scan_batch(ctx, g, tgt, wins, pas, local=local, inside=bool(inside),
                                                         cutoff=0.0)
                for k, (i0, j0, strand) in enumerate(wins):
                    want = o.trscan(rc if strand else dsq, i0, j0, inside, pas, local, cutoff=0.0)
                    got = dict(vsc=vsc[k], vmode=int(vm[k]), bsc_full=bf[k], hits=[h[1:] for h in hits if h[0] == k])
                    _same(got, want, (local, pas, inside, k))
        g.free()
    tgt.free()
