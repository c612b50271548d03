"""GPU parity of the HMM-banded CYK / Inside scanners against the CPU oracle,
through the C ABI: vsc_root, the envelope bounds and the root deck must be
bit-identical for CYK and for Inside (same float operations in the same
order, same FLogsum table), and the greedy hit lists identical."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S
from oracle_cm_bind import OracleCM

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _case(ss, seed, flank=25, r=8):
    cm = S.build_cm(ss, seed)
    hom = S.emit(cm, seed + 100)
    rng = np.random.default_rng(seed)
    left, right = rng.integers(0, 4, flank), rng.integers(0, 4, flank)
    seq = np.concatenate([left, hom, right]).astype(np.uint8)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cm, dsq, L, flank + 1, flank + len(hom)


STRUCTS = [("<<<..>>>", 1), ("<.>.<.>", 2), ("<<..>>.<<...>>..", 3), (S.cloverleaf(), 4),
           (S.multibranch(120, 3, 4), 5), (S.random_nested(90, 6), 6), (S.random_nested(200, 7), 7),
           (S.multibranch(300, 12, 8), 8)]


@pytest.mark.parametrize("ss,seed", STRUCTS)
def test_cyk_inside_parity(ctx, ss, seed):
    cm, dsq, L, h0, h1 = _case(ss, seed)
    o = OracleCM(cm)
    g = ixg.CM(ctx, cm)
    for r, (i0, j0) in [(3, (h0, h1)), (8, (h0, h1)), (12, (max(1, h0 - 10), min(L, h1 + 10)))]:
        b = S.bands(cm, i0, j0, r=r, seed=seed + r)
        for inside in (0, 1):
            _say("cm", len(ss), seed, r, inside)
            want = o.scan(b, dsq, i0, j0, inside, env_cutoff=-5.0)
            if inside == 0:
                vsc, ei, ej = g.cyk_scan(b, dsq, L, i0, j0, env_cutoff=-5.0)
                hits = None
            else:
                vsc, ei, ej, hits = g.scan_hits(b, dsq, L, i0, j0, inside=True, env_cutoff=-5.0, cutoff=-10.0)
            assert np.float32(vsc).tobytes() == np.float32(want["vsc"]).tobytes(), (vsc, want["vsc"])
            assert (ei, ej) == (want["envi"], want["envj"])
            if hits is not None:
                wh = o.hits(b, dsq, i0, j0, want["root"], -10.0, True)
                assert [(h[0], h[1]) for h in hits] == [(h[0], h[1]) for h in wh]
                assert np.array_equal(np.float32([h[2] for h in hits]), np.float32([h[2] for h in wh]))
                assert np.array_equal(np.float32([h[3] for h in hits]), np.float32([h[3] for h in wh]))
    g.free()


@pytest.mark.parametrize("ss,seed", [(S.cloverleaf(), 21), (S.random_nested(90, 22), 22)])
def test_insert_emissions(ctx, ss, seed):
    cm, dsq, L, h0, h1 = _case(ss, seed)
    cm = S.skew_inserts(cm, seed)
    o = OracleCM(cm)
    g = ixg.CM(ctx, cm)
    for r in (3, 12):
        b = S.bands(cm, h0, h1, r=r, seed=seed + r)
        for inside in (0, 1):
            _say("inserts", len(ss), seed, r, inside)
            want = o.scan(b, dsq, h0, h1, inside, env_cutoff=-5.0)
            vsc, ei, ej = g.cyk_scan(b, dsq, L, h0, h1, env_cutoff=-5.0) if inside == 0 else \
                g.scan_hits(b, dsq, L, h0, h1, inside=True, env_cutoff=-5.0, cutoff=-10.0)[:3]
            assert np.float32(vsc).tobytes() == np.float32(want["vsc"]).tobytes(), (vsc, want["vsc"])
            assert (ei, ej) == (want["envi"], want["envj"])
    g.free()


def test_wide_bands_small(ctx):
    cm, dsq, L, h0, h1 = _case("<<.>.<..>>", 9, flank=3)
    o = OracleCM(cm)
    g = ixg.CM(ctx, cm)
    b = S.wide_bands(cm, 1, L)
    for inside in (0, 1):
        want = o.scan(b, dsq, 1, L, inside)
        vsc, ei, ej, hits = g.scan_hits(b, dsq, L, 1, L, inside=bool(inside), cutoff=-20.0)
        assert np.float32(vsc).tobytes() == np.float32(want["vsc"]).tobytes()
        wh = o.hits(b, dsq, 1, L, want["root"], -20.0, True)
        assert [(h[0], h[1], np.float32(h[2])) for h in hits] == [(h[0], h[1], np.float32(h[2])) for h in wh]
    g.free()


def test_batch_both_strands(ctx):
    cm = S.build_cm(S.cloverleaf(), 11)
    rng = np.random.default_rng(3)
    L = 6000
    seq = rng.integers(0, 4, L).astype(np.uint8)
    starts = []
    for k in range(8):
        h = S.emit(cm, 200 + k)
        s = 200 + k * 700
        if k % 2:
            h = np.where(h < 4, 3 - h, h)[::-1]
        seq[s:s + len(h)] = h
        starts.append((s + 1, s + len(h), k % 2))
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    o = OracleCM(cm)
    g = ixg.CM(ctx, cm)
    tgt = ixg.Target(ctx, dsq, L)
    envs = []
    for (a, b_, strand) in starts:
        if strand:  # coordinates on the reverse complement
            a, b_ = L - b_ + 1, L - a + 1
        i0, j0 = max(1, a - 15), min(L, b_ + 15)
        envs.append((i0, j0, strand, S.bands(cm, i0, j0, r=9, seed=a)))
    for inside in (0, 1):
        _say("batch", inside)
        vsc, ei, ej, hits, cells = ixg.hb_scan_batch(ctx, g, tgt, envs, inside=bool(inside), env_cutoff=0.0,
                                                     report=True, cutoff=0.0)
        assert cells > 0
        for k, (i0, j0, strand, b) in enumerate(envs):
            d = rc if strand else dsq
            want = o.scan(b, d, i0, j0, inside, env_cutoff=0.0)
            assert np.float32(vsc[k]).tobytes() == np.float32(want["vsc"]).tobytes(), k
            assert (ei[k], ej[k]) == (want["envi"], want["envj"])
            wh = o.hits(b, d, i0, j0, want["root"], 0.0, True)
            got = [(h[1], h[2], np.float32(h[3])) for h in hits if h[0] == k]
            assert got == [(h[0], h[1], np.float32(h[2])) for h in wh], k
        assert (vsc > 0).sum() >= 1  # planted homologs are found in some envelopes
    tgt.free()
    g.free()


def test_multi_cm_queue_equals_per_model_batches(ctx):
    # envelopes of three CMs of different sizes interleaved in one persistent
    # launch (ixg_hb_scan_multi) give each envelope exactly what its own
    # model's batch gives, and the oracle's scores and hits
    cms = [S.build_cm(S.cloverleaf(), 11), S.build_cm(S.multibranch(150, 4, 5), 12),
           S.build_cm(S.random_nested(90, 13), 13)]
    rng = np.random.default_rng(21)
    L = 12000
    seq = rng.integers(0, 4, L).astype(np.uint8)
    spans = []
    for k in range(18):
        m = k % 3
        h = S.emit(cms[m], 300 + k)
        s = 100 + k * 640
        if k % 4 == 1:
            h = np.where(h < 4, 3 - h, h)[::-1]
        seq[s:s + len(h)] = h[:600]
        spans.append((m, s + 1, s + min(len(h), 600), int(k % 4 == 1)))
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    gs = [ixg.CM(ctx, cm) for cm in cms]
    tgt = ixg.Target(ctx, dsq, L)
    envs, env_cm = [], []
    for m, a, b_, strand in spans:
        if strand:
            a, b_ = L - b_ + 1, L - a + 1
        i0, j0 = max(1, a - 12), min(L, b_ + 12)
        envs.append((i0, j0, strand, S.bands(cms[m], i0, j0, r=8, seed=a)))
        env_cm.append(m)
    for inside in (0, 1):
        _say("multi", inside)
        vsc, ei, ej, hits, cells = ixg.hb_scan_multi(ctx, gs, env_cm, tgt, envs, inside=bool(inside),
                                                     report=True, cutoff=0.0)
        tot = 0
        for m in range(3):
            idx = [k for k in range(len(envs)) if env_cm[k] == m]
            v1, e1, f1, h1, c1 = ixg.hb_scan_batch(ctx, gs[m], tgt, [envs[k] for k in idx], inside=bool(inside),
                                                  report=True, cutoff=0.0)
            tot += c1
            for t, k in enumerate(idx):
                assert np.float32(vsc[k]).tobytes() == np.float32(v1[t]).tobytes(), (m, k)
                assert (ei[k], ej[k]) == (e1[t], f1[t]), (m, k)
                got = [(h[1], h[2], np.float32(h[3]), np.float32(h[4])) for h in hits if h[0] == k]
                want = [(h[1], h[2], np.float32(h[3]), np.float32(h[4])) for h in h1 if h[0] == t]
                assert got == want, (m, k)
        assert cells == tot
        o = OracleCM(cms[1])
        k = env_cm.index(1)
        i0, j0, strand, b = envs[k]
        want = o.scan(b, rc if strand else dsq, i0, j0, inside, env_cutoff=0.0)
        assert np.float32(vsc[k]).tobytes() == np.float32(want["vsc"]).tobytes()
        assert (vsc > 0).sum() >= 1
    tgt.free()
    for g in gs:
        g.free()
