"""GPU parity of the STD-pass search, F1 through F7 (ixg_search_std), with
the oracle's restatement of FaithfulSearcher::search (oracle/pipeline.c):
the reported hits (genome coordinates, strand, score, null3 bias, E-value,
order) and every accounting counter (residues searched, the twelve F1-F3b
counters, the eight F4-F5b counters, the CYK filter's) identical, on the
config-1 model (72-column cloverleaf CM) over targets of one and of three
100 kb chunks per strand with planted homologs on both strands; with the
CYK envelope redefinition off; with a matrix size limit small enough that
the tau iteration runs to its ceiling and drops envelopes; and with a CYK
final round."""
import numpy as np
import pytest

from infernox_amd import ixg, models
from oracle_pipe_bind import OracleSearch

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _opts(p7, cm, cmev, L, **kw):
    W = int(cm["W"])
    f3 = ixg.F3Opts()
    f3.f1, f3.f1b, f3.f2, f3.f2b, f3.f3, f3.f3b = 0.35, 1.0, 1.0, 1.0, 0.005, 0.005
    f3.cmw, f3.maxw, f3.opts = W, W, ixg.DO_MSV | ixg.DO_FWD | ixg.DO_FWDBIAS
    env = ixg.env_opts(f4=0.005, f4b=0.005, f5=0.005, gfmu=p7["gfmu"], gflambda=p7["gflambda"])
    ez = 2.0 * L / (2 * W)
    return ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"] if not kw.get("do_cyk") else cmev["lc"], ez, W=W,
                           **kw)


def _same_acct(a, b):
    assert a.nres == b.nres
    for f, _ in ixg.Acct._fields_:
        assert getattr(a.f, f) == getattr(b.f, f), f
    for f, _ in ixg.EnvAcct._fields_:
        assert getattr(a.e, f) == getattr(b.e, f), f
    assert (a.n_past_cyk, a.pos_past_cyk, a.n_hits) == (b.n_past_cyk, b.pos_past_cyk, b.n_hits)


@pytest.mark.parametrize("L0,nhom,seed", [(60000, 3, 2), (250000, 6, 3)])
def test_search_std_parity(ctx, L0, nhom, seed):
    p7, cm, cp9, cmev = models.config1_models()
    dsq, L, placed = models.target_with_homologs(cm, L0, nhom, seed)
    W = int(cm["W"])
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    tgt = ixg.Target(ctx, dsq, L)
    osr = OracleSearch(p7, cm, cp9, W)
    for kw in (dict(), dict(do_fcykenv=False), dict(do_cyk=True)):
        _say("search", L0, kw)
        o = _opts(p7, cm, cmev, L, **kw)
        got, ga = ixg.search_std(ctx, g7, gcm, g9, hb, tgt, o)
        want, wa = osr.search(dsq, L, o)
        assert [h[:3] for h in got] == [h[:3] for h in want], (kw, got, want)
        assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want]))
        assert np.array_equal(np.float32([h[4] for h in got]), np.float32([h[4] for h in want]))
        assert [h[5] for h in got] == [h[5] for h in want]
        _same_acct(ga, wa)
        if not kw:
            found = sum(any(min(h[0], h[1]) <= e and max(h[0], h[1]) >= s and h[2] == st for h in got)
                        for s, e, st in placed)
            assert found >= nhom - 1, (placed, got)
    # a size limit the planted envelopes exceed: the tau iteration reaches
    # maxtau and drops what still does not fit
    o = _opts(p7, cm, cmev, L)
    o.size_limit = 0.02
    _say("search small size limit")
    got, ga = ixg.search_std(ctx, g7, gcm, g9, hb, tgt, o)
    want, wa = osr.search(dsq, L, o)
    assert [h[:3] for h in got] == [h[:3] for h in want]
    _same_acct(ga, wa)
    tgt.free()
    g9.free()
    gcm.free()
    g7.free()
    hb.free()


def _trunc_target(cm, body_len, seed, ends):
    """A random target whose ends hold truncated homologs: ends is a list of
    (end, strand, frac) — the homolog sampled from the CM, cut to frac of
    its length at the target's 5' (end 0) or 3' (end 1) end, on the given
    strand (1: its reverse complement)."""
    from infernox_amd import synth_cm as S
    r = np.random.default_rng(seed)
    left, right = [], []
    for k, (end, strand, frac) in enumerate(ends):
        h = S.emit(cm, seed * 100 + k)
        n = max(8, int(len(h) * frac))
        if strand:
            h = np.where(h < 4, 3 - h, h)[::-1]
        # on its own strand the homolog is cut at the end that the target's
        # end is for that strand
        cut_front = (end == 0) != bool(strand)
        part = h[len(h) - n:] if cut_front else h[:n]
        (left if end == 0 else right).append(part)
    body = r.integers(0, 4, body_len).astype(np.uint8)
    seq = np.concatenate(left + [body] + right).astype(np.uint8)
    return np.concatenate([[255], seq, [255]]).astype(np.uint8), len(seq)


def _same_tr_acct(a, b):
    _same_acct(a, b)
    for q in range(3):
        assert a.tr_nres[q] == b.tr_nres[q], q
        for f, _ in ixg.Acct._fields_:
            assert getattr(a.tr_f[q], f) == getattr(b.tr_f[q], f), (q, f)
        for f, _ in ixg.EnvAcct._fields_:
            assert getattr(a.tr_e[q], f) == getattr(b.tr_e[q], f), (q, f)
        assert (a.tr_n_past_cyk[q], a.tr_pos_past_cyk[q], a.tr_hits[q]) == \
            (b.tr_n_past_cyk[q], b.tr_pos_past_cyk[q], b.tr_hits[q]), q


@pytest.mark.parametrize("body,seed,ends", [
    (3000, 21, [(0, 0, 0.6), (1, 0, 0.6)]),
    (3000, 22, [(0, 1, 0.5), (1, 1, 0.7)]),
    (20, 23, [(0, 0, 0.7)]),
    (230000, 24, [(0, 0, 0.6), (1, 1, 0.6), (1, 0, 0.5)]),
])
def test_search_truncated_passes_parity(ctx, body, seed, ends):
    from infernox_amd import synth_cp9 as P
    p7, cm, cp9, cmev = models.config1_models()
    trdicts = P.build_trunc_cp9s(cm, 1)
    dsq, L = _trunc_target(cm, body, seed, ends)
    W = int(cm["W"])
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    g, l, ig, il = ixg.cm_trpenalties(cm)
    gcm.set_trpenalties(g, l)
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    tr9 = [ixg.CP9(ctx, h) for h in trdicts]
    trhb = [ixg.HbMap(cm, h) for h in trdicts]
    tgt = ixg.Target(ctx, dsq, L)
    osr = OracleSearch(p7, cm, cp9, W, trcp9=trdicts)
    for kw in (dict(trunc=True), dict(trunc=True, do_fcykenv=False), dict(trunc=True, do_cyk=True)):
        _say("search truncated", body, seed, kw)
        o = _opts(p7, cm, cmev, L, **kw)
        o.env.lftau, o.env.lflambda, o.env.clen = p7["lftau"], p7["lflambda"], int(cm["clen"])
        got, ga = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, o)
        want, wa = osr.search(dsq, L, o)
        assert [h[:3] + h[6:] for h in got] == [h[:3] + h[6:] for h in want], (kw, got, want)
        assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want]))
        assert np.array_equal(np.float32([h[4] for h in got]), np.float32([h[4] for h in want]))
        assert [h[5] for h in got] == [h[5] for h in want]
        _same_tr_acct(ga, wa)
        if kw == dict(trunc=True):
            assert sum(wa.tr_hits) >= 1, want  # a truncated pass reported a planted end
    for x in [tgt, g9, gcm, g7, hb] + tr9 + trhb:
        x.free()


def _multi_set(cm, case):
    from test_oracle_pipeline import _multi_seqs
    if case == "small":
        return _multi_seqs(cm, 40)
    # a chromosome-sized sequence of three chunks among many short ones
    big, L, _ = models.target_with_homologs(cm, 210000, 4, 9)
    seqs = _multi_seqs(cm, 50)
    r = np.random.default_rng(51)
    for k in range(24):
        n = int(np.clip(r.lognormal(np.log(600), 0.8), 20, 5000))
        seqs.append(np.concatenate([[255], r.integers(0, 4, n).astype(np.uint8), [255]]).astype(np.uint8))
    return seqs[:2] + [big] + seqs[2:]


@pytest.mark.parametrize("case", ["small", "mixed"])
def test_search_multi_sequence_parity(ctx, case):
    # one ixg_search over a multi-sequence target against the oracle's
    # search over the same sequences (hits with their sequence index, order,
    # scores, E-values, every counter), and against one ixg_search per
    # sequence at the same E-value scale (the union, in the same order)
    from infernox_amd import synth_cp9 as P
    from test_oracle_pipeline import _fin_key
    p7, cm, cp9, cmev = models.config1_models()
    trdicts = P.build_trunc_cp9s(cm, 1)
    seqs = _multi_set(cm, case)
    tot = sum(len(d) - 2 for d in seqs)
    W = int(cm["W"])
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    g, l, ig, il = ixg.cm_trpenalties(cm)
    gcm.set_trpenalties(g, l)
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    tr9 = [ixg.CP9(ctx, h) for h in trdicts]
    trhb = [ixg.HbMap(cm, h) for h in trdicts]
    tgt = ixg.MultiTarget(ctx, seqs)
    osr = OracleSearch(p7, cm, cp9, W, trcp9=trdicts)
    o = _opts(p7, cm, cmev, tot, trunc=True)
    o.env.lftau, o.env.lflambda, o.env.clen = p7["lftau"], p7["lflambda"], int(cm["clen"])
    _say("search multi", case, len(seqs), tot)
    got, ga = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, o, with_seq=True)
    want, wa = osr.search_multi(seqs, o, threads=8)
    assert [h[:3] + h[6:] for h in got] == [h[:3] + h[6:] for h in want], (got, want)
    assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want]))
    assert np.array_equal(np.float32([h[4] for h in got]), np.float32([h[4] for h in want]))
    assert [h[5] for h in got] == [h[5] for h in want]
    _same_tr_acct(ga, wa)
    assert len({h[7] for h in got}) >= 3, got
    union = []
    for s, d in enumerate(seqs):
        if len(d) == 2:
            continue  # an empty sequence has no tasks
        one = ixg.Target(ctx, d, len(d) - 2)
        hs, _ = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, one, o)
        union += [h + (s,) for h in hs]
        one.free()
    assert [h[:3] + h[6:] for h in got] == [h[:3] + h[6:] for h in sorted(union, key=_fin_key)]
    for x in [tgt, g9, gcm, g7, hb] + tr9 + trhb:
        x.free()


def test_search_one_query_many_cms(ctx):
    # cmscan-style: one query searched against 8 CMs of 50-400 consensus
    # columns (random nested structures, 40-60% paired), each with the
    # default pipeline, against the oracle's search (hits with their pass,
    # scores, E-values and every counter)
    from infernox_amd import synth_cm as S, synth_cp9 as P
    r = np.random.default_rng(81)
    cols = [50, 90, 140, 200, 250, 300, 350, 400]
    pool = []
    for k, clen in enumerate(cols):
        ss = S.random_nested(clen, 800 + k)
        cm = models.cm_from_alignment(ss, models.alignment(ss, 30, 800 + k), 800 + k)
        pool.append((cm, P.build_cp9_like(cm, 1, el=True), models.p7_from_cm(cm, models.CONFIG1_EV)))
    # the query: random sequence with a homolog of three of the models, one cut at the 3' end
    parts = [r.integers(0, 4, 3000).astype(np.uint8)]
    for k in (1, 4, 7):
        parts += [S.emit(pool[k][0], 900 + k).astype(np.uint8), r.integers(0, 4, 2000).astype(np.uint8)]
    h = S.emit(pool[2][0], 999).astype(np.uint8)
    parts.append(h[: 2 * len(h) // 3])
    seq = np.concatenate(parts)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    L = len(seq)
    tgt = ixg.Target(ctx, dsq, L)
    nhits = 0
    for k, (cm, cp9, p7) in enumerate(pool):
        W = int(cm["W"])
        trd = P.build_trunc_cp9s(cm, 1)
        g7, gcm, g9, hb = ixg.P7(ctx, p7, W), ixg.CM(ctx, cm), ixg.CP9(ctx, cp9), ixg.HbMap(cm, cp9)
        gcm.set_trpenalties(*ixg.cm_trpenalties(cm)[:2])
        tr9 = [ixg.CP9(ctx, x) for x in trd]
        trhb = [ixg.HbMap(cm, x) for x in trd]
        o = _opts(p7, cm, models.CONFIG1_CM_EV, L, trunc=True)
        o.env.lftau, o.env.lflambda, o.env.clen = p7["lftau"], p7["lflambda"], int(cm["clen"])
        _say("one query, CM", k, int(cm["clen"]), int(cm["M"]))
        got, ga = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, o)
        want, wa = OracleSearch(p7, cm, cp9, W, trcp9=trd).search(dsq, L, o, threads=8)
        assert [h[:3] + h[6:] for h in got] == [h[:3] + h[6:] for h in want], (k, got, want)
        assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want]))
        assert np.array_equal(np.float32([h[4] for h in got]), np.float32([h[4] for h in want]))
        assert [h[5] for h in got] == [h[5] for h in want]
        _same_tr_acct(ga, wa)
        nhits += len(got)
        for x in [g9, gcm, g7, hb] + tr9 + trhb:
            x.free()
    tgt.free()
    assert nhits >= 3
