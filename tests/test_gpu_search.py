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
