"""Serving loop (ixg_server_*): each request's result in a mix of concurrent
requests equals the same search run alone.

A request is the filter stage of one (model, sequence) search on both strands:
the (chunk x strand) tasks of cm_search.rs:1774-1808 filtered with
f3_filter_sequence (cm_pipeline.rs:1863). Run alone it is ixg_window_tasks +
ixg_f3_filter_tasks, whose parity with the oracle test_gpu_p7.py pins; here
the served result must be the same bit for bit (windows, float32 scores, and
the per-task F1F3Acct) whatever batch or stream it ran in.
"""
import time

import numpy as np
import pytest

from infernox_amd import ixg, synth

pytestmark = pytest.mark.gpu

EV = dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69)


def _say(*a):
    print(*a, flush=True)


def _opts(M, maxw):
    o = ixg.F3Opts()
    o.f1, o.f1b, o.f2, o.f2b, o.f3, o.f3b = 0.35, 1.0, 1.0, 1.0, 0.02, 0.02
    o.cmw, o.maxw, o.opts = M, maxw, 7
    return o


def _models(ctx, sizes):
    out = []
    for k, M in enumerate(sizes):
        p = synth.consensus_profile(M, 100 + k)
        p.update(EV)
        maxw = int(1.25 * M) + 20
        out.append((p, ixg.P7(ctx, p, maxw), _opts(M, maxw)))
    return out


def _requests(models, lengths, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for n, L in enumerate(lengths):
        m = int(rng.integers(len(models)))
        d = synth.iid_dsq(int(L), seed * 1000 + n)
        if L > 2000:
            synth.plant(d, int(L), models[m][0], 1 + int(L) // 40000, seed + n)
        reqs.append((m, d, int(L)))
    return reqs


def _alone(ctx, models, m, d, L):
    _, g, o = models[m]
    tasks, nt = ixg.window_tasks(L, o.maxw)
    if nt == 0:
        return [], [], []
    tgt = ixg.Target(ctx, d, L)
    wins, accts = ixg.f3_filter_tasks(ctx, g, tgt, tasks, nt, o)
    tgt.free()
    tl = [(tasks[i].start, tasks[i].end, tasks[i].strand) for i in range(nt)]
    return tl, wins, [accts[i] for i in range(nt)]


def _same(got, want, what):
    tl, w, acct, _ = got
    wtl, ww, wacct = want
    assert tl == wtl, what
    assert [x[:3] for x in w] == [x[:3] for x in ww], what
    assert np.array_equal(np.float32([x[3] for x in w]), np.float32([x[3] for x in ww])), what
    for a, b in zip(acct, wacct):
        for f, _ in ixg.Acct._fields_:
            assert getattr(a, f) == getattr(b, f), (what, f)


def test_served_equals_alone_burst_and_staggered(ctx):
    models = _models(ctx, [72, 120, 200])
    lengths = [0, 1, 37, 150, 900, 2500, 7000, 12000, 30000, 64000, 99000, 180000, 250000] + [3000] * 12 + \
              [20000] * 6
    reqs = _requests(models, lengths, 5)
    _say("serve: computing", len(reqs), "requests alone")
    want = [_alone(ctx, models, m, d, L) for m, d, L in reqs]
    assert sum(len(w[1]) for w in want) > 0

    srv = ixg.Server(ctx, nstreams=4, max_inflight=1 << 22, max_batch=1 << 20)
    try:
        _say("serve: burst of", len(reqs))
        tickets = [srv.submit(models[m][1], d, L, models[m][2]) for m, d, L in reqs]
        for k, t in enumerate(tickets):
            _same(srv.result(t), want[k], ("burst", k, reqs[k][2]))
        st = srv.stats()
        assert st["requests"] == len(reqs)
        assert st["largest_batch"] > 1, st  # same-model requests were batched

        _say("serve: staggered arrivals")
        order = np.random.default_rng(7).permutation(len(reqs))
        tickets = {}
        for n, k in enumerate(order):
            m, d, L = reqs[k]
            tickets[k] = srv.submit(models[m][1], d, L, models[m][2])
            if n % 5 == 4:
                time.sleep(0.002)
            if n % 7 == 6:  # collect some results while later requests still arrive
                done = [j for j in tickets if srv.done(tickets[j])]
                for j in done:
                    _same(srv.result(tickets.pop(j)), want[j], ("staggered", j))
        for j, t in tickets.items():
            _same(srv.result(t), want[j], ("staggered", j))
    finally:
        srv.free()
        for _, g, _ in models:
            g.free()


def test_served_small_inflight_budget_serializes(ctx):
    # an in-flight budget smaller than one request still runs it (alone)
    models = _models(ctx, [90])
    reqs = _requests(models, [5000, 40000, 1200, 40000], 9)
    want = [_alone(ctx, models, m, d, L) for m, d, L in reqs]
    srv = ixg.Server(ctx, nstreams=2, max_inflight=10000, max_batch=10000)
    try:
        tickets = [srv.submit(models[m][1], d, L, models[m][2]) for m, d, L in reqs]
        for k, t in enumerate(tickets):
            _same(srv.result(t), want[k], ("budget", k))
    finally:
        srv.free()
        models[0][1].free()


def _bundle(ctx, which):
    """A resident search bundle of a workload model: (SearchModel, opts,
    handles to free, cm dict)."""
    from infernox_amd import models, synth_cp9
    p7, cm, cp9, cmev = models.config1_models() if which == 1 else models.config3_models()
    W = int(cm["W"])
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    gcm.set_trpenalties(*ixg.cm_trpenalties(cm)[:2])
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    trd = synth_cp9.build_trunc_cp9s(cm, 1)
    tr9 = [ixg.CP9(ctx, h) for h in trd]
    trhb = [ixg.HbMap(cm, h) for h in trd]
    f3 = ixg.F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.02, f3b=0.02, cmw=W, maxw=W, opts=7)
    env = ixg.env_opts(f4=0.02, f4b=0.02, f5=0.02, gfmu=p7["gfmu"], gflambda=p7["gflambda"], lftau=p7["lftau"],
                       lflambda=p7["lflambda"], clen=int(cm["clen"]))
    # a fixed E-value scale (a 2 Mb search space), as a server fixes it
    o = ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"], 2.0e6 / (2 * W), W=W, trunc=True)
    return ixg.SearchModel(g7, gcm, g9, hb, tr9, trhb), o, [g7, gcm, g9, hb] + tr9 + trhb, cm


def _search_requests(cm, n, seed):
    """Sequences of 0-20 kb: random, with a CM-sampled homolog inside, or
    with one cut at an end (the truncated passes)."""
    from infernox_amd import synth_cm as S
    r = np.random.default_rng(seed)
    out = []
    for k in range(n):
        L0 = int(np.clip(r.lognormal(np.log(2000), 1.2), 0, 20000)) if k else 0
        body = r.integers(0, 4, L0).astype(np.uint8)
        kind = k % 3
        if kind and L0 > 0:
            h = S.emit(cm, seed * 1000 + k).astype(np.uint8)
            if kind == 1:
                at = int(r.integers(0, L0 + 1))
                body = np.concatenate([body[:at], h, body[at:]])
            else:
                body = np.concatenate([body, h[: max(8, 2 * len(h) // 3)]])
        out.append(np.concatenate([[255], body, [255]]).astype(np.uint8))
    return out


def test_served_search_equals_alone(ctx):
    # whole searches (sequence, CM) served in a mix of two models of 230 and
    # 374 states, in a burst and staggered with results collected while
    # later requests arrive: every request's hits (coordinates, strand,
    # float score and bias, E-value, pass, order) equal ixg_search on the
    # sequence alone
    bundles = [_bundle(ctx, 1), _bundle(ctx, 3)]
    reqs = []
    for b, (bm, o, _, cm) in enumerate(bundles):
        reqs += [(b, d) for d in _search_requests(cm, 18, 61 + b)]
    order = np.random.default_rng(7).permutation(len(reqs))
    want = {}
    for q in order:
        b, d = reqs[q]
        bm, o, hs, _ = bundles[b]
        t = ixg.Target(ctx, d, len(d) - 2)
        want[q], _ = ixg.search(ctx, *[x for x in hs[:3]], hs[3], hs[4:7], hs[7:10], t, o)
        t.free()
    srv = ixg.Server(ctx, nstreams=3, max_inflight=1 << 22, max_batch=1 << 20)
    for phase in ("burst", "staggered"):
        _say("served search", phase, len(reqs))
        tickets = {}
        for n_sub, q in enumerate(order):
            b, d = reqs[q]
            tickets[q] = srv.submit_search(bundles[b][0], d, len(d) - 2, bundles[b][1])
            if phase == "staggered" and n_sub % 6 == 5:
                time.sleep(0.002)
        batched = 0
        for q in order:
            got, times = srv.search_result(tickets[q])
            batched = max(batched, times[3])
            assert [h[:3] + h[6:] for h in got] == [h[:3] + h[6:] for h in want[q]], (phase, q)
            assert np.array_equal(np.float32([h[3] for h in got]), np.float32([h[3] for h in want[q]]))
            assert np.array_equal(np.float32([h[4] for h in got]), np.float32([h[4] for h in want[q]]))
            assert [h[5] for h in got] == [h[5] for h in want[q]]
        if phase == "burst":
            assert batched > 1  # requests of one model shared a search
    assert sum(len(v) for v in want.values()) >= 8  # the planted homologs are reported
    srv.free()
    for _, _, hs, _ in bundles:
        for x in hs:
            x.free()
