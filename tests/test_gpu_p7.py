"""GPU parity of the p7 filter kernels against the CPU oracle, through the C
ABI. Integer results (SSV diagonals, Viterbi/MSV scores, window lists,
accounting) must be identical; Forward and bias scores are compared bit for
bit as well (the tolerance below is 0 ulp; see _close)."""
import numpy as np
import pytest

from infernox_amd import ixg, synth
from oracle_bind import Oracle, F3Opts as OF3

pytestmark = pytest.mark.gpu

EV = dict(lmmu=-12.0, lmlambda=0.69, lvmu=-11.0, lvlambda=0.69, lftau=-6.0, lflambda=0.69)
FWD_ULP = 0  # Forward/bias tolerance in float32 ulps


def _close(a, b, ulp=FWD_ULP):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    return np.all(np.abs(ia - ib) <= ulp)


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("M", [1, 2, 7, 8, 9, 16, 17, 63, 64, 65, 100, 128, 200, 301, 400])
def test_ssv_diagonals_parity(ctx, M):
    rng = synth.Lcg(1000 + M)
    p = synth.lcg_profile(rng, M, EV)
    o = Oracle(p, 4 * M + 40)
    g = ixg.P7(ctx, p, 4 * M + 40)
    tot = 0
    for L in (1, 5, 63, 64, 65, 1000, 4096, 4097 + M, 9001):
        d = synth.lcg_dsq(rng, L) if L < 2000 else synth.iid_dsq(L, L + M)
        for f1 in (0.35, 0.9, 0.02):
            _say("ssv", M, L, f1)
            want = o.ssv_diagonals(d, L, f1)
            got = g.ssv_diagonals(d, L, f1)
            assert [(a, b, c) for a, b, c, _ in got] == [(a, b, c) for a, b, c, _ in want], (M, L, f1)
            assert np.array_equal(np.float32([x[3] for x in got]), np.float32([x[3] for x in want]))
            tot += len(want)
    g.free()
    assert tot > 0


@pytest.mark.parametrize("M", [72, 300])
def test_ssv_chunk_with_homologs(ctx, M):
    p = synth.consensus_profile(M, M)
    p.update(EV)
    p["lmmu"] = -8.0
    L = 100000 + 4 * M
    d = synth.iid_dsq(L, 77)
    planted = synth.plant(d, L, p, 12, 5)
    o = Oracle(p, int(1.25 * M) + 10)
    g = ixg.P7(ctx, p, int(1.25 * M) + 10)
    for f1 in (0.35, 0.02):
        _say("ssv chunk", M, f1)
        want = o.ssv_diagonals(d, L, f1)
        got = g.ssv_diagonals(d, L, f1)
        assert len(want) > len(planted) // 2
        assert [w[:3] for w in got] == [w[:3] for w in want]
        assert np.array_equal(np.float32([x[3] for x in got]), np.float32([x[3] for x in want]))
    g.free()


def _windows(rng, L, n):
    offs, lens = [], []
    for _ in range(n):
        ln = int(rng.integers(1, min(L, 900) + 1))
        s = int(rng.integers(0, L - ln + 1))
        offs.append(s)
        lens.append(ln)
    return np.array(offs, np.int64), np.array(lens, np.int64)


@pytest.mark.parametrize("M", [1, 2, 3, 8, 9, 17, 33, 64, 65, 99, 100, 101, 128, 129, 150, 200, 256, 257,
                               300, 400])
def test_window_scores_parity(ctx, M):
    rng = synth.Lcg(7 + M)
    p = synth.lcg_profile(rng, M, EV)
    o = Oracle(p, 200)
    g = ixg.P7(ctx, p, 200)
    L = 6000
    seq = synth.lcg_dsq(rng, L)
    seq[0] = 255
    r = np.random.default_rng(M)
    offs, lens = _windows(r, L, 96)
    _say("scores", M)
    fw = g.fwd_scores(seq, offs, lens)
    bs = g.bias_scores(seq, offs, lens)
    vs, vo = g.vit_scores(seq, offs, lens)
    ms, mo = g.msv_scores(seq, offs, lens)
    for i, (s, ln) in enumerate(zip(offs, lens)):
        sub = np.concatenate([[255], seq[s + 1:s + ln + 1], [255]]).astype(np.uint8)
        assert _close(fw[i], o.fwd(sub, ln)), ("fwd", M, i, fw[i], o.fwd(sub, ln))
        assert _close(bs[i], o.bias_score(sub, ln)), ("bias", M, i)
        v = o.vit_scalar(sub, ln)
        assert bool(vo[i]) == (v is None)
        if v is not None:
            assert np.float32(vs[i]).tobytes() == np.float32(v).tobytes(), ("vit", M, i, vs[i], v)
        mv = o.msv_score(sub, ln, ln)
        assert bool(mo[i]) == (mv is None)
        if mv is not None:
            assert np.float32(ms[i]).tobytes() == np.float32(mv).tobytes(), ("msv", M, i)
    g.free()


def _opts(cls, f1, f3, cmw, maxw, opts, f2=0.15):
    o = cls()
    o.f1, o.f1b, o.f2, o.f2b, o.f3, o.f3b = f1, 0.35, f2, f2, f3, f3
    o.cmw, o.maxw, o.opts = cmw, maxw, opts
    return o


@pytest.mark.parametrize("M,flags", [(72, 7), (300, 7), (120, 7 | 8 | 16), (72, 7 | 32), (72, 6)])
def test_f3_filter_sequence_parity(ctx, M, flags):
    p = synth.consensus_profile(M, 3 * M)
    p.update(dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69))
    maxw = int(1.25 * M) + 20
    L = 100000 + maxw
    d = synth.iid_dsq(L, M + flags)
    synth.plant(d, L, p, 20, M)
    o = Oracle(p, maxw)
    g = ixg.P7(ctx, p, maxw)
    _say("f3", M, flags)
    want, wa = o.f3_filter_sequence(d, L, _opts(OF3, 0.35, 0.02, M + 10, maxw, flags))
    got, ga = g.f3_filter_sequence(d, L, _opts(ixg.F3Opts, 0.35, 0.02, M + 10, maxw, flags))
    assert [w[:2] for w in got] == [w[:2] for w in want]
    assert np.array_equal(np.float32([w[2] for w in got]), np.float32([w[2] for w in want]))
    for f, _ in ixg.Acct._fields_:
        assert getattr(ga, f) == getattr(wa, f), f
    g.free()


def test_f3_filter_tasks_both_strands(ctx):
    M = 300
    p = synth.consensus_profile(M, 11)
    p.update(dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69))
    maxw = 400
    L = 260000
    d = synth.markov2_dsq(L, 3)
    synth.plant(d, L, p, 30, 4)
    g = ixg.P7(ctx, p, maxw)
    o = Oracle(p, maxw)
    tgt = ixg.Target(ctx, d, L)
    tasks, nt = ixg.window_tasks(L, maxw)
    opts = _opts(ixg.F3Opts, 0.35, 0.005, M, maxw, 7)
    wins, accts = ixg.f3_filter_tasks(ctx, g, tgt, tasks, nt, opts)
    for t in range(nt):
        tk = tasks[t]
        seg = d[tk.start:tk.end + 1]
        if tk.strand:
            seg = np.where(seg < 4, 3 - seg, seg)[::-1]
        sub = np.concatenate([[255], seg, [255]]).astype(np.uint8)
        _say("task", t, tk.start, tk.end, tk.strand)
        want, wa = o.f3_filter_sequence(sub, tk.end - tk.start + 1, _opts(OF3, 0.35, 0.005, M, maxw, 7))
        got = [(s, e, sc) for (tt, s, e, sc) in wins if tt == t]
        assert [w[:2] for w in got] == [w[:2] for w in want], t
        assert np.array_equal(np.float32([w[2] for w in got]), np.float32([w[2] for w in want]))
        for f, _ in ixg.Acct._fields_:
            assert getattr(accts[t], f) == getattr(wa, f), (t, f)
    tgt.free()
    g.free()


@pytest.mark.parametrize("M", [72, 300, 400])
def test_fwd_launch_shapes_identical(ctx, M):
    # the F3 Forward's launch shapes give the same scores bit for bit, and the
    # autotuner (which runs each applicable shape once per window-count
    # class and keeps the fastest) gives them too
    rng = synth.Lcg(11 + M)
    p = synth.lcg_profile(rng, M, EV)
    g = ixg.P7(ctx, p, 200)
    L = 40000
    seq = synth.lcg_dsq(rng, L)
    seq[0] = 255
    offs, lens = _windows(np.random.default_rng(M + 5), L, 700)
    nq = max(2, (M - 1) // 4 + 1)
    shapes = [v for v, ok in ((0, nq <= 80), (1, nq <= 100), (2, True)) if ok]
    _say("fwd shapes", M, shapes)
    got = []
    for v in shapes:
        ctx.set_autotune(2 + v)
        got.append(np.float32(g.fwd_scores(seq, offs, lens)))
    ctx.set_autotune(1)
    tuned = [np.float32(g.fwd_scores(seq, offs, lens)) for _ in range(2)]
    choices = ctx.autotune_choices()
    ctx.set_autotune(1)
    for x in got[1:] + tuned:
        assert np.array_equal(x.view(np.int32), got[0].view(np.int32))
    assert len(choices) == 1 and choices[0][0] == nq and choices[0][2] in shapes, choices
    g.free()
