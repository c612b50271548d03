"""GPU parity of the F4/F5 envelope definition against the oracle, through
the C ABI: the glocal Forward (scores, every special-state row, the full
matrix) and Backward (scores, special rows) bit-identical at M = 72, 120,
300 in the multihit and unihit configurations, for items configured for
their own length and for another one; then the STD-pass envelope definition
(envelopes in order and all eight accounting counters) on chunks with
planted copies, including regions that go through the stochastic-trace
clustering (rt3 = 0, and merged regions with rt2 = 0.9) and the F5b gate."""
import numpy as np
import pytest

from infernox_amd import ixg, synth
from oracle_env_bind import OracleEnv, opts as oopts

pytestmark = pytest.mark.gpu

EV = dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69)


def _say(*a):
    print(*a, flush=True)


def _profile(M, seed):
    p = synth.consensus_profile(M, seed)
    p.update(EV)
    return p


def _cons(p):
    M = int(p["M"])
    return np.argmax(np.asarray(p["mat"]).reshape(M + 1, 4)[1:], axis=1).astype(np.uint8)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("M", [72, 120, 300])
def test_gfb_parity(ctx, M):
    p = _profile(M, 100 + M)
    o = OracleEnv(p, 100)
    g = ixg.P7(ctx, p, int(1.25 * M) + 20)
    r = np.random.default_rng(M)
    L = 6000
    seq = synth.iid_dsq(L, M)
    cons = _cons(p)
    for s in (500, 2500, 4100):
        seq[s:s + M] = cons
    offs, lens = [], []
    for s in (480, 2470, 4080):
        offs.append(s)
        lens.append(M + 60)
    for _ in range(13):
        ln = int(r.integers(2, 2 * M + 40))
        offs.append(int(r.integers(0, L - ln)))
        lens.append(ln)
    offs, lens = np.array(offs, np.int64), np.array(lens, np.int64)
    for unihit in (False, True):
        for cfg in (None, lens + 17):
            _say("gfb", M, unihit, cfg is not None)
            fs, fx, fm = ixg.gfb_scores(g, seq, offs, lens, cfg_len=cfg, unihit=unihit, matrix=True)
            bs, bx, _ = ixg.gfb_scores(g, seq, offs, lens, cfg_len=cfg, unihit=unihit, backward=True)
            for i, (s, ln) in enumerate(zip(offs, lens)):
                sub = np.concatenate([[255], seq[s + 1:s + ln + 1], [255]]).astype(np.uint8)
                o.config(int(ln if cfg is None else cfg[i]), unihit)
                wf, wfx, wfm = o.gforward(sub, int(ln))
                wb, wbx, _ = o.gbackward(sub, int(ln))
                assert _bits(fs[i]) == _bits(wf), ("fwd score", i, fs[i], wf)
                assert np.array_equal(_bits(fx[i]), _bits(wfx)), ("fwd specials", i)
                assert np.array_equal(_bits(fm[i]), _bits(wfm)), ("fwd matrix", i)
                assert _bits(bs[i]) == _bits(wb), ("bwd score", i, bs[i], wb)
                assert np.array_equal(_bits(bx[i]), _bits(wbx)), ("bwd specials", i)
    g.free()


def _chunk(p, L, seed, plants):
    d = synth.iid_dsq(L, seed)
    cons = _cons(p)
    r = np.random.default_rng(seed)
    for s in plants:
        c = cons.copy()
        mut = r.random(len(c)) < 0.1  # 10% point substitutions
        c[mut] = r.integers(0, 4, mut.sum())
        d[s:s + len(c)] = c
    return d


@pytest.mark.parametrize("M", [72, 120])
def test_env_def_parity(ctx, M):
    p = _profile(M, 200 + M)
    o = OracleEnv(p, 100)
    g = ixg.P7(ctx, p, int(1.25 * M) + 20)
    L = 20000
    plants = [1000, 5000, 5000 + M + 3, 9000, 14000, 14000 + M + 40]
    d = _chunk(p, L, M, plants)
    W = int(1.3 * M)
    wins = [(900, 900 + 2 * W), (4950, 4950 + 2 * W + M), (8980, 8980 + W), (13950, 13950 + 2 * W + 60),
            (17000, 17000 + 2 * W), (2000, 2000 + W)]
    for kw in (dict(), dict(rt3=0.0), dict(rt2=0.9, rt3=0.0), dict(do_edefbias=True, f5b=0.005),
               dict(do_gfwdbias=False, f4=0.05)):
        _say("env_def", M, kw)
        want, wa = o.env_def(d, L, wins, oopts(**kw))
        got, ga = ixg.env_def(g, d, L, wins, ixg.env_opts(**kw))
        assert got == want, (kw, got, want)
        for f in ("n_past_gfwd", "pos_past_gfwd", "n_past_gfwdbias", "pos_past_gfwdbias", "n_past_edef",
                  "pos_past_edef", "n_past_edefbias", "pos_past_edefbias"):
            assert getattr(ga, f) == getattr(wa, f), (kw, f)
        assert len(want) >= 3  # planted copies are found
    g.free()


@pytest.mark.parametrize("M", [72, 120])
def test_env_def_truncated_passes_parity(ctx, M):
    # trunc_env_def with the 5' (Rgm), 3' (Lgm) and 5'&3' (Tgm) profiles: the
    # local-tail statistics, the 3' pass's ln(1/clen) correction, the
    # enforced first / final residue, and the profile a multidomain region
    # leaves behind (unihit at that window's length) carried into the later
    # windows of the chunk (rt3 = 0 makes every region multidomain)
    p = _profile(M, 300 + M)
    p.update(gfmu=-9.0, gflambda=0.69)
    o = OracleEnv(p, 100)
    g = ixg.P7(ctx, p, int(1.25 * M) + 20)
    L = 4 * M + 80
    d = _chunk(p, L, M + 7, [10, M + 30, 3 * M + 40])
    W = int(1.3 * M)
    cases = {2: [(1, L - 5), (1, 2 * M + 10), (1, M + 40), (3, M + 40)],
             3: [(5, L), (M, L), (2 * M + 20, L), (M, L - 1)],
             4: [(1, L), (2, L), (1, L)]}
    for pass_idx, wins in cases.items():
        for kw in (dict(), dict(rt3=0.0), dict(rt2=0.9, rt3=0.0), dict(do_edefbias=True, f5b=0.005)):
            kw = dict(kw, lftau=-5.5, lflambda=0.69, clen=M, gfmu=-9.0, gflambda=0.69, f4=0.05, f5=0.05)
            _say("env_def truncated", M, pass_idx, kw)
            want, wa = o.env_def(d, L, wins, oopts(**kw), pass_idx=pass_idx)
            got, ga = ixg.env_def(g, d, L, wins, ixg.env_opts(**kw), pass_idx=pass_idx)
            assert got == want, (pass_idx, kw, got, want)
            for f, _ in ixg.EnvAcct._fields_:
                assert getattr(ga, f) == getattr(wa, f), (pass_idx, kw, f)
            if "rt3" not in kw and pass_idx != 4:
                assert len(want) >= 1, (pass_idx, kw)
    g.free()
