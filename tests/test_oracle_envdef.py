"""Pins the oracle's F4/F5 envelope definition (oracle/envdef.c):

* the nats FLogsum table against its closed form log(1 + exp(-i/1000));
* the glocal Forward against a float64 Forward over the model written out
  as an explicit state graph (Plan7 glocal: S N B M I D E J C T, N/J/C
  emitting on their self loops, the left wing retracted into B->M_k entry
  probabilities), in probability space, multihit and unihit;
* Forward total == Backward total (the two walk the matrix in opposite
  directions, so a wrong index or a dropped term breaks the equality);
* posterior decoding invariants (mocc in [0, 1], btot / etot nondecreasing,
  expected begins == expected ends);
* domain definition on windows with one and with two planted copies of the
  consensus (two domains come out of the stochastic-trace clustering of a
  multidomain region) and on a random window (only negative-scoring domains).
"""
import math

import numpy as np
import pytest

from infernox_amd import synth
from oracle_env_bind import OracleEnv, lib, opts

H_MM, H_MI, H_MD, H_IM, H_II, H_DM, H_DD = range(7)


def test_p7_logsum_table_closed_form():
    L = lib()
    for i in (0, 1, 2, 10, 100, 693, 1000, 5000, 15699, 15999):
        assert abs(L.ixo_p7_flogsum_table(i) - math.log(1 + math.exp(-i / 1000.0))) < 1e-6
    r = np.random.default_rng(0)
    for _ in range(500):
        a, b = (float(x) for x in r.normal(0, 5, 2))
        want = max(a, b) + math.log1p(math.exp(-abs(a - b)))
        assert abs(L.ixo_p7_flogsum(a, b) - want) < 1.1e-3
    assert L.ixo_p7_flogsum(-math.inf, 3.0) == 3.0
    assert L.ixo_p7_flogsum(20.0, 2.0) == 20.0  # 18 >= 15.7: the smaller term is dropped


def graph_forward(p, dsq, L, cfg_len, unihit):
    """log P(seq | glocal profile) / P(seq | background), float64, from an
    explicit state graph. cfg_len: the length the N/C/J loops are set for."""
    M = int(p["M"])
    t = np.asarray(p["trans"], np.float64).reshape(M + 1, 7)
    mat = np.asarray(p["mat"], np.float64).reshape(M + 1, 4)
    nj = 0.0 if unihit else 1.0
    pmove = (2.0 + nj) / (cfg_len + 2.0 + nj)
    ploop = 1.0 - pmove
    e_move, e_loop = (1.0, 0.0) if unihit else (0.5, 0.5)
    # B -> M_k: enter M_1 directly, or walk D_1..D_{k-1} and leave to M_k
    entry = np.zeros(M + 1)
    entry[1] = 1.0 - t[0, H_MD]
    walk = t[0, H_MD]
    for k in range(2, M + 1):
        entry[k] = walk * t[k - 1, H_DM]
        walk *= t[k - 1, H_DD]
    odds = lambda k, x: mat[k, x] / 0.25 if x < 4 else 0.0
    Mp, Ip, Dp = np.zeros(M + 2), np.zeros(M + 2), np.zeros(M + 2)
    N, B, J, C = 1.0, pmove, 0.0, 0.0
    for i in range(1, L + 1):
        x = int(dsq[i])
        Mc, Ic, Dc = np.zeros(M + 2), np.zeros(M + 2), np.zeros(M + 2)
        for k in range(1, M + 1):  # consuming transitions (from row i-1)
            into = B * entry[k]
            if k > 1:
                into += Mp[k - 1] * t[k - 1, H_MM] + Ip[k - 1] * t[k - 1, H_IM] + Dp[k - 1] * t[k - 1, H_DM]
            Mc[k] = into * odds(k, x)
            if k < M:
                Ic[k] = (Mp[k] * t[k, H_MI] + Ip[k] * t[k, H_II]) * (1.0 if x < 4 else 0.0)
        for k in range(2, M + 1):  # silent D chain
            Dc[k] = Mc[k - 1] * t[k - 1, H_MD] + Dc[k - 1] * t[k - 1, H_DD]
        E = Mc[M] + Dc[M]
        J = J * ploop + E * e_loop
        C = C * ploop + E * e_move
        N = N * ploop
        B = N * pmove + J * pmove
        Mp, Ip, Dp = Mc, Ic, Dc
    return math.log(C * pmove) if C > 0 else -math.inf


EV = dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69)


def _profile(M, seed):
    p = synth.consensus_profile(M, seed)
    p.update(EV)
    return p


@pytest.mark.parametrize("M,L,seed", [(2, 3, 1), (3, 6, 2), (5, 9, 3), (8, 20, 4), (20, 40, 5), (72, 120, 6)])
@pytest.mark.parametrize("unihit", [False, True])
def test_gforward_equals_float64_graph_forward(M, L, seed, unihit):
    p = _profile(M, seed)
    o = OracleEnv(p, L)
    r = np.random.default_rng(seed)
    dsq = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
    if L > 10:  # a planted piece of the consensus so the score is not all noise
        cons = np.argmax(np.asarray(p["mat"]).reshape(M + 1, 4)[1:], axis=1)
        n = min(M, L - 4)
        dsq[3:3 + n] = cons[:n]
    o.config(L, unihit)
    sc, _, _ = o.gforward(dsq, L)
    want = graph_forward(p, dsq, L, L, unihit)
    tol = 2e-3 * (L + M) ** 0.5 + 1e-3
    assert abs(sc - want) < tol, (sc, want)


@pytest.mark.parametrize("M,L", [(4, 12), (30, 80), (72, 200), (150, 220)])
def test_gforward_total_equals_gbackward_total(M, L):
    p = _profile(M, M + 7)
    o = OracleEnv(p, L)
    d = synth.iid_dsq(L, M)
    synth.plant(d, L, p, 1, M)
    for unihit in (False, True):
        o.config(L, unihit)
        f, fx, _ = o.gforward(d, L)
        b, bx, _ = o.gbackward(d, L)
        assert abs(f - b) < 2e-3 * max(1.0, abs(f)), (f, b)
        if not unihit:
            btot, etot, mocc = o.decoding(L, fx, bx)
            assert np.all(np.diff(btot) >= -1e-6) and np.all(np.diff(etot) >= -1e-6)
            assert np.all(mocc >= -1e-3) and np.all(mocc <= 1 + 1e-3)
            assert abs(btot[L] - etot[L]) < 2e-2 * max(1.0, btot[L])


def _consensus(p):
    M = int(p["M"])
    return np.argmax(np.asarray(p["mat"]).reshape(M + 1, 4)[1:], axis=1).astype(np.uint8)


def _window_with_copies(p, L, starts, seed):
    """Random window with full consensus copies of the model at `starts`."""
    r = np.random.default_rng(seed)
    d = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
    cons = _consensus(p)
    for s in starts:
        n = min(len(cons), L + 1 - s)
        d[s:s + n] = cons[:n]
    return d


def test_domaindef_one_and_two_domains():
    M = 40
    p = _profile(M, 11)
    o = OracleEnv(p, 100)
    # one copy
    L = 100
    d = _window_with_copies(p, L, [30], 1)
    o.config(L)
    f, fx, _ = o.gforward(d, L)
    b, bx, _ = o.gbackward(d, L)
    doms = o.domaindef(d, L, fx, bx)
    assert len(doms) == 1 and doms[0][0] <= 34 and doms[0][1] >= 64, doms
    # two copies separated by a short gap: one multidomain region resolved
    # into two domains by the sampled traces
    L = 130
    d = _window_with_copies(p, L, [10, 62], 2)
    o.config(L)
    f, fx, _ = o.gforward(d, L)
    b, bx, _ = o.gbackward(d, L)
    doms = o.domaindef(d, L, fx, bx)
    assert len(doms) == 2, doms
    assert doms[0][1] < doms[1][0] + 5
    # rt3 = 0 sends every region through the stochastic-trace clustering;
    # with rt2 = 0.9 the two copies share one region that the clustering
    # splits into two domains
    doms_tr = o.domaindef(d, L, fx, bx, rt3=0.0)
    assert [x[:2] for x in doms_tr] == [x[:2] for x in doms]
    doms_m = o.domaindef(d, L, fx, bx, rt2=0.9, rt3=0.0)
    assert len(doms_m) == 2 and abs(doms_m[0][0] - doms[0][0]) <= 2 and abs(doms_m[1][1] - doms[1][1]) <= 2
    # random sequence: a glocal alignment is forced somewhere, but its
    # envelope score is far below what the F5 gate passes
    L = 120
    d = np.concatenate([[255], np.random.default_rng(3).integers(0, 4, L), [255]]).astype(np.uint8)
    o.config(L)
    f, fx, _ = o.gforward(d, L)
    b, bx, _ = o.gbackward(d, L)
    assert all(sc < 0 for _, _, sc in o.domaindef(d, L, fx, bx))


def test_env_def_std_gates():
    M = 40
    p = _profile(M, 11)
    p["gfmu"], p["gflambda"] = -5.0, 0.69
    o = OracleEnv(p, 100)
    L = 2000
    d = synth.iid_dsq(L, 5)
    d[400:400 + 45] = _window_with_copies(p, 60, [5], 9)[5:50]
    wins = [(380, 480), (1000, 1100), (1500, 1620)]
    envs, acct = o.env_def(d, L, wins, opts())
    assert len(envs) == 1 and envs[0][0] >= 380 and envs[0][1] <= 480 and envs[0][0] < 410, envs
    assert acct.n_past_gfwd >= 1 and acct.n_past_edef == 1
    # a strict F4 threshold stops everything before the envelope stage
    envs, acct = o.env_def(d, L, wins, opts(f4=1e-30))
    assert envs == [] and acct.n_past_gfwd == 0
