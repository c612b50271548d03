"""Pins the CPU oracle for the p7 filter path against facts that do not
depend on it: the reference's own striped == scalar Viterbi sweep, closed-form
(segment-sum) SSV row maxima, a float64 natural-order Forward computed
straight from the HMM probabilities, and a float64 2-state bias HMM."""
import math

import numpy as np
import pytest

from infernox_amd import synth
from oracle_bind import Oracle

LN2 = math.log(2.0)


def test_vit_striped_matches_scalar():
    # p7_vitfilter.rs:603-635 sse_matches_scalar, restated over both oracle forms
    rng = synth.Lcg(0x1234_5678_9abc_def0)
    ms = [1, 2, 3, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 40, 63, 64, 65, 100, 128, 200]
    ls = [1, 2, 5, 17, 50, 200, 500]
    n = 0
    for m in ms:
        for _ in range(8):
            o = Oracle(synth.lcg_profile(rng, m), 100)
            for L in ls:
                d = synth.lcg_dsq(rng, L)
                a, b = o.vit_scalar(d, L), o.vit_striped(d, L)
                assert (a is None) == (b is None)
                if a is not None:
                    assert np.float32(a).tobytes() == np.float32(b).tobytes(), (m, L, a, b)
                n += 1
    assert n > 1000


def _ssv_rowmax_closed_form(f, dsq, L):
    """xe(i) = max(0, xb + best ungapped segment sum of (bias - cost)) over
    diagonals ending at row i; exact when no cell reaches the 255 ceiling."""
    M = len(f["prefix"]) - 1
    rbv = f["rbv"].reshape(M + 1, 18)
    xb = max(0, int(f["base_b"]) - min(255, int(f["tjb_b"]) + int(f["tbm_b"])))
    bias = int(f["bias_b"])
    out = np.zeros(L, np.int64)
    for i in range(1, L + 1):
        best = 0
        for k in range(1, M + 1):                    # diagonal cell (i, k)
            for s in range(max(1, k - i + 1), k + 1):  # segment start, target >= 1
                tot = xb
                for t in range(s, k + 1):
                    tot += bias - int(rbv[t, dsq[i - k + t]])
                best = max(best, tot)
        out[i - 1] = best
    return out


@pytest.mark.parametrize("M,L,seed", [(1, 12, 1), (3, 20, 2), (5, 24, 3), (8, 30, 4)])
def test_ssv_rowmax_closed_form(M, L, seed):
    p = synth.consensus_profile(M, seed)
    p.update(dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    o = Oracle(p, 50)
    f = o.msv_fields()
    d = synth.iid_dsq(L, seed + 10)
    got = o.ssv_rowmax(d, L).astype(np.int64)
    want = _ssv_rowmax_closed_form(f, d, L)
    assert want.max() < 255
    np.testing.assert_array_equal(got, want)


def test_ssv_first_trigger_row():
    # Before the first trigger the scan is reset-free, so the first emitted
    # diagonal must end its backward walk at or before the first row whose
    # closed-form maximum reaches the threshold.
    p = synth.consensus_profile(6, 5)
    p.update(dict(lmmu=-11.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    o = Oracle(p, 30)
    d = synth.iid_dsq(60, 9)
    thr = o.ssv_thresh(0.35)
    xe = _ssv_rowmax_closed_form(o.msv_fields(), d, 60)
    rows = np.nonzero(xe >= thr)[0]
    wins = o.ssv_diagonals(d, 60, 0.35)
    assert len(rows) > 0 and len(wins) > 0
    first_row = rows[0] + 1
    n, length, k, _ = wins[0]
    assert n <= first_row and 1 <= k <= 6 and length >= 1


def _forward_f64(p, dsq, L):
    """Local multihit Forward in float64, natural order, straight from the
    HMM probabilities (no odds tables, no striping, no rescaling)."""
    M = p["M"]
    mat = p["mat"].astype(np.float64)
    t = p["trans"].astype(np.float64)
    occ = np.zeros(M + 1)
    occ[1] = t[0, 1] + t[0, 0]
    for k in range(2, M + 1):
        occ[k] = occ[k - 1] * (t[k - 1, 0] + t[k - 1, 1]) + (1 - occ[k - 1]) * t[k - 1, 5]
    Z = sum(occ[k] * (M - k + 1) for k in range(1, M + 1))
    tbm = occ / Z
    pmove = 3.0 / (L + 3.0)
    ploop = 1 - pmove
    Mp, Ip, Dp = np.zeros(M + 1), np.zeros(M + 1), np.zeros(M + 1)
    xN, xJ, xC, xB = 1.0, 0.0, 0.0, pmove
    for i in range(1, L + 1):
        x = dsq[i]
        Mc, Ic, Dc = np.zeros(M + 1), np.zeros(M + 1), np.zeros(M + 1)
        for k in range(1, M + 1):
            odds = mat[k, x] / 0.25 if x < 4 else 0.0
            Mc[k] = odds * (xB * tbm[k] + Mp[k - 1] * t[k - 1, 0] + Ip[k - 1] * t[k - 1, 3] + Dp[k - 1] * t[k - 1, 5])
            if k < M:
                Ic[k] = Mp[k] * t[k, 1] + Ip[k] * t[k, 4]
        for k in range(2, M + 1):
            Dc[k] = Mc[k - 1] * t[k - 1, 2] + Dc[k - 1] * t[k - 1, 6]
        xE = Mc.sum() + Dc.sum()
        xN = xN * ploop
        xC = xC * ploop + xE * 0.5
        xJ = xJ * ploop + xE * 0.5
        xB = xJ * pmove + xN * pmove
        Mp, Ip, Dp = Mc, Ic, Dc
    return math.log(xC * pmove)


@pytest.mark.parametrize("M", [1, 3, 9, 40, 120])
def test_forward_matches_float64_model(M):
    p = synth.consensus_profile(M, 100 + M)
    p.update(dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    o = Oracle(p, 100)
    r = np.random.default_rng(M)
    for L in (1, 7, 60):
        d = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
        want = _forward_f64(p, d, L)
        got = o.fwd(d, L)
        assert abs(got - want) < 1e-4 * max(1.0, abs(want)), (M, L, got, want)
        # the scalar (--strict) order agrees to float rounding
        assert abs(o.fwd_scalar(d, L) - got) < 1e-4 * max(1.0, abs(want))


def _bias_f64(compo, M, dsq, L):
    l1 = M / 8.0
    p1 = L / (L + 1.0)
    t = np.array([[p1, 1 - p1, 1.0], [1 / (l1 + 1), l1 / (l1 + 1), 1.0]])
    e1 = np.asarray(compo, np.float64) / 0.25
    f = np.array([0.999, 0.001]) * np.array([1.0, e1[dsq[1]]])
    for i in range(2, L + 1):
        f = (f @ t[:, :2]) * np.array([1.0, e1[dsq[i]]])
    return math.log(f[0] * t[0, 2] + f[1] * t[1, 2]) + L * math.log(p1) + math.log(1 - p1)


@pytest.mark.parametrize("L", [1, 2, 17, 300])
def test_bias_matches_float64_model(L):
    p = synth.consensus_profile(50, 7, gc=0.7)
    p.update(dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    o = Oracle(p, 100)
    d = synth.iid_dsq(L, 3, gc=0.7)
    assert abs(o.bias_score(d, L) - _bias_f64(p["compo"], 50, d, L)) < 1e-4 * max(1.0, L / 50)


def test_ilogsum_matches_c_dump():
    # cp9.rs:91-101 ilogsum_table_matches_c_dump (golden values from C Infernal)
    from oracle_bind import lib
    for i, v in [(0, 1000), (1, 1000), (2, 999), (10, 995), (100, 951), (693, 695), (1000, 585),
                 (5000, 44), (10000, 1), (19999, 0)]:
        assert lib().ixo_ilogsum_table(i) == v
    assert lib().ixo_ilogsum(-987654321, 1234) == 1234       # -INFTY sentinel (cp9.rs:81-88)
    assert lib().ixo_ilogsum(0, -25000) == 0                 # beyond the table
    assert lib().ixo_ilogsum(500, 500) == 1500


def test_flogsum_closed_form():
    from oracle_bind import lib
    for i in (0, 1, 500, 1000, 5000, 22999):
        assert abs(lib().ixo_flogsum_table(i) - math.log2(1 + 2 ** (-i / 1000))) < 2e-6
    for a, b in [(0.0, 0.0), (1.5, -2.25), (-3.0, 10.0), (5.0, 5.0001)]:
        assert abs(lib().ixo_flogsum(a, b) - math.log2(2 ** a + 2 ** b)) < 2e-3
    assert lib().ixo_flogsum(2.0, float("-inf")) == 2.0


def test_gumbel_exp_stats():
    from oracle_bind import lib
    L = lib()
    for p in (0.35, 0.02, 1e-3):
        x = L.ixo_gumbel_invsurv(p, -9.0, 0.7)
        assert abs(L.ixo_gumbel_surv(x, -9.0, 0.7) - p) < 1e-9
    assert L.ixo_exp_surv(-5.0, -4.0, 0.7) == 1.0
    assert abs(L.ixo_exp_surv(-2.0, -4.0, 0.7) - math.exp(-1.4)) < 1e-12
