"""Pins the oracle's SSV diagonal search (oracle/p7f.c ixo_ssv_diagonals,
following cm_pipeline.rs:521-717) by brute force on tiny profiles
(M <= 4, L <= 8, and M <= 12, L <= 16 so that the forward walk's patience
of five positions is reached), every byte parameter drawn at random.

The brute force does not run the row recurrence. The saturating steps
v -> sat(sat(max(v, xB) + bias) - cost) are monotone, so the recurrence's
value at (i, k) is the maximum, over every diagonal start s, of the chain
started fresh from xB at row s; a reset after a reported diagonal only
removes the starts before it. The walk back is the largest start whose
undone score is at or below the baseline; the walk forward is the last
running-maximum record before five positions in a row fall short of it.
A dropped saturation, a wrong diagonal index, an off-by-one in the start,
the end or the skip, or a lost reset each changes some window here."""
import ctypes as C

import numpy as np

from oracle_bind import Msv, RawWin, lib, u8

KP = 18


def _sat(v):
    return max(0, min(255, v))


def brute_ssv(rbv, m, bias, base, tjb, tbm, scale, thresh, dsq, L, stats=None):
    xb = _sat(base - _sat(tjb + tbm))
    baseline = base - tjb - tbm

    def cost(k, x):
        return int(rbv[k, x]) if x < KP else 255

    def chain(s, i, k):
        """Value at (i, k) of the diagonal entered fresh (from xB) at row s."""
        v = None
        for t in range(s, i + 1):
            kt = k - (i - t)
            v = _sat(_sat((xb if v is None else max(v, xb)) + bias) - cost(kt, dsq[t]))
        return v

    wins = []
    reset = 1
    i = 1
    while i <= L:
        vals = [0] * (m + 1)
        for k in range(1, m + 1):
            lo = max(reset, i - k + 1)
            vals[k] = max(chain(s, i, k) for s in range(lo, i + 1))
        xe = max(vals[1:])
        if xe >= thresh and xe > 0:
            end_k = vals.index(xe)  # leftmost argmax
            # start: the largest row s with xe - sum_{t=s..i} (bias - cost) <= baseline
            lo = max(1, i - end_k + 1)
            if xe <= baseline:
                s_star = i + 1
            else:
                s_star = lo
                for s in range(i, lo - 1, -1):
                    r = xe - sum(bias - cost(end_k - (i - t), dsq[t]) for t in range(s, i + 1))
                    if r <= baseline:
                        s_star = s
                        break
            # end: running-maximum records (ties count) before 5 misses in a row
            sc, best, best_n, since = xe, xe, i, 0
            n = i + 1
            while end_k + (n - i) < m and n <= L:
                sc += bias - cost(end_k + (n - i), dsq[n])
                if sc >= best:
                    best, best_n, since = sc, n, 0
                else:
                    since += 1
                    if since == 5:
                        if stats is not None:
                            stats["patience"] = stats.get("patience", 0) + 1
                        break
                n += 1
            end_model = end_k + (best_n - i)
            start_model = end_k - (i - s_star)
            score = (np.float32(best - tjb) - np.float32(base)) / np.float32(scale) - np.float32(3.0)
            wins.append((s_star, end_model - start_model + 1, end_model, np.float32(score)))
            reset = best_n + 1
            i = best_n + 1
            continue
        i += 1
    return wins


def oracle_ssv(rbv, m, bias, base, tjb, tbm, tec, scale, dsq, L, mmu, mlambda, f1):
    f = Msv()
    f.m, f.scale_b = m, scale
    f.base_b, f.bias_b, f.tbm_b, f.tec_b, f.tjb_b = base, bias, tbm, tec, tjb
    flat = np.ascontiguousarray(rbv, np.uint8).ravel()
    f.rbv = flat.ctypes.data_as(C.POINTER(C.c_uint8))
    f.mmu, f.mlambda, f.max_length = mmu, mlambda, 10
    thresh = lib().ixo_ssv_thresh(C.byref(f), f1)
    n = lib().ixo_ssv_diagonals(C.byref(f), u8(dsq), L, f1, None, 0)
    out = (RawWin * max(n, 1))()
    lib().ixo_ssv_diagonals(C.byref(f), u8(dsq), L, f1, out, n)
    return thresh, [(w.n, w.length, w.k, np.float32(w.score)) for w in out[:n]]


def _cases(rng, n, mmax, lmax):
    for case in range(n):
        m = int(rng.integers(1, mmax + 1))
        L = int(rng.integers(1, lmax + 1))
        bias = int(rng.integers(0, 40))
        base = int(rng.integers(60, 200))
        tjb, tbm, tec = (int(x) for x in rng.integers(0, 30, 3))
        scale = float(rng.uniform(1.0, 4.0))
        rbv = rng.integers(0, 60, (m + 1, KP)).astype(np.uint8)
        # residues mostly ACGU, some degenerate codes, and now and then a
        # code past the alphabet (cost 255)
        seq = rng.choice(np.arange(20), L, p=[0.2] * 4 + [0.01] * 14 + [0.03] * 2)
        dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
        # a threshold that some rows reach: low mu, steep lambda
        mmu = float(rng.uniform(-12.0, -2.0))
        yield case, m, L, bias, base, tjb, tbm, tec, scale, rbv, dsq, mmu


def test_ssv_walks_brute_force():
    rng = np.random.default_rng(20)
    for mmax, lmax, n, min_trig, min_multi, min_pat in ((4, 8, 3000, 300, 100, 0), (12, 16, 600, 100, 50, 20)):
        ntrig = nmulti = 0
        stats = {}
        for case, m, L, bias, base, tjb, tbm, tec, scale, rbv, dsq, mmu in _cases(rng, n, mmax, lmax):
            thresh, got = oracle_ssv(rbv, m, bias, base, tjb, tbm, tec, scale, dsq, L, mmu, 2.0, 0.5)
            want = brute_ssv(rbv, m, bias, base, tjb, tbm, scale, thresh, dsq, L, stats)
            assert got == want, (case, m, L, bias, base, tjb, tbm, thresh, dsq.tolist(), got, want)
            ntrig += len(got) > 0
            nmulti += len(got) > 1
        # the cases exercise triggers, the reset-and-skip and the patience stop
        assert ntrig > min_trig and nmulti > min_multi, (mmax, ntrig, nmulti)
        assert stats.get("patience", 0) >= min_pat, (mmax, stats)
