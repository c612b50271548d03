"""Pins the oracle's full-matrix Forward (p7_fwdback.rs p7_forward +
forward_score): in the multihit length model its score equals the striped
Forward filter score bit for bit over model sizes around the DD early-exit
boundary (the reference's own equality, p7_fwdback.rs:606-632); in both the
multihit and unihit models every row of the retained matrices, multiplied by
the scale factors applied up to it, equals a float64 natural-order Forward
computed straight from the HMM probabilities, and so do the specials."""
import math

import numpy as np
import pytest

from infernox_amd import synth
from oracle_bind import Oracle


def _oracle(M, seed):
    p = synth.consensus_profile(M, seed)
    p.update(dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    return p, Oracle(p, 400)


@pytest.mark.parametrize("M", [2, 3, 7, 33, 98, 99, 100, 101, 103, 150])
def test_multihit_score_equals_striped_filter(M):
    p, o = _oracle(M, 7 * M + 1)
    r = np.random.default_rng(M)
    for L in (1, 10, 57, 210):
        d = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
        full = o.p7_forward(d, L)
        assert np.float32(full["score"]).tobytes() == np.float32(o.fwd(d, L)).tobytes(), (M, L)


def _forward_f64(p, dsq, L, unihit):
    """Natural-order Forward in float64 from the probabilities: rows M, I, D and
    specials (E N J B C) for i = 0..L."""
    M = p["M"]
    mat = p["mat"].astype(np.float64)
    t = p["trans"].astype(np.float64)
    occ = np.zeros(M + 1)
    occ[1] = t[0, 1] + t[0, 0]
    for k in range(2, M + 1):
        occ[k] = occ[k - 1] * (t[k - 1, 0] + t[k - 1, 1]) + (1 - occ[k - 1]) * t[k - 1, 5]
    Z = sum(occ[k] * (M - k + 1) for k in range(1, M + 1))
    tbm = occ / Z
    nj = 0.0 if unihit else 1.0
    pmove = (2.0 + nj) / (L + 2.0 + nj)
    ploop = 1 - pmove
    emove, eloop = (1.0, 0.0) if unihit else (0.5, 0.5)
    Ms, Is, Ds = np.zeros((L + 1, M + 1)), np.zeros((L + 1, M + 1)), np.zeros((L + 1, M + 1))
    X = np.zeros((L + 1, 5))
    xN, xJ, xC, xB = 1.0, 0.0, 0.0, pmove
    X[0] = (0.0, xN, xJ, xB, xC)
    for i in range(1, L + 1):
        x = dsq[i]
        for k in range(1, M + 1):
            odds = mat[k, x] / 0.25 if x < 4 else 0.0
            Ms[i, k] = odds * (xB * tbm[k] + Ms[i - 1, k - 1] * t[k - 1, 0] + Is[i - 1, k - 1] * t[k - 1, 3] +
                               Ds[i - 1, k - 1] * t[k - 1, 5])
            if k < M:
                Is[i, k] = Ms[i - 1, k] * t[k, 1] + Is[i - 1, k] * t[k, 4]
        for k in range(2, M + 1):
            Ds[i, k] = Ms[i, k - 1] * t[k - 1, 2] + Ds[i, k - 1] * t[k - 1, 6]
        xE = Ms[i].sum() + Ds[i].sum()
        xN = xN * ploop
        xC = xC * ploop + xE * emove
        xJ = xJ * ploop + xE * eloop
        xB = xJ * pmove + xN * pmove
        X[i] = (xE, xN, xJ, xB, xC)
    return Ms, Is, Ds, X, math.log(xC * pmove)


@pytest.mark.parametrize("unihit", [False, True])
@pytest.mark.parametrize("M", [4, 30, 120])
def test_rows_and_specials_equal_float64_model(M, unihit):
    p, o = _oracle(M, 300 + M)
    r = np.random.default_rng(M)
    # two homologs between random flanks, so that rows with xE > 1e4 are rescaled
    seq = np.concatenate([r.integers(0, 4, 20), synth.emit_p7(p, r), r.integers(0, 4, 10), synth.emit_p7(p, r),
                          r.integers(0, 4, 15)])
    L = len(seq)
    d = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    full = o.p7_forward(d, L, unihit=unihit)
    Ms, Is, Ds, X, sc = _forward_f64(p, d, L, unihit)
    assert abs(full["score"] - sc) < 1e-4 * max(1.0, abs(sc))
    cum = np.cumprod(full["xmx"][:, 5].astype(np.float64))  # scale applied up to and including each row
    if M >= 120:
        assert (full["xmx"][:, 5] > 1).any()  # the rescaling path ran
    for name, ref in (("mmx", Ms), ("imx", Is), ("dmx", Ds)):
        got = full[name].astype(np.float64) * cum[:, None]
        assert np.allclose(got, ref, rtol=2e-4, atol=1e-30), name
    for c in range(5):
        got = full["xmx"][:, c].astype(np.float64) * cum
        assert np.allclose(got, X[:, c], rtol=2e-4, atol=1e-30), c
