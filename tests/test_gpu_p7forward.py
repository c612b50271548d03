"""GPU parity of the full-matrix Forward (p7_fwdback.rs:301 p7_forward and
:529 forward_score) against the CPU oracle, through the C ABI: every retained
row of the M, I and D matrices, the specials E N J B C SCALE, totscale and
the score are identical bit for bit, in the multihit and unihit length
models, for model sizes on both sides of the DD early-exit threshold (M=100)
and sequences long enough to trigger rescaling; the batched entry point
equals the single-region one region by region."""
import numpy as np
import pytest

from infernox_amd import ixg, synth
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

EV = dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.0, lvlambda=0.7, lftau=-4.0, lflambda=0.7)


def _say(*a):
    print(*a, flush=True)


def _same(got, want, tag):
    for key in ("mmx", "imx", "dmx", "xmx"):
        g, w = np.asarray(got[key], np.float32), np.asarray(want[key], np.float32)
        assert g.shape == w.shape, (tag, key)
        bad = np.argwhere(g.view(np.int32) != w.view(np.int32))
        assert len(bad) == 0, (tag, key, bad[:5], g[tuple(bad[0])], w[tuple(bad[0])])
    assert np.float32(got["totscale"]).tobytes() == np.float32(want["totscale"]).tobytes(), tag
    assert np.float32(got["score"]).tobytes() == np.float32(want["score"]).tobytes(), tag


def _seq(p, L, seed):
    """iid background with a planted homolog (consensus emission) when it fits."""
    r = np.random.default_rng(seed)
    s = r.integers(0, 4, L)
    M = p["M"]
    if L >= M + 4:
        s[2:2 + M] = np.argmax(p["mat"][1:, :4], axis=1)
    return np.concatenate([[255], s, [255]]).astype(np.uint8)


@pytest.mark.parametrize("M", [2, 3, 7, 33, 98, 99, 100, 101, 103, 150, 400])
def test_p7_forward_parity(ctx, M):
    p = synth.consensus_profile(M, 7 * M + 1)
    p.update(EV)
    o = Oracle(p, 400)
    g = ixg.P7(ctx, p, 400)
    rescaled = 0
    for L in (1, 10, 57, 210, M + 60):
        d = _seq(p, L, 3 * M + L)
        for unihit in (False, True):
            tag = (M, L, unihit)
            _say("p7_forward", *tag)
            want = o.p7_forward(d, L, unihit)
            got = g.p7_forward(d, L, unihit)
            _same(got, want, tag)
            rescaled += int(np.any(want["xmx"][:, 5] != 1.0))
    g.free()
    if M >= 150:
        assert rescaled > 0  # the planted homolog drives xE above the rescale threshold


def test_p7_forward_batch_equals_single(ctx):
    M = 120
    p = synth.consensus_profile(M, 5)
    p.update(EV)
    o = Oracle(p, 400)
    g = ixg.P7(ctx, p, 400)
    r = np.random.default_rng(9)
    seq = np.concatenate([[255], r.integers(0, 4, 2000), [255]]).astype(np.uint8)
    seq[300:300 + M] = np.argmax(p["mat"][1:, :4], axis=1)
    regions = [(0, 1), (10, 37), (250, 400), (1000, 999), (5, 2), (1990, 10)] + [(40 * k, 41) for k in range(30)]
    for unihit in (False, True):
        _say("p7_forward_batch", unihit, len(regions))
        got = g.p7_forward_batch(seq, regions, unihit)
        for (off, L), gr in zip(regions, got):
            d = np.concatenate([[255], seq[off + 1:off + L + 1], [255]]).astype(np.uint8)
            _same(gr, o.p7_forward(d, L, unihit), (off, L, unihit))
    g.free()
