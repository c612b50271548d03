"""The F3 Forward entry point that allocates no device memory
(ixg_fwd_scores_noalloc) against ixg_fwd_scores, through the C ABI: two
calls in a row on one context, the second smaller, each bit-identical to
ixg_fwd_scores on the same windows, and the context's device memory in use
unchanged by the calls. Sizes past the stated limits are refused."""
import ctypes as C

import numpy as np
import pytest

from infernox_amd import ixg, synth

pytestmark = pytest.mark.gpu

EV = dict(lmmu=-12.0, lmlambda=0.69, lvmu=-11.0, lvlambda=0.69, lftau=-6.0, lflambda=0.69)


def _windows(r, L, n, maxlen):
    lens = r.integers(1, maxlen + 1, n).astype(np.int64)
    offs = np.array([int(r.integers(0, L - ln + 1)) for ln in lens], np.int64)
    return offs, lens


def test_noalloc_fwd_equals_fwd_scores():
    with ixg.Context(0) as ctx:
        rng = synth.Lcg(91)
        p = synth.lcg_profile(rng, 150, EV)
        g = ixg.P7(ctx, p, 200)
        r = np.random.default_rng(5)
        for L, n, maxlen in ((20000, 700, 900), (3000, 40, 300)):  # the second call smaller
            seq = synth.lcg_dsq(rng, L)
            seq[0] = 255
            offs, lens = _windows(r, L, n, maxlen)
            print("noalloc call", L, n, flush=True)
            want = g.fwd_scores(seq, offs, lens)
            before = ctx.mem_in_use()
            got = g.fwd_scores_noalloc(seq, offs, lens)
            assert ctx.mem_in_use() == before
            assert got.tobytes() == want.tobytes()
        # past the limits: refused before any device work
        big = np.zeros((16 << 20) + 8, np.uint8)
        big[0] = 255
        sc = np.zeros(1, np.float32)
        rc = ixg.lib().ixg_fwd_scores_noalloc(ctx.h, g.h, ixg.u8ptr(big), len(big),
                                               ixg.i64ptr(np.zeros(1, np.int64)), ixg.i64ptr(np.ones(1, np.int64)),
                                               1, ixg.f32ptr(sc))
        assert rc == ixg.IXG_ERANGE
        g.free()
