"""GPU parity of the CP9 banding DPs (cp9.rs:1275 cp9_forward, :1995
cp9_backward, :1517 cp9_fb2hmmbands) against the CPU oracle, through the C
ABI: the six Forward/Backward matrices, both best scores (bits) and the
posterior HMM bands identical, in search (scan) and alignment mode, with and
without local ends, single window and batched over both strands of a target
(windows of many lengths, several waves)."""
import numpy as np
import pytest

from infernox_amd import ixg, synth_cm as S, synth_cp9 as P
from oracle_cp9_bind import OracleCP9

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _bits(x):
    return np.float32(x).tobytes()


def _model(ss, seed, el=True):
    cm = S.build_cm(ss, seed)
    return cm, P.build_cp9(cm, seed, el=el)


def _window(cm, seed, flank):
    r = np.random.default_rng(seed)
    h = S.emit(cm, seed + 70)
    seq = np.concatenate([r.integers(0, 4, flank), h, r.integers(0, 4, flank)])
    return np.concatenate([[255], seq, [255]]).astype(np.uint8), len(seq)


MODELS = [(".<<<..>>>..<<..>>.", 1, True), (S.cloverleaf(), 2, True), (S.cloverleaf(), 3, False),
          (S.multibranch(120, 3, 4), 4, True)]


@pytest.mark.parametrize("ss,seed,el", MODELS)
@pytest.mark.parametrize("scan", [True, False])
def test_cp9_single_window_matrices_scores_bands(ctx, ss, seed, el, scan):
    cm, h = _model(ss, seed, el)
    dsq, L = _window(cm, seed, 15)
    o = OracleCP9(h)
    g = ixg.CP9(ctx, h)
    try:
        for i0, j0 in ((1, L), (4, L - 3), (7, 7)):
            _say("cp9 single", ss[:20], seed, el, scan, i0, j0)
            f, b = o.forward(dsq, i0, j0, scan, not scan), o.backward(dsq, i0, j0, scan, not scan)
            for tau in (1e-7, 1e-3):
                got = g.bands(dsq, L, i0, j0, tau, scan, matrices=True)
                want = o.bands(dsq, i0, j0, tau, f, b, scan)
                for k, m in (("fm", f["mmx"]), ("fi", f["imx"]), ("fd", f["dmx"]), ("bm", b["mmx"]),
                             ("bi", b["imx"]), ("bd", b["dmx"])):
                    assert np.array_equal(got[k], m), (k, i0, j0, tau)
                assert _bits(got["fsc"]) == _bits(f["sc"]) and _bits(got["bsc"]) == _bits(b["sc"]), (i0, j0)
                assert np.array_equal(got["pn"], want), (i0, j0, tau)
    finally:
        g.free()


def test_cp9_batch_both_strands(ctx):
    cm, h = _model(S.cloverleaf(), 5, True)
    r = np.random.default_rng(11)
    parts, wins, pos = [], [], 1
    for k in range(150):  # three waves of windows of many lengths
        hh = S.emit(cm, 500 + k)[: 20 + (k * 7) % 60]
        fl = int(r.integers(0, 30))
        part = np.concatenate([r.integers(0, 4, fl), hh]).astype(np.uint8)
        parts.append(part)
        wins.append((pos, pos + len(part) - 1, k % 2))
        pos += len(part)
    seq = np.concatenate(parts)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    rc = np.concatenate([[255], np.where(seq < 4, 3 - seq, seq)[::-1], [255]]).astype(np.uint8)
    tgt = ixg.Target(ctx, dsq, L)
    g = ixg.CP9(ctx, h)
    o = OracleCP9(h)
    try:
        for scan in (True, False):
            _say("cp9 batch", scan)
            pn, fs, bs = ixg.cp9_bands_batch(ctx, g, tgt, wins, 1e-7, scan)
            for k, (i0, j0, strand) in enumerate(wins):
                d = rc if strand else dsq
                f, b = o.forward(d, i0, j0, scan, not scan), o.backward(d, i0, j0, scan, not scan)
                assert np.array_equal(pn[k], o.bands(d, i0, j0, 1e-7, f, b, scan)), (k, scan)
                assert _bits(fs[k]) == _bits(f["sc"]) and _bits(bs[k]) == _bits(b["sc"]), (k, scan)
    finally:
        g.free()
        tgt.free()


def test_cp9_create_rejects_a_local_end_that_does_not_run_forward(ctx):
    """A local end whose source node is not below its end node has no
    counterpart in a CP9 built from a CM; the library refuses the model
    instead of computing its Backward local-end cells in another order."""
    cm, h = _model(S.cloverleaf(), 6, True)
    M = h["M"]
    ct, off, idx = h["el_from_ct"].copy(), h["el_from_off"].copy(), h["el_from_idx"].copy()
    k = int(np.nonzero(ct[1:M + 1])[0][0]) + 1  # the first end node with a source
    idx[off[k]] = k  # the source becomes the end node itself
    bad = dict(h, el_from_idx=idx)
    _say("cp9 create bad local end", k)
    with pytest.raises(ixg.IxgError, match="precede"):
        ixg.CP9(ctx, bad)
    g = ixg.CP9(ctx, h)  # the untouched model is accepted
    g.free()


@pytest.mark.parametrize("mode", [ixg.IXG_CP9_FWD_SCAN, ixg.IXG_CP9_BCK_SCAN, 0,
                                  ixg.IXG_CP9_FWD_SCAN | ixg.IXG_CP9_ALIGN])
def test_cp9_truncated_pass_modes(ctx, mode):
    """cp9_seq2bands_trunc (cp9.rs:4130-4156) scans the Forward unless the pass
    forces the first residue and the Backward unless it forces the final one,
    and always reports best-over-position scores; every combination of the
    mode bits matches the oracle run with the same flags."""
    fscan, bscan = bool(mode & ixg.IXG_CP9_FWD_SCAN), bool(mode & ixg.IXG_CP9_BCK_SCAN)
    align = bool(mode & ixg.IXG_CP9_ALIGN)
    for ss, seed, el in ((S.cloverleaf(), 7, True), (".<<<..>>>..<<..>>.", 8, False)):
        cm, h = _model(ss, seed, el)
        dsq, L = _window(cm, seed, 12)
        o = OracleCP9(h)
        g = ixg.CP9(ctx, h)
        try:
            for i0, j0 in ((1, L), (5, L - 2)):
                _say("cp9 mode", mode, seed, i0, j0)
                f, b = o.forward(dsq, i0, j0, fscan, align), o.backward(dsq, i0, j0, bscan, align)
                got = g.bands(dsq, L, i0, j0, 1e-7, matrices=True, mode=mode)
                for k, m in (("fm", f["mmx"]), ("fi", f["imx"]), ("fd", f["dmx"]), ("bm", b["mmx"]),
                             ("bi", b["imx"]), ("bd", b["dmx"])):
                    assert np.array_equal(got[k], m), (k, mode, i0, j0)
                assert _bits(got["fsc"]) == _bits(f["sc"]) and _bits(got["bsc"]) == _bits(b["sc"]), (mode, i0)
                assert np.array_equal(got["pn"], o.bands(dsq, i0, j0, 1e-7, f, b, fscan)), (mode, i0, j0)
        finally:
            g.free()


def _slots_needed(h):
    """The wavefront kernels' LDS slots for this HMM (the host's interval
    colouring): Forward, over local-end sources; Backward, over end nodes."""
    M, ct, off, idx = h["M"], h["el_from_ct"], h["el_from_off"], h["el_from_idx"]
    last, low = {}, {}
    for k in range(1, M + 2):
        for c in range(ct[k]):
            y = int(idx[off[k] + c])
            last[y] = max(last.get(y, -1), k)
            if k < M:
                low[k] = min(low.get(k, 1 << 30), y)

    def colour(items, free):
        ends = []
        for a, b in items:
            for q, e in enumerate(ends):
                if free(e, a):
                    ends[q] = b
                    break
            else:
                ends.append(b)
        return len(ends)
    nsf = colour(sorted(last.items()), lambda e, a: e + 2 <= a)
    nsb = colour(sorted(low.items(), reverse=True), lambda e, a: e - 2 >= a)
    return nsf, nsb


def test_cp9_row_parallel_fallback_on_deeply_nested_local_ends(ctx):
    """A 200-pair stem nests 200 local ends inside one another, more than the
    wavefront kernels' LDS slots hold in either direction, so the
    row-parallel kernels (lanes over nodes, serial delete chain) run; they
    match the oracle in search and alignment mode, single and batched."""
    ss = "<" * 200 + "." * 6 + ">" * 200 + "." * 30
    cm, h = _model(ss, 9, True)
    M = h["M"]
    nsf, nsb = _slots_needed(h)
    nz, nlist = 11495, len(h["el_from_idx"])
    fixed = 4 * (M + 1) * 4 + 2 * (M + 2) * 4 + nlist * 4 + ((nz * 2 + 15) // 16) * 16
    assert nsf * 65 * 4 + fixed > 65472 and nsb * 65 * 4 + fixed > 65472, (nsf, nsb)
    dsq, L = _window(cm, 9, 10)
    o = OracleCP9(h)
    g = ixg.CP9(ctx, h)
    try:
        for scan in (True, False):
            _say("cp9 fallback", M, nsf, nsb, scan)
            f, b = o.forward(dsq, 1, L, scan, not scan), o.backward(dsq, 1, L, scan, not scan)
            got = g.bands(dsq, L, 1, L, 1e-7, scan, matrices=True)
            for k, m in (("fm", f["mmx"]), ("fi", f["imx"]), ("fd", f["dmx"]), ("bm", b["mmx"]),
                         ("bi", b["imx"]), ("bd", b["dmx"])):
                assert np.array_equal(got[k], m), (k, scan)
            assert _bits(got["fsc"]) == _bits(f["sc"]) and _bits(got["bsc"]) == _bits(b["sc"]), scan
            assert np.array_equal(got["pn"], o.bands(dsq, 1, L, 1e-7, f, b, scan)), scan
        tgt = ixg.Target(ctx, dsq, L)
        try:
            wins = [(1, L, 0), (5, L - 9, 1), (30, L - 40, 0)]
            rc = np.concatenate([[255], np.where(dsq[1:L + 1] < 4, 3 - dsq[1:L + 1], dsq[1:L + 1])[::-1], [255]])
            rc = rc.astype(np.uint8)
            pn, fs, bs = ixg.cp9_bands_batch(ctx, g, tgt, wins, 1e-7, True)
            for k, (i0, j0, strand) in enumerate(wins):
                d = rc if strand else dsq
                f, b = o.forward(d, i0, j0, True, False), o.backward(d, i0, j0, True, False)
                assert np.array_equal(pn[k], o.bands(d, i0, j0, 1e-7, f, b, True)), k
                assert _bits(fs[k]) == _bits(f["sc"]) and _bits(bs[k]) == _bits(b["sc"]), k
        finally:
            tgt.free()
    finally:
        g.free()
