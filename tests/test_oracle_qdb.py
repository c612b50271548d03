"""Pins the QDB / non-banded scanners of the CPU oracle (cm_nohmm.rs
generic_scan): in the global configuration the non-banded CYK root scores
equal a textbook unbanded CYK over (state, i, j) bit for bit, the integer
Inside root scores equal a float64 log-sum Inside within the integer
rounding, query-dependent bands only lower CYK, and the HB scanner with
bands covering everything gives the same root scores; cyk_seq_filter's
envelopes cover every padded CYK hit and are merged and ordered."""
import numpy as np
import pytest

from infernox_amd import synth_cm as S
from oracle_cm_bind import OracleCM, OracleQdbCM, bands_struct
from test_oracle_cm import textbook

NEG = -1e36
STRUCTS = ["<<..>>.", "<.>.<.>", ".<<...>.>.", "<<.>.<..>>"]


def _case(ss, seed, L=10):
    cm = S.build_cm(ss, seed)
    r = np.random.default_rng(seed)
    dsq = np.concatenate([[255], r.integers(0, 4, L), [255]]).astype(np.uint8)
    return cm, dsq, L


@pytest.mark.parametrize("ss,seed", [(s, k) for k, s in enumerate(STRUCTS)])
def test_nonbanded_global_cyk_and_inside_equal_textbook(ss, seed):
    cm, dsq, L = _case(ss, seed)
    o = OracleQdbCM(cm, local=False)
    assert np.all(o.endsc <= NEG) and np.all(o.ibeginsc == -987654321)
    W = int(cm["W"])
    assert L <= W
    for inside in (0, 1):
        _, root = o.scan(dsq, 1, L, inside, want_root=True, cutoff=1e30)
        want = textbook(cm, o, dsq, 1, L, inside)[0]
        n = 0
        for j in range(1, L + 1):
            for d in range(1, j + 1):
                got, w = float(root[j - 1, d]), want[(j - d + 1, j)]
                if w <= -1e35:
                    assert got <= -1e35, (j, d, got, w)
                elif inside:
                    assert abs(got - w) < 0.02, (j, d, got, w)  # integer scores and ILogsum round to 1/1000 bit
                else:
                    assert np.float32(got).tobytes() == np.float32(w).tobytes(), (j, d, got, w)
                n += 1
        assert n == L * (L + 1) // 2


def _qdb(cm, slack):
    """Per-state d bands around the state's consensus subtree length."""
    lp, rp = S.state_spans(cm)
    span = np.maximum(0, rp - lp + 1)
    return (np.maximum(0, span - slack).astype(np.int32), (span + slack).astype(np.int32))


@pytest.mark.parametrize("clen,seed", [(40, 1), (60, 2)])
def test_qdb_only_lowers_cyk_and_inside_ge_cyk(clen, seed):
    cm = S.build_cm(S.random_nested(clen, seed), seed)
    seq = S.emit(cm, seed)
    r = np.random.default_rng(seed)
    seq = np.concatenate([r.integers(0, 4, 8), seq, r.integers(0, 4, 8)])
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    for local in (False, True):
        o = OracleQdbCM(cm, local=local)
        _, full = o.scan(dsq, 1, L, 0, want_root=True, cutoff=1e30)
        dmin, dmax = _qdb(cm, 6)
        _, band = o.scan(dsq, 1, L, 0, dmin, dmax, want_root=True, cutoff=1e30)
        assert np.all(band <= full)
        _, ins = o.scan(dsq, 1, L, 1, want_root=True, cutoff=1e30)
        ok = full > -1e35
        assert np.all(ins[ok] >= full[ok] - 0.05)  # Inside >= CYK up to the integer rounding
        assert full.max() > 5.0  # the planted homolog scores


@pytest.mark.parametrize("ss,seed", [(s, k) for k, s in enumerate(STRUCTS)])
def test_hb_with_covering_bands_equals_nonbanded_cyk(ss, seed):
    cm, dsq, L = _case(ss, seed, L=12)
    oq = OracleQdbCM(cm, local=False)
    _, root = oq.scan(dsq, 1, L, 0, want_root=True, cutoff=1e30)
    ohb = OracleCM(cm)
    # the HB scan with the global tables: overwrite its configured scores
    import ctypes as C
    for name in ("tsc", "oesc", "beginsc", "endsc"):
        src = getattr(oq, name).astype(np.float32).ravel()
        dst = np.ctypeslib.as_array(getattr(ohb.sc, name), shape=src.shape)
        dst[:] = src
    ohb.sc.el_selfsc = oq.el_selfsc
    ohb.sc.local_begin = 0
    b = S.wide_bands(cm, 1, L)
    got = ohb.scan(b, dsq, 1, L, 0)
    off = 0
    for j in range(b["jmin"][0], b["jmax"][0] + 1):
        jp = j - b["jmin"][0]
        lo, hi = b["hdmin"][b["hd_off"][0] + jp], b["hdmax"][b["hd_off"][0] + jp]
        for d in range(lo, hi + 1):
            if d >= 1 and j >= 1:
                hv, qv = float(got["root"][off + d - lo]), float(root[j - 1, d])
                if hv <= -1e35:
                    assert qv <= -1e35
                else:
                    assert np.float32(hv).tobytes() == np.float32(qv).tobytes(), (j, d, hv, qv)
        off += max(0, hi - lo + 1)


def test_cyk_seq_filter_envelopes_cover_hits():
    cm = S.build_cm(S.cloverleaf(), 3)
    r = np.random.default_rng(3)
    parts = [r.integers(0, 4, 150)]
    for k in range(3):
        parts += [S.emit(cm, 30 + k), r.integers(0, 4, 120)]
    seq = np.concatenate(parts)
    n = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    o = OracleQdbCM(cm, local=True)
    dmin, dmax = _qdb(cm, 10)
    hits = o.scan(dsq, 1, n, 0, dmin, dmax, cutoff=5.0)
    env = o.seq_filter(dsq, n, dmin, dmax, 5.0)
    W = int(cm["W"])
    assert len(hits) >= 3 and len(env) >= 1
    for (i, j, _, _) in hits:
        a, b = max(1, j - (W - 1)), min(n, i + (W - 1))
        assert any(es <= a and b <= ee for es, ee in env)
    for k in range(1, len(env)):
        assert env[k - 1][1] < env[k][0]
