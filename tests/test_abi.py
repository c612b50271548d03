"""CPU checks of the C ABI library: it loads, exports every entry point the
header declares, and its host-built profile tables equal the oracle's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from infernox_amd import ixg, synth
from oracle_bind import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(ixg_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    L = ixg.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in sorted(syms):
        assert hasattr(L, s), s
    assert set(ixg.EXPORTS) <= syms


def test_window_tasks_match_reference_layout():
    # cm_search.rs:1774-1808: chunks of 100000 new residues plus maxW context
    arr, n = ixg.window_tasks(250000, 400)
    got = [(arr[i].start, arr[i].end, arr[i].strand, arr[i].ctx) for i in range(n)]
    assert got == [(1, 100000, 0, 0), (1, 100000, 1, 0), (99601, 200000, 0, 400), (99601, 200000, 1, 400),
                   (199601, 250000, 0, 400), (199601, 250000, 1, 400)]
    arr, n = ixg.window_tasks(50, 400, True, False)
    assert n == 1 and (arr[0].start, arr[0].end) == (1, 50)


def _host_p7(p, maxlen):
    s = ixg.p7_struct(p)
    h = C.c_void_p()
    ixg.check(ixg.lib().ixg_p7_create(None, C.byref(s), maxlen, C.byref(h)))
    return s, h


@pytest.mark.parametrize("M,seed", [(1, 1), (2, 2), (7, 3), (33, 4), (100, 5), (301, 6)])
def test_host_tables_equal_oracle(M, seed):
    rng = synth.Lcg(seed)
    p = synth.lcg_profile(rng, M)
    s, h = _host_p7(p, 137)
    L = ixg.lib()
    o = Oracle(p, 137)
    # MSV
    rbv = np.zeros((M + 1) * 18, np.uint8)
    b = np.zeros(9, np.uint8)
    pre = np.zeros(M + 1, np.float32)
    suf = np.zeros(M + 1, np.float32)
    ixg.check(L.ixg_p7_msv_tables(h, ixg.u8ptr(rbv), ixg.u8ptr(b), ixg.f32ptr(pre), ixg.f32ptr(suf)))
    f = o.msv_fields()
    np.testing.assert_array_equal(rbv, f["rbv"])
    assert b[4:].tolist() == [190, f["bias_b"], f["tbm_b"], f["tec_b"], f["tjb_b"]]
    assert pre.tobytes() == f["prefix"].tobytes() and suf.tobytes() == f["suffix"].tobytes()
    # Viterbi
    P = C.POINTER(C.c_int16)
    rwv = np.zeros((M + 1) * 18, np.int16)
    tr = np.zeros(8 * (M + 1), np.int16)
    dd = np.zeros(1, np.int16)
    ixg.check(L.ixg_p7_vit_tables(h, rwv.ctypes.data_as(P), tr.ctypes.data_as(P), dd.ctypes.data_as(P)))
    v = o.svit
    np.testing.assert_array_equal(rwv, np.ctypeslib.as_array(v.rwv, shape=((M + 1) * 18,)))
    for j, nm in enumerate(("tmm", "tim", "tdm", "vbm", "tmi", "tii", "tmd", "tdd")):
        np.testing.assert_array_equal(tr[j * (M + 1):(j + 1) * (M + 1)],
                                      np.ctypeslib.as_array(getattr(v, nm), shape=(M + 1,)))
    assert dd[0] == v.ddbound_w
    # Forward
    nq = C.c_int32()
    ixg.check(L.ixg_p7_fwd_tables(h, C.byref(nq), None, None, None))
    q = nq.value
    assert q == o.sfwd.nq
    tm, td, rf = np.zeros(q * 28, np.float32), np.zeros(q * 4, np.float32), np.zeros(18 * q * 4, np.float32)
    ixg.check(L.ixg_p7_fwd_tables(h, C.byref(nq), ixg.f32ptr(tm), ixg.f32ptr(td), ixg.f32ptr(rf)))
    assert tm.tobytes() == np.ctypeslib.as_array(o.sfwd.tmain, shape=(q * 28,)).tobytes()
    assert td.tobytes() == np.ctypeslib.as_array(o.sfwd.tddv, shape=(q * 4,)).tobytes()
    assert rf.tobytes() == np.ctypeslib.as_array(o.sfwd.rfvv, shape=(18 * q * 4,)).tobytes()
    ixg.check(L.ixg_p7_free(None, h))


def test_bad_arguments_are_rejected():
    L = ixg.lib()
    h = C.c_void_p()
    s = ixg.P7Hmm()
    s.M = 0
    assert L.ixg_p7_create(None, C.byref(s), 10, C.byref(h)) == ixg.IXG_EINVAL
    assert L.ixg_p7_free(None, None) == ixg.IXG_EINVAL


@pytest.mark.parametrize("ss,seed", [("<<..>>.", 1), ("<.>.<.>", 2), (None, 3)])
def test_cm_host_tables_equal_oracle(ss, seed):
    from infernox_amd import synth_cm as S
    from oracle_cm_bind import OracleCM
    cm = S.build_cm(ss if ss else S.cloverleaf(), seed)
    g = ixg.CM(None, cm)
    tsc, oesc, bs, es, el = g.tables()
    o = OracleCM(cm)
    assert tsc.tobytes() == o.tsc.tobytes()
    assert oesc.tobytes() == o.oesc.tobytes()
    assert bs.tobytes() == o.beginsc.tobytes() and es.tobytes() == o.endsc.tobytes()
    assert np.float32(el).tobytes() == np.float32(o.el_selfsc).tobytes()
    g.free()


@pytest.mark.parametrize("ss,seed", [("<<..>>.", 1), ("<.>.<.>", 2), (None, 3)])
def test_cm_truncation_tables_equal_oracle(ss, seed):
    from infernox_amd import synth_cm as S
    from oracle_cm_bind import OracleTrCM, OracleTrQdbCM
    cm = S.build_cm(ss if ss else S.multibranch(120, 4, seed), seed)
    g = ixg.CM(None, cm)
    lm, rm, lp, rp = g.trtables()
    o = OracleTrCM(cm)
    assert lm.tobytes() == o.lmesc.tobytes() and rm.tobytes() == o.rmesc.tobytes()
    sl, sr = S.state_spans(cm)
    assert np.array_equal(lp, sl) and np.array_equal(rp, sr)
    gp, lpn = S.tr_penalties(cm, seed)
    g.set_trpenalties(gp, lpn)  # host-only CM: stores the tables
    ilm, irm = g.imtables()
    oq = OracleTrQdbCM(cm)
    assert ilm.tobytes() == oq.ilmesc.tobytes() and irm.tobytes() == oq.irmesc.tobytes()
    g.free()


def test_truncated_scan_rejects_bad_arguments():
    from infernox_amd import synth_cm as S
    L = ixg.lib()
    cm = S.build_cm("<<..>>.", 1)
    g = ixg.CM(None, cm)
    vsc, vm, ei, ej, n = C.c_float(), C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
    # no context / host-only CM
    rc = L.ixg_trcyk_scan_hb(None, g.h, None, None, None, 0, 1, 1, 0.0, 0, 2, 0.0, 1, C.byref(vsc), C.byref(vm),
                             C.byref(ei), C.byref(ej), None, 0, C.byref(n))
    assert rc == ixg.IXG_EINVAL
    assert L.ixg_cm_set_trpenalties(None, None, None, None) == ixg.IXG_EINVAL
    assert L.ixg_cm_set_itrpenalties(None, g.h, None, None) == ixg.IXG_EINVAL
    f, m = C.c_float(), C.c_int32()
    rc = L.ixg_trqdb_scan(None, g.h, None, 0, 1, 1, None, 0, 2, 1, 0.0, 0, C.byref(f), C.byref(m), None, None, 0,
                          C.byref(n))
    assert rc == ixg.IXG_EINVAL
    # a node list that is not a guide tree (bifurcation child node before its parent)
    bad = dict(cm)
    bad["ndtype"] = np.array([S.ROOT_ND] + [S.BIF_ND] * (cm["nodes"] - 1), np.int8)
    p = ixg.cm_params(bad)
    h = C.c_void_p()
    assert L.ixg_cm_create(None, C.byref(p), C.byref(h)) == ixg.IXG_EINVAL
    g.free()


@pytest.mark.parametrize("local", [True, False])
def test_cm_integer_and_global_tables_equal_oracle(local):
    from infernox_amd import synth_cm as S
    from oracle_cm_bind import OracleQdbCM
    cm = S.build_cm(S.multibranch(120, 4, 2), 2)
    g = ixg.CM(None, cm, local=local)
    tsc, oesc, bs, es, el = g.tables()
    itsc, ioesc, ib, ie, iel = g.itables()
    o = OracleQdbCM(cm, local=local)
    assert tsc.tobytes() == o.tsc.tobytes() and oesc.tobytes() == o.oesc.tobytes()
    assert bs.tobytes() == o.beginsc.tobytes() and es.tobytes() == o.endsc.tobytes()
    assert np.array_equal(itsc, o.itsc) and np.array_equal(ioesc, o.ioesc)
    assert np.array_equal(ib, o.ibeginsc) and np.array_equal(ie, o.iendsc) and iel == o.iel_selfsc
    g.free()


def test_cp9_entry_points_reject_bad_arguments():
    from infernox_amd import synth_cm as S, synth_cp9 as P
    L = ixg._cp9_lib()
    cm = S.build_cm(S.cloverleaf(), 1)
    h = P.build_cp9(cm, 1)
    keep = {k: np.ascontiguousarray(h[k], np.uint8 if k == "has_el" else np.int32)
            for k in ("otsc", "msc", "isc", "has_el", "el_from_ct", "el_from_off", "el_from_idx")}
    p = ixg.Cp9Params()
    p.M, p.el, p.el_selfsc = h["M"], 1, h["el_selfsc"]
    for k, v in keep.items():
        setattr(p, k, v.ctypes.data_as(C.POINTER(C.c_uint8 if k == "has_el" else C.c_int32)))
    out = C.c_void_p()
    assert L.ixg_cp9_create(None, C.byref(p), C.byref(out)) == ixg.IXG_EINVAL  # no context
    p.M = 0
    assert L.ixg_cp9_create(None, C.byref(p), C.byref(out)) == ixg.IXG_EINVAL
    assert L.ixg_cp9_free(None, None) == ixg.IXG_OK
    pn = np.zeros(6 * 80, np.int32)
    assert L.ixg_cp9_bands(None, None, None, 0, 1, 1, 1e-7, 1, None, None, None, None, None, None, None, None,
                           pn.ctypes.data_as(C.POINTER(C.c_int32))) == ixg.IXG_EINVAL
    assert L.ixg_cp9_bands_batch(None, None, None, None, 1, 1e-7, 1, None, None, None) == ixg.IXG_EINVAL
