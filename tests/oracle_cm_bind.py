"""ctypes binding of the CM part of the CPU oracle (oracle/liboracle.so).
Test infrastructure only."""
import ctypes as C

import numpy as np

from oracle_bind import INIT_LOCK, lib as _lib

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)


class Cm(C.Structure):
    _fields_ = [("M", C.c_int), ("nodes", C.c_int), ("clen", C.c_int), ("W", C.c_int), ("null", C.c_float * 4),
                ("el_selfsc", C.c_float), ("pbegin", C.c_float), ("pend", C.c_float), ("n3_omega", C.c_double),
                ("sttype", _i8p), ("ndtype", _i8p), ("nodemap", _i32p), ("cfirst", _i32p), ("cnum", _i32p),
                ("t", _f32p), ("e", _f32p)]


class CmSc(C.Structure):
    _fields_ = [("M", C.c_int), ("local_begin", C.c_int), ("el_selfsc", C.c_float), ("tsc", _f32p),
                ("oesc", _f32p), ("beginsc", _f32p), ("endsc", _f32p)]


class CmISc(C.Structure):
    _fields_ = [("itsc", _i32p), ("ioesc", _i32p), ("ibeginsc", _i32p), ("iendsc", _i32p), ("iel_selfsc", C.c_int32)]


class Bands(C.Structure):
    _fields_ = [("jmin", _i32p), ("jmax", _i32p), ("hd_off", _i32p), ("hdmin", _i32p), ("hdmax", _i32p)]


class Hit(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32), ("score", C.c_float), ("bias", C.c_float)]


_u8p = C.POINTER(C.c_uint8)


class TrValid(C.Structure):
    _fields_ = [("jvalid", _u8p), ("lvalid", _u8p), ("rvalid", _u8p), ("tvalid", _u8p)]


class TrHit(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32), ("score", C.c_float), ("bias", C.c_float),
                ("mode", C.c_int32), ("root", C.c_int32)]


_ready = False


def lib():
    if _ready:
        return _lib()
    with INIT_LOCK:
        return _init()


def _init():
    global _ready
    L = _lib()
    if not _ready:
        L.ixo_cm_configure.argtypes = [C.POINTER(Cm), C.POINTER(CmSc)]
        L.ixo_cm_free_scores.argtypes = [C.POINTER(CmSc)]
        L.ixo_hb_ncells.argtypes = [C.POINTER(Bands), C.c_int]
        L.ixo_hb_ncells.restype = C.c_int64
        L.ixo_hb_scan.argtypes = [C.POINTER(Cm), C.POINTER(CmSc), C.POINTER(Bands), C.POINTER(C.c_uint8), C.c_int,
                                  C.c_int, C.c_int, C.c_float, _f32p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  _f32p, _f32p]
        L.ixo_hb_scan.restype = C.c_int64
        L.ixo_hb_report_hits.argtypes = [C.POINTER(Cm), C.POINTER(Bands), C.POINTER(C.c_uint8), C.c_int, C.c_int,
                                         _f32p, C.c_float, C.c_int, C.POINTER(Hit), C.c_int64]
        L.ixo_hb_report_hits.restype = C.c_int64
        L.ixo_cm_marginals.argtypes = [C.POINTER(Cm), _f32p, _f32p]
        L.ixo_emit_map.argtypes = [C.POINTER(Cm), _i32p, _i32p]
        L.ixo_tr_validity.argtypes = [C.POINTER(Cm), _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _u8p, _u8p, _u8p, _u8p]
        L.ixo_trhb_scan.argtypes = [C.POINTER(Cm), C.POINTER(CmSc), _f32p, _f32p, _f32p, C.POINTER(Bands),
                                    C.POINTER(TrValid), _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                    C.c_int, C.c_float, _f32p, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), C.POINTER(TrHit), C.c_int64, _f32p]
        L.ixo_trhb_scan.restype = C.c_int64
        L.ixo_cm_configure_full.argtypes = [C.POINTER(Cm), C.c_int, C.POINTER(CmSc), C.POINTER(CmISc)]
        L.ixo_cm_free_iscores.argtypes = [C.POINTER(CmISc)]
        L.ixo_qdb_scan.argtypes = [C.POINTER(Cm), C.POINTER(CmSc), C.POINTER(CmISc), _i32p, _i32p, _u8p, C.c_int,
                                   C.c_int, C.c_int, C.c_float, C.c_int, C.POINTER(Hit), C.c_int64, _f32p]
        L.ixo_qdb_scan.restype = C.c_int64
        L.ixo_cyk_seq_filter.argtypes = [C.POINTER(Cm), C.POINTER(CmSc), _i32p, _i32p, _u8p, C.c_int, C.c_float,
                                         C.c_int, _i32p, _i32p, C.c_int64]
        L.ixo_cyk_seq_filter.restype = C.c_int64
        L.ixo_cm_imarginals.argtypes = [C.POINTER(Cm), _i32p, _i32p]
        L.ixo_trqdb_scan.argtypes = [C.POINTER(Cm), C.POINTER(CmSc), C.POINTER(CmISc), _f32p, _f32p, _i32p, _i32p,
                                     _f32p, _i32p, _i32p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_float, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float), C.POINTER(TrHit), C.c_int64, _f32p]
        L.ixo_trqdb_scan.restype = C.c_int64
        L.ixo_cm_psi.argtypes = [C.POINTER(Cm), C.POINTER(C.c_double)]
        L.ixo_trpenalties.argtypes = [C.POINTER(Cm), _f32p, _f32p, _i32p, _i32p]
        _ready = True
    return L


def cm_psi(cm):
    """cm_expected_state_occupancy of the un-localized CM (float64 [M])."""
    s = cm_struct(cm)
    psi = np.zeros(cm["M"], np.float64)
    lib().ixo_cm_psi(C.byref(s), psi.ctypes.data_as(C.POINTER(C.c_double)))
    return psi


def trpenalties(cm):
    """TrPenalties::new: (g, l, ig, il), each [3, M]."""
    s = cm_struct(cm)
    M = cm["M"]
    g, l = np.zeros((3, M), np.float32), np.zeros((3, M), np.float32)
    ig, il = np.zeros((3, M), np.int32), np.zeros((3, M), np.int32)
    lib().ixo_trpenalties(C.byref(s), g.ctypes.data_as(_f32p), l.ctypes.data_as(_f32p), ig.ctypes.data_as(_i32p),
                          il.ctypes.data_as(_i32p))
    return g, l, ig, il


def cm_struct(cm):
    keep = {k: np.ascontiguousarray(cm[k]) for k in ("sttype", "ndtype", "nodemap", "cfirst", "cnum", "t", "e")}
    s = Cm()
    s.M, s.nodes, s.clen, s.W = int(cm["M"]), int(cm["nodes"]), int(cm["clen"]), int(cm["W"])
    for i in range(4):
        s.null[i] = float(cm["null"][i])
    s.el_selfsc, s.pbegin, s.pend, s.n3_omega = float(cm["el_selfsc"]), float(cm["pbegin"]), float(cm["pend"]), \
        float(cm["n3_omega"])
    s.sttype = keep["sttype"].astype(np.int8).ctypes.data_as(_i8p)
    keep["sttype8"] = keep["sttype"].astype(np.int8)
    s.sttype = keep["sttype8"].ctypes.data_as(_i8p)
    keep["ndtype8"] = keep["ndtype"].astype(np.int8)
    s.ndtype = keep["ndtype8"].ctypes.data_as(_i8p)
    for k in ("nodemap", "cfirst", "cnum"):
        keep[k] = keep[k].astype(np.int32)
        setattr(s, k, keep[k].ctypes.data_as(_i32p))
    keep["t"] = keep["t"].astype(np.float32)
    keep["e"] = keep["e"].astype(np.float32)
    s.t = keep["t"].ctypes.data_as(_f32p)
    s.e = keep["e"].ctypes.data_as(_f32p)
    s._keep = keep
    return s


def bands_struct(b):
    keep = {k: np.ascontiguousarray(b[k], np.int32) for k in ("jmin", "jmax", "hd_off", "hdmin", "hdmax")}
    s = Bands()
    for k, v in keep.items():
        setattr(s, k, v.ctypes.data_as(_i32p))
    s._keep = keep
    return s


class OracleCM:
    def __init__(self, cm):
        self.cm = cm
        self.s = cm_struct(cm)
        self.sc = CmSc()
        lib().ixo_cm_configure(C.byref(self.s), C.byref(self.sc))
        M = cm["M"]
        self.tsc = np.ctypeslib.as_array(self.sc.tsc, shape=(M, 6)).copy()
        self.oesc = np.ctypeslib.as_array(self.sc.oesc, shape=(M, 18 * 18)).copy()
        self.beginsc = np.ctypeslib.as_array(self.sc.beginsc, shape=(M,)).copy()
        self.endsc = np.ctypeslib.as_array(self.sc.endsc, shape=(M,)).copy()
        self.el_selfsc = self.sc.el_selfsc

    def __del__(self):
        try:
            lib().ixo_cm_free_scores(C.byref(self.sc))
        except Exception:
            pass

    def ncells(self, b):
        return lib().ixo_hb_ncells(C.byref(bands_struct(b)), self.cm["M"])

    def scan(self, b, dsq, i0, j0, inside, env_cutoff=0.0, want_alpha=False):
        bs = bands_struct(b)
        n = self.ncells(b)
        root = np.zeros(max(1, n), np.float32)
        alpha = np.zeros(max(1, n), np.float32) if want_alpha else None
        vsc = C.c_float()
        ei, ej = C.c_int64(), C.c_int64()
        d = np.ascontiguousarray(dsq, np.uint8)
        nr = lib().ixo_hb_scan(C.byref(self.s), C.byref(self.sc), C.byref(bs), d.ctypes.data_as(C.POINTER(C.c_uint8)),
                               i0, j0, int(inside), env_cutoff, C.byref(vsc), C.byref(ei), C.byref(ej),
                               root.ctypes.data_as(_f32p), alpha.ctypes.data_as(_f32p) if want_alpha else None)
        out = dict(vsc=vsc.value, envi=ei.value, envj=ej.value, root=root[:nr].copy())
        if want_alpha:
            out["alpha"] = alpha[:n].copy()
        return out

    def hits(self, b, dsq, i0, j0, root, cutoff, do_null3=True):
        bs = bands_struct(b)
        d = np.ascontiguousarray(dsq, np.uint8)
        r = np.ascontiguousarray(root, np.float32)
        cap = 4096
        out = (Hit * cap)()
        n = lib().ixo_hb_report_hits(C.byref(self.s), C.byref(bs), d.ctypes.data_as(C.POINTER(C.c_uint8)), i0, j0,
                                     r.ctypes.data_as(_f32p), cutoff, int(do_null3), out, cap)
        return [(out[k].i, out[k].j, out[k].score, out[k].bias) for k in range(min(n, cap))]


def valid_struct(val):
    keep = {k: np.ascontiguousarray(val[k], np.uint8) for k in ("jvalid", "lvalid", "rvalid", "tvalid")}
    s = TrValid()
    for k, v in keep.items():
        setattr(s, k, v.ctypes.data_as(_u8p))
    s._keep = keep
    return s


class OracleTrCM(OracleCM):
    """OracleCM plus the truncated scanners' inputs and scan. `scores`
    (optional) replaces the configured score tables by hand-made ones:
    dict with tsc [M,6], oesc [M,324], beginsc, endsc [M], el_selfsc."""

    def __init__(self, cm, scores=None):
        super().__init__(cm)
        M = cm["M"]
        if scores is not None:
            self._own = {k: np.ascontiguousarray(scores[k], np.float32) for k in ("tsc", "oesc", "beginsc", "endsc")}
            self.sc2 = CmSc()
            self.sc2.M, self.sc2.local_begin, self.sc2.el_selfsc = M, 1, float(scores["el_selfsc"])
            for k, v in self._own.items():
                setattr(self.sc2, k, v.ctypes.data_as(_f32p))
        else:
            self.sc2 = self.sc
        self.lmesc = np.zeros((M, 18), np.float32)
        self.rmesc = np.zeros((M, 18), np.float32)
        lib().ixo_cm_marginals(C.byref(self.s), self.lmesc.ctypes.data_as(_f32p), self.rmesc.ctypes.data_as(_f32p))

    def emit_map(self):
        n = self.cm["nodes"]
        lp = np.zeros(n, np.int32)
        rp = np.zeros(n, np.int32)
        lib().ixo_emit_map(C.byref(self.s), lp.ctypes.data_as(_i32p), rp.ctypes.data_as(_i32p))
        return lp, rp

    def validity(self, pass_idx, sp1, sp2, ep1, ep2):
        lp, rp = self.emit_map()
        M = self.cm["M"]
        out = {k: np.zeros(M, np.uint8) for k in ("jvalid", "lvalid", "rvalid", "tvalid")}
        lib().ixo_tr_validity(C.byref(self.s), lp.ctypes.data_as(_i32p), rp.ctypes.data_as(_i32p), pass_idx, sp1,
                              sp2, ep1, ep2, *[out[k].ctypes.data_as(_u8p) for k in ("jvalid", "lvalid", "rvalid",
                                                                                      "tvalid")])
        return out

    def trscan(self, b, val, pty, dsq, i0, j0, pass_idx, local, inside, cutoff=0.0, do_null3=True, env_cutoff=0.0,
               want_planes=False, lmesc=None, rmesc=None):
        bs = bands_struct(b)
        vs = valid_struct(val)
        n = self.ncells(b)
        planes = np.zeros(max(4, 4 * n), np.float32) if want_planes else None
        p = np.ascontiguousarray(pty, np.float32)
        lm = np.ascontiguousarray(self.lmesc if lmesc is None else lmesc, np.float32)
        rm = np.ascontiguousarray(self.rmesc if rmesc is None else rmesc, np.float32)
        d = np.ascontiguousarray(dsq, np.uint8)
        vsc, vmode = C.c_float(), C.c_int()
        ei, ej = C.c_int64(), C.c_int64()
        cap = 8192
        hits = (TrHit * cap)()
        nh = lib().ixo_trhb_scan(C.byref(self.s), C.byref(self.sc2), lm.ctypes.data_as(_f32p),
                                 rm.ctypes.data_as(_f32p), p.ctypes.data_as(_f32p), C.byref(bs), C.byref(vs),
                                 d.ctypes.data_as(_u8p), i0, j0, pass_idx, int(local), int(inside), cutoff,
                                 int(do_null3), env_cutoff, C.byref(vsc), C.byref(vmode), C.byref(ei), C.byref(ej),
                                 hits, cap, planes.ctypes.data_as(_f32p) if want_planes else None)
        out = dict(vsc=vsc.value, vmode=vmode.value, envi=ei.value, envj=ej.value,
                   hits=[(hits[k].i, hits[k].j, hits[k].score, hits[k].bias, hits[k].mode, hits[k].root)
                         for k in range(min(nh, cap))])
        if want_planes:
            out["planes"] = planes[:4 * n].reshape(4, n).copy()
        return out


class OracleQdbCM:
    """Float and integer scores of a CM in local or global configuration and
    the QDB / non-banded scanners."""

    def __init__(self, cm, local=False):
        self.cm = cm
        self.s = cm_struct(cm)
        self.fs, self.isc = CmSc(), CmISc()
        lib().ixo_cm_configure_full(C.byref(self.s), int(local), C.byref(self.fs), C.byref(self.isc))
        M = cm["M"]
        self.tsc = np.ctypeslib.as_array(self.fs.tsc, shape=(M, 6)).copy()
        self.oesc = np.ctypeslib.as_array(self.fs.oesc, shape=(M, 324)).copy()
        self.beginsc = np.ctypeslib.as_array(self.fs.beginsc, shape=(M,)).copy()
        self.endsc = np.ctypeslib.as_array(self.fs.endsc, shape=(M,)).copy()
        self.el_selfsc = self.fs.el_selfsc
        self.itsc = np.ctypeslib.as_array(self.isc.itsc, shape=(M, 6)).copy()
        self.ioesc = np.ctypeslib.as_array(self.isc.ioesc, shape=(M, 324)).copy()
        self.ibeginsc = np.ctypeslib.as_array(self.isc.ibeginsc, shape=(M,)).copy()
        self.iendsc = np.ctypeslib.as_array(self.isc.iendsc, shape=(M,)).copy()
        self.iel_selfsc = self.isc.iel_selfsc

    def __del__(self):
        try:
            lib().ixo_cm_free_scores(C.byref(self.fs))
            lib().ixo_cm_free_iscores(C.byref(self.isc))
        except Exception:
            pass

    def scan(self, dsq, i0, j0, inside, dmin=None, dmax=None, cutoff=0.0, do_null3=True, want_root=False):
        d = np.ascontiguousarray(dsq, np.uint8)
        W = int(self.cm["W"])
        root = np.zeros((j0 - i0 + 1) * (W + 1), np.float32) if want_root else None
        keep = [np.ascontiguousarray(x, np.int32) if x is not None else None for x in (dmin, dmax)]
        cap = 1 << 16
        out = (Hit * cap)()
        n = lib().ixo_qdb_scan(C.byref(self.s), C.byref(self.fs), C.byref(self.isc),
                               keep[0].ctypes.data_as(_i32p) if keep[0] is not None else None,
                               keep[1].ctypes.data_as(_i32p) if keep[1] is not None else None,
                               d.ctypes.data_as(_u8p), i0, j0, int(inside), cutoff, int(do_null3), out, cap,
                               root.ctypes.data_as(_f32p) if want_root else None)
        hits = [(out[k].i, out[k].j, out[k].score, out[k].bias) for k in range(min(n, cap))]
        if want_root:
            return hits, root.reshape(j0 - i0 + 1, W + 1)
        return hits

    def seq_filter(self, dsq, n, dmin, dmax, cutoff, do_null3=True):
        d = np.ascontiguousarray(dsq, np.uint8)
        keep = [np.ascontiguousarray(x, np.int32) if x is not None else None for x in (dmin, dmax)]
        cap = 4096
        es, ee = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        k = lib().ixo_cyk_seq_filter(C.byref(self.s), C.byref(self.fs),
                                     keep[0].ctypes.data_as(_i32p) if keep[0] is not None else None,
                                     keep[1].ctypes.data_as(_i32p) if keep[1] is not None else None,
                                     d.ctypes.data_as(_u8p), n, cutoff, int(do_null3), es.ctypes.data_as(_i32p),
                                     ee.ctypes.data_as(_i32p), cap)
        return list(zip(es[:k].tolist(), ee[:k].tolist()))


PTY_INDEX = {2: 1, 3: 2, 4: 0, 5: 0}  # pass -> penalty row (cm_trunc.rs:99-107)


class OracleTrQdbCM(OracleQdbCM):
    """OracleQdbCM plus the QDB / non-banded truncated scanners
    (generic_tr_scan): float and integer marginal emissions and the
    truncated-begin penalties (float g/l and integer ig/il, [3, M])."""

    def __init__(self, cm, local=False, seed=0, pen=None):
        super().__init__(cm, local=local)
        M = cm["M"]
        self.lmesc = np.zeros((M, 18), np.float32)
        self.rmesc = np.zeros((M, 18), np.float32)
        self.ilmesc = np.zeros((M, 18), np.int32)
        self.irmesc = np.zeros((M, 18), np.int32)
        lib().ixo_cm_marginals(C.byref(self.s), self.lmesc.ctypes.data_as(_f32p), self.rmesc.ctypes.data_as(_f32p))
        lib().ixo_cm_imarginals(C.byref(self.s), self.ilmesc.ctypes.data_as(_i32p), self.irmesc.ctypes.data_as(_i32p))
        if pen is None:
            from infernox_amd import synth_cm as S
            g, l = S.tr_penalties(cm, seed)
            ig, il = S.tr_ipenalties(cm, seed)
            pen = dict(g=g, l=l, ig=ig, il=il)
        self.pen = pen

    def trscan(self, dsq, i0, j0, inside, pass_idx, local, dmax=None, cutoff=0.0, do_null3=True, want_best=False):
        d = np.ascontiguousarray(dsq, np.uint8)
        W = int(self.cm["W"])
        row = PTY_INDEX[pass_idx]
        pty = np.ascontiguousarray(self.pen["l" if local else "g"][row], np.float32)
        ipty = np.ascontiguousarray(self.pen["il" if local else "ig"][row], np.int32)
        dm = np.ascontiguousarray(dmax, np.int32) if dmax is not None else None
        best = np.zeros((j0 - i0 + 1) * (W + 1), np.float32) if want_best else None
        vsc, vmode, bfull = C.c_float(), C.c_int(), C.c_float()
        cap = 1 << 14
        hits = (TrHit * cap)()
        n = lib().ixo_trqdb_scan(C.byref(self.s), C.byref(self.fs), C.byref(self.isc),
                                 self.lmesc.ctypes.data_as(_f32p), self.rmesc.ctypes.data_as(_f32p),
                                 self.ilmesc.ctypes.data_as(_i32p), self.irmesc.ctypes.data_as(_i32p),
                                 pty.ctypes.data_as(_f32p), ipty.ctypes.data_as(_i32p),
                                 dm.ctypes.data_as(_i32p) if dm is not None else None, d.ctypes.data_as(_u8p),
                                 int(i0), int(j0), int(inside), int(pass_idx), int(local), float(cutoff),
                                 int(do_null3), C.byref(vsc), C.byref(vmode), C.byref(bfull), hits, cap,
                                 best.ctypes.data_as(_f32p) if want_best else None)
        out = dict(vsc=vsc.value, vmode=vmode.value, bsc_full=bfull.value,
                   hits=[(hits[k].i, hits[k].j, hits[k].score, hits[k].bias, hits[k].mode, hits[k].root)
                         for k in range(min(n, cap))])
        if want_best:
            out["best"] = best.reshape(j0 - i0 + 1, W + 1)
        return out
