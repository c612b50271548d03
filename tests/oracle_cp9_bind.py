"""ctypes binding of the CP9 banding DPs of the CPU oracle
(oracle/liboracle.so: cp9_forward, cp9_backward, cp9_fb2hmmbands).
Test infrastructure only."""
import ctypes as C

import numpy as np

from oracle_bind import INIT_LOCK, lib as _lib

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class Cp9(C.Structure):
    _fields_ = [("M", C.c_int), ("el", C.c_int), ("el_selfsc", C.c_int32), ("otsc", _i32p), ("msc", _i32p),
                ("isc", _i32p), ("has_el", _u8p), ("el_from_ct", _i32p), ("el_from_off", _i32p),
                ("el_from_idx", _i32p)]


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
        dp = [C.POINTER(Cp9), _u8p, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i64p]
        for f in (L.ixo_cp9_forward, L.ixo_cp9_backward):
            f.argtypes = dp
            f.restype = C.c_float
        L.ixo_cp9_band_thresh.argtypes = [C.c_double]
        L.ixo_cp9_band_thresh.restype = C.c_int32
        L.ixo_cp9_fb2hmmbands.argtypes = [C.POINTER(Cp9), _u8p] + [_i32p] * 6 + [C.c_int, C.c_int, C.c_double,
                                                                                   C.c_int, _i32p]
        L.ixo_cp9_occupancy.argtypes = [C.POINTER(Cp9), _u8p] + [_i32p] * 4 + [C.c_int, C.c_int, C.c_int, _i32p]
        _ready = True
    return L


def _p(a, t=_i32p):
    return a.ctypes.data_as(t)


class OracleCP9:
    def __init__(self, h):
        self.h = h
        self.keep = {k: np.ascontiguousarray(h[k], np.uint8 if k == "has_el" else np.int32)
                     for k in ("otsc", "msc", "isc", "has_el", "el_from_ct", "el_from_off", "el_from_idx")}
        s = Cp9()
        s.M, s.el, s.el_selfsc = int(h["M"]), int(h["el"]), int(h["el_selfsc"])
        for k, v in self.keep.items():
            setattr(s, k, _p(v, _u8p if k == "has_el" else _i32p))
        self.s = s

    def _dp(self, f, dsq, i0, j0, scan, align):
        M, L = self.h["M"], j0 - i0 + 1
        mx = [np.zeros((L + 1) * (M + 1), np.int32) for _ in range(4)]
        sca = np.zeros(L + 1, np.int32)
        bp = C.c_int64()
        d = np.ascontiguousarray(dsq, np.uint8)
        sc = f(C.byref(self.s), _p(d, _u8p), int(i0), int(j0), int(scan), int(align), *(_p(m) for m in mx),
               _p(sca), C.byref(bp))
        shp = (L + 1, M + 1)
        return dict(sc=sc, pos=bp.value, mmx=mx[0].reshape(shp), imx=mx[1].reshape(shp), dmx=mx[2].reshape(shp),
                    elmx=mx[3].reshape(shp), sca=sca)

    def forward(self, dsq, i0, j0, scan=True, align=False):
        return self._dp(lib().ixo_cp9_forward, dsq, i0, j0, scan, align)

    def backward(self, dsq, i0, j0, scan=True, align=False):
        return self._dp(lib().ixo_cp9_backward, dsq, i0, j0, scan, align)

    def bands(self, dsq, i0, j0, tau, f=None, b=None, scan=True):
        """HMM bands [6, M+1]: min_m, max_m, min_i, max_i, min_d, max_d (-1 unset)."""
        f = f or self.forward(dsq, i0, j0, scan, not scan)
        b = b or self.backward(dsq, i0, j0, scan, not scan)
        pn = np.zeros(6 * (self.h["M"] + 1), np.int32)
        d = np.ascontiguousarray(dsq, np.uint8)
        mats = [np.ascontiguousarray(x) for x in (f["mmx"], f["imx"], f["dmx"], b["mmx"], b["imx"], b["dmx"])]
        lib().ixo_cp9_fb2hmmbands(C.byref(self.s), _p(d, _u8p), *(_p(m) for m in mats), int(i0), int(j0),
                                  1.0 - tau, int(scan), _p(pn))
        return pn.reshape(6, self.h["M"] + 1)


    def trunc_pass_bands(self, dsq, i0, j0, tau, pass_idx):
        """cp9_seq2bands_trunc's DPs for a truncated pass: (pn [6, M+1], occ
        [M+1]); the Forward / Backward scan unless the pass forces the first
        / final residue."""
        fscan = pass_idx not in (2, 4)
        bscan = pass_idx not in (3, 4)
        f = self.forward(dsq, i0, j0, fscan, False)
        b = self.backward(dsq, i0, j0, bscan, False)
        pn = np.zeros(6 * (self.h["M"] + 1), np.int32)
        d = np.ascontiguousarray(dsq, np.uint8)
        mats = [np.ascontiguousarray(x) for x in (f["mmx"], f["imx"], f["dmx"], b["mmx"], b["imx"], b["dmx"])]
        lib().ixo_cp9_fb2hmmbands(C.byref(self.s), _p(d, _u8p), *(_p(m) for m in mats), int(i0), int(j0),
                                  1.0 - tau, int(fscan), _p(pn))
        occ = np.zeros(self.h["M"] + 1, np.int32)
        lib().ixo_cp9_occupancy(C.byref(self.s), _p(d, _u8p), _p(mats[0]), _p(mats[2]), _p(mats[3]), _p(mats[5]),
                                int(i0), int(j0), int(fscan), _p(occ))
        return pn.reshape(6, self.h["M"] + 1), occ


class TrBand(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("sp1", "sp2", "ep1", "ep2", "rmarg_imin", "rmarg_imax", "lmarg_jmin",
                                         "lmarg_jmax")] + [(n, _u8p) for n in ("jvalid", "lvalid", "rvalid", "tvalid")]


# ---- HMM -> CM band conversion (oracle/hmmband.c) ----------------------------

class HbCm(C.Structure):
    _fields_ = [("M", C.c_int), ("nodes", C.c_int), ("sttype", C.POINTER(C.c_int8)), ("ndtype", C.POINTER(C.c_int8)),
                ("nodemap", _i32p), ("cfirst", _i32p), ("cnum", _i32p), ("local_begin", C.c_int),
                ("local_end", C.c_int)]


class HbCp9(C.Structure):
    _fields_ = [("hmm_m", C.c_int), ("local_begin", C.c_int), ("local_end", C.c_int), ("el", C.c_int),
                ("esc", _i32p), ("tdm", _i32p), ("tim", _i32p), ("el_from_ct", _i32p), ("el_from_off", _i32p),
                ("el_from_idx", _i32p)]


class OracleHb:
    """The oracle's band conversion for a CM dict and a CP9 dict."""

    def __init__(self, cm, h, cm_local=True, cp9_local=True):
        L = lib()
        L.ixo_hmm2ij_bands.argtypes = [C.POINTER(HbCm), C.POINTER(HbCp9), _i32p, C.c_int, C.c_int, C.c_int] + [_i32p] * 4
        L.ixo_hb_ij2d.argtypes = [C.POINTER(HbCm)] + [_i32p] * 4 + [C.c_int, _i32p, _i32p, _i32p, C.c_int64]
        L.ixo_hb_ij2d.restype = C.c_int64
        L.ixo_hb_mx_mb.argtypes = [C.POINTER(HbCm)] + [_i32p] * 5 + [C.c_int, C.POINTER(C.c_int64)]
        L.ixo_hb_mx_mb.restype = C.c_double
        self.M = int(cm["M"])
        k = self.keep = dict(sttype=np.ascontiguousarray(cm["sttype"], np.int8),
                             ndtype=np.ascontiguousarray(cm["ndtype"], np.int8),
                             nodemap=np.ascontiguousarray(cm["nodemap"], np.int32),
                             cfirst=np.ascontiguousarray(cm["cfirst"], np.int32),
                             cnum=np.ascontiguousarray(cm["cnum"], np.int32))
        c = HbCm()
        c.M, c.nodes = self.M, int(cm["nodes"])
        c.sttype = k["sttype"].ctypes.data_as(C.POINTER(C.c_int8))
        c.ndtype = k["ndtype"].ctypes.data_as(C.POINTER(C.c_int8))
        for f in ("nodemap", "cfirst", "cnum"):
            setattr(c, f, _p(k[f]))
        c.local_begin = c.local_end = int(cm_local)
        otsc = np.asarray(h["otsc"], np.int32).reshape(-1, 12)
        for f, col in (("esc", 10), ("tdm", 2), ("tim", 1)):
            k[f] = np.ascontiguousarray(otsc[:, col], np.int32)
        for f in ("el_from_ct", "el_from_off", "el_from_idx"):
            k[f] = np.ascontiguousarray(h[f], np.int32)
        p = HbCp9()
        p.hmm_m = int(h["M"])
        p.local_begin = p.local_end = int(cp9_local)
        p.el = int(bool(h["el"]))
        for f in ("esc", "tdm", "tim", "el_from_ct", "el_from_off", "el_from_idx"):
            setattr(p, f, _p(k[f]))
        self.c, self.p = c, p

    def hmm2ij(self, pn, i0, j0, search=True):
        pn = np.ascontiguousarray(pn, np.int32).copy()
        out = [np.zeros(self.M, np.int32) for _ in range(4)]
        lib().ixo_hmm2ij_bands(C.byref(self.c), C.byref(self.p), _p(pn), int(i0), int(j0), int(search),
                               *[_p(a) for a in out])
        return (*out, pn)

    def hmm2ij_trunc(self, pn, occ, i0, j0, pass_idx):
        """cp9_seq2bands_trunc's steps 3-4: (imin, imax, jmin, jmax, jvalid,
        lvalid, rvalid, tvalid, [sp1 sp2 ep1 ep2 rmarg_imin rmarg_imax
        lmarg_jmin lmarg_jmax])."""
        L = lib()
        L.ixo_tr_predict_start_end.argtypes = [_i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                               C.POINTER(TrBand)]
        L.ixo_tr_candidates.argtypes = [C.POINTER(HbCm), C.c_int, C.POINTER(TrBand)]
        L.ixo_hmm2ij_bands_tr.argtypes = [C.POINTER(HbCm), C.POINTER(HbCp9), _i32p, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(TrBand)] + [_i32p] * 4
        pn = np.ascontiguousarray(pn, np.int32).copy()
        occ = np.ascontiguousarray(occ, np.int32)
        val = [np.zeros(self.M, np.uint8) for _ in range(4)]
        tb = TrBand()
        tb.jvalid, tb.lvalid, tb.rvalid, tb.tvalid = [_p(v, _u8p) for v in val]
        L.ixo_tr_predict_start_end(_p(pn), _p(occ), int(self.p.hmm_m), int(i0), int(j0), 0.01, 0.98, C.byref(tb))
        L.ixo_tr_candidates(C.byref(self.c), int(pass_idx), C.byref(tb))
        out = [np.zeros(self.M, np.int32) for _ in range(4)]
        L.ixo_hmm2ij_bands_tr(C.byref(self.c), C.byref(self.p), _p(pn), int(i0), int(j0), 1, C.byref(tb),
                              *[_p(a) for a in out])
        spep = [tb.sp1, tb.sp2, tb.ep1, tb.ep2, tb.rmarg_imin, tb.rmarg_imax, tb.lmarg_jmin, tb.lmarg_jmax]
        return (*out, *val, spep)

    def ij2d(self, imin, imax, jmin, jmax, trunc=False):
        a = [np.ascontiguousarray(x, np.int32) for x in (imin, imax, jmin, jmax)]
        hd_off = np.zeros(self.M + 1, np.int32)
        n = lib().ixo_hb_ij2d(C.byref(self.c), *[_p(x) for x in a], int(trunc), _p(hd_off), None, None, 0)
        hdmin, hdmax = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        lib().ixo_hb_ij2d(C.byref(self.c), *[_p(x) for x in a], int(trunc), _p(hd_off), _p(hdmin), _p(hdmax), n)
        return dict(jmin=a[2], jmax=a[3], hd_off=hd_off, hdmin=hdmin[:n], hdmax=hdmax[:n])

    def mx_mb(self, b, L):
        cells = C.c_int64()
        mb = lib().ixo_hb_mx_mb(C.byref(self.c), *[_p(np.ascontiguousarray(b[f], np.int32))
                                                   for f in ("jmin", "jmax", "hd_off", "hdmin", "hdmax")], int(L),
                                C.byref(cells))
        return mb, cells.value
