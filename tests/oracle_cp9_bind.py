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
