"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test
infrastructure only: the checker, never the thing measured."""
import ctypes as C
import os
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")


class P7(C.Structure):
    _fields_ = [("M", C.c_int), ("mat", C.POINTER(C.c_float)), ("trans", C.POINTER(C.c_float)),
                ("compo", C.c_float * 4), ("lmmu", C.c_double), ("lmlambda", C.c_double),
                ("lvmu", C.c_double), ("lvlambda", C.c_double), ("lftau", C.c_double), ("lflambda", C.c_double)]


class Msv(C.Structure):
    _fields_ = [("m", C.c_int), ("scale_b", C.c_float), ("base_b", C.c_uint8), ("bias_b", C.c_uint8),
                ("tbm_b", C.c_uint8), ("tec_b", C.c_uint8), ("tjb_b", C.c_uint8), ("rbv", C.POINTER(C.c_uint8)),
                ("mmu", C.c_double), ("mlambda", C.c_double), ("prefix", C.POINTER(C.c_float)),
                ("suffix", C.POINTER(C.c_float)), ("max_length", C.c_int)]


_i16p = C.POINTER(C.c_int16)


class Vit(C.Structure):
    _fields_ = [("m", C.c_int), ("scale_w", C.c_float), ("base_w", C.c_int16), ("rwv", _i16p),
                ("tmm", _i16p), ("tim", _i16p), ("tdm", _i16p), ("vbm", _i16p), ("tmi", _i16p), ("tii", _i16p),
                ("tmd", _i16p), ("tdd", _i16p), ("xw_e_move", C.c_int16), ("xw_e_loop", C.c_int16),
                ("q", C.c_int), ("tw", _i16p), ("rw", _i16p), ("ddbound_w", C.c_int16)]


_f32p = C.POINTER(C.c_float)


class Fwd(C.Structure):
    _fields_ = [("m", C.c_int), ("rfv", _f32p), ("tbm", _f32p), ("amm", _f32p), ("aim", _f32p), ("adm", _f32p),
                ("tmi", _f32p), ("tii", _f32p), ("tmd", _f32p), ("tdd", _f32p), ("ftau", C.c_double),
                ("flambda", C.c_double), ("nq", C.c_int), ("tmain", _f32p), ("tddv", _f32p), ("rfvv", _f32p)]


class Bias(C.Structure):
    _fields_ = [("eo", (C.c_float * 2) * 18), ("t1", C.c_float * 3), ("pi", C.c_float * 2)]


class RawWin(C.Structure):
    _fields_ = [("n", C.c_int64), ("length", C.c_int64), ("k", C.c_int64), ("score", C.c_float)]


class Window(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64), ("score", C.c_float)]


class Acct(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_past_msv", "pos_past_msv", "n_past_msvbias", "pos_past_msvbias",
        "n_past_vit", "pos_past_vit", "n_past_vitbias", "pos_past_vitbias",
        "n_past_fwd", "pos_past_fwd", "n_past_fwdbias", "pos_past_fwdbias")]


class F3Opts(C.Structure):
    _fields_ = [("f1", C.c_double), ("f1b", C.c_double), ("f2", C.c_double), ("f2b", C.c_double),
                ("f3", C.c_double), ("f3b", C.c_double), ("cmw", C.c_int64), ("maxw", C.c_int64), ("opts", C.c_int)]


_lib = None
# Lazy set-up runs once under this lock: ctypes creates a function object on
# first attribute access, so two threads configuring at once can leave a
# caller holding one whose argtypes are not yet set.
INIT_LOCK = threading.RLock()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with INIT_LOCK:
        return _init()


def _init():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        u8p, i64 = C.POINTER(C.c_uint8), C.c_int64
        L.ixo_build_msv.argtypes = [C.POINTER(P7), C.c_int, C.POINTER(Msv)]
        L.ixo_build_vit.argtypes = [C.POINTER(P7), C.POINTER(Vit)]
        L.ixo_build_fwd.argtypes = [C.POINTER(P7), C.POINTER(Fwd)]
        L.ixo_build_bias.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(Bias)]
        L.ixo_free_msv.argtypes = [C.POINTER(Msv)]
        L.ixo_free_vit.argtypes = [C.POINTER(Vit)]
        L.ixo_free_fwd.argtypes = [C.POINTER(Fwd)]
        L.ixo_ssv_diagonals.argtypes = [C.POINTER(Msv), u8p, i64, C.c_double, C.POINTER(RawWin), i64]
        L.ixo_ssv_diagonals.restype = i64
        L.ixo_ssv_thresh.argtypes = [C.POINTER(Msv), C.c_double]
        L.ixo_ssv_rowmax.argtypes = [C.POINTER(Msv), u8p, i64, u8p]
        L.ixo_f1_msv_windows.argtypes = [C.POINTER(Msv), u8p, i64, C.c_double, i64, C.POINTER(Window), i64]
        L.ixo_f1_msv_windows.restype = i64
        L.ixo_msv_score.argtypes = [C.POINTER(Msv), u8p, i64, C.c_int, _f32p]
        L.ixo_vit_score_scalar.argtypes = [C.POINTER(Vit), u8p, i64, _f32p]
        L.ixo_vit_score_striped.argtypes = [C.POINTER(Vit), u8p, i64, _f32p]
        L.ixo_fwd_score_striped.argtypes = [C.POINTER(Fwd), u8p, i64]
        L.ixo_fwd_score_striped.restype = C.c_float
        L.ixo_fwd_score_scalar.argtypes = [C.POINTER(Fwd), u8p, i64]
        L.ixo_fwd_score_scalar.restype = C.c_float
        f32p = C.POINTER(C.c_float)
        L.ixo_p7_forward.argtypes = [C.POINTER(Fwd), u8p, i64, C.c_int, f32p, f32p, f32p, f32p, f32p]
        L.ixo_p7_forward.restype = C.c_float
        L.ixo_bias_score.argtypes = [C.POINTER(Bias), u8p, i64]
        L.ixo_bias_score.restype = C.c_float
        L.ixo_f3_filter_sequence.argtypes = [C.POINTER(Msv), C.POINTER(Vit), C.POINTER(Fwd), C.POINTER(Bias),
                                             C.POINTER(P7), u8p, i64, C.POINTER(F3Opts), C.POINTER(Window), i64,
                                             C.POINTER(Acct)]
        L.ixo_f3_filter_sequence.restype = i64
        for nm in ("ixo_gumbel_invsurv", "ixo_gumbel_surv", "ixo_exp_surv"):
            getattr(L, nm).argtypes = [C.c_double, C.c_double, C.c_double]
            getattr(L, nm).restype = C.c_double
        L.ixo_ilogsum.argtypes = [C.c_int, C.c_int]
        L.ixo_ilogsum_table.argtypes = [C.c_int]
        L.ixo_flogsum.argtypes = [C.c_float, C.c_float]
        L.ixo_flogsum.restype = C.c_float
        L.ixo_flogsum_table.argtypes = [C.c_int]
        L.ixo_flogsum_table.restype = C.c_float
        _lib = L
    return _lib


def u8(a):
    return np.ascontiguousarray(a, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))


class Oracle:
    """All p7 filter structures of the oracle for one profile."""

    def __init__(self, p, max_length):
        self.mat = np.ascontiguousarray(p["mat"], np.float32)
        self.trans = np.ascontiguousarray(p["trans"], np.float32)
        s = P7()
        s.M = int(p["M"])
        s.mat = self.mat.ctypes.data_as(_f32p)
        s.trans = self.trans.ctypes.data_as(_f32p)
        for i in range(4):
            s.compo[i] = float(p["compo"][i])
        for k in ("lmmu", "lmlambda", "lvmu", "lvlambda", "lftau", "lflambda"):
            setattr(s, k, float(p[k]))
        self.p7 = s
        L = lib()
        self.smsv, self.svit, self.sfwd, self.sbias = Msv(), Vit(), Fwd(), Bias()
        L.ixo_build_msv(C.byref(s), int(max_length), C.byref(self.smsv))
        L.ixo_build_vit(C.byref(s), C.byref(self.svit))
        L.ixo_build_fwd(C.byref(s), C.byref(self.sfwd))
        comp = np.ascontiguousarray(p["compo"], np.float32)
        L.ixo_build_bias(comp.ctypes.data_as(_f32p), s.M, C.byref(self.sbias))
        self.M = s.M

    def __del__(self):
        try:
            L = lib()
            L.ixo_free_msv(C.byref(self.smsv))
            L.ixo_free_vit(C.byref(self.svit))
            L.ixo_free_fwd(C.byref(self.sfwd))
        except Exception:
            pass

    def ssv_diagonals(self, dsq, L, f1):
        n = lib().ixo_ssv_diagonals(C.byref(self.smsv), u8(dsq), L, f1, None, 0)
        out = (RawWin * max(n, 1))()
        lib().ixo_ssv_diagonals(C.byref(self.smsv), u8(dsq), L, f1, out, n)
        return [(out[i].n, out[i].length, out[i].k, out[i].score) for i in range(n)]

    def ssv_rowmax(self, dsq, L):
        xe = np.zeros(L + 2, np.uint8)
        lib().ixo_ssv_rowmax(C.byref(self.smsv), u8(dsq), L, xe.ctypes.data_as(C.POINTER(C.c_uint8)))
        return xe[1:L + 1]

    def msv_fields(self):
        m = self.smsv
        M = m.m
        rbv = np.ctypeslib.as_array(m.rbv, shape=((M + 1) * 18,)).copy()
        pre = np.ctypeslib.as_array(m.prefix, shape=(M + 1,)).copy()
        suf = np.ctypeslib.as_array(m.suffix, shape=(M + 1,)).copy()
        return dict(scale_b=m.scale_b, base_b=m.base_b, bias_b=m.bias_b, tbm_b=m.tbm_b, tec_b=m.tec_b,
                    tjb_b=m.tjb_b, rbv=rbv, prefix=pre, suffix=suf)

    def ssv_thresh(self, f1):
        return lib().ixo_ssv_thresh(C.byref(self.smsv), f1)

    def msv_score(self, dsq, L, wlen=0):
        v = C.c_float()
        r = lib().ixo_msv_score(C.byref(self.smsv), u8(dsq), L, wlen, C.byref(v))
        return None if r else v.value

    def vit_scalar(self, dsq, L):
        v = C.c_float()
        r = lib().ixo_vit_score_scalar(C.byref(self.svit), u8(dsq), L, C.byref(v))
        return None if r else v.value

    def vit_striped(self, dsq, L):
        v = C.c_float()
        r = lib().ixo_vit_score_striped(C.byref(self.svit), u8(dsq), L, C.byref(v))
        return None if r else v.value

    def fwd(self, dsq, L):
        return lib().ixo_fwd_score_striped(C.byref(self.sfwd), u8(dsq), L)

    def p7_forward(self, dsq, L, unihit=False):
        """Full-matrix Forward: dict(score, totscale, mmx, imx, dmx [(L+1), M+1], xmx [(L+1), 6])."""
        M = self.M
        f32p = C.POINTER(C.c_float)
        mm, im, dm = (np.zeros((L + 1) * (M + 1), np.float32) for _ in range(3))
        xm = np.zeros((L + 1) * 6, np.float32)
        ts = C.c_float()
        sc = lib().ixo_p7_forward(C.byref(self.sfwd), u8(dsq), L, int(unihit), mm.ctypes.data_as(f32p),
                                  im.ctypes.data_as(f32p), dm.ctypes.data_as(f32p), xm.ctypes.data_as(f32p),
                                  C.byref(ts))
        shp = (L + 1, M + 1)
        return dict(score=sc, totscale=ts.value, mmx=mm.reshape(shp), imx=im.reshape(shp), dmx=dm.reshape(shp),
                    xmx=xm.reshape(L + 1, 6))

    def fwd_scalar(self, dsq, L):
        return lib().ixo_fwd_score_scalar(C.byref(self.sfwd), u8(dsq), L)

    def bias_score(self, dsq, L):
        return lib().ixo_bias_score(C.byref(self.sbias), u8(dsq), L)

    def f3_filter_sequence(self, dsq, L, opts):
        n = lib().ixo_f3_filter_sequence(C.byref(self.smsv), C.byref(self.svit), C.byref(self.sfwd),
                                         C.byref(self.sbias), C.byref(self.p7), u8(dsq), L, C.byref(opts),
                                         None, 0, None)
        out = (Window * max(n, 1))()
        acct = Acct()
        lib().ixo_f3_filter_sequence(C.byref(self.smsv), C.byref(self.svit), C.byref(self.sfwd), C.byref(self.sbias),
                                     C.byref(self.p7), u8(dsq), L, C.byref(opts), out, n, C.byref(acct))
        return [(out[i].start, out[i].end, out[i].score) for i in range(n)], acct
