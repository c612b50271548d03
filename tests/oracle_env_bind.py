"""ctypes binding of the oracle's F4/F5 envelope definition (oracle/envdef.c).
Test infrastructure only: the checker, never the thing measured."""
import ctypes as C

import numpy as np

import oracle_bind as OB

_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)


class Gm(C.Structure):
    _fields_ = [("M", C.c_int), ("tsc", _f32p), ("rsc", _f32p), ("xsc", (C.c_float * 2) * 4), ("nj", C.c_float),
                ("L", C.c_int), ("local_end", C.c_int)]


class Domain(C.Structure):
    _fields_ = [("ienv", C.c_int), ("jenv", C.c_int), ("envsc", C.c_float), ("domcorrection", C.c_float)]


class EnvOpts(C.Structure):
    _fields_ = [("f4", C.c_double), ("f4b", C.c_double), ("f5", C.c_double), ("f5b", C.c_double),
                ("gfmu", C.c_double), ("gflambda", C.c_double), ("rt1", C.c_float), ("rt2", C.c_float),
                ("rt3", C.c_float), ("ns", C.c_int), ("do_gfwdbias", C.c_int), ("do_edefbias", C.c_int),
                ("lftau", C.c_double), ("lflambda", C.c_double), ("clen", C.c_int32), ("pad", C.c_int32)]


class EnvAcct(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_past_gfwd", "pos_past_gfwd", "n_past_gfwdbias", "pos_past_gfwdbias",
        "n_past_edef", "pos_past_edef", "n_past_edefbias", "pos_past_edefbias")]


_ready = False


def lib():
    global _ready
    L = OB.lib()
    with OB.INIT_LOCK:
        if not _ready:
            i64 = C.c_int64
            L.ixo_p7_flogsum.argtypes = [C.c_float, C.c_float]
            L.ixo_p7_flogsum.restype = C.c_float
            L.ixo_p7_flogsum_table.argtypes = [C.c_int]
            L.ixo_p7_flogsum_table.restype = C.c_float
            L.ixo_gm_build.argtypes = [C.POINTER(OB.P7), C.c_int, C.POINTER(Gm)]
            L.ixo_gm_free.argtypes = [C.POINTER(Gm)]
            for nm in ("ixo_gm_reconfig_length", "ixo_gm_reconfig_unihit", "ixo_gm_reconfig_multihit"):
                getattr(L, nm).argtypes = [C.POINTER(Gm), C.c_int]
            for nm in ("ixo_gforward", "ixo_gbackward"):
                getattr(L, nm).argtypes = [_u8p, C.c_int, C.POINTER(Gm), _f32p, _f32p]
                getattr(L, nm).restype = C.c_float
            L.ixo_gdomain_decoding.argtypes = [C.POINTER(Gm), C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p]
            L.ixo_domaindef_glocal.argtypes = [C.POINTER(Gm), _u8p, C.c_int, _f32p, _f32p, C.c_float, C.c_float,
                                               C.c_float, C.c_int, C.c_int, C.POINTER(Domain), C.c_int]
            L.ixo_env_def_std.argtypes = [C.POINTER(Gm), C.POINTER(OB.Bias), _u8p, i64, C.POINTER(OB.Window), i64,
                                          C.POINTER(EnvOpts), C.POINTER(C.c_int32), C.POINTER(C.c_int32), i64,
                                          C.POINTER(EnvAcct)]
            L.ixo_env_def_std.restype = i64
            L.ixo_env_def_pass.argtypes = [C.POINTER(Gm), C.c_int, C.POINTER(OB.Bias), _u8p, i64,
                                           C.POINTER(OB.Window), i64, C.POINTER(EnvOpts), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32), i64, C.POINTER(EnvAcct)]
            L.ixo_env_def_pass.restype = i64
            L.ixo_gm_build_trunc.argtypes = [C.POINTER(OB.P7), C.c_int, C.POINTER(Gm)]
            _ready = True
    return L


def opts(f4=0.005, f4b=0.005, f5=0.005, f5b=0.005, gfmu=-5.0, gflambda=0.69, rt1=0.25, rt2=0.10, rt3=0.20, ns=200,
         do_gfwdbias=True, do_edefbias=False, lftau=-4.0, lflambda=0.69, clen=1):
    o = EnvOpts()
    o.f4, o.f4b, o.f5, o.f5b, o.gfmu, o.gflambda = f4, f4b, f5, f5b, gfmu, gflambda
    o.rt1, o.rt2, o.rt3, o.ns = rt1, rt2, rt3, ns
    o.do_gfwdbias, o.do_edefbias = int(bool(do_gfwdbias)), int(bool(do_edefbias))
    o.lftau, o.lflambda, o.clen = lftau, lflambda, int(clen)
    return o


class OracleEnv:
    """The glocal profile of a p7 HMM dict and the oracle's F4/F5 functions."""

    def __init__(self, p, L=100):
        self.base = OB.Oracle(p, max(L, 10))
        self.gm = Gm()
        lib().ixo_gm_build(C.byref(self.base.p7), int(L), C.byref(self.gm))
        self.M = int(p["M"])

    def __del__(self):
        try:
            lib().ixo_gm_free(C.byref(self.gm))
        except Exception:
            pass

    def config(self, L, unihit=False):
        if unihit:
            lib().ixo_gm_reconfig_unihit(C.byref(self.gm), int(L))
        else:
            lib().ixo_gm_reconfig_multihit(C.byref(self.gm), int(L))

    def xsc(self):
        return np.array([[self.gm.xsc[s][t] for t in range(2)] for s in range(4)], np.float32)

    def tables(self):
        M = self.M
        tsc = np.ctypeslib.as_array(self.gm.tsc, shape=((M + 1) * 8,)).reshape(M + 1, 8).copy()
        rsc = np.ctypeslib.as_array(self.gm.rsc, shape=(18 * (M + 1) * 2,)).reshape(18, M + 1, 2).copy()
        return tsc, rsc

    def _run(self, fn, dsq, L):
        M = self.M
        dp = np.zeros((L + 1) * (M + 1) * 3, np.float32)
        xm = np.zeros((L + 1) * 5, np.float32)
        sc = fn(OB.u8(dsq), int(L), C.byref(self.gm), dp.ctypes.data_as(_f32p), xm.ctypes.data_as(_f32p))
        return sc, xm.reshape(L + 1, 5), dp.reshape(L + 1, 3, M + 1)

    def gforward(self, dsq, L):
        return self._run(lib().ixo_gforward, dsq, L)

    def gbackward(self, dsq, L):
        return self._run(lib().ixo_gbackward, dsq, L)

    def decoding(self, L, fx, bx):
        out = [np.zeros(L + 1, np.float32) for _ in range(3)]
        fx = np.ascontiguousarray(fx, np.float32)
        bx = np.ascontiguousarray(bx, np.float32)
        lib().ixo_gdomain_decoding(C.byref(self.gm), int(L), fx.ctypes.data_as(_f32p), bx.ctypes.data_as(_f32p),
                                   *[a.ctypes.data_as(_f32p) for a in out])
        return out

    def domaindef(self, dsq, L, fx, bx, rt1=0.25, rt2=0.10, rt3=0.20, ns=200, unihit=False):
        fx = np.ascontiguousarray(fx, np.float32)
        bx = np.ascontiguousarray(bx, np.float32)
        out = (Domain * 256)()
        n = lib().ixo_domaindef_glocal(C.byref(self.gm), OB.u8(dsq), int(L), fx.ctypes.data_as(_f32p),
                                       bx.ctypes.data_as(_f32p), rt1, rt2, rt3, ns, int(unihit), out, 256)
        return [(out[i].ienv, out[i].jenv, out[i].envsc) for i in range(min(n, 256))]

    def env_def(self, dsq, L, windows, o, pass_idx=1):
        """trunc_env_def of one chunk for pass 1 (STD, gm) or a truncated
        pass 2 / 3 / 4 (Rgm / Lgm / Tgm built here)."""
        n = len(windows)
        arr = (OB.Window * max(n, 1))()
        for k, (a, b) in enumerate(windows):
            arr[k].start, arr[k].end, arr[k].score = int(a), int(b), 0.0
        cap = 64 + 8 * n
        es, ee = (C.c_int32 * cap)(), (C.c_int32 * cap)()
        acct = EnvAcct()
        gm = self.gm
        if pass_idx != 1:
            gm = Gm()
            lib().ixo_gm_build_trunc(C.byref(self.base.p7), int(pass_idx) - 2, C.byref(gm))
        ne = lib().ixo_env_def_pass(C.byref(gm), int(pass_idx), C.byref(self.base.sbias), OB.u8(dsq), int(L), arr, n,
                                    C.byref(o), es, ee, cap, C.byref(acct))
        if pass_idx != 1:
            lib().ixo_gm_free(C.byref(gm))
        return [(es[i], ee[i]) for i in range(min(ne, cap))], acct
