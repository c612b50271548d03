"""ctypes binding of libixg.so (the C ABI declared in include/ixg.h).

Thin plumbing for tests and the benchmark: every call goes straight to the
HIP library. There is no CPU fallback; loading fails loudly when the built
library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libixg.so")

IXG_OK, IXG_EINVAL, IXG_ENOMEM, IXG_EDEVICE, IXG_ERANGE = 0, -1, -2, -3, -4
DO_MSV, DO_FWD, DO_FWDBIAS, DO_VIT, DO_VITBIAS, DO_MSVBIAS = 1, 2, 4, 8, 16, 32


class P7Hmm(C.Structure):
    _fields_ = [("M", C.c_int32), ("mat", C.POINTER(C.c_float)), ("trans", C.POINTER(C.c_float)),
                ("compo", C.c_float * 4), ("lmmu", C.c_double), ("lmlambda", C.c_double),
                ("lvmu", C.c_double), ("lvlambda", C.c_double), ("lftau", C.c_double),
                ("lflambda", C.c_double)]


class RawWin(C.Structure):
    _fields_ = [("n", C.c_int64), ("length", C.c_int64), ("k", C.c_int64), ("score", C.c_float),
                ("task", C.c_int32)]


class Window(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64), ("score", C.c_float), ("task", C.c_int32)]


class Acct(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_past_msv", "pos_past_msv", "n_past_msvbias", "pos_past_msvbias",
        "n_past_vit", "pos_past_vit", "n_past_vitbias", "pos_past_vitbias",
        "n_past_fwd", "pos_past_fwd", "n_past_fwdbias", "pos_past_fwdbias")]


class F3Opts(C.Structure):
    _fields_ = [("f1", C.c_double), ("f1b", C.c_double), ("f2", C.c_double), ("f2b", C.c_double),
                ("f3", C.c_double), ("f3b", C.c_double), ("cmw", C.c_int64), ("maxw", C.c_int64),
                ("opts", C.c_int32)]


class Task(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64), ("strand", C.c_int32), ("ctx", C.c_int32)]


# every symbol include/ixg.h declares
EXPORTS = [
    "ixg_init", "ixg_shutdown", "ixg_last_error", "ixg_mem_in_use", "ixg_p7_create", "ixg_p7_free",
    "ixg_p7_msv_tables", "ixg_p7_vit_tables", "ixg_p7_fwd_tables", "ixg_ssv_diagonals",
    "ixg_vit_scores", "ixg_fwd_scores", "ixg_bias_scores", "ixg_msv_scores",
    "ixg_f3_filter_sequence", "ixg_target_create", "ixg_target_free", "ixg_window_tasks",
    "ixg_f3_filter_tasks", "ixg_last_timings", "ixg_cm_create", "ixg_cm_free", "ixg_cm_tables",
    "ixg_cyk_scan_hb", "ixg_inside_scan_hb", "ixg_hb_scan_batch", "ixg_cm_set_trpenalties", "ixg_cm_trtables",
    "ixg_trcyk_scan_hb", "ixg_trinside_scan_hb", "ixg_trhb_scan_batch", "ixg_kernel_timings",
    "ixg_copy_probe", "ixg_cm_create_cfg", "ixg_cm_itables", "ixg_qdb_scan", "ixg_qdb_scan_batch",
    "ixg_cyk_seq_filter", "ixg_p7_forward", "ixg_p7_forward_batch", "ixg_cm_set_itrpenalties", "ixg_cm_imtables",
    "ixg_trqdb_scan", "ixg_trqdb_scan_batch", "ixg_cp9_create", "ixg_cp9_free", "ixg_cp9_bands",
    "ixg_cp9_bands_batch", "ixg_server_create", "ixg_server_submit", "ixg_server_wait", "ixg_server_result",
    "ixg_server_stats", "ixg_server_free", "ixg_hb_scan_multi", "ixg_fwd_scores_noalloc", "ixg_gfb_scores",
    "ixg_env_def", "ixg_hbmap_create", "ixg_hbmap_free", "ixg_hmm2ij_bands", "ixg_ij2d_bands", "ixg_search_std", "ixg_search_stats",
]
KERNELS = ["ssv_sweep", "ssv_resolve", "fwd", "bias", "vit", "msv", "hb_scan", "trhb_scan", "copy_probe",
           "qdb_scan", "fwd_full", "trqdb_scan", "cp9", "gfwd", "gbwd", "gfb"]


class CmParams(C.Structure):
    _fields_ = [("M", C.c_int32), ("nodes", C.c_int32), ("clen", C.c_int32), ("W", C.c_int32),
                ("null", C.c_float * 4), ("el_selfsc", C.c_float), ("pbegin", C.c_float), ("pend", C.c_float),
                ("n3_omega", C.c_double), ("sttype", C.POINTER(C.c_int8)), ("ndtype", C.POINTER(C.c_int8)),
                ("nodemap", C.POINTER(C.c_int32)), ("cfirst", C.POINTER(C.c_int32)), ("cnum", C.POINTER(C.c_int32)),
                ("t", C.POINTER(C.c_float)), ("e", C.POINTER(C.c_float))]


class Bands(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_int32)) for n in ("jmin", "jmax", "hd_off", "hdmin", "hdmax")]


class Hit(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32), ("score", C.c_float), ("bias", C.c_float), ("env", C.c_int32)]


class Envelope(C.Structure):
    _fields_ = [("i0", C.c_int32), ("j0", C.c_int32), ("strand", C.c_int32), ("seq", C.c_int32), ("bands", Bands)]


class QdbWin(C.Structure):
    _fields_ = [("i0", C.c_int32), ("j0", C.c_int32), ("strand", C.c_int32), ("seq", C.c_int32)]


class TrValid(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_uint8)) for n in ("jvalid", "lvalid", "rvalid", "tvalid")]


class TrHit(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32), ("score", C.c_float), ("bias", C.c_float), ("mode", C.c_int32),
                ("root", C.c_int32), ("env", C.c_int32)]


class TrEnvelope(C.Structure):
    _fields_ = [("i0", C.c_int32), ("j0", C.c_int32), ("strand", C.c_int32), ("seq", C.c_int32), ("bands", Bands),
                ("valid", TrValid)]

_lib = None


def lib():
    """Load libixg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, f32p, u8p = C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        i64p = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        L.ixg_last_error.restype = C.c_char_p
        L.ixg_init.argtypes = [C.c_int, C.c_size_t, C.POINTER(vp)]
        L.ixg_shutdown.argtypes = [vp]
        L.ixg_mem_in_use.argtypes = [vp]
        L.ixg_mem_in_use.restype = C.c_size_t
        L.ixg_p7_create.argtypes = [vp, C.POINTER(P7Hmm), i32, C.POINTER(vp)]
        L.ixg_p7_free.argtypes = [vp, vp]
        L.ixg_p7_msv_tables.argtypes = [vp, u8p, u8p, f32p, f32p]
        L.ixg_p7_vit_tables.argtypes = [vp, C.POINTER(C.c_int16), C.POINTER(C.c_int16), C.POINTER(C.c_int16)]
        L.ixg_p7_fwd_tables.argtypes = [vp, C.POINTER(C.c_int32), f32p, f32p, f32p]
        L.ixg_ssv_diagonals.argtypes = [vp, vp, u8p, i64, C.c_double, C.POINTER(RawWin), i64, i64p]
        for nm in ("ixg_vit_scores", "ixg_msv_scores"):
            getattr(L, nm).argtypes = [vp, vp, u8p, i64, i64p, i64p, i64, f32p, u8p]
        for nm in ("ixg_fwd_scores", "ixg_bias_scores", "ixg_fwd_scores_noalloc"):
            getattr(L, nm).argtypes = [vp, vp, u8p, i64, i64p, i64p, i64, f32p]
        L.ixg_f3_filter_sequence.argtypes = [vp, vp, u8p, i64, C.POINTER(F3Opts), C.POINTER(Window), i64, i64p,
                                             C.POINTER(Acct)]
        L.ixg_target_create.argtypes = [vp, u8p, i64, C.POINTER(vp)]
        L.ixg_target_free.argtypes = [vp, vp]
        L.ixg_window_tasks.argtypes = [i64, i64, C.c_int, C.c_int, C.POINTER(Task), i64]
        L.ixg_window_tasks.restype = i64
        L.ixg_f3_filter_tasks.argtypes = [vp, vp, vp, C.POINTER(Task), i64, C.POINTER(F3Opts), C.POINTER(Window),
                                          i64, i64p, C.POINTER(Acct)]
        L.ixg_last_timings.argtypes = [vp, C.POINTER(C.c_double), i64p]
        L.ixg_kernel_timings.argtypes = [vp, C.POINTER(C.c_double), i64p, C.c_int]
        L.ixg_copy_probe.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.ixg_p7_forward.argtypes = [vp, vp, u8p, i64, C.c_int, f32p, f32p, f32p, f32p, f32p, f32p]
        L.ixg_p7_forward_batch.argtypes = [vp, vp, u8p, i64, i64p, i64p, i64, C.c_int, f32p, f32p, f32p, f32p]
        L.ixg_cm_create_cfg.argtypes = [vp, C.POINTER(CmParams), C.c_int, C.POINTER(vp)]
        L.ixg_cm_itables.argtypes = [vp, i32p, i32p, i32p, i32p, i32p]
        L.ixg_qdb_scan.argtypes = [vp, vp, u8p, i64, i32, i32, i32p, i32p, C.c_int, C.c_float, C.c_int,
                                   C.POINTER(Hit), i64, i64p]
        L.ixg_qdb_scan_batch.argtypes = [vp, vp, vp, C.POINTER(QdbWin), i64, i32p, i32p, C.c_int, C.c_float,
                                         C.c_int, C.POINTER(Hit), i64, i64p]
        L.ixg_cyk_seq_filter.argtypes = [vp, vp, u8p, i64, i32p, i32p, C.c_float, C.c_int, i32p, i32p, i64, i64p]
        L.ixg_cm_create.argtypes = [vp, C.POINTER(CmParams), C.POINTER(vp)]
        L.ixg_cm_free.argtypes = [vp, vp]
        L.ixg_cm_tables.argtypes = [vp, f32p, f32p, f32p, f32p, f32p]
        L.ixg_cyk_scan_hb.argtypes = [vp, vp, C.POINTER(Bands), u8p, i64, i32, i32, C.c_float, f32p, i64p, i64p]
        L.ixg_inside_scan_hb.argtypes = [vp, vp, C.POINTER(Bands), u8p, i64, i32, i32, C.c_int, C.c_float, C.c_float,
                                         C.c_int, f32p, i64p, i64p, C.POINTER(Hit), i64, i64p]
        L.ixg_hb_scan_batch.argtypes = [vp, vp, vp, C.POINTER(Envelope), i64, C.c_int, C.c_float, C.c_int,
                                        C.c_float, C.c_int, f32p, i64p, i64p, C.POINTER(Hit), i64, i64p, i64p]
        L.ixg_cm_set_trpenalties.argtypes = [vp, vp, f32p, f32p]
        L.ixg_cm_set_itrpenalties.argtypes = [vp, vp, i32p, i32p]
        L.ixg_cm_imtables.argtypes = [vp, i32p, i32p]
        L.ixg_trqdb_scan.argtypes = [vp, vp, u8p, i64, i32, i32, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                     f32p, i32p, f32p, C.POINTER(TrHit), i64, i64p]
        L.ixg_trqdb_scan_batch.argtypes = [vp, vp, vp, C.POINTER(QdbWin), i64, i32p, C.c_int, C.c_int, C.c_int,
                                           C.c_float, C.c_int, f32p, i32p, f32p, C.POINTER(TrHit), i64, i64p]
        L.ixg_cm_trtables.argtypes = [vp, f32p, f32p, i32p, i32p]
        for nm in ("ixg_trcyk_scan_hb", "ixg_trinside_scan_hb"):
            getattr(L, nm).argtypes = [vp, vp, C.POINTER(Bands), C.POINTER(TrValid), u8p, i64, i32, i32, C.c_float,
                                       C.c_int, C.c_int, C.c_float, C.c_int, f32p, i32p, i64p, i64p,
                                       C.POINTER(TrHit), i64, i64p]
        L.ixg_trhb_scan_batch.argtypes = [vp, vp, vp, C.POINTER(TrEnvelope), i64, C.c_int, C.c_int, C.c_int,
                                          C.c_float, C.c_float, C.c_int, f32p, i32p, i64p, i64p, C.POINTER(TrHit),
                                          i64, i64p, i64p]
        L.ixg_hb_scan_multi.argtypes = [vp, C.POINTER(vp), C.c_int32, i32p, vp, C.POINTER(Envelope), i64, C.c_int,
                                        C.c_float, C.c_int, C.c_float, C.c_int, f32p, i64p, i64p, C.POINTER(Hit),
                                        i64, i64p, i64p]
        L.ixg_server_create.argtypes = [vp, C.c_int, C.c_size_t, i64, i64, C.POINTER(vp)]
        L.ixg_server_submit.argtypes = [vp, vp, u8p, i64, C.POINTER(F3Opts), C.POINTER(C.c_uint64)]
        L.ixg_server_wait.argtypes = [vp, C.c_uint64, i64, C.POINTER(C.c_int)]
        L.ixg_server_result.argtypes = [vp, C.c_uint64, C.POINTER(Task), i64, i64p, C.POINTER(Window), i64, i64p,
                                        C.POINTER(Acct), C.POINTER(C.c_double)]
        L.ixg_server_stats.argtypes = [vp, i64p]
        L.ixg_server_free.argtypes = [vp]
        _lib = L
    return _lib


class IxgError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != IXG_OK:
        raise IxgError(f"{what}: status {rc}: {lib().ixg_last_error().decode()}")


def u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def f32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def i64ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Context:
    """One device, one HIP stream (ixg_init / ixg_shutdown)."""

    def __init__(self, device=0, mem_budget=0):
        self.h = C.c_void_p()
        check(lib().ixg_init(device, mem_budget, C.byref(self.h)), "ixg_init")

    def close(self):
        if self.h:
            check(lib().ixg_shutdown(self.h), "ixg_shutdown")
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def mem_in_use(self):
        return lib().ixg_mem_in_use(self.h)

    def set_autotune(self, mode):
        """ixg_set_autotune: 1 autotune the F3 Forward's launch shape, 0 fixed
        order, 2 + v force shape v (0 stripe per wave, 1 four lanes per
        window, 2 generic)."""
        check(lib().ixg_set_autotune(self.h, int(mode)), "ixg_set_autotune")

    def autotune_choices(self):
        """[(stripe count, log2 window-count class, shape)] chosen so far."""
        q, a, b, n = (C.c_int32 * 256)(), (C.c_int32 * 256)(), (C.c_int32 * 256)(), C.c_int64()
        i32p = C.POINTER(C.c_int32)
        lib().ixg_autotune_choices.argtypes = [C.c_void_p, i32p, i32p, i32p, C.c_int64, C.POINTER(C.c_int64)]
        check(lib().ixg_autotune_choices(self.h, q, a, b, 256, C.byref(n)), "ixg_autotune_choices")
        return [(q[i], a[i], b[i]) for i in range(min(n.value, 256))]

    def timings(self):
        ms = (C.c_double * 5)()
        cells = (C.c_int64 * 5)()
        check(lib().ixg_last_timings(self.h, ms, cells), "timings")
        return list(ms), list(cells)

    def copy_probe(self, nbytes, reps=3):
        """Mean milliseconds of a plain device copy of nbytes."""
        ms = C.c_double()
        check(lib().ixg_copy_probe(self.h, int(nbytes), int(reps), C.byref(ms)), "copy probe")
        return ms.value

    def kernel_timings(self, reset=False):
        """{kernel: (total ms, launches)} since the last reset."""
        ms = (C.c_double * len(KERNELS))()
        n = (C.c_int64 * len(KERNELS))()
        check(lib().ixg_kernel_timings(self.h, ms, n, int(reset)), "kernel timings")
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNELS)}


def p7_struct(p7):
    """ctypes P7Hmm view of a dict {M, mat (M+1,4), trans (M+1,7), compo, ev...}; keeps arrays alive."""
    mat = np.ascontiguousarray(p7["mat"], dtype=np.float32)
    trans = np.ascontiguousarray(p7["trans"], dtype=np.float32)
    s = P7Hmm()
    s.M = int(p7["M"])
    s.mat = f32ptr(mat)
    s.trans = f32ptr(trans)
    for i in range(4):
        s.compo[i] = float(p7["compo"][i])
    for k in ("lmmu", "lmlambda", "lvmu", "lvlambda", "lftau", "lflambda"):
        setattr(s, k, float(p7[k]))
    s._keep = (mat, trans)
    return s


class P7:
    """Device-resident p7 filter profile (ixg_p7_create)."""

    def __init__(self, ctx, p7, max_length):
        self.ctx = ctx
        self.M = int(p7["M"])
        self.s = p7_struct(p7)
        self.h = C.c_void_p()
        check(lib().ixg_p7_create(ctx.h, C.byref(self.s), int(max_length), C.byref(self.h)), "ixg_p7_create")

    def free(self):
        if self.h:
            check(lib().ixg_p7_free(self.ctx.h, self.h), "ixg_p7_free")
            self.h = C.c_void_p()

    def msv_tables(self):
        M = self.M
        rbv = np.zeros((M + 1) * 18, np.uint8)
        b = np.zeros(9, np.uint8)
        pre = np.zeros(M + 1, np.float32)
        suf = np.zeros(M + 1, np.float32)
        check(lib().ixg_p7_msv_tables(self.h, u8ptr(rbv), u8ptr(b), f32ptr(pre), f32ptr(suf)))
        return rbv, b, pre, suf

    def vit_tables(self):
        M = self.M
        rwv = np.zeros((M + 1) * 18, np.int16)
        tr = np.zeros(8 * (M + 1), np.int16)
        ddb = np.zeros(1, np.int16)
        P = C.POINTER(C.c_int16)
        check(lib().ixg_p7_vit_tables(self.h, rwv.ctypes.data_as(P), tr.ctypes.data_as(P), ddb.ctypes.data_as(P)))
        return rwv, tr, int(ddb[0])

    def fwd_tables(self):
        nq = C.c_int32()
        check(lib().ixg_p7_fwd_tables(self.h, C.byref(nq), None, None, None))
        n = nq.value
        tm = np.zeros(n * 28, np.float32)
        td = np.zeros(n * 4, np.float32)
        rf = np.zeros(18 * n * 4, np.float32)
        check(lib().ixg_p7_fwd_tables(self.h, C.byref(nq), f32ptr(tm), f32ptr(td), f32ptr(rf)))
        return n, tm, td, rf

    def ssv_diagonals(self, dsq, L, f1, cap=None):
        """dsq: uint8 array, 1-indexed (len >= L+2). Returns list of (n, length, k, score)."""
        dsq = np.ascontiguousarray(dsq, np.uint8)
        cap = cap if cap is not None else max(16, L // 4 + 16)
        while True:
            out = (RawWin * cap)()
            n = C.c_int64()
            rc = lib().ixg_ssv_diagonals(self.ctx.h, self.h, u8ptr(dsq), int(L), float(f1), out, cap, C.byref(n))
            if rc == IXG_ERANGE and n.value > cap:
                cap = n.value
                continue
            check(rc, "ixg_ssv_diagonals")
            return [(out[i].n, out[i].length, out[i].k, out[i].score) for i in range(n.value)]

    def _scores(self, name, seq, offs, lens, with_ovf):
        seq = np.ascontiguousarray(seq, np.uint8)
        offs = np.ascontiguousarray(offs, np.int64)
        lens = np.ascontiguousarray(lens, np.int64)
        n = len(offs)
        sc = np.zeros(n, np.float32)
        args = [self.ctx.h, self.h, u8ptr(seq), len(seq), i64ptr(offs), i64ptr(lens), n, f32ptr(sc)]
        ovf = np.zeros(n, np.uint8)
        if with_ovf:
            args.append(u8ptr(ovf))
        check(getattr(lib(), name)(*args), name)
        return (sc, ovf) if with_ovf else sc

    def fwd_scores(self, seq, offs, lens):
        return self._scores("ixg_fwd_scores", seq, offs, lens, False)

    def fwd_scores_noalloc(self, seq, offs, lens):
        """fwd_scores through the entry point that allocates no device memory."""
        return self._scores("ixg_fwd_scores_noalloc", seq, offs, lens, False)

    def bias_scores(self, seq, offs, lens):
        return self._scores("ixg_bias_scores", seq, offs, lens, False)

    def vit_scores(self, seq, offs, lens):
        return self._scores("ixg_vit_scores", seq, offs, lens, True)

    def msv_scores(self, seq, offs, lens):
        return self._scores("ixg_msv_scores", seq, offs, lens, True)

    def p7_forward(self, dsq, L, unihit=False):
        """Full-matrix Forward: dict(score, totscale, mmx, imx, dmx [(L+1), M+1], xmx [(L+1), 6])."""
        M = self.M
        d = np.ascontiguousarray(dsq, np.uint8)
        mm, im, dm = (np.zeros((L + 1) * (M + 1), np.float32) for _ in range(3))
        xm = np.zeros((L + 1) * 6, np.float32)
        ts, sc = np.zeros(1, np.float32), np.zeros(1, np.float32)
        check(lib().ixg_p7_forward(self.ctx.h, self.h, u8ptr(d), int(L), int(unihit), f32ptr(mm), f32ptr(im),
                                   f32ptr(dm), f32ptr(xm), f32ptr(ts), f32ptr(sc)), "ixg_p7_forward")
        shp = (L + 1, M + 1)
        return dict(score=float(sc[0]), totscale=float(ts[0]), mmx=mm.reshape(shp), imx=im.reshape(shp),
                    dmx=dm.reshape(shp), xmx=xm.reshape(L + 1, 6))

    def p7_forward_batch(self, seq, regions, unihit=False):
        """Full-matrix Forward over regions [(off, len)] of one buffer: list of
        dicts as p7_forward returns."""
        M = self.M
        s = np.ascontiguousarray(seq, np.uint8)
        off = np.int64([r[0] for r in regions])
        ln = np.int64([r[1] for r in regions])
        n = len(regions)
        mat = np.zeros(max(1, int(np.sum(3 * (ln + 1) * (M + 1)))), np.float32)
        xm = np.zeros(max(1, int(np.sum(6 * (ln + 1)))), np.float32)
        sc, ts = np.zeros(max(1, n), np.float32), np.zeros(max(1, n), np.float32)
        check(lib().ixg_p7_forward_batch(self.ctx.h, self.h, u8ptr(s), len(s), i64ptr(off), i64ptr(ln), n,
                                         int(unihit), f32ptr(mat), f32ptr(xm), f32ptr(sc), f32ptr(ts)),
              "ixg_p7_forward_batch")
        out, mo, xo = [], 0, 0
        for r in range(n):
            L = int(ln[r])
            pl = (L + 1) * (M + 1)
            shp = (L + 1, M + 1)
            out.append(dict(score=float(sc[r]), totscale=float(ts[r]), mmx=mat[mo:mo + pl].reshape(shp),
                            imx=mat[mo + pl:mo + 2 * pl].reshape(shp), dmx=mat[mo + 2 * pl:mo + 3 * pl].reshape(shp),
                            xmx=xm[xo:xo + 6 * (L + 1)].reshape(L + 1, 6)))
            mo += 3 * pl
            xo += 6 * (L + 1)
        return out

    def f3_filter_sequence(self, dsq, L, opts, cap=None):
        dsq = np.ascontiguousarray(dsq, np.uint8)
        cap = cap if cap is not None else max(64, L // 16 + 64)
        while True:
            out = (Window * cap)()
            n = C.c_int64()
            acct = Acct()
            rc = lib().ixg_f3_filter_sequence(self.ctx.h, self.h, u8ptr(dsq), int(L), C.byref(opts), out, cap,
                                              C.byref(n), C.byref(acct))
            if rc == IXG_ERANGE and n.value > cap:
                cap = n.value
                continue
            check(rc, "ixg_f3_filter_sequence")
            return [(out[i].start, out[i].end, out[i].score) for i in range(n.value)], acct


class Target:
    """Device-resident digitized sequence with its reverse complement."""

    def __init__(self, ctx, dsq, L):
        self.ctx = ctx
        self.L = int(L)
        self.dsq = np.ascontiguousarray(dsq, np.uint8)
        self.h = C.c_void_p()
        check(lib().ixg_target_create(ctx.h, u8ptr(self.dsq), self.L, C.byref(self.h)), "ixg_target_create")

    def free(self):
        if self.h:
            check(lib().ixg_target_free(self.ctx.h, self.h), "ixg_target_free")
            self.h = C.c_void_p()


class MultiTarget(Target):
    """Several device-resident digitized sequences (a database), each with its
    reverse complement; seqs is a list of 1-indexed dsq arrays."""

    def __init__(self, ctx, seqs):
        self.ctx = ctx
        self.dsqs = [np.ascontiguousarray(d, np.uint8) for d in seqs]
        self.lens = [len(d) - 2 for d in self.dsqs]
        self.L = self.lens[0]
        n = len(self.dsqs)
        ptrs = (C.POINTER(C.c_uint8) * n)(*[u8ptr(d) for d in self.dsqs])
        lens = (C.c_int64 * n)(*self.lens)
        self.h = C.c_void_p()
        L = lib()
        L.ixg_target_create_multi.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_int64),
                                              C.c_int64, C.POINTER(C.c_void_p)]
        check(L.ixg_target_create_multi(ctx.h, ptrs, lens, n, C.byref(self.h)), "ixg_target_create_multi")


def window_tasks(L, maxw, do_top=True, do_bot=True):
    n = lib().ixg_window_tasks(int(L), int(maxw), int(do_top), int(do_bot), None, 0)
    arr = (Task * max(n, 1))()
    lib().ixg_window_tasks(int(L), int(maxw), int(do_top), int(do_bot), arr, n)
    return arr, n


def f3_filter_tasks(ctx, p7, tgt, tasks, ntasks, opts, cap=None):
    cap = cap if cap is not None else max(1024, tgt.L // 8)
    while True:
        out = (Window * cap)()
        n = C.c_int64()
        accts = (Acct * max(ntasks, 1))()
        rc = lib().ixg_f3_filter_tasks(ctx.h, p7.h, tgt.h, tasks, int(ntasks), C.byref(opts), out, cap, C.byref(n),
                                       accts)
        if rc == IXG_ERANGE and n.value > cap:
            cap = n.value
            continue
        check(rc, "ixg_f3_filter_tasks")
        break
    wins = [(out[i].task, out[i].start, out[i].end, out[i].score) for i in range(n.value)]
    return wins, accts


class Server:
    """Continuous-batching request loop over nstreams worker streams (ixg_server_*).

    submit() queues (model, sequence) and returns a ticket; result(ticket)
    waits and returns (tasks, windows, accts, times_ms) where windows are
    (task, start, end, score) with the task index local to the request and
    times_ms = (queue wait, run, latency, requests in the batch)."""

    def __init__(self, ctx, nstreams=4, mem_budget=0, max_inflight=1 << 26, max_batch=1 << 24):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().ixg_server_create(ctx.h, int(nstreams), int(mem_budget), int(max_inflight), int(max_batch),
                                      C.byref(self.h)), "ixg_server_create")

    def submit(self, p7, dsq, L, opts):
        d = np.ascontiguousarray(dsq, np.uint8)
        t = C.c_uint64()
        check(lib().ixg_server_submit(self.h, p7.h, u8ptr(d), int(L), C.byref(opts), C.byref(t)),
              "ixg_server_submit")
        return t.value

    def done(self, ticket, timeout_us=0):
        d = C.c_int()
        check(lib().ixg_server_wait(self.h, ticket, int(timeout_us), C.byref(d)), "ixg_server_wait")
        return bool(d.value)

    def result(self, ticket, tcap=64, wcap=1024):
        nt, nw = C.c_int64(), C.c_int64()
        while True:
            tasks, wins, accts = (Task * tcap)(), (Window * wcap)(), (Acct * tcap)()
            times = (C.c_double * 4)()
            rc = lib().ixg_server_result(self.h, ticket, tasks, tcap, C.byref(nt), wins, wcap, C.byref(nw), accts,
                                         times)
            if rc == IXG_ERANGE:
                tcap, wcap = max(tcap, nt.value), max(wcap, nw.value)
                continue
            check(rc, "ixg_server_result")
            break
        w = [(wins[i].task, wins[i].start, wins[i].end, wins[i].score) for i in range(nw.value)]
        tl = [(tasks[i].start, tasks[i].end, tasks[i].strand) for i in range(nt.value)]
        return tl, w, [accts[i] for i in range(nt.value)], tuple(times)

    def submit_search(self, bundle, dsq, L, opts):
        """A whole-search request against a SearchModel bundle (ixg_server_submit_search)."""
        d = np.ascontiguousarray(dsq, np.uint8)
        t = C.c_uint64()
        L_ = _search_lib()
        L_.ixg_server_submit_search.argtypes = [C.c_void_p, C.POINTER(SearchModel.S), C.POINTER(C.c_uint8), C.c_int64,
                                                C.POINTER(SearchOpts), C.POINTER(C.c_uint64)]
        check(L_.ixg_server_submit_search(self.h, C.byref(bundle.s), u8ptr(d), int(L), C.byref(opts), C.byref(t)),
              "ixg_server_submit_search")
        return t.value

    def search_result(self, ticket, cap=64):
        """([(start, stop, strand, score, bias, evalue, pass)], times_ms) of a search request."""
        L_ = _search_lib()
        L_.ixg_server_search_result.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(SearchHit), C.c_int64,
                                                C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        n = C.c_int64()
        while True:
            out, times = (SearchHit * max(cap, 1))(), (C.c_double * 4)()
            rc = L_.ixg_server_search_result(self.h, ticket, out, cap, C.byref(n), times)
            if rc == IXG_ERANGE:
                cap = n.value
                continue
            check(rc, "ixg_server_search_result")
            break
        return [(out[i].start, out[i].stop, out[i].strand, out[i].score, out[i].bias, out[i].evalue,
                 out[i].pass_idx) for i in range(n.value)], tuple(times)

    def stats(self):
        s = (C.c_int64 * 4)()
        check(lib().ixg_server_stats(self.h, s), "ixg_server_stats")
        return dict(batches=s[0], requests=s[1], largest_batch=s[2], residues=s[3])

    def free(self):
        if self.h:
            check(lib().ixg_server_free(self.h), "ixg_server_free")
            self.h = C.c_void_p()


def cm_params(cm):
    """ctypes CmParams view of a synth_cm CM dict (keeps its arrays alive)."""
    keep = dict(sttype=np.ascontiguousarray(cm["sttype"], np.int8), ndtype=np.ascontiguousarray(cm["ndtype"], np.int8),
                nodemap=np.ascontiguousarray(cm["nodemap"], np.int32), cfirst=np.ascontiguousarray(cm["cfirst"], np.int32),
                cnum=np.ascontiguousarray(cm["cnum"], np.int32), t=np.ascontiguousarray(cm["t"], np.float32),
                e=np.ascontiguousarray(cm["e"], np.float32))
    s = CmParams()
    s.M, s.nodes, s.clen, s.W = int(cm["M"]), int(cm["nodes"]), int(cm["clen"]), int(cm["W"])
    for i in range(4):
        s.null[i] = float(cm["null"][i])
    s.el_selfsc, s.pbegin, s.pend, s.n3_omega = float(cm["el_selfsc"]), float(cm["pbegin"]), float(cm["pend"]), \
        float(cm["n3_omega"])
    s.sttype = keep["sttype"].ctypes.data_as(C.POINTER(C.c_int8))
    s.ndtype = keep["ndtype"].ctypes.data_as(C.POINTER(C.c_int8))
    for k in ("nodemap", "cfirst", "cnum"):
        setattr(s, k, keep[k].ctypes.data_as(C.POINTER(C.c_int32)))
    s.t = f32ptr(keep["t"])
    s.e = f32ptr(keep["e"])
    s._keep = keep
    return s


def cm_trpenalties(cm):
    """TrPenalties::new of the CM (ixg_cm_trpenalties, host only): (g, l,
    ig, il), each [3, M]."""
    p = cm_params(cm)
    M = int(cm["M"])
    g, l = np.zeros((3, M), np.float32), np.zeros((3, M), np.float32)
    ig, il = np.zeros((3, M), np.int32), np.zeros((3, M), np.int32)
    i32p = C.POINTER(C.c_int32)
    check(lib().ixg_cm_trpenalties(C.byref(p), f32ptr(g), f32ptr(l), ig.ctypes.data_as(i32p), il.ctypes.data_as(i32p)),
          "ixg_cm_trpenalties")
    return g, l, ig, il


def bands_struct(b):
    keep = {k: np.ascontiguousarray(b[k], np.int32) for k in ("jmin", "jmax", "hd_off", "hdmin", "hdmax")}
    s = Bands()
    for k, v in keep.items():
        setattr(s, k, v.ctypes.data_as(C.POINTER(C.c_int32)))
    s._keep = keep
    return s


class CM:
    """Device-resident configured CM (ixg_cm_create); ctx=None builds host tables only."""

    def __init__(self, ctx, cm, local=True):
        self.ctx = ctx
        self.M = int(cm["M"])
        self.p = cm_params(cm)
        self.h = C.c_void_p()
        check(lib().ixg_cm_create_cfg(ctx.h if ctx else None, C.byref(self.p), int(local), C.byref(self.h)),
              "ixg_cm_create_cfg")

    def free(self):
        if self.h:
            check(lib().ixg_cm_free(self.ctx.h if self.ctx else None, self.h), "ixg_cm_free")
            self.h = C.c_void_p()

    def tables(self):
        M = self.M
        tsc, oesc = np.zeros(M * 6, np.float32), np.zeros(M * 324, np.float32)
        bs, es, el = np.zeros(M, np.float32), np.zeros(M, np.float32), np.zeros(1, np.float32)
        check(lib().ixg_cm_tables(self.h, f32ptr(tsc), f32ptr(oesc), f32ptr(bs), f32ptr(es), f32ptr(el)))
        return tsc.reshape(M, 6), oesc.reshape(M, 324), bs, es, float(el[0])

    def cyk_scan(self, bands, dsq, L, i0, j0, env_cutoff=0.0):
        d = np.ascontiguousarray(dsq, np.uint8)
        vsc, ei, ej = C.c_float(), C.c_int64(), C.c_int64()
        b = bands_struct(bands)
        check(lib().ixg_cyk_scan_hb(self.ctx.h, self.h, C.byref(b), u8ptr(d), int(L), int(i0), int(j0),
                                    float(env_cutoff), C.byref(vsc), C.byref(ei), C.byref(ej)), "ixg_cyk_scan_hb")
        return vsc.value, ei.value, ej.value

    def scan_hits(self, bands, dsq, L, i0, j0, inside=True, env_cutoff=0.0, cutoff=0.0, null3=True, cap=4096):
        d = np.ascontiguousarray(dsq, np.uint8)
        vsc, ei, ej, n = C.c_float(), C.c_int64(), C.c_int64(), C.c_int64()
        b = bands_struct(bands)
        hits = (Hit * cap)()
        check(lib().ixg_inside_scan_hb(self.ctx.h, self.h, C.byref(b), u8ptr(d), int(L), int(i0), int(j0), int(inside),
                                       float(env_cutoff), float(cutoff), int(null3), C.byref(vsc), C.byref(ei),
                                       C.byref(ej), hits, cap, C.byref(n)), "ixg_inside_scan_hb")
        return vsc.value, ei.value, ej.value, [(hits[k].i, hits[k].j, hits[k].score, hits[k].bias)
                                               for k in range(n.value)]

    def itables(self):
        M = self.M
        i32p = C.POINTER(C.c_int32)
        itsc, ioesc = np.zeros(M * 6, np.int32), np.zeros(M * 324, np.int32)
        ib, ie, el = np.zeros(M, np.int32), np.zeros(M, np.int32), np.zeros(1, np.int32)
        check(lib().ixg_cm_itables(self.h, *[x.ctypes.data_as(i32p) for x in (itsc, ioesc, ib, ie, el)]))
        return itsc.reshape(M, 6), ioesc.reshape(M, 324), ib, ie, int(el[0])

    def qdb_scan(self, dsq, L, i0, j0, inside=False, dmin=None, dmax=None, cutoff=0.0, null3=True, cap=1 << 16):
        """QDB (dmin/dmax) or non-banded CYK / integer Inside of one window: hits (i, j, score, bias)."""
        d = np.ascontiguousarray(dsq, np.uint8)
        i32p = C.POINTER(C.c_int32)
        keep = [np.ascontiguousarray(x, np.int32) if x is not None else None for x in (dmin, dmax)]
        ptr = [k.ctypes.data_as(i32p) if k is not None else None for k in keep]
        hits = (Hit * cap)()
        n = C.c_int64()
        check(lib().ixg_qdb_scan(self.ctx.h, self.h, u8ptr(d), int(L), int(i0), int(j0), ptr[0], ptr[1], int(inside),
                                 float(cutoff), int(null3), hits, cap, C.byref(n)), "ixg_qdb_scan")
        return [(hits[k].i, hits[k].j, hits[k].score, hits[k].bias) for k in range(n.value)]

    def cyk_seq_filter(self, dsq, n, dmin, dmax, cutoff, null3=True, cap=4096):
        d = np.ascontiguousarray(dsq, np.uint8)
        i32p = C.POINTER(C.c_int32)
        keep = [np.ascontiguousarray(x, np.int32) if x is not None else None for x in (dmin, dmax)]
        ptr = [k.ctypes.data_as(i32p) if k is not None else None for k in keep]
        es, ee = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        k = C.c_int64()
        check(lib().ixg_cyk_seq_filter(self.ctx.h, self.h, u8ptr(d), int(n), ptr[0], ptr[1], float(cutoff),
                                       int(null3), es.ctypes.data_as(i32p), ee.ctypes.data_as(i32p), cap,
                                       C.byref(k)), "ixg_cyk_seq_filter")
        return list(zip(es[:k.value].tolist(), ee[:k.value].tolist()))

    def set_trpenalties(self, g, l):
        """Truncated-begin penalties g[3, M], l[3, M] (log2)."""
        self._trp = (np.ascontiguousarray(g, np.float32), np.ascontiguousarray(l, np.float32))
        check(lib().ixg_cm_set_trpenalties(self.ctx.h if self.ctx else None, self.h, f32ptr(self._trp[0]),
                                           f32ptr(self._trp[1])), "ixg_cm_set_trpenalties")

    def set_itrpenalties(self, ig, il):
        """Integer truncated-begin penalties ig[3, M], il[3, M]."""
        self._itrp = (np.ascontiguousarray(ig, np.int32), np.ascontiguousarray(il, np.int32))
        i32p = C.POINTER(C.c_int32)
        check(lib().ixg_cm_set_itrpenalties(self.ctx.h, self.h, self._itrp[0].ctypes.data_as(i32p),
                                            self._itrp[1].ctypes.data_as(i32p)), "ixg_cm_set_itrpenalties")

    def imtables(self):
        M = self.M
        lm, rm = np.zeros(M * 18, np.int32), np.zeros(M * 18, np.int32)
        i32p = C.POINTER(C.c_int32)
        check(lib().ixg_cm_imtables(self.h, lm.ctypes.data_as(i32p), rm.ctypes.data_as(i32p)))
        return lm.reshape(M, 18), rm.reshape(M, 18)

    def trqdb_scan(self, dsq, L, i0, j0, pass_idx, local=True, inside=False, dmax=None, cutoff=0.0, null3=True,
                   cap=1 << 14):
        """QDB / non-banded TrCYK (float) or TrInside (integer) of one window:
        dict(vsc, vmode, bsc_full, hits [(i, j, score, bias, mode, root)])."""
        d = np.ascontiguousarray(dsq, np.uint8)
        i32p = C.POINTER(C.c_int32)
        dm = np.ascontiguousarray(dmax, np.int32) if dmax is not None else None
        vsc, vm, bf, n = np.zeros(1, np.float32), np.zeros(1, np.int32), np.zeros(1, np.float32), C.c_int64()
        hits = (TrHit * cap)()
        check(lib().ixg_trqdb_scan(self.ctx.h, self.h, u8ptr(d), int(L), int(i0), int(j0),
                                   dm.ctypes.data_as(i32p) if dm is not None else None, int(inside), int(pass_idx),
                                   int(local), float(cutoff), int(null3), f32ptr(vsc), vm.ctypes.data_as(i32p),
                                   f32ptr(bf), hits, cap, C.byref(n)), "ixg_trqdb_scan")
        return dict(vsc=float(vsc[0]), vmode=int(vm[0]), bsc_full=float(bf[0]),
                    hits=[h[1:] for h in _trhits(hits, n.value)])

    def trtables(self):
        M = self.M
        lm, rm = np.zeros(M * 18, np.float32), np.zeros(M * 18, np.float32)
        lp, rp = np.zeros(M, np.int32), np.zeros(M, np.int32)
        i32p = C.POINTER(C.c_int32)
        check(lib().ixg_cm_trtables(self.h, f32ptr(lm), f32ptr(rm), lp.ctypes.data_as(i32p),
                                    rp.ctypes.data_as(i32p)))
        return lm.reshape(M, 18), rm.reshape(M, 18), lp, rp

    def trscan(self, bands, valid, dsq, L, i0, j0, pass_idx, local=True, inside=False, cutoff=0.0, null3=True,
               env_cutoff=0.0, cap=4096):
        """TrCYK (inside=False) / FTrInside of one envelope: (vsc, vmode, envi, envj, hits)."""
        d = np.ascontiguousarray(dsq, np.uint8)
        vsc, vm, ei, ej, n = C.c_float(), C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        b, v = bands_struct(bands), valid_struct(valid)
        hits = (TrHit * cap)()
        fn = lib().ixg_trinside_scan_hb if inside else lib().ixg_trcyk_scan_hb
        check(fn(self.ctx.h, self.h, C.byref(b), C.byref(v), u8ptr(d), int(L), int(i0), int(j0), float(cutoff),
                 int(null3), int(pass_idx), float(env_cutoff), int(local), C.byref(vsc), C.byref(vm), C.byref(ei),
                 C.byref(ej), hits, cap, C.byref(n)), "truncated scan")
        return vsc.value, vm.value, ei.value, ej.value, [h[1:] for h in _trhits(hits, n.value)]


def valid_struct(v):
    keep = {k: np.ascontiguousarray(v[k], np.uint8) for k in ("jvalid", "lvalid", "rvalid", "tvalid")}
    s = TrValid()
    for k, a in keep.items():
        setattr(s, k, a.ctypes.data_as(C.POINTER(C.c_uint8)))
    s._keep = keep
    return s


def _trhits(hits, n):
    return [(hits[k].env, hits[k].i, hits[k].j, hits[k].score, hits[k].bias, hits[k].mode, hits[k].root)
            for k in range(n)]


class EnvBatch:
    """The ixg_envelope (or ixg_trenvelope) array of a list of envelopes,
    marshalled once so that repeated scans of the same batch pass it as is:
    (i0, j0, strand, bands dict) or, truncated, (..., validity dict)."""

    def __init__(self, envs, trunc=False):
        self.n, self.trunc = len(envs), trunc
        self.arr = ((TrEnvelope if trunc else Envelope) * max(self.n, 1))()
        self.keep = []
        for k, e in enumerate(envs):
            i0, j0, strand, b = e[:4]
            a = self.arr[k]
            a.i0, a.j0, a.strand, a.seq = int(i0), int(j0), int(strand), 0
            bs = bands_struct(b)  # holds the arrays its pointers refer to
            a.bands = bs
            self.keep.append(bs)
            if trunc:
                vs = valid_struct(e[4])
                a.valid = vs
                self.keep.append(vs)

    def __len__(self):
        return self.n


def hb_scan_batch(ctx, cm, tgt, envs, inside=False, env_cutoff=0.0, report=False, cutoff=0.0, null3=True):
    """envs: list of (i0, j0, strand, bands dict), or an EnvBatch of them.
    Returns (vsc, envi, envj, hits, cells)."""
    batch = envs if isinstance(envs, EnvBatch) else EnvBatch(envs)
    n, arr = batch.n, batch.arr
    vsc = np.zeros(max(n, 1), np.float32)
    ei = np.zeros(max(n, 1), np.int64)
    ej = np.zeros(max(n, 1), np.int64)
    cap = 1 << 16
    hits = (Hit * cap)()
    nh, cells = C.c_int64(), C.c_int64()
    check(lib().ixg_hb_scan_batch(ctx.h, cm.h, tgt.h, arr, n, int(inside), float(env_cutoff), int(report),
                                  float(cutoff), int(null3), f32ptr(vsc), i64ptr(ei), i64ptr(ej), hits, cap,
                                  C.byref(nh), C.byref(cells)), "ixg_hb_scan_batch")
    hl = [(hits[k].env, hits[k].i, hits[k].j, hits[k].score, hits[k].bias) for k in range(nh.value)] if report else []
    return vsc[:n], ei[:n], ej[:n], hl, cells.value


def hb_scan_multi(ctx, cms, env_cm, tgt, envs, inside=False, env_cutoff=0.0, report=False, cutoff=0.0, null3=True):
    """Envelopes of several CMs in one persistent launch: envs as for
    hb_scan_batch, env_cm[k] = index into cms of envelope k's model.
    Returns (vsc, envi, envj, hits, cells)."""
    batch = envs if isinstance(envs, EnvBatch) else EnvBatch(envs)
    n, arr = batch.n, batch.arr
    ecm = np.ascontiguousarray(env_cm, np.int32)
    assert len(ecm) == n
    hs = (C.c_void_p * len(cms))(*[c.h.value for c in cms])
    vsc = np.zeros(max(n, 1), np.float32)
    ei = np.zeros(max(n, 1), np.int64)
    ej = np.zeros(max(n, 1), np.int64)
    cap = 1 << 16
    hits = (Hit * cap)()
    nh, cells = C.c_int64(), C.c_int64()
    check(lib().ixg_hb_scan_multi(ctx.h, hs, len(cms), ecm.ctypes.data_as(C.POINTER(C.c_int32)), tgt.h, arr, n,
                                  int(inside), float(env_cutoff), int(report), float(cutoff), int(null3), f32ptr(vsc),
                                  i64ptr(ei), i64ptr(ej), hits, cap, C.byref(nh), C.byref(cells)),
          "ixg_hb_scan_multi")
    hl = [(hits[k].env, hits[k].i, hits[k].j, hits[k].score, hits[k].bias) for k in range(nh.value)] if report else []
    return vsc[:n], ei[:n], ej[:n], hl, cells.value


def trhb_scan_batch(ctx, cm, tgt, envs, pass_idx, local=True, inside=False, env_cutoff=0.0, cutoff=0.0, null3=True):
    """envs: list of (i0, j0, strand, bands dict, validity dict), or an
    EnvBatch(..., trunc=True) of them. Returns
    (vsc, vmode, envi, envj, hits (env, i, j, score, bias, mode, root), cells)."""
    batch = envs if isinstance(envs, EnvBatch) else EnvBatch(envs, trunc=True)
    n, arr = batch.n, batch.arr
    vsc = np.zeros(max(n, 1), np.float32)
    vm = np.zeros(max(n, 1), np.int32)
    ei = np.zeros(max(n, 1), np.int64)
    ej = np.zeros(max(n, 1), np.int64)
    cap = 1 << 16
    hits = (TrHit * cap)()
    nh, cells = C.c_int64(), C.c_int64()
    check(lib().ixg_trhb_scan_batch(ctx.h, cm.h, tgt.h, arr, n, int(inside), int(pass_idx), int(local),
                                    float(env_cutoff), float(cutoff), int(null3), f32ptr(vsc),
                                    vm.ctypes.data_as(C.POINTER(C.c_int32)), i64ptr(ei), i64ptr(ej), hits, cap,
                                    C.byref(nh), C.byref(cells)), "ixg_trhb_scan_batch")
    return vsc[:n], vm[:n], ei[:n], ej[:n], _trhits(hits, nh.value), cells.value


def trqdb_scan_batch(ctx, cm, tgt, wins, pass_idx, local=True, inside=False, dmax=None, cutoff=0.0, null3=True):
    """wins: list of (i0, j0, strand). Returns (vsc, vmode, bsc_full, hits (env, i, j, score, bias, mode, root))."""
    n = len(wins)
    arr = (QdbWin * max(n, 1))()
    for k, (i0, j0, strand) in enumerate(wins):
        arr[k].i0, arr[k].j0, arr[k].strand, arr[k].seq = int(i0), int(j0), int(strand), 0
    i32p = C.POINTER(C.c_int32)
    dm = np.ascontiguousarray(dmax, np.int32) if dmax is not None else None
    vsc, vm, bf = np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.float32)
    cap = 1 << 16
    hits = (TrHit * cap)()
    nh = C.c_int64()
    check(lib().ixg_trqdb_scan_batch(ctx.h, cm.h, tgt.h, arr, n, dm.ctypes.data_as(i32p) if dm is not None else None,
                                     int(inside), int(pass_idx), int(local), float(cutoff), int(null3), f32ptr(vsc),
                                     vm.ctypes.data_as(i32p), f32ptr(bf), hits, cap, C.byref(nh)),
          "ixg_trqdb_scan_batch")
    return vsc[:n], vm[:n], bf[:n], _trhits(hits, nh.value)


def qdb_scan_batch(ctx, cm, tgt, wins, inside=False, dmin=None, dmax=None, cutoff=0.0, null3=True):
    """wins: list of (i0, j0, strand). Returns hits (env, i, j, score, bias)."""
    n = len(wins)
    arr = (QdbWin * max(n, 1))()
    for k, (i0, j0, strand) in enumerate(wins):
        arr[k].i0, arr[k].j0, arr[k].strand, arr[k].seq = int(i0), int(j0), int(strand), 0
    i32p = C.POINTER(C.c_int32)
    keep = [np.ascontiguousarray(x, np.int32) if x is not None else None for x in (dmin, dmax)]
    ptr = [k.ctypes.data_as(i32p) if k is not None else None for k in keep]
    cap = 1 << 16
    hits = (Hit * cap)()
    nh = C.c_int64()
    check(lib().ixg_qdb_scan_batch(ctx.h, cm.h, tgt.h, arr, n, ptr[0], ptr[1], int(inside), float(cutoff), int(null3),
                                   hits, cap, C.byref(nh)), "ixg_qdb_scan_batch")
    return [(hits[k].env, hits[k].i, hits[k].j, hits[k].score, hits[k].bias) for k in range(nh.value)]


# ---- CP9 HMM banding (cp9.rs:1275 cp9_forward, :1995 cp9_backward, :1517 cp9_fb2hmmbands) ----

class Cp9Params(C.Structure):
    _fields_ = [("M", C.c_int32), ("el", C.c_int32), ("el_selfsc", C.c_int32), ("pad", C.c_int32),
                ("otsc", C.POINTER(C.c_int32)), ("msc", C.POINTER(C.c_int32)), ("isc", C.POINTER(C.c_int32)),
                ("has_el", C.POINTER(C.c_uint8)), ("el_from_ct", C.POINTER(C.c_int32)),
                ("el_from_off", C.POINTER(C.c_int32)), ("el_from_idx", C.POINTER(C.c_int32))]


class Cp9Win(C.Structure):
    _fields_ = [("i0", C.c_int32), ("j0", C.c_int32), ("strand", C.c_int32), ("seq", C.c_int32)]


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


_cp9_ready = False


def _cp9_lib():
    global _cp9_ready
    L = lib()
    if not _cp9_ready:
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        i32p, f32p, u8p = C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        L.ixg_cp9_create.argtypes = [vp, C.POINTER(Cp9Params), C.POINTER(vp)]
        L.ixg_cp9_free.argtypes = [vp, vp]
        L.ixg_cp9_bands.argtypes = [vp, vp, u8p, i64, i32, i32, C.c_double, C.c_int] + [i32p] * 6 + [f32p, f32p, i32p]
        L.ixg_cp9_bands_batch.argtypes = [vp, vp, vp, C.POINTER(Cp9Win), i64, C.c_double, C.c_int, f32p, f32p, i32p]
        _cp9_ready = True
    return L


class CP9:
    """Device-resident configured CP9 HMM (ixg_cp9_create) from a dict with
    M, el, el_selfsc, otsc, msc, isc, has_el, el_from_ct, el_from_off,
    el_from_idx (synth_cp9.build_cp9 makes one)."""

    def __init__(self, ctx, h):
        self.ctx, self.M = ctx, int(h["M"])
        self.keep = {k: np.ascontiguousarray(h[k], np.uint8 if k == "has_el" else np.int32)
                     for k in ("otsc", "msc", "isc", "has_el", "el_from_ct", "el_from_off", "el_from_idx")}
        p = Cp9Params()
        p.M, p.el, p.el_selfsc, p.pad = self.M, int(bool(h["el"])), int(h["el_selfsc"]), 0
        for k, v in self.keep.items():
            setattr(p, k, v.ctypes.data_as(C.POINTER(C.c_uint8 if k == "has_el" else C.c_int32)))
        self.h = C.c_void_p()
        check(_cp9_lib().ixg_cp9_create(ctx.h, C.byref(p), C.byref(self.h)), "ixg_cp9_create")

    def free(self):
        if self.h:
            check(_cp9_lib().ixg_cp9_free(self.ctx.h, self.h), "ixg_cp9_free")
            self.h = C.c_void_p()

    def bands(self, dsq, L, i0, j0, tau, scan=True, matrices=False, mode=None):
        """One window: dict(pn [6, M+1], fsc, bsc[, fm, fi, fd, bm, bi, bd [(L'+1), M+1]]).
        mode: IXG_CP9_* bits; by default search (both scans) or, with
        scan=False, alignment."""
        M1, Lw = self.M + 1, j0 - i0 + 1
        d = np.ascontiguousarray(dsq, np.uint8)
        mats = [np.zeros((Lw + 1) * M1, np.int32) for _ in range(6)] if matrices else None
        pn = np.zeros(6 * M1, np.int32)
        fs, bs = C.c_float(), C.c_float()
        mp = [(_i32p(m) if matrices else None) for m in (mats or [None] * 6)]
        check(_cp9_lib().ixg_cp9_bands(self.ctx.h, self.h, u8ptr(d), int(L), int(i0), int(j0), float(tau),
                                  cp9_mode(scan, mode), *mp, C.byref(fs), C.byref(bs), _i32p(pn)), "ixg_cp9_bands")
        out = dict(pn=pn.reshape(6, M1), fsc=fs.value, bsc=bs.value)
        if matrices:
            for k, m in zip(("fm", "fi", "fd", "bm", "bi", "bd"), mats):
                out[k] = m.reshape(Lw + 1, M1)
        return out


IXG_CP9_FWD_SCAN, IXG_CP9_BCK_SCAN, IXG_CP9_ALIGN = 1, 2, 4


def cp9_mode(scan=True, mode=None):
    """The mode bits of cp9_seq2bands: search scans both DPs, alignment
    reports whole-window scores; an explicit mode wins."""
    if mode is not None:
        return int(mode)
    return IXG_CP9_FWD_SCAN | IXG_CP9_BCK_SCAN if scan else IXG_CP9_ALIGN


def cp9_bands_batch(ctx, cp9, tgt, wins, tau, scan=True, mode=None):
    """wins: list of (i0, j0, strand). Returns (pn [n, 6, M+1], fsc [n], bsc [n])."""
    n, M1 = len(wins), cp9.M + 1
    arr = (Cp9Win * max(n, 1))()
    for k, (i0, j0, strand) in enumerate(wins):
        arr[k].i0, arr[k].j0, arr[k].strand, arr[k].seq = int(i0), int(j0), int(strand), 0
    pn = np.zeros(max(n, 1) * 6 * M1, np.int32)
    fs, bs = np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.float32)
    check(_cp9_lib().ixg_cp9_bands_batch(ctx.h, cp9.h, tgt.h, arr, n, float(tau), cp9_mode(scan, mode), f32ptr(fs), f32ptr(bs),
                                    _i32p(pn)), "ixg_cp9_bands_batch")
    return pn[:n * 6 * M1].reshape(n, 6, M1), fs[:n], bs[:n]


# ---- F4/F5 envelope definition ----------------------------------------------

IXG_GFB_BACKWARD, IXG_GFB_UNIHIT, IXG_GFB_MATRIX = 1, 2, 4


class EnvOpts(C.Structure):
    _fields_ = [("f4", C.c_double), ("f4b", C.c_double), ("f5", C.c_double), ("f5b", C.c_double),
                ("gfmu", C.c_double), ("gflambda", C.c_double), ("rt1", C.c_float), ("rt2", C.c_float),
                ("rt3", C.c_float), ("ns", C.c_int32), ("do_gfwdbias", C.c_int32), ("do_edefbias", C.c_int32),
                ("lftau", C.c_double), ("lflambda", C.c_double), ("clen", C.c_int32), ("pad", C.c_int32)]


class EnvAcct(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "n_past_gfwd", "pos_past_gfwd", "n_past_gfwdbias", "pos_past_gfwdbias",
        "n_past_edef", "pos_past_edef", "n_past_edefbias", "pos_past_edefbias")]


def env_opts(f4=0.005, f4b=0.005, f5=0.005, f5b=0.005, gfmu=-5.0, gflambda=0.69, rt1=0.25, rt2=0.10, rt3=0.20,
             ns=200, do_gfwdbias=True, do_edefbias=False, lftau=-4.0, lflambda=0.69, clen=1):
    o = EnvOpts()
    o.f4, o.f4b, o.f5, o.f5b, o.gfmu, o.gflambda = f4, f4b, f5, f5b, gfmu, gflambda
    o.rt1, o.rt2, o.rt3, o.ns = rt1, rt2, rt3, ns
    o.do_gfwdbias, o.do_edefbias = int(bool(do_gfwdbias)), int(bool(do_edefbias))
    o.lftau, o.lflambda, o.clen = lftau, lflambda, int(clen)
    return o


_env_ready = False


def _env_lib():
    global _env_ready
    L = lib()
    if not _env_ready:
        vp, i64 = C.c_void_p, C.c_int64
        i32p, i64p, f32p, u8p = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        L.ixg_gfb_scores.argtypes = [vp, vp, u8p, i64, i64p, i64p, i32p, i64, C.c_int, f32p, f32p, f32p]
        L.ixg_env_def.argtypes = [vp, vp, u8p, i64, C.POINTER(Window), i64, C.POINTER(EnvOpts), i32p, i32p, i64, i64p,
                                  C.POINTER(EnvAcct)]
        L.ixg_env_def_pass.argtypes = [vp, vp, u8p, i64, C.POINTER(Window), i64, C.c_int, C.POINTER(EnvOpts), i32p,
                                       i32p, i64, i64p, C.POINTER(EnvAcct)]
        _env_ready = True
    return L


def gfb_scores(p7, seq, offs, lens, cfg_len=None, backward=False, unihit=False, matrix=False):
    """Glocal Forward / Backward over items seq[off+1 .. off+len] (ixg_gfb_scores).
    Returns (scores, [specials (len+1, 5) per item], [matrix (len+1, 3, M+1) per item] or None)."""
    seq = np.ascontiguousarray(seq, np.uint8)
    offs = np.ascontiguousarray(offs, np.int64)
    lens = np.ascontiguousarray(lens, np.int64)
    n, M = len(offs), p7.M
    sc = np.zeros(max(n, 1), np.float32)
    xm = np.zeros(max(int(5 * (lens + 1).sum()), 1), np.float32)
    mat = np.zeros(max(int(3 * (M + 1) * (lens + 1).sum()), 1), np.float32) if matrix else None
    cl = None if cfg_len is None else np.ascontiguousarray(cfg_len, np.int32)
    mode = (IXG_GFB_BACKWARD if backward else 0) | (IXG_GFB_UNIHIT if unihit else 0) | (IXG_GFB_MATRIX if matrix else 0)
    check(_env_lib().ixg_gfb_scores(p7.ctx.h, p7.h, u8ptr(seq), len(seq), i64ptr(offs), i64ptr(lens),
                                    None if cl is None else cl.ctypes.data_as(C.POINTER(C.c_int32)), n, mode,
                                    f32ptr(sc), f32ptr(xm), None if mat is None else f32ptr(mat)), "ixg_gfb_scores")
    xs, ms, xo, mo = [], [], 0, 0
    for ln in lens:
        xs.append(xm[xo:xo + 5 * (ln + 1)].reshape(ln + 1, 5))
        xo += 5 * (ln + 1)
        if matrix:
            ms.append(mat[mo:mo + 3 * (M + 1) * (ln + 1)].reshape(ln + 1, 3, M + 1))
            mo += 3 * (M + 1) * (ln + 1)
    return sc[:n], xs, (ms if matrix else None)


def env_def(p7, dsq, L, windows, opts, pass_idx=1):
    """Envelope definition of one chunk (ixg_env_def_pass; pass 1 STD, 2 5',
    3 3', 4 5'&3'): windows [(start, end)] -> ([(es, ee)], EnvAcct)."""
    d = np.ascontiguousarray(dsq, np.uint8)
    n = len(windows)
    arr = (Window * max(n, 1))()
    for k, (a, b) in enumerate(windows):
        arr[k].start, arr[k].end, arr[k].score, arr[k].task = int(a), int(b), 0.0, 0
    cap = 64 + 8 * n
    while True:
        es, ee = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        nenv = C.c_int64()
        acct = EnvAcct()
        rc = _env_lib().ixg_env_def_pass(p7.ctx.h, p7.h, u8ptr(d), int(L), arr, n, int(pass_idx), C.byref(opts),
                                         es.ctypes.data_as(C.POINTER(C.c_int32)),
                                         ee.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(nenv), C.byref(acct))
        if rc == IXG_ERANGE and nenv.value > cap:
            cap = nenv.value
            continue
        check(rc, "ixg_env_def")
        return [(int(es[i]), int(ee[i])) for i in range(nenv.value)], acct


# ---- HMM -> CM band conversion ------------------------------------------------

def cp9_params(h):
    """Cp9Params view of a CP9 dict (keeps its arrays alive)."""
    keep = {k: np.ascontiguousarray(h[k], np.uint8 if k == "has_el" else np.int32)
            for k in ("otsc", "msc", "isc", "has_el", "el_from_ct", "el_from_off", "el_from_idx")}
    p = Cp9Params()
    p.M, p.el, p.el_selfsc, p.pad = int(h["M"]), int(bool(h["el"])), int(h["el_selfsc"]), 0
    for k, v in keep.items():
        setattr(p, k, v.ctypes.data_as(C.POINTER(C.c_uint8 if k == "has_el" else C.c_int32)))
    p._keep = keep
    return p


_hb_ready = False


def _hb_lib():
    global _hb_ready
    L = lib()
    if not _hb_ready:
        vp, i32p, i64p = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        L.ixg_hbmap_create.argtypes = [C.POINTER(CmParams), C.c_int, C.POINTER(Cp9Params), C.c_int, C.POINTER(vp)]
        L.ixg_hbmap_free.argtypes = [vp]
        L.ixg_hmm2ij_bands.argtypes = [vp, i32p, C.c_int32, C.c_int32, C.c_int, i32p, i32p, i32p, i32p]
        L.ixg_ij2d_bands.argtypes = [vp, i32p, i32p, i32p, i32p, C.c_int, i32p, i32p, i32p, C.c_int64, i64p]
        u8p = C.POINTER(C.c_uint8)
        L.ixg_hmm2ij_bands_trunc.argtypes = [vp, i32p, i32p, C.c_int32, C.c_int32, C.c_int, i32p, i32p, i32p, i32p,
                                             u8p, u8p, u8p, u8p, i32p]
        _hb_ready = True
    return L


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class HbMap:
    """Host map of a CM and its CP9 for the HMM -> CM band conversion."""

    def __init__(self, cm, cp9, cm_local=True, cp9_local=True):
        self.M = int(cm["M"])
        self._c, self._p = cm_params(cm), cp9_params(cp9)
        self.h = C.c_void_p()
        check(_hb_lib().ixg_hbmap_create(C.byref(self._c), int(cm_local), C.byref(self._p), int(cp9_local),
                                         C.byref(self.h)), "ixg_hbmap_create")

    def free(self):
        if self.h:
            _hb_lib().ixg_hbmap_free(self.h)
            self.h = C.c_void_p()

    def hmm2ij(self, pn, i0, j0, search=True):
        """Returns (imin, imax, jmin, jmax, pn after doctoring)."""
        pn = np.ascontiguousarray(pn, np.int32).copy()
        out = [np.zeros(self.M, np.int32) for _ in range(4)]
        check(_hb_lib().ixg_hmm2ij_bands(self.h, _i32(pn), int(i0), int(j0), int(search), *[_i32(a) for a in out]),
              "ixg_hmm2ij_bands")
        return (*out, pn)

    def hmm2ij_trunc(self, pn, occ, i0, j0, pass_idx):
        """ixg_hmm2ij_bands_trunc: (imin, imax, jmin, jmax, jvalid, lvalid,
        rvalid, tvalid, [sp1 sp2 ep1 ep2 rmarg_imin rmarg_imax lmarg_jmin
        lmarg_jmax])."""
        L = _hb_lib()
        pn = np.ascontiguousarray(pn, np.int32).copy()
        occ = np.ascontiguousarray(occ, np.int32)
        out = [np.zeros(self.M, np.int32) for _ in range(4)]
        val = [np.zeros(self.M, np.uint8) for _ in range(4)]
        spep = np.zeros(8, np.int32)
        check(L.ixg_hmm2ij_bands_trunc(self.h, _i32(pn), _i32(occ), int(i0), int(j0), int(pass_idx),
                                       *[_i32(a) for a in out], *[u8ptr(v) for v in val], _i32(spep)),
              "ixg_hmm2ij_bands_trunc")
        return (*out, *val, spep.tolist())

    def ij2d(self, imin, imax, jmin, jmax, trunc=False):
        """Bands dict in the scanners' layout (jmin, jmax, hd_off, hdmin, hdmax)."""
        a = [np.ascontiguousarray(x, np.int32) for x in (imin, imax, jmin, jmax)]
        hd_off = np.zeros(self.M + 1, np.int32)
        n = C.c_int64()
        _hb_lib().ixg_ij2d_bands(self.h, *[_i32(x) for x in a], int(trunc), _i32(hd_off), None, None, 0, C.byref(n))
        hdmin, hdmax = np.zeros(max(n.value, 1), np.int32), np.zeros(max(n.value, 1), np.int32)
        check(_hb_lib().ixg_ij2d_bands(self.h, *[_i32(x) for x in a], int(trunc), _i32(hd_off), _i32(hdmin),
                                       _i32(hdmax), n.value, C.byref(n)), "ixg_ij2d_bands")
        return dict(jmin=a[2], jmax=a[3], hd_off=hd_off, hdmin=hdmin[:n.value], hdmax=hdmax[:n.value])


# ---- the STD pass of a search, F1 through F7 ------------------------------------

class SearchOpts(C.Structure):
    _fields_ = [("f3", F3Opts), ("env", EnvOpts), ("f6", C.c_double), ("cyk_mu", C.c_double),
                ("cyk_lambda", C.c_double), ("cyk_env_cutoff", C.c_float), ("pli_t", C.c_float),
                ("fin_mu", C.c_double), ("fin_lambda", C.c_double), ("ez", C.c_double), ("e_report", C.c_double),
                ("fcyk_tau", C.c_double), ("final_tau", C.c_double), ("maxtau", C.c_double),
                ("size_limit", C.c_float), ("do_fcyk", C.c_int32), ("do_fcykenv", C.c_int32),
                ("do_null3", C.c_int32), ("do_cyk", C.c_int32), ("do_top", C.c_int32), ("do_bot", C.c_int32),
                ("trunc_mode", C.c_int32), ("pad", C.c_int32)]


class SearchHit(C.Structure):
    _fields_ = [("start", C.c_int64), ("stop", C.c_int64), ("score", C.c_float), ("bias", C.c_float),
                ("evalue", C.c_double), ("strand", C.c_int32), ("pass_idx", C.c_int32), ("seq", C.c_int64)]


class SearchAcct(C.Structure):
    _fields_ = [("nres", C.c_uint64), ("f", Acct), ("e", EnvAcct), ("n_past_cyk", C.c_uint64),
                ("pos_past_cyk", C.c_uint64), ("n_hits", C.c_uint64), ("tr_nres", C.c_uint64 * 3),
                ("tr_f", Acct * 3), ("tr_e", EnvAcct * 3), ("tr_n_past_cyk", C.c_uint64 * 3),
                ("tr_pos_past_cyk", C.c_uint64 * 3), ("tr_hits", C.c_uint64 * 3)]


def search_opts(f3, env, f6, cyk, fin, ez, e_report=10.0, pli_t=None, cykenvx=10, fcyk_tau=1e-4, final_tau=5e-6,
                maxtau=0.05, W=None, do_fcyk=True, do_fcykenv=True, do_null3=True, do_cyk=False, top=True, bot=True,
                trunc=False):
    """Options of ixg_search_std. cyk / fin: (mu, lambda) exponential tails of
    the local CYK filter and the final round; the CYK envelope cutoff and the
    reporting threshold follow them as cm_search.rs:1656-1721 derives them
    (pli_t from e_report and ez unless given); the size limit from W
    (cm_search.rs:890-896)."""
    o = SearchOpts()
    o.f3, o.env = f3, env
    o.f6, o.cyk_mu, o.cyk_lambda = f6, cyk[0], cyk[1]
    f6env = min(1.0, float(np.float32(f6) * np.float32(cykenvx)))
    o.cyk_env_cutoff = float(np.float32(cyk[0] + np.log(f6env) / (-1.0 * cyk[1])))
    o.fin_mu, o.fin_lambda, o.ez, o.e_report = fin[0], fin[1], ez, e_report
    o.pli_t = float(np.float32(fin[0] + np.log(e_report / ez) / (-1.0 * fin[1]))) if pli_t is None else pli_t
    o.fcyk_tau, o.final_tau, o.maxtau = fcyk_tau, final_tau, maxtau
    w = float(W if W is not None else 1000)
    o.size_limit = float(min(1024.0, max(256.0, (w - 1000.0) / (3000.0 - 1000.0) * (1024.0 - 256.0) + 256.0)))
    o.do_fcyk, o.do_fcykenv, o.do_null3, o.do_cyk = int(do_fcyk), int(do_fcykenv), int(do_null3), int(do_cyk)
    o.do_top, o.do_bot = int(top), int(bot)
    o.trunc_mode = int(bool(trunc))
    return o


_search_ready = False


def _search_lib():
    global _search_ready
    L = lib()
    if not _search_ready:
        vp = C.c_void_p
        L.ixg_search_std.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(SearchOpts), C.POINTER(SearchHit), C.c_int64,
                                     C.POINTER(C.c_int64), C.POINTER(SearchAcct)]
        L.ixg_search.argtypes = [vp, vp, vp, vp, vp, C.POINTER(vp), C.POINTER(vp), vp, C.POINTER(SearchOpts),
                                 C.POINTER(SearchHit), C.c_int64, C.POINTER(C.c_int64), C.POINTER(SearchAcct)]
        _search_ready = True
    return L


def _search_call(call, name, with_pass=False, with_seq=False):
    cap = 256
    while True:
        out = (SearchHit * cap)()
        n = C.c_int64()
        acct = SearchAcct()
        rc = call(out, cap, n, acct)
        if rc == IXG_ERANGE and n.value > cap:
            cap = n.value
            continue
        check(rc, name)
        return [(out[i].start, out[i].stop, out[i].strand, out[i].score, out[i].bias, out[i].evalue) +
                ((out[i].pass_idx,) if with_pass else ()) + ((out[i].seq,) if with_seq else ())
                for i in range(n.value)], acct


def search_std(ctx, p7, cm, cp9, hbmap, tgt, opts):
    """ixg_search_std: ([(start, stop, strand, score, bias, evalue)], SearchAcct)."""
    L = _search_lib()
    return _search_call(lambda out, cap, n, acct: L.ixg_search_std(ctx.h, p7.h, cm.h, cp9.h, hbmap.h, tgt.h,
                                                                   C.byref(opts), out, cap, C.byref(n), C.byref(acct)),
                        "ixg_search_std")


def search(ctx, p7, cm, cp9, hbmap, trcp9, trmap, tgt, opts, with_seq=False):
    """ixg_search with the truncated passes' CP9s and band maps (5', 3',
    5'&3'): ([(start, stop, strand, score, bias, evalue, pass)], SearchAcct);
    with_seq appends each hit's sequence index (a MultiTarget's)."""
    L = _search_lib()
    vp = C.c_void_p
    tc = (vp * 3)(*[x.h.value if hasattr(x.h, 'value') else x.h for x in trcp9])
    tm = (vp * 3)(*[x.h.value if hasattr(x.h, 'value') else x.h for x in trmap])
    return _search_call(lambda out, cap, n, acct: L.ixg_search(ctx.h, p7.h, cm.h, cp9.h, hbmap.h, tc, tm, tgt.h,
                                                               C.byref(opts), out, cap, C.byref(n), C.byref(acct)),
                        "ixg_search", with_pass=True, with_seq=with_seq)


class SearchModel:
    """A model bundle for served searches (ixg_search_model): the handles stay
    owned by the caller and must outlive the server."""

    class S(C.Structure):
        _fields_ = [("p7", C.c_void_p), ("cm", C.c_void_p), ("cp9", C.c_void_p), ("map", C.c_void_p),
                    ("trcp9", C.c_void_p * 3), ("trmap", C.c_void_p * 3)]

    def __init__(self, p7, cm, cp9, hbmap, trcp9=None, trmap=None):
        def hv(x):
            return x.h.value if hasattr(x.h, "value") else x.h
        self.s = SearchModel.S()
        self.s.p7, self.s.cm, self.s.cp9, self.s.map = hv(p7), hv(cm), hv(cp9), hv(hbmap)
        for q in range(3):
            if trcp9:
                self.s.trcp9[q] = hv(trcp9[q])
                self.s.trmap[q] = hv(trmap[q])


SEARCH_STAGES = ("f1_f3b", "f4_f5", "f6_bands", "f6_cyk", "f7_bands", "f7_inside", "finalize", "total",
                 "truncated_passes")


def search_stats(ctx):
    """Stage wall times (ms, by SEARCH_STAGES name) and the CYK / Inside HB
    cell counts of the context's last search_std call."""
    ms = (C.c_double * len(SEARCH_STAGES))()
    cells = (C.c_int64 * 2)()
    check(lib().ixg_search_stats(ctx.h, ms, cells), "ixg_search_stats")
    return dict(zip(SEARCH_STAGES, list(ms))), (cells[0], cells[1])
