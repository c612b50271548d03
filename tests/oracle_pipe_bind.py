"""ctypes binding of the oracle's STD-pass search (oracle/pipeline.c).
Test infrastructure only: the checker, never the thing measured."""
import ctypes as C

import numpy as np

import oracle_bind as OB
import oracle_cm_bind as OC
import oracle_cp9_bind as O9
from oracle_env_bind import Gm, OracleEnv
from infernox_amd import ixg


class Pipe(C.Structure):
    _fields_ = [("msv", C.POINTER(OB.Msv)), ("vit", C.POINTER(OB.Vit)), ("fwd", C.POINTER(OB.Fwd)),
                ("bias", C.POINTER(OB.Bias)), ("p7", C.POINTER(OB.P7)), ("gm", C.POINTER(Gm)),
                ("cm", C.POINTER(OC.Cm)), ("sc", C.POINTER(OC.CmSc)), ("cp9", C.POINTER(O9.Cp9)),
                ("hbcm", C.POINTER(O9.HbCm)), ("hbcp9", C.POINTER(O9.HbCp9)), ("gmt", C.POINTER(Gm) * 3),
                ("trcp9", C.POINTER(O9.Cp9) * 3), ("trhbcp9", C.POINTER(O9.HbCp9) * 3),
                ("lmesc", C.POINTER(C.c_float)), ("rmesc", C.POINTER(C.c_float)), ("trp_l", C.POINTER(C.c_float))]


class OracleSearch:
    """The oracle's search for a p7 HMM dict, a CM dict and a CP9 dict (and,
    for the truncated passes, the three truncated CP9 dicts). The options are
    ixg.SearchOpts (the oracle's struct has the same layout)."""

    def __init__(self, p7, cm, cp9, max_length, trcp9=None):
        self.env = OracleEnv(p7, max_length)
        self.base = self.env.base
        self.ocm = OC.OracleCM(cm)
        self.o9 = O9.OracleCP9(cp9)
        self.hb = O9.OracleHb(cm, cp9)
        P = Pipe()
        b = self.base
        P.msv, P.vit, P.fwd, P.bias, P.p7 = (C.pointer(b.smsv), C.pointer(b.svit), C.pointer(b.sfwd),
                                             C.pointer(b.sbias), C.pointer(b.p7))
        P.gm = C.pointer(self.env.gm)
        P.cm, P.sc = C.pointer(self.ocm.s), C.pointer(self.ocm.sc)
        P.cp9 = C.pointer(self.o9.s)
        P.hbcm, P.hbcp9 = C.pointer(self.hb.c), C.pointer(self.hb.p)
        if trcp9 is not None:
            L0 = OB.lib()
            L0.ixo_gm_build_trunc.argtypes = [C.POINTER(OB.P7), C.c_int, C.POINTER(Gm)]
            self.gmt = [Gm() for _ in range(3)]
            for q in range(3):
                L0.ixo_gm_build_trunc(C.byref(b.p7), q, C.byref(self.gmt[q]))
                P.gmt[q] = C.pointer(self.gmt[q])
            self.tr9 = [O9.OracleCP9(h) for h in trcp9]
            self.trhb = [O9.OracleHb(cm, h) for h in trcp9]
            for q in range(3):
                P.trcp9[q] = C.pointer(self.tr9[q].s)
                P.trhbcp9[q] = C.pointer(self.trhb[q].p)
            tcm = OC.OracleTrCM(cm)
            self.tcm = tcm
            self.lmesc, self.rmesc = tcm.lmesc, tcm.rmesc
            self.trp_l = np.ascontiguousarray(OC.trpenalties(cm)[1], np.float32)
            fp = C.POINTER(C.c_float)
            P.lmesc, P.rmesc = self.lmesc.ctypes.data_as(fp), self.rmesc.ctypes.data_as(fp)
            P.trp_l = self.trp_l.ctypes.data_as(fp)
        self.P = P
        L = OB.lib()
        L.ixo_search_std.argtypes = [C.POINTER(Pipe), C.POINTER(ixg.SearchOpts), C.POINTER(C.c_uint8), C.c_int64,
                                     C.POINTER(ixg.SearchHit), C.c_int64, C.POINTER(ixg.SearchAcct)]
        L.ixo_search_std.restype = C.c_int64
        L.ixo_search_std_mt.argtypes = L.ixo_search_std.argtypes + [C.c_int]
        L.ixo_search_std_mt.restype = C.c_int64
        L.ixo_search_multi_mt.argtypes = [C.POINTER(Pipe), C.POINTER(ixg.SearchOpts),
                                          C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_int64), C.c_int64,
                                          C.POINTER(ixg.SearchHit), C.c_int64, C.POINTER(ixg.SearchAcct), C.c_int]
        L.ixo_search_multi_mt.restype = C.c_int64

    def search(self, dsq, L, opts, threads=1):
        """Hits (start, stop, strand, score, bias, evalue) and the counters;
        threads > 1 runs the tasks on that many threads (same result)."""
        d = np.ascontiguousarray(dsq, np.uint8)
        cap = 4096
        out = (ixg.SearchHit * cap)()
        acct = ixg.SearchAcct()
        n = OB.lib().ixo_search_std_mt(C.byref(self.P), C.byref(opts), d.ctypes.data_as(C.POINTER(C.c_uint8)),
                                       int(L), out, cap, C.byref(acct), int(threads))
        return [(out[i].start, out[i].stop, out[i].strand, out[i].score, out[i].bias, out[i].evalue) +
                ((out[i].pass_idx,) if opts.trunc_mode else ()) for i in range(min(n, cap))], acct

    def search_multi(self, seqs, opts, threads=1):
        """The search over several sequences (1-indexed dsq arrays): hits
        (start, stop, strand, score, bias, evalue, pass, seq) and the counters."""
        ds = [np.ascontiguousarray(d, np.uint8) for d in seqs]
        n = len(ds)
        ptrs = (C.POINTER(C.c_uint8) * n)(*[d.ctypes.data_as(C.POINTER(C.c_uint8)) for d in ds])
        lens = (C.c_int64 * n)(*[len(d) - 2 for d in ds])
        cap = 8192
        out = (ixg.SearchHit * cap)()
        acct = ixg.SearchAcct()
        k = OB.lib().ixo_search_multi_mt(C.byref(self.P), C.byref(opts), ptrs, lens, n, out, cap, C.byref(acct),
                                         int(threads))
        return [(out[i].start, out[i].stop, out[i].strand, out[i].score, out[i].bias, out[i].evalue,
                 out[i].pass_idx, out[i].seq) for i in range(min(k, cap))], acct
