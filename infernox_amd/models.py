"""Models and targets of the search workloads, built the way the workloads
name them, with stand-ins where the reference's own tools are not part of
this build:

* a seeded WUSS-structured alignment (cloverleaf: stems 7/4/5/5 bp, loops
  7-9 nt; pairs drawn GC/CG .3/.3, AU/UA .15/.15, GU/UG .05/.05; loops
  uniform; 10% point mutations);
* a CM on the structure's guide tree (synth_cm's node/state layout) whose
  emissions are the alignment's counts plus one pseudocount per cell (the
  reference's cmbuild applies Dirichlet-mixture priors and entropy weighting
  instead, cm_modelmaker.rs / prior.rs / eweight.rs, not reproduced);
* its CP9 and p7 filter HMM from the CM's consensus columns
  (synth_cp9.build_cp9_like; the p7 takes the same match emissions with
  mostly match-to-match transitions), and E-value parameters calibrated on
  random sequence (tests/golden/calibrate_config1.py, values held here);
* the target: iid uniform ACGU with CM-sampled homologs inserted on random
  strands.
"""
import numpy as np

from . import synth_cm as S
from . import synth_cp9 as P

PAIRS = [(2, 1), (1, 2), (0, 3), (3, 0), (2, 3), (3, 2)]  # GC CG AU UA GU UG
PAIR_W = [0.3, 0.3, 0.15, 0.15, 0.05, 0.05]


def cloverleaf_ss():
    """72-column tRNA-like cloverleaf: acceptor 7 bp, D arm 4 bp / 8-nt loop,
    anticodon arm 5 bp / 7-nt loop, T arm 5 bp / 7-nt loop."""
    return S.cloverleaf(stems=(7, 4, 5, 5), loops=(8, 7, 7), clen=72)


def alignment(ss, nseq, seed, mut=0.10):
    """nseq x len(ss) digitized alignment: a sampled consensus (pairs from the
    pair weights, unpaired uniform) with mut point mutations per position."""
    r = np.random.default_rng(seed)
    ct = S.wuss_to_ct(ss)
    n = len(ss)
    cons = np.zeros(n, np.int64)
    for i in range(1, n + 1):
        j = ct[i]
        if j == 0:
            cons[i - 1] = r.integers(0, 4)
        elif i < j:
            a, b = PAIRS[int(r.choice(6, p=PAIR_W))]
            cons[i - 1], cons[j - 1] = a, b
    aln = np.tile(cons, (nseq, 1))
    m = r.random(aln.shape) < mut
    aln[m] = r.integers(0, 4, int(m.sum()))
    return aln.astype(np.uint8)


def cm_from_alignment(ss, aln, seed):
    """CM on the guide tree of ss with emissions estimated from the
    alignment (counts + 1 per cell); transitions as synth_cm.build_cm."""
    cm = S.build_cm(ss, seed)
    lpos, rpos = P.emit_map(cm)
    e = cm["e"].copy()
    for nd in range(cm["nodes"]):
        t, v = cm["ndtype"][nd], cm["nodemap"][nd]
        if t == S.MATP_ND:
            i, j = lpos[nd] - 1, rpos[nd] - 1
            cnt = np.ones(16)
            for a, b in zip(aln[:, i], aln[:, j]):
                cnt[a * 4 + b] += 1
            e[v] = cnt / cnt.sum()
            for k, col in ((1, i), (2, j)):  # MATP ML / MR: the column marginal
                c = np.ones(4)
                for a in aln[:, col]:
                    c[a] += 1
                e[v + k, :4] = c / c.sum()
        elif t in (S.MATL_ND, S.MATR_ND):
            col = (lpos[nd] if t == S.MATL_ND else rpos[nd]) - 1
            c = np.ones(4)
            for a in aln[:, col]:
                c[a] += 1
            e[v, :4] = c / c.sum()
    cm["e"] = e.astype(np.float32)
    return cm


def p7_from_cm(cm, ev):
    """p7 filter HMM over the CM's consensus columns (match emissions from the
    CM's column marginals) with the given E-value parameters."""
    h = P.build_cp9_like(cm, 0)
    lpos, rpos = P.emit_map(cm)
    M = int(h["M"])
    mat = np.full((M + 1, 4), 0.25)
    for nd in range(cm["nodes"]):
        t, v = cm["ndtype"][nd], cm["nodemap"][nd]
        e = np.asarray(cm["e"][v], np.float64)
        if t == S.MATL_ND:
            mat[lpos[nd]] = e[:4]
        elif t == S.MATR_ND:
            mat[rpos[nd]] = e[:4]
        elif t == S.MATP_ND:
            pe = e.reshape(4, 4)
            mat[lpos[nd]] = pe.sum(axis=1)
            mat[rpos[nd]] = pe.sum(axis=0)
    mat /= mat.sum(axis=1, keepdims=True)
    trans = np.zeros((M + 1, 7))
    for k in range(M + 1):
        trans[k] = [0.96, 0.02, 0.02, 0.7, 0.3, 0.8, 0.2]
    trans[M] = [1, 0, 0, 1, 0, 1, 0]
    p = dict(M=M, mat=mat.astype(np.float32), trans=trans.astype(np.float32),
             compo=mat[1:].mean(axis=0).astype(np.float32))
    p.update(ev)
    return p


# E-value parameters of the config-1 model (cloverleaf, alignment seed 1),
# from tests/golden/calibrate_config1.py: p7 local MSV / Viterbi Gumbel mu,
# local Forward and glocal Forward exponential tau (lambda ln 2); CM local
# CYK and Inside exponential tails (mu, lambda) in bits.
CONFIG1_EV = dict(lmmu=-7.7994, lmlambda=0.693147, lvmu=-8.7676, lvlambda=0.693147, lftau=-4.1874,
                  lflambda=0.693147, gfmu=-18.0086, gflambda=0.693147)
CONFIG1_CM_EV = dict(lc=(-1.4198, 0.693147), li=(-0.1401, 0.693147))


def config1_models(ev=None, cm_ev=None):
    """(p7 dict, CM dict, CP9 dict, cm_ev) of the config-1 model."""
    ss = cloverleaf_ss()
    aln = alignment(ss, 40, seed=1)
    cm = cm_from_alignment(ss, aln, seed=1)
    cp9 = P.build_cp9_like(cm, 1, el=True)
    p7 = p7_from_cm(cm, ev or CONFIG1_EV)
    return p7, cm, cp9, (cm_ev or CONFIG1_CM_EV)


def target_with_homologs(cm, L, n, seed, gc=0.5):
    """iid target of L residues (ACGU with the given GC) with n CM-sampled
    homologs inserted at random positions on random strands. Returns (dsq,
    total length, [(start, end, strand)]) with 1-based inclusive coordinates."""
    r = np.random.default_rng(seed)
    p = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
    seq = list(r.choice(4, size=L, p=p).astype(np.uint8))
    ins = sorted(int(x) for x in r.integers(0, L, n))
    out, placed, shift = [], [], 0
    last = 0
    for k, pos in enumerate(ins):
        h = S.emit(cm, seed * 1000 + k)
        strand = int(r.integers(0, 2))
        if strand:
            h = np.where(h < 4, 3 - h, h)[::-1]
        out.extend(seq[last:pos])
        start = len(out) + 1
        out.extend(int(x) for x in h)
        placed.append((start, len(out), strand))
        last = pos
    out.extend(seq[last:])
    Lt = len(out)
    dsq = np.concatenate([[255], np.array(out, np.uint8), [255]]).astype(np.uint8)
    return dsq, Lt, placed
