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
  random sequence (tests/golden/calibrate_models.py, values held here);
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
# from tests/golden/calibrate_models.py 1: p7 local MSV / Viterbi Gumbel mu,
# local Forward and glocal Forward exponential tau (lambda ln 2); CM local
# CYK and Inside exponential tails (mu, lambda) in bits.
CONFIG1_EV = dict(lmmu=-7.737, lmlambda=0.693147, lvmu=-8.7349, lvlambda=0.693147, lftau=-3.7808,
                  lflambda=0.693147, gfmu=-17.6875, gflambda=0.693147)
CONFIG1_CM_EV = dict(lc=(-1.0937, 0.693147), li=(0.1358, 0.693147))


def config1_models(ev=None, cm_ev=None):
    """(p7 dict, CM dict, CP9 dict, cm_ev) of the config-1 model."""
    ss = cloverleaf_ss()
    aln = alignment(ss, 40, seed=1)
    cm = cm_from_alignment(ss, aln, seed=1)
    cp9 = P.build_cp9_like(cm, 1, el=True)
    p7 = p7_from_cm(cm, ev or CONFIG1_EV)
    return p7, cm, cp9, (cm_ev or CONFIG1_CM_EV)


def three_hairpin_ss():
    """120-column structure of three hairpins in a row: stems of 8, 10 and
    9 bp closing loops of 10, 12 and 8 nt, with 8 / 7 / 9 / 12-nt unpaired
    stretches before, between and after them."""
    return ("." * 8 + "<" * 8 + "." * 10 + ">" * 8 + "." * 7 + "<" * 10 + "." * 12 + ">" * 10 + "." * 9 + "<" * 9 +
            "." * 8 + ">" * 9 + "." * 12)


# the same for the three-hairpin model (calibrate_models.py 3)
CONFIG3_EV = dict(lmmu=-8.4042, lmlambda=0.693147, lvmu=-9.323, lvlambda=0.693147, lftau=-3.805,
                  lflambda=0.693147, gfmu=-36.98, gflambda=0.693147)
CONFIG3_CM_EV = dict(lc=(-1.2905, 0.693147), li=(-0.1442, 0.693147))


def config3_models(ev=None, cm_ev=None):
    """(p7 dict, CM dict, CP9 dict, cm_ev) of the batched-workload model: the
    three-hairpin structure, a 40-sequence alignment (seed 4) and the same
    builders as config1_models."""
    ss = three_hairpin_ss()
    aln = alignment(ss, 40, seed=4)
    cm = cm_from_alignment(ss, aln, seed=4)
    cp9 = P.build_cp9_like(cm, 1, el=True)
    p7 = p7_from_cm(cm, ev or CONFIG3_EV)
    return p7, cm, cp9, (cm_ev or CONFIG3_CM_EV)


# the same for the 300-column model (calibrate_models.py 2 50: 50 CM tail
# samples, the non-banded scans of its 959 states being the slow part)
CONFIG2_EV = dict(lmmu=-9.8474, lmlambda=0.693147, lvmu=-10.4678, lvlambda=0.693147, lftau=-6.026,
                  lflambda=0.693147, gfmu=-119.6546, gflambda=0.693147)
CONFIG2_CM_EV = dict(lc=(-0.8797, 0.693147), li=(1.1846, 0.693147))


def config2_models(ev=None, cm_ev=None):
    """(p7 dict, CM dict, CP9 dict, cm_ev) of the genome-scale model: a
    300-column structure of 12 hairpins in multibranch loops (seed 3), a
    60-sequence alignment (seed 3) and the same builders as config1_models."""
    ss = S.multibranch(300, 12, 3)
    aln = alignment(ss, 60, seed=3)
    cm = cm_from_alignment(ss, aln, seed=3)
    cp9 = P.build_cp9_like(cm, 1, el=True)
    p7 = p7_from_cm(cm, ev or CONFIG2_EV)
    return p7, cm, cp9, (cm_ev or CONFIG2_CM_EV)


def markov_genome_with_homologs(cm, L, n, seed, gc=0.508):
    """An order-2 Markov sequence of L residues at the given GC (seed) with n
    CM-sampled homologs inserted at random positions on random strands.
    Returns (dsq, total length, [(start, end, strand)])."""
    from . import synth
    base = synth.markov2_dsq(L, seed, gc=gc)[1:L + 1]
    r = np.random.default_rng(seed + 1)
    ins = sorted(int(x) for x in r.integers(0, L, n))
    out, placed, last = [], [], 0
    for k, pos in enumerate(ins):
        h = S.emit(cm, seed * 1000 + k).astype(np.uint8)
        if int(r.integers(0, 2)):
            h = np.where(h < 4, 3 - h, h)[::-1].astype(np.uint8)
            strand = 1
        else:
            strand = 0
        out.append(base[last:pos])
        start = sum(len(x) for x in out) + 1
        out.append(h)
        placed.append((start, start + len(h) - 1, strand))
        last = pos
    out.append(base[last:])
    seq = np.concatenate(out).astype(np.uint8)
    return np.concatenate([[255], seq, [255]]).astype(np.uint8), len(seq), placed


def lognormal_database(n, median, sigma, lo, hi, gc, seed, cm=None, n_hom=0):
    """n iid sequences (ACGU at the given GC) with log-normal lengths (the
    given median and sigma, clipped to [lo, hi]); with a CM, n_hom
    CM-sampled homologs inserted into randomly chosen sequences on random
    strands. Returns (list of 1-indexed dsq arrays, [(seq, start, end,
    strand)])."""
    r = np.random.default_rng(seed)
    lens = np.clip(np.rint(r.lognormal(np.log(median), sigma, n)), lo, hi).astype(np.int64)
    p = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
    seqs = [r.choice(4, size=int(k), p=p).astype(np.uint8) for k in lens]
    placed = []
    if cm is not None:
        for k in range(n_hom):
            s = int(r.integers(0, n))
            h = S.emit(cm, seed * 100000 + k)
            strand = int(r.integers(0, 2))
            if strand:
                h = np.where(h < 4, 3 - h, h)[::-1]
            at = int(r.integers(0, len(seqs[s]) + 1))
            seqs[s] = np.concatenate([seqs[s][:at], h.astype(np.uint8), seqs[s][at:]])
            placed = [(q, a + len(h) if q == s and a > at else a, b + len(h) if q == s and a > at else b, sd)
                      for q, a, b, sd in placed]
            placed.append((s, at + 1, at + len(h), strand))
    dsqs = [np.concatenate([[255], x, [255]]).astype(np.uint8) for x in seqs]
    return dsqs, placed


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
