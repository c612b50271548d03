"""Synthetic inputs of the benchmark workloads' shapes.

* ``lcg_profile`` / ``lcg_dsq``: the seeded random p7 profile and digitized
  sequence generator of the reference's own kernel test
  (p7_vitfilter.rs:547-597: a 64-bit LCG, uniform [0.05, 1.05) draws
  normalized per distribution, residue codes 0..15).
* ``consensus_profile``: an RNA-family-like p7 filter HMM (a consensus residue
  per column, mostly match-to-match transitions) for throughput runs.
* ``iid_dsq`` / ``markov2_dsq``: genome-like targets; ``plant``: inserts
  homologs sampled from a p7 profile on random strands.

Digitized alphabet: 0..3 = A,C,G,U, 4 gap, 5..15 IUPAC, 255 sentinel.
"""
import numpy as np

MASK64 = (1 << 64) - 1
LN2 = float(np.log(2.0))


class Lcg:
    """The reference test's LCG (p7_vitfilter.rs:547-560)."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u32(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK64
        return self.s >> 33

    def unit(self):
        # 0.05 + (x as f32 / u32::MAX as f32), all in f32
        x = np.float32(self.next_u32())
        return np.float32(np.float32(0.05) + np.float32(x / np.float32(4294967295.0)))


def lcg_profile(rng, M, ev=None):
    """p7_vitfilter.rs:563-585 random_profile: returns a profile dict."""
    mat = np.zeros((M + 1, 4), np.float32)
    mat[0, :] = 0.25
    for k in range(1, M + 1):
        e = [rng.unit() for _ in range(4)]
        s = np.float32(0.0)
        for x in range(4):
            s = np.float32(s + e[x])
        mat[k] = [np.float32(e[x] / s) for x in range(4)]
    trans = np.zeros((M + 1, 7), np.float32)
    for k in range(M + 1):
        mm, mi, md = rng.unit(), rng.unit(), rng.unit()
        ms = np.float32(np.float32(mm + mi) + md)
        im, ii = rng.unit(), rng.unit()
        is_ = np.float32(im + ii)
        dm, dd = rng.unit(), rng.unit()
        ds = np.float32(dm + dd)
        trans[k] = [mm / ms, mi / ms, md / ms, im / is_, ii / is_, dm / ds, dd / ds]
    p = dict(M=M, mat=mat, trans=trans, compo=mat[1:].mean(axis=0).astype(np.float32))
    p.update(ev or dict(lmmu=-9.0, lmlambda=0.7, lvmu=-9.5, lvlambda=0.7, lftau=-4.0, lflambda=0.7))
    return p


def lcg_dsq(rng, L):
    """p7_vitfilter.rs:587-596 random_dsq (codes 0..15; sentinels 0)."""
    d = np.zeros(L + 2, np.uint8)
    for i in range(1, L + 1):
        d[i] = rng.next_u32() % 16
    return d


def consensus_profile(M, seed, gc=0.5):
    """An RNA-family-like p7 profile: consensus residue per column with
    match probability 0.55-0.9, mostly M->M transitions."""
    r = np.random.default_rng(seed)
    base_p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    cons = r.choice(4, size=M + 1, p=base_p)
    mat = np.zeros((M + 1, 4), np.float64)
    mat[0] = 0.25
    for k in range(1, M + 1):
        pc = r.uniform(0.55, 0.9)
        rest = r.dirichlet(np.ones(3) * 2.0) * (1 - pc)
        row = np.insert(rest, cons[k], pc) if cons[k] < 3 else np.append(rest, pc)
        mat[k] = row
    trans = np.zeros((M + 1, 7), np.float64)
    for k in range(M + 1):
        mi = r.uniform(0.004, 0.03)
        md = r.uniform(0.004, 0.03)
        im = r.uniform(0.55, 0.85)
        dm = r.uniform(0.55, 0.9)
        trans[k] = [1 - mi - md, mi, md, im, 1 - im, dm, 1 - dm]
    trans[M] = [1, 0, 0, 1, 0, 1, 0]
    mat = mat.astype(np.float32)
    return dict(M=M, mat=mat, trans=trans.astype(np.float32),
                compo=mat[1:].mean(axis=0).astype(np.float32), cons=cons)


def iid_dsq(L, seed, gc=0.5):
    r = np.random.default_rng(seed)
    p = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
    d = np.empty(L + 2, np.uint8)
    d[0] = d[L + 1] = 255
    d[1:L + 1] = r.choice(4, size=L, p=p).astype(np.uint8)
    return d


def markov2_dsq(L, seed, gc=0.508):
    """Order-2 Markov chain: P(x | previous two) = the GC-matched base
    composition scaled by a seeded per-context factor, renormalized."""
    r = np.random.default_rng(seed)
    base = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    P = base[None, :] * r.uniform(0.7, 1.3, size=(16, 4))
    P /= P.sum(axis=1, keepdims=True)
    cum = np.cumsum(P, axis=1)
    u = r.random(L)
    out = np.empty(L + 2, np.uint8)
    out[0] = out[L + 1] = 255
    a, b = 0, 1
    # vectorized by blocks is not possible for a Markov chain; keep it simple
    cuml = cum.tolist()
    ul = u.tolist()
    res = bytearray(L)
    for i in range(L):
        row = cuml[a * 4 + b]
        x = 0 if ul[i] < row[0] else (1 if ul[i] < row[1] else (2 if ul[i] < row[2] else 3))
        res[i] = x
        a, b = b, x
    out[1:L + 1] = np.frombuffer(bytes(res), np.uint8)
    return out


def emit_p7(p, r):
    """Samples one local homolog from a p7 profile: a random model span
    k1..k2 traversed with the profile's M/I/D transitions."""
    M = p["M"]
    mat, trans = p["mat"].astype(np.float64), p["trans"].astype(np.float64)
    k1 = int(r.integers(1, max(2, M // 4)))
    k2 = int(r.integers(max(k1 + 1, 3 * M // 4), M + 1))
    out = []
    k, st = k1, 0  # 0 match, 1 insert, 2 delete
    while k <= k2:
        if st == 0:
            out.append(int(r.choice(4, p=mat[k] / mat[k].sum())))
            t = trans[k][:3]
        elif st == 1:
            out.append(int(r.integers(0, 4)))
            t = np.array([trans[k][3], trans[k][4], 0.0])
        else:
            t = np.array([trans[k][5], 0.0, trans[k][6]])
        t = t / t.sum()
        nxt = int(r.choice(3, p=t))
        if nxt == 1 and st != 2:
            st = 1
            continue
        k += 1
        st = 0 if nxt == 0 else 2
    return np.array(out, np.uint8)


def plant(dsq, L, p, n, seed):
    """Overwrites n positions of dsq with sampled homologs on random strands.
    Returns the planted (start, end, strand) list (1-based, inclusive)."""
    r = np.random.default_rng(seed)
    where = []
    for _ in range(n):
        h = emit_p7(p, r)
        if len(h) == 0 or len(h) >= L:
            continue
        s = int(r.integers(1, L - len(h) + 1))
        strand = int(r.integers(0, 2))
        if strand:
            h = np.where(h < 4, 3 - h, h)[::-1]
        dsq[s:s + len(h)] = h
        where.append((s, s + len(h) - 1, strand))
    return where


# E-value parameters of the benchmark profiles consensus_profile(M, seed),
# calibrated with tests/golden/calibrate_synth.py (lambda = ln 2; MSV/Viterbi
# mu by known-lambda Gumbel ML over 200 iid 200-mers; Forward tau from the 4%
# tail over 200 iid 100-mers).
BENCH_EV = {
    (72, 1): {'lmmu': -8.1547, 'lmlambda': 0.693147, 'lvmu': -8.8831, 'lvlambda': 0.693147, 'lftau': -3.4469,
              'lflambda': 0.693147},
    (300, 3): {'lmmu': -10.042, 'lmlambda': 0.693147, 'lvmu': -10.884, 'lvlambda': 0.693147, 'lftau': -5.2848,
               'lflambda': 0.693147},
    (120, 4): {'lmmu': -8.8689, 'lmlambda': 0.693147, 'lvmu': -9.5312, 'lvlambda': 0.693147, 'lftau': -4.6316,
               'lflambda': 0.693147},
}


def bench_profile(M, seed):
    p = consensus_profile(M, seed)
    p.update(BENCH_EV[(M, seed)])
    return p
