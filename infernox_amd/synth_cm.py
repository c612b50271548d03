"""Synthetic covariance models and HMM bands of the benchmark shapes.

A CM is built from a consensus secondary structure the way cmbuild lays it
out: the structure becomes a guide tree (left-unpaired -> MATL, right-unpaired
-> MATR, outer pair -> MATP, otherwise a bifurcation into BEGL/BEGR subtrees),
and each node expands into its states with the child/parent connection rules
of cm_modelmaker.rs:979-1105 (cm_from_guide). Transition and emission
probabilities are sampled around a consensus (mostly match-to-match, base
pairs drawn GC/CG .3/.3, AU/UA .15/.15, GU/UG .05/.05).

Bands are shaped like CP9 bands: each state gets an i band and a j band
around its node's consensus span mapped onto the envelope, and the d bands
follow from them exactly as cp9.rs:3815 ij2d_bands derives them.
"""
import numpy as np

D_ST, MP_ST, ML_ST, MR_ST, IL_ST, IR_ST, S_ST, E_ST, B_ST = range(9)
BIF_ND, MATP_ND, MATL_ND, MATR_ND, BEGL_ND, BEGR_ND, ROOT_ND, END_ND = range(8)
CHILD_COUNT = [1, 4, 2, 2, 1, 1, 0, 1]      # cm_modelmaker.rs:982
PAIRS = [(2, 1), (1, 2), (0, 3), (3, 0), (2, 3), (3, 2)]  # GC CG AU UA GU UG
PAIR_W = [0.3, 0.3, 0.15, 0.15, 0.05, 0.05]


def wuss_to_ct(ss):
    """WUSS / bracket string -> 1-based pair table (0 = unpaired)."""
    ct = [0] * (len(ss) + 1)
    stack = {}
    opens, closes = "<([{", ">)]}"
    for i, c in enumerate(ss, 1):
        if c in opens:
            stack.setdefault(opens.index(c), []).append(i)
        elif c in closes:
            j = stack[closes.index(c)].pop()
            ct[i], ct[j] = j, i
    return ct


def cloverleaf(stems=(7, 4, 5, 5), loops=(7, 7, 7), clen=72):
    """tRNA-like cloverleaf of clen columns: acceptor stem, then D, anticodon
    and T arms (stems 7/4/5/5 bp, loops of 7-9 nt) joined by 2-, 1- and
    4-nt linkers, then an unpaired 3' tail."""
    acc, d, ac, t = stems
    arms = ["<" * s + "." * lp + ">" * s for s, lp in ((d, loops[0]), (ac, loops[1]), (t, loops[2]))]
    core = "<" * acc + ".." + arms[0] + "." + arms[1] + "...." + arms[2] + ">" * acc
    tail = clen - len(core)
    if tail < 0:
        raise ValueError("cloverleaf does not fit")
    return core + "." * tail


def multibranch(clen, nstems, seed, stem_range=(4, 9), loop_range=(4, 9)):
    """A nested structure of clen columns with nstems hairpins hung off
    multibranch loops (used for the 300-column, 12-stem model)."""
    r = np.random.default_rng(seed)
    stems = [int(r.integers(*stem_range)) for _ in range(nstems)]
    loops = [int(r.integers(*loop_range)) for _ in range(nstems)]
    hairpins = ["<" * s + "." * lp + ">" * s for s, lp in zip(stems, loops)]
    used = sum(len(h) for h in hairpins)
    # an enclosing stem around groups of hairpins to make multibranch loops
    free = clen - used
    if free < 2 * nstems:
        raise ValueError("structure does not fit")
    enclose = min(6, max(0, (free - nstems) // 4))
    linkers = free - 2 * enclose
    per = [linkers // (nstems + 1)] * (nstems + 1)
    per[-1] += linkers - sum(per)
    body = "".join("." * per[k] + hairpins[k] for k in range(nstems)) + "." * per[-1]
    return "<" * enclose + body + ">" * enclose


def random_nested(clen, seed, paired_frac=(0.4, 0.6)):
    """Random nested structure with a paired fraction in paired_frac."""
    r = np.random.default_rng(seed)
    target = float(r.uniform(*paired_frac))
    for _ in range(200):
        ss = ["."] * clen
        npair = 0
        tries = 0
        while npair < target * clen / 2 and tries < 4 * clen:
            tries += 1
            s = int(r.integers(3, 9))
            lp = int(r.integers(3, 9))
            span = 2 * s + lp
            if span >= clen:
                continue
            i = int(r.integers(0, clen - span + 1))
            seg = ss[i:i + span]
            # place only inside an unpaired stretch or fully enclosing a balanced block
            inner = "".join(ss[i + s:i + s + lp])
            if any(c != "." for c in seg[:s] + seg[s + lp:]) or inner.count("<") != inner.count(">"):
                continue
            if any(c != "." for c in seg[s:s + lp]) and "".join(seg[s:s + lp]).find(">") < \
                    "".join(seg[s:s + lp]).find("<"):
                continue
            for k in range(s):
                ss[i + k] = "<"
                ss[i + span - 1 - k] = ">"
            npair += s
        st = "".join(ss)
        ct = wuss_to_ct(st)
        if all(ct[ct[i]] == i for i in range(1, clen + 1) if ct[i]):
            return st
    return st


class Node:
    __slots__ = ("type", "a", "b", "left", "right", "prv")

    def __init__(self, t, a, b):
        self.type, self.a, self.b = t, a, b
        self.left = self.right = None
        self.prv = None


def guide_tree(ct, clen):
    """Guide tree nodes in the order cm_from_guide visits them (preorder,
    left subtree of a bifurcation first). Node .a/.b = consensus span."""
    root = Node(ROOT_ND, 1, clen)

    def build(i, j, parent):
        if i > j:
            n = Node(END_ND, i, j)
        elif ct[i] == 0:
            n = Node(MATL_ND, i, j)
            n.left = build(i + 1, j, n)
        elif ct[j] == 0:
            n = Node(MATR_ND, i, j)
            n.left = build(i, j - 1, n)
        elif ct[i] == j:
            n = Node(MATP_ND, i, j)
            n.left = build(i + 1, j - 1, n)
        else:
            n = Node(BIF_ND, i, j)
            k = ct[i]
            bl = Node(BEGL_ND, i, k)
            bl.left = build(i, k, bl)
            bl.left.prv = bl
            br = Node(BEGR_ND, k + 1, j)
            br.left = build(k + 1, j, br)
            br.left.prv = br
            n.left, n.right = bl, br
            bl.prv = br.prv = n
        n.prv = parent
        return n

    root.left = build(1, clen, root)
    order = []

    def visit(n):
        order.append(n)
        if n.type == BIF_ND:
            visit(n.left)
            visit(n.right)
        elif n.left is not None:
            visit(n.left)

    visit(root)
    return order


def build_cm(ss, seed, W=None, insert_self=(0.15, 0.45)):
    """Probability CM from a WUSS structure (dict of numpy arrays)."""
    clen = len(ss)
    ct = wuss_to_ct(ss)
    nodes = guide_tree(ct, clen)
    r = np.random.default_rng(seed)
    sttype, ndidx, cfirst, cnum, nodemap, ndtype = [], [], [], [], [], []
    node_lr = []
    bif_stack = []
    for nd_i, n in enumerate(nodes):
        s = len(sttype)
        nodemap.append(s)
        ndtype.append(n.type)
        node_lr.append((n.a, n.b))
        nxt = n.left.type if (n.left is not None and n.type != BIF_ND) else None
        cc = CHILD_COUNT[nxt] if nxt is not None else 0

        def add(t, cf, cn):
            sttype.append(t)
            ndidx.append(nd_i)
            cfirst.append(cf)
            cnum.append(cn)

        if n.type == BIF_ND:
            add(B_ST, s + 1, -1)
            bif_stack.append(s)
        elif n.type == MATP_ND:
            for k, t in enumerate((MP_ST, ML_ST, MR_ST, D_ST)):
                add(t, s + 4, 2 + cc)
            add(IL_ST, s + 4, 2 + cc)
            add(IR_ST, s + 5, 1 + cc)
        elif n.type == MATL_ND:
            add(ML_ST, s + 2, 1 + cc)
            add(D_ST, s + 2, 1 + cc)
            add(IL_ST, s + 2, 1 + cc)
        elif n.type == MATR_ND:
            add(MR_ST, s + 2, 1 + cc)
            add(D_ST, s + 2, 1 + cc)
            add(IR_ST, s + 2, 1 + cc)
        elif n.type == BEGL_ND:
            add(S_ST, s + 1, cc)
        elif n.type == BEGR_ND:
            b = bif_stack.pop()
            cnum[b] = s
            add(S_ST, s + 1, 1 + cc)
            add(IL_ST, s + 1, 1 + cc)
        elif n.type == ROOT_ND:
            add(S_ST, s + 1, 2 + cc)
            add(IL_ST, s + 1, 2 + cc)
            add(IR_ST, s + 2, 1 + cc)
        else:
            add(E_ST, -1, 0)
    M = len(sttype)
    sttype = np.array(sttype, np.int8)
    cfirst = np.array(cfirst, np.int32)
    cnum = np.array(cnum, np.int32)
    t = np.zeros((M, 6), np.float32)
    e = np.zeros((M, 16), np.float32)
    for v in range(M):
        st = sttype[v]
        if st in (B_ST, E_ST):
            continue
        n = cnum[v]
        kids = [cfirst[v] + k for k in range(n)]
        w = np.full(n, 0.02)
        # the consensus successor: the first non-insert child
        main = next((k for k, y in enumerate(kids) if sttype[y] not in (IL_ST, IR_ST)), n - 1)
        w[main] = 0.9
        if st in (IL_ST, IR_ST):
            w[0] = r.uniform(*insert_self)  # self loop
        w = w * r.uniform(0.7, 1.3, size=n)
        t[v, :n] = w / w.sum()
        if st == MP_ST:
            pr = r.uniform(0.6, 0.85)
            a, b = PAIRS[int(r.choice(6, p=PAIR_W))]
            rest = r.dirichlet(np.ones(15)) * (1 - pr)
            em = np.insert(rest, a * 4 + b, pr)
            e[v] = em
        elif st in (ML_ST, MR_ST):
            pc = r.uniform(0.55, 0.9)
            c = int(r.integers(0, 4))
            rest = r.dirichlet(np.ones(3) * 2) * (1 - pc)
            e[v, :4] = np.insert(rest, c, pc)
        elif st in (IL_ST, IR_ST):
            e[v, :4] = 0.25
    if W is None:
        W = int(round(1.3 * clen)) + 8
    return dict(M=M, nodes=len(nodes), clen=clen, W=W, null=np.full(4, 0.25, np.float32),
                el_selfsc=np.float32(-0.08926734), pbegin=np.float32(0.05), pend=np.float32(0.05),
                n3_omega=0.03125, sttype=sttype, ndtype=np.array(ndtype, np.int8),
                nodemap=np.array(nodemap, np.int32), ndidx=np.array(ndidx, np.int32), cfirst=cfirst, cnum=cnum,
                t=t, e=e, node_lr=np.array(node_lr, np.int32), ss=ss)


def state_delta(st):
    return 2 if st == MP_ST else (1 if st in (ML_ST, MR_ST, IL_ST, IR_ST) else 0)


def state_left_delta(st):
    return 1 if st in (MP_ST, ML_ST, IL_ST) else 0


def state_right_delta(st):
    return 1 if st in (MP_ST, MR_ST, IR_ST) else 0


def ij_bands(cm, i0, j0, r, seed, jitter=2, c1=1, c2=None):
    """CP9-shaped i/j bands of every state for the envelope [i0, j0]: the
    node's consensus span [a, b] mapped linearly onto the envelope, widened by
    r (+ a seeded jitter of 0..jitter). With c1/c2 only consensus columns
    c1..c2 map onto the envelope (a truncated alignment); spans outside are
    clipped to the envelope's ends. Returns imin, imax, jmin, jmax (int32)."""
    rng = np.random.default_rng(seed)
    M = cm["M"]
    c2 = cm["clen"] if c2 is None else c2
    scale = (j0 - i0 + 1) / (c2 - c1 + 1)
    lr = cm["node_lr"][cm["ndidx"]]
    ic = i0 + np.round((lr[:, 0] - c1) * scale).astype(np.int64)
    jc = i0 + np.round((lr[:, 1] - c1 + 1) * scale).astype(np.int64) - 1
    ic = np.clip(ic, i0, j0 + 1)
    jc = np.clip(jc, i0 - 1, j0)
    ri = r + rng.integers(0, jitter + 1, size=M)
    rj = r + rng.integers(0, jitter + 1, size=M)
    imin = np.maximum(i0, ic - ri).astype(np.int32)
    imax = np.minimum(j0 + 1, ic + ri).astype(np.int32)
    jmin = np.maximum(i0 - 1, jc - rj).astype(np.int32)
    jmax = np.minimum(j0, jc + rj).astype(np.int32)
    return imin, imax, jmin, jmax


def ij2d(cm, imin, imax, jmin, jmax, do_trunc=False):
    """cp9.rs:3815-3850 ij2d_bands: per (v, j) d ranges, flattened."""
    M = cm["M"]
    st = np.asarray(cm["sttype"], np.int64)
    n = np.maximum(0, np.asarray(jmax, np.int64) - np.asarray(jmin, np.int64) + 1)
    hd_off = np.zeros(M + 1, np.int32)
    hd_off[1:] = np.cumsum(n)
    v = np.repeat(np.arange(M), n)
    jp = np.arange(int(n.sum())) - np.repeat(hd_off[:-1].astype(np.int64), n)
    j = jp + np.asarray(jmin, np.int64)[v]
    sd = np.where(st == MP_ST, 2, np.where((st >= ML_ST) & (st <= IR_ST), 1, 0))
    sl = np.isin(st, (MP_ST, ML_ST, IL_ST)).astype(np.int64)
    sr = np.isin(st, (MP_ST, MR_ST, IR_ST)).astype(np.int64)
    dn = (np.maximum(sl, sr) if do_trunc else sd)[v]
    hdn = j - np.asarray(imax, np.int64)[v] + 1
    hdx = j - np.asarray(imin, np.int64)[v] + 1
    hmn = np.where(hdx < dn, -1, np.maximum(hdn, dn))
    hmx = np.where(hdx < dn, -2, hdx)
    e = st[v] == E_ST
    hmn = np.where(e, 0, hmn)
    hmx = np.where(e, 0, hmx)
    return dict(jmin=np.asarray(jmin, np.int32), jmax=np.asarray(jmax, np.int32), hd_off=hd_off,
                hdmin=hmn.astype(np.int32), hdmax=hmx.astype(np.int32))


def bands(cm, i0, j0, r=8, seed=0, do_trunc=False):
    return ij2d(cm, *ij_bands(cm, i0, j0, r, seed), do_trunc=do_trunc)


def wide_bands(cm, i0, j0):
    """Bands covering every (j, d) with i0 <= i <= j+1 <= j0+1: the banded
    scanner then equals the unbanded one."""
    M = cm["M"]
    imin = np.full(M, i0, np.int32)
    imax = np.full(M, j0 + 1, np.int32)
    jmin = np.full(M, i0 - 1, np.int32)
    jmax = np.full(M, j0, np.int32)
    return ij2d(cm, imin, imax, jmin, jmax)


def hb_cells(b):
    w = b["hdmax"] - b["hdmin"] + 1
    return int(w[w > 0].sum())


def skew_inserts(cm, seed, alpha=0.7):
    """A copy of cm whose IL / IR emissions are drawn from a Dirichlet(alpha)
    instead of the background, so that insert emission scores are not 0."""
    rng = np.random.default_rng(seed)
    e = cm["e"].copy()
    for v in np.flatnonzero(np.isin(cm["sttype"], (IL_ST, IR_ST))):
        e[v, :4] = rng.dirichlet(np.full(4, alpha))
    return dict(cm, e=e.astype(np.float32))


def emit(cm, seed):
    """Samples one sequence from the CM from its root (global parse)."""
    r = np.random.default_rng(seed)
    st, cf, cn, t, e = cm["sttype"], cm["cfirst"], cm["cnum"], cm["t"], cm["e"]

    def go(v):
        s = st[v]
        if s == E_ST:
            return []
        if s == B_ST:
            return go(cf[v]) + go(cn[v])
        left, right = [], []
        if s == MP_ST:
            x = int(r.choice(16, p=e[v] / e[v].sum()))
            left, right = [x // 4], [x % 4]
        elif s in (ML_ST, IL_ST):
            left = [int(r.choice(4, p=e[v, :4] / e[v, :4].sum()))]
        elif s in (MR_ST, IR_ST):
            right = [int(r.choice(4, p=e[v, :4] / e[v, :4].sum()))]
        n = cn[v]
        p = t[v, :n].astype(np.float64)
        y = cf[v] + int(r.choice(n, p=p / p.sum()))
        return left + go(y) + right

    import sys
    sys.setrecursionlimit(100000)
    return np.array(go(0), np.uint8)


# ---- truncated scanning inputs ----

PASS_STD_ANY, PASS_5P_ONLY, PASS_3P_ONLY, PASS_5P_AND_3P_FORCE, PASS_5P_AND_3P_ANY = 1, 2, 3, 4, 5
TRMODE_T, TRMODE_R, TRMODE_L, TRMODE_J, TRMODE_UNKNOWN = 0, 1, 2, 3, 4


def fill_flags(pass_idx):
    """(fill L, fill R, fill T) of a pipeline pass (cm_dpsearch_trunc.rs:184-199)."""
    both = pass_idx in (PASS_5P_AND_3P_FORCE, PASS_5P_AND_3P_ANY)
    return (both or pass_idx == PASS_3P_ONLY, both or pass_idx == PASS_5P_ONLY, both)


def state_spans(cm):
    """Consensus span (first, last column) of every state's subtree: the
    node spans of the guide tree, which is what the emit map gives after its
    per-node-type offsets (cp9.rs:144-204, 1934-1943)."""
    lr = cm["node_lr"][cm["ndidx"]]
    return lr[:, 0].astype(np.int64), lr[:, 1].astype(np.int64)


def _tr_qualifies(cm):
    st, nd = cm["sttype"], cm["ndtype"][cm["ndidx"]]
    return ((st == MP_ST) | ((st == ML_ST) & (nd == MATL_ND)) | ((st == MR_ST) & (nd == MATR_ND)) |
            (st == B_ST) | np.isin(st, (IL_ST, IR_ST)))


def _tr_penalty_probs(cm, seed):
    """Truncated-begin probabilities pg[3, M] (global) and pl[3, M] (local) as
    float32, 0 where the entry is not allowed; index 0 = 5'&3', 1 = 5' only,
    2 = 3' only. Shaped like cm_trunc.rs:192-384 (global: 2/(clen(clen+1))
    or 1/clen spread over the qualifying states of a node; local: the local
    begin probability over the subtree's fragment count), with a seeded
    jitter: the scanners treat them as given."""
    rng = np.random.default_rng(seed)
    M, clen = cm["M"], cm["clen"]
    q = _tr_qualifies(cm)
    lp, rp = state_spans(cm)
    pg = np.zeros((3, M), np.float64)
    pl = np.zeros((3, M), np.float64)
    jit = rng.uniform(0.8, 1.0, size=M)
    g53 = 2.0 / (clen * (clen + 1))
    pg[0, q] = g53 * jit[q]
    m5 = q & (rp == clen)
    m3 = q & (lp == 1)
    pg[1, m5] = jit[m5] / clen
    pg[2, m3] = jit[m3] / clen
    sub = np.maximum(1, rp - lp + 1).astype(np.float64)
    pb = 0.05 / max(1, int(q.sum()))
    pl[0, q] = pb / (sub[q] * (sub[q] + 1) / 2) * jit[q]
    pl[1, q] = pb / sub[q] * jit[q]
    pl[2, q] = pb / sub[q] * jit[q]
    return pg.astype(np.float32), pl.astype(np.float32)


def tr_penalties(cm, seed=0):
    """Truncated-begin penalties g[3, M] (global) and l[3, M] (local), log2
    of the probabilities of _tr_penalty_probs, -1e36 where the entry is not
    allowed (cm_trunc.rs:349-361 converts with sreLOG2)."""
    out = []
    for p in _tr_penalty_probs(cm, seed):
        lg = np.full(p.shape, -1e36, np.float32)
        ok = p > 0
        lg[ok] = (np.log(p[ok].astype(np.float64)) * 1.44269504).astype(np.float32)
        out.append(lg)
    return out[0], out[1]


def tr_ipenalties(cm, seed=0):
    """Integer truncated-begin penalties ig[3, M] and il[3, M] of the same
    probabilities: Prob2Score(p, 1) = floor(0.5 + 1000 log2 p), -INFTY where
    not allowed (cm_trunc.rs:349-374)."""
    out = []
    for p in _tr_penalty_probs(cm, seed):
        ip = np.full(p.shape, -987654321, np.int32)
        ok = p > 0
        ip[ok] = np.floor(0.5 + 1000.0 * np.log2(p[ok].astype(np.float64))).astype(np.int32)
        out.append(ip)
    return out[0], out[1]


def tr_validity(cm, pass_idx, sp1, sp2, ep1, ep2):
    """jvalid/lvalid/rvalid/tvalid (uint8[M]) of an envelope whose alignment
    may start in consensus columns sp1..sp2 and end in ep1..ep2
    (cp9.rs:1925-1990)."""
    al, ar, at = fill_flags(pass_idx)
    st = cm["sttype"]
    lp, rp = state_spans(cm)
    jv = (lp >= sp1) & (rp <= ep1)
    lv = al & (lp >= sp1) & (lp <= ep1) & (rp > ep2)
    rv = ar & (rp <= ep1) & (rp >= sp1) & (lp < sp2)
    tv = (st == B_ST) & at & (lp < sp2) & (rp > ep2)
    pj = np.isin(st, (B_ST, MP_ST, ML_ST, IL_ST, MR_ST, IR_ST))
    pl = np.isin(st, (B_ST, MP_ST, ML_ST, IL_ST))
    pr = np.isin(st, (B_ST, MP_ST, MR_ST, IR_ST))
    jv[0] |= bool((pj & jv).any())
    lv[0] |= bool((pl & lv).any())
    rv[0] |= bool((pr & rv).any())
    tv[0] |= bool(((st == B_ST) & tv).any())
    return dict(jvalid=jv.astype(np.uint8), lvalid=lv.astype(np.uint8), rvalid=rv.astype(np.uint8),
                tvalid=tv.astype(np.uint8))


def tr_bands(cm, i0, j0, r=8, seed=0, c1=1, c2=None):
    """Bands of a truncated scan: d bands start at max(left, right) emission
    count instead of the full count (cp9.rs:3832-3836); consensus columns
    c1..c2 are the part of the model present in the envelope."""
    return ij2d(cm, *ij_bands(cm, i0, j0, r, seed, c1=c1, c2=c2), do_trunc=True)


def wide_tr_bands(cm, i0, j0):
    M = cm["M"]
    return ij2d(cm, np.full(M, i0, np.int32), np.full(M, j0 + 1, np.int32), np.full(M, i0 - 1, np.int32),
                np.full(M, j0, np.int32), do_trunc=True)
