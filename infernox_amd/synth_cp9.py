"""Synthetic CP9 HMMs (the band-producing HMM of a CM) for tests and the
bench: integer log-odds scores in the DP's transition order and the EL
(local end) connectivity derived from a CM's guide tree, as the reference
lays them out (cp9.rs:871-918 CP9 and the cp9O order; cp9.rs:1042-1076
cp9_init_el; cp9.rs:144-207 create_emit_map). Probabilities are random
(seeded); the CM -> CP9 construction itself is not reproduced."""
import numpy as np

from . import synth_cm as S

INFTY = 987654321
# cp9O transition order (cp9.rs:906-918)
MM, IM, DM, BM, MI, II, DI, MD, ID, DD, ME, MEL = range(12)


def emit_map(cm):
    """Left and right consensus positions of every node (cp9.rs:144-207)."""
    nodes, nt, nodemap = cm["nodes"], cm["ndtype"], cm["nodemap"]
    ndidx = np.zeros(cm["M"], np.int32)
    for nd in range(nodes):
        ndidx[nodemap[nd]:(nodemap[nd + 1] if nd + 1 < nodes else cm["M"])] = nd
    lpos, rpos = np.full(nodes, -1, np.int32), np.full(nodes, -1, np.int32)
    cpos, pda = 0, [(0, 0)]
    while pda:
        right, nd = pda.pop()
        t = nt[nd]
        if right:
            rpos[nd] = cpos + 1
            if t in (S.MATP_ND, S.MATR_ND):
                cpos += 1
        else:
            if t in (S.MATP_ND, S.MATL_ND):
                cpos += 1
            lpos[nd] = cpos
            pda.append((1, nd))
            if t == S.BIF_ND:
                v = nodemap[nd]
                pda.append((0, int(ndidx[cm["cnum"][v]])))
                pda.append((0, int(ndidx[cm["cfirst"][v]])))
            elif t != S.END_ND:
                pda.append((0, nd + 1))
    return lpos, rpos


def prob2score(p, null=1.0):
    """cp9.rs:977 Prob2Score: floor(0.5 + 1000 log2(p / null)), -INFTY at 0."""
    p = np.asarray(p, np.float32)
    with np.errstate(divide="ignore"):
        r = (p / np.float32(null)).astype(np.float32)
        v = np.floor(0.5 + 1000.0 * np.where(r > 0, np.log(r.astype(np.float64)) * 1.44269504, -1e36))
    return np.where(p == 0, -INFTY, np.maximum(v, -INFTY)).astype(np.int32)


def _nexits(cm):
    """Nodes with a local end (cp9.rs:1121-1130)."""
    nt = cm["ndtype"]
    return sum(1 for nd in range(1, cm["nodes"])
               if nt[nd] in (S.MATP_ND, S.MATL_ND, S.MATR_ND, S.BEGL_ND, S.BEGR_ND) and nt[nd + 1] != S.END_ND)


def build_cp9(cm, seed=0, el=True, pbegin=0.05, pend=0.05, gap=0.05, trunc=None):
    """A configured CP9 for the CM's consensus (M = clen): dict with M, el,
    el_selfsc, otsc [(M+1), 12], msc / isc [(M+1), 18], has_el [M+1],
    el_from_ct [M+2], el_from_off [M+3], el_from_idx.

    trunc = (pentry, pexit) configures the begins and ends of the same draws
    as cp9_sw_config does for the truncated passes' CP9s (cp9.rs:1097-1119:
    begin[1] = (1 - pentry)(1 - s), begin[k >= 2] = pentry (1 - s) / (M - 1),
    end[k < M] = b / (1 - b (k - 1)) with b = pexit / (M - 1), the match
    transitions rescaled to the rest; cp9.rs:2235-2330), then local ends
    (cp9.rs:1121-1152, pend over the exit nodes) when el is set."""
    rng = np.random.default_rng(seed)
    lpos, rpos = emit_map(cm)
    M = int(rpos[0] - 1)
    t = np.zeros((M + 1, 12), np.float64)
    for k in range(M + 1):
        # from M_k: M, I, D, E, EL; from I_k and D_k: M, I, D
        m = rng.dirichlet([40.0, 1.0, 1.0])
        end = pend if k >= 1 else 0.0
        elp = 0.02 if (el and k >= 1) else 0.0
        if trunc is not None:
            end, elp = 0.0, 0.0
        t[k, MM], t[k, MI], t[k, MD] = m * (1 - end - elp)
        t[k, ME], t[k, MEL] = end, elp
        t[k, IM], t[k, II], t[k, ID] = rng.dirichlet([6.0, 3.0, 0.5])
        t[k, DM], t[k, DI], t[k, DD] = rng.dirichlet([6.0, 0.5, 2.0])
        t[k, BM] = pbegin if k >= 1 else 0.0
    # the begin state reaches M_1 only through its begin transition, as a
    # locally configured CP9 does (the M_0 -> M_1 transition is impossible)
    t[0, BM] = 0.0
    t[0, MI], t[0, MD] = t[0, MI] / (t[0, MI] + t[0, MD] + 1e-9) * 0.1, t[0, MD] / (t[0, MI] + t[0, MD] + 1e-9) * 0.1
    t[0, MM] = 0.0
    if M >= 1:
        t[M, MM] = t[M, MD] = t[M, ID] = t[M, DD] = 0.0
        t[M, ME] = 1.0 - t[M, MI] - t[M, MEL]
    if trunc is not None and M >= 2:
        pentry, pexit = trunc
        s0 = t[0, MI] + t[0, MD]
        t[1:, BM] = pentry * (1.0 - s0) / (M - 1)
        t[1, BM] = (1.0 - pentry) * (1.0 - s0)
        b = pexit / (M - 1)
        for k in range(1, M):
            t[k, ME] = b / (1.0 - b * (k - 1))
            tot = t[k, MM] + t[k, MI] + t[k, MD] + t[k, MEL]
            for x in (MM, MI, MD, MEL):
                t[k, x] *= (1.0 - t[k, ME]) / tot
        if el:
            to_el = float(cm["pend"]) / max(1, _nexits(cm))
            for nd in range(cm["nodes"]):
                if cm["ndtype"][nd] in (S.MATP_ND, S.MATL_ND, S.MATR_ND, S.BEGL_ND, S.BEGR_ND) and \
                        cm["ndtype"][nd + 1] != S.END_ND:
                    k = int(lpos[nd])
                    if 1 <= k <= M and t[k, MEL] == 0.0:
                        t[k, MEL] = to_el
                        f = 1.0 - to_el / (1.0 - t[k, ME])
                        for x in (MM, MI, MD):
                            t[k, x] *= f
    otsc = prob2score(t.astype(np.float32))
    msc = np.zeros((M + 1, 18), np.int32)
    isc = np.zeros((M + 1, 18), np.int32)
    for k in range(M + 1):
        me = rng.dirichlet([1.0] * 4) if k >= 1 else np.full(4, 0.25)
        ie = rng.dirichlet([8.0] * 4)
        for e, out in ((me, msc), (ie, isc)):
            sc = prob2score(e.astype(np.float32), 0.25)
            out[k, :4] = sc
            out[k, 4:] = int(np.mean(sc))  # degenerate codes: the mean score
            out[k, 4] = out[k, 16] = out[k, 17] = -INFTY  # gap, non-residue, missing
    if M >= 1:
        msc[0, :] = -INFTY
    has_el = np.zeros(M + 1, np.uint8)
    srcs = [[] for _ in range(M + 2)]
    if el:
        for nd in range(cm["nodes"]):
            if cm["ndtype"][nd] in (S.MATP_ND, S.MATL_ND, S.MATR_ND, S.BEGL_ND, S.BEGR_ND) and \
                    cm["ndtype"][nd + 1] != S.END_ND:
                srcs[rpos[nd]].append(int(lpos[nd]))
                has_el[lpos[nd]] = 1
    ct = np.array([len(x) for x in srcs], np.int32)
    off = np.concatenate([[0], np.cumsum(ct)]).astype(np.int32)
    idx = np.array([y for x in srcs for y in x] or [0], np.int32)
    return dict(M=M, el=bool(el), el_selfsc=int(prob2score(np.float32(0.94))), otsc=otsc, msc=msc, isc=isc,
                has_el=has_el, el_from_ct=ct, el_from_off=off, el_from_idx=idx)


def build_cp9_like(cm, seed=0, el=True, pbegin=0.05, pend=0.05, trunc=None):
    """A configured CP9 whose match emissions follow the CM's consensus
    columns (the marginal of the CM state that emits each column: MATL ML,
    MATR MR, or the left / right marginal of MATP MP) and whose transitions
    mostly stay on the match path, so its posterior bands track the CM's
    homologs. A stand-in for the reference's CM -> CP9 construction
    (cp9.rs:2212), which is not reproduced; the local-end lists are the
    guide tree's, as in build_cp9."""
    h = build_cp9(cm, seed=seed, el=el, pbegin=pbegin, pend=pend, trunc=trunc)
    lpos, rpos = emit_map(cm)
    M = h["M"]
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
    msc = h["msc"].copy()
    for k in range(1, M + 1):
        p = mat[k] / mat[k].sum()
        sc = prob2score(p.astype(np.float32), 0.25)
        msc[k, :4] = sc
        msc[k, 5:16] = int(np.mean(sc))
    h["msc"] = msc
    return h


def build_trunc_cp9s(cm, seed=0, el=True):
    """The truncated passes' CP9s (Rcp9, Lcp9, Tcp9 of cp9.rs:2235-2330) for
    the CM: the same draws as build_cp9_like configured with (pentry, pexit)
    = ((M-1)/M, 0) for the 5' pass, (0, (M-1)/M) for the 3' pass, both for
    5'&3'; local ends when el is set (the local search's configuration)."""
    lpos, rpos = emit_map(cm)
    M = int(rpos[0] - 1)
    full = float(np.float32(M - 1) / np.float32(M))
    return [build_cp9_like(cm, seed, el=el, trunc=cfg) for cfg in ((full, 0.0), (0.0, full), (full, full))]
