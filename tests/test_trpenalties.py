"""Truncated-begin penalties (cm_trunc.rs:192 TrPenalties::new) and the
expected state occupancy they are built on (cp9.rs:752): the oracle is
pinned by the occupancy's node sums, by Monte Carlo parse sampling, by the
conservation of the global truncated-begin mass per node and by closed forms
on single-chain models; the product's host builder equals the oracle bit for
bit."""
import numpy as np
import pytest

from infernox_amd import ixg, models, synth_cm as S
from oracle_cm_bind import cm_psi, trpenalties

CASES = [("clover", lambda: models.config1_models()[1]),
         ("multi300", lambda: S.build_cm(S.multibranch(300, 12, 8), 8)),
         ("nested200", lambda: S.build_cm(S.random_nested(200, 7), 7)),
         ("nested90", lambda: S.build_cm(S.random_nested(90, 6), 6))]


def _sample_visits(cm, n, seed):
    """Visits per state over n parses sampled from the CM's transitions."""
    r = np.random.default_rng(seed)
    st, cf, cn, t = cm["sttype"], cm["cfirst"], cm["cnum"], cm["t"]
    cnt = np.zeros(cm["M"])
    for _ in range(n):
        stack = [0]
        while stack:
            v = stack.pop()
            cnt[v] += 1
            if st[v] == S.E_ST:
                continue
            if st[v] == S.B_ST:
                stack += [int(cn[v]), int(cf[v])]
                continue
            p = t[v, :cn[v]].astype(np.float64)
            stack.append(int(cf[v] + r.choice(cn[v], p=p / p.sum())))
    return cnt / n


@pytest.mark.parametrize("name,mk", CASES)
def test_psi_node_sums(name, mk):
    cm = mk()
    psi = cm_psi(cm)
    for n in range(cm["nodes"]):
        v = np.where(cm["ndidx"] == n)[0]
        split = [x for x in v if cm["sttype"][x] not in (S.IL_ST, S.IR_ST)]
        assert abs(psi[split].sum() - 1.0) < 1e-6, (n, psi[split].sum())


def test_psi_matches_sampled_parses():
    cm = S.build_cm(S.random_nested(40, 3), 3)
    psi = cm_psi(cm)
    emp = _sample_visits(cm, 20000, 1)
    assert np.all(np.abs(psi - emp) < 0.03 + 0.05 * psi), np.abs(psi - emp).max()


@pytest.mark.parametrize("name,mk", CASES)
def test_global_mass_per_node(name, mk):
    cm = mk()
    g, _, ig, _ = trpenalties(cm)
    clen = cm["clen"]
    pg = np.where(g > -1e30, np.exp2(g.astype(np.float64)), 0.0)
    starts = [n for n in range(cm["nodes"]) if cm["ndtype"][n] in (S.MATP_ND, S.MATL_ND, S.MATR_ND, S.BIF_ND)]
    # every truncated begin into a node's split state or the inserts before
    # it shares the node's 2/(clen(clen+1)): the total is that times the nodes
    assert abs(pg[0].sum() - 2.0 / (clen * (clen + 1)) * len(starts)) < 1e-6
    # integer scores are the rounded log2 of the same probabilities
    ok = g > -1e30
    assert np.all(np.abs(ig[ok] - 1000.0 * g[ok]) <= 1.0)


def _chain(clen, kind):
    """A CM of clen MATL (kind 'L') or MATR ('R') nodes: '.' * clen."""
    cm = S.build_cm("." * clen, 5)
    return cm


def test_local_chain_closed_form():
    # an all-MATL chain: the 5' truncated begin at node k carries the sum of
    # every earlier node's begin over its fragment count; the 3' one does not
    clen = 12
    cm = _chain(clen, "L")
    _, l, _, _ = trpenalties(cm)
    nd = [n for n in range(cm["nodes"]) if cm["ndtype"][n] == S.MATL_ND]
    pbegin = np.float32(cm["pbegin"])
    begin = [np.float32(1.0 - float(pbegin))] + [pbegin / np.float32(len(nd) - 1)] * (len(nd) - 1)
    psi = cm_psi(cm)
    run5 = np.float32(0.0)
    for k, n in enumerate(nd):
        m = cm["nodemap"][n]
        sub = clen - k
        run5 = np.float32(begin[k] / np.float32(sub) + run5)
        cur3 = np.float32(begin[k] / np.float32(sub))
        # the node's own share of the mass (the insert before it takes the rest)
        ins = [cm["nodemap"][n - 1] + (2 if cm["ndtype"][n - 1] == S.MATL_ND else 1)]
        if cm["ndtype"][n - 1] == S.ROOT_ND:
            ins = [cm["nodemap"][n - 1] + 1, cm["nodemap"][n - 1] + 2]
        share = np.float32(psi[m]) / np.float32(np.float32(psi[m]) + sum(np.float32(psi[i]) for i in ins))
        assert abs(2.0 ** float(l[1, m]) - float(share) * float(run5)) < 1e-6 * float(run5), k
        assert abs(2.0 ** float(l[2, m]) - float(share) * float(cur3)) < 1e-6 * float(cur3), k


@pytest.mark.parametrize("name,mk", CASES)
def test_product_equals_oracle(name, mk):
    cm = mk()
    for a, b in zip(trpenalties(cm), ixg.cm_trpenalties(cm)):
        assert np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))
