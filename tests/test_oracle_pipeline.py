"""The oracle's STD-pass search (oracle/pipeline.c) on the CPU: the planted
homologs of a 60 kb config-1 target are reported; the finalize invariants
hold (E-values ascending and at most the reporting threshold, no two
reported hits overlap on one strand, coordinates inside the target and
oriented by strand); the threaded run equals the single-threaded one hit for
hit and counter for counter; the accounting is consistent (residues counted
once per strand, F1 passes >= F3 passes >= CYK passes >= reported hits)."""
import numpy as np

from infernox_amd import ixg, models
from oracle_pipe_bind import OracleSearch


def _opts(p7, cm, cmev, L):
    W = int(cm["W"])
    f3 = ixg.F3Opts()
    f3.f1, f3.f1b, f3.f2, f3.f2b, f3.f3, f3.f3b = 0.35, 1.0, 1.0, 1.0, 0.005, 0.005
    f3.cmw, f3.maxw, f3.opts = W, W, ixg.DO_MSV | ixg.DO_FWD | ixg.DO_FWDBIAS
    env = ixg.env_opts(f4=0.005, f4b=0.005, f5=0.005, gfmu=p7["gfmu"], gflambda=p7["gflambda"])
    return ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"], 2.0 * L / (2 * W), W=W)


def test_oracle_search_invariants():
    p7, cm, cp9, cmev = models.config1_models()
    dsq, L, placed = models.target_with_homologs(cm, 60000, 3, 2)
    osr = OracleSearch(p7, cm, cp9, int(cm["W"]))
    o = _opts(p7, cm, cmev, L)
    hits, a = osr.search(dsq, L, o)
    found = sum(any(min(h[0], h[1]) <= e and max(h[0], h[1]) >= s and h[2] == st for h in hits)
                for s, e, st in placed)
    assert found == len(placed), (placed, hits)
    ev = [h[5] for h in hits]
    assert ev == sorted(ev) and all(x <= o.e_report for x in ev)
    for h in hits:
        assert 1 <= min(h[0], h[1]) and max(h[0], h[1]) <= L
        assert (h[0] <= h[1]) if h[2] == 0 else (h[0] >= h[1])
    for x in range(len(hits)):
        for y in range(x + 1, len(hits)):
            a0, b0 = sorted(hits[x][:2])
            a1, b1 = sorted(hits[y][:2])
            assert hits[x][2] != hits[y][2] or b0 < a1 or b1 < a0
    assert a.nres == 2 * L
    assert a.f.n_past_msv >= a.f.n_past_fwd >= a.e.n_past_gfwd >= 0
    assert a.n_past_cyk >= a.n_hits == len(hits)
    h4, a4 = osr.search(dsq, L, o, threads=4)
    assert h4 == hits
    assert bytes(a4) == bytes(a)


def test_oracle_truncated_passes():
    # homologs cut at the target's ends on both strands: the STD pass alone
    # misses them, the terminal truncated passes (5' on the first maxW
    # residues of the 5'-end chunk, 3' on the last maxW of the 3'-end one)
    # report them with their pass; residues are charged per pass as
    # min(maxW, chunk) (cm_search.rs:2283-2300); a target of at most maxW
    # residues also runs the 5'&3' pass; the threaded run is identical
    from infernox_amd import synth_cm as S, synth_cp9 as P
    p7, cm, cp9, cmev = models.config1_models()
    W = int(cm["W"])
    tr = P.build_trunc_cp9s(cm, 1)
    osr = OracleSearch(p7, cm, cp9, W, trcp9=tr)
    r = np.random.default_rng(7)
    h1, h2 = S.emit(cm, 11), S.emit(cm, 12)
    seq = np.concatenate([h1[len(h1) // 3:], r.integers(0, 4, 3000), h2[:2 * len(h2) // 3]]).astype(np.uint8)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    o = _opts(p7, cm, cmev, L)
    o.env.lftau, o.env.lflambda, o.env.clen = p7["lftau"], p7["lflambda"], int(cm["clen"])
    std, _ = osr.search(dsq, L, o)
    o.trunc_mode = 1
    hits, a = osr.search(dsq, L, o)
    assert std == []
    assert sorted(h[6] for h in hits) == [2, 3], hits
    five = [h for h in hits if h[6] == 2][0]
    three = [h for h in hits if h[6] == 3][0]
    assert five[0] == 1 and five[2] == 0 and three[1] == L and three[2] == 0
    assert list(a.tr_nres) == [2 * W, 2 * W, 0]  # both strands hold both ends of the one chunk
    assert list(a.tr_hits[:2]) == [1, 1]
    h4, a4 = osr.search(dsq, L, o, threads=4)
    assert h4 == hits and bytes(a4) == bytes(a)
    # a target shorter than maxW: 5', 3' and 5'&3' passes on the whole chunk
    part = h1[: min(len(h1), W - 10)]
    short = np.concatenate([[255], part, [255]]).astype(np.uint8)
    Ls = len(part)
    hs, as_ = osr.search(short, Ls, o)
    assert list(as_.tr_nres) == [2 * Ls] * 3


def _multi_seqs(cm, seed):
    """Sequences of uneven length (one empty): homologs cut at ends, a full
    homolog inside, random filler."""
    from infernox_amd import synth_cm as S
    r = np.random.default_rng(seed)
    h1, h2, h3 = S.emit(cm, seed + 1), S.emit(cm, seed + 2), S.emit(cm, seed + 3)
    parts = [
        np.concatenate([h1[len(h1) // 3:], r.integers(0, 4, 900)]),
        r.integers(0, 4, 0),
        np.concatenate([r.integers(0, 4, 400), h2, r.integers(0, 4, 500)]),
        np.concatenate([r.integers(0, 4, 1200), h3[:2 * len(h3) // 3]]),
        r.integers(0, 4, 60),
    ]
    return [np.concatenate([[255], p.astype(np.uint8), [255]]).astype(np.uint8) for p in parts]


def _fin_key(h):
    # hit_sorter_by_evalue (cm_search.rs:1922-1931): E, score desc, sequence,
    # start, pass desc, stop
    return (h[5], -h[3], h[7], h[0], -h[6], h[1])


def test_oracle_multi_sequence_search():
    # a search over several sequences is the union of the per-sequence
    # searches at the same E-value scale (the task list is flat and overlap
    # removal runs per sequence and strand), in the five-key E-value order;
    # counters add up; the threaded run is identical
    from infernox_amd import synth_cp9 as P
    p7, cm, cp9, cmev = models.config1_models()
    W = int(cm["W"])
    osr = OracleSearch(p7, cm, cp9, W, trcp9=P.build_trunc_cp9s(cm, 1))
    seqs = _multi_seqs(cm, 40)
    tot = sum(len(d) - 2 for d in seqs)
    o = _opts(p7, cm, cmev, tot)
    o.env.lftau, o.env.lflambda, o.env.clen = p7["lftau"], p7["lflambda"], int(cm["clen"])
    o.trunc_mode = 1
    hits, a = osr.search_multi(seqs, o)
    union, nres, tr_nres = [], 0, [0, 0, 0]
    for s, d in enumerate(seqs):
        hs, as_ = osr.search(d, len(d) - 2, o)
        union += [h + (s,) for h in hs]
        nres += as_.nres
        tr_nres = [x + y for x, y in zip(tr_nres, as_.tr_nres)]
    assert hits == sorted(union, key=_fin_key)
    assert a.nres == nres and list(a.tr_nres) == tr_nres
    assert {h[7] for h in hits} >= {0, 2, 3}, hits  # the planted homologs are found in their sequences
    h4, a4 = osr.search_multi(seqs, o, threads=4)
    assert h4 == hits and bytes(a4) == bytes(a)
