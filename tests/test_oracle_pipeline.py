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
