"""Calibrates the E-value parameters of the synthetic benchmark profiles
(infernox_amd.synth.BENCH_PROFILES) with the CPU oracle, in the manner of
HMMER's p7_Calibrate: lambda fixed at ln 2; MSV and Viterbi Gumbel mu by
maximum likelihood with known lambda over 200 iid sequences of length 200;
Forward tau from the 4% tail of 200 iid sequences of length 100.
Prints the dict literal that synth.py holds."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from infernox_amd import synth  # noqa: E402
from oracle_bind import Oracle  # noqa: E402

LAM = math.log(2.0)


def null1(L):
    p1 = L / (L + 1.0)
    return L * math.log(p1) + math.log(1 - p1)


def calibrate(M, seed, gc=0.5):
    p = synth.consensus_profile(M, seed, gc)
    p.update(dict(lmmu=0, lmlambda=LAM, lvmu=0, lvlambda=LAM, lftau=0, lflambda=LAM))
    o = Oracle(p, 200)
    r = np.random.default_rng(seed + 1000)
    ms, vs, fs = [], [], []
    for n in range(200):
        d = np.concatenate([[255], r.integers(0, 4, 200), [255]]).astype(np.uint8)
        m = o.msv_score(d, 200, 200)
        v = o.vit_scalar(d, 200)
        ms.append((m - null1(200)) / LAM)
        vs.append((v - null1(200)) / LAM)
        d2 = np.concatenate([[255], r.integers(0, 4, 100), [255]]).astype(np.uint8)
        fs.append((o.fwd(d2, 100) - null1(100)) / LAM)
    mu_m = -math.log(np.mean(np.exp(-LAM * np.array(ms)))) / LAM
    mu_v = -math.log(np.mean(np.exp(-LAM * np.array(vs)))) / LAM
    tau = float(np.quantile(fs, 1 - 0.04)) + math.log(0.04) / LAM
    return dict(lmmu=round(mu_m, 4), lmlambda=round(LAM, 6), lvmu=round(mu_v, 4), lvlambda=round(LAM, 6),
                lftau=round(tau, 4), lflambda=round(LAM, 6))


if __name__ == "__main__":
    out = {}
    for M, seed in ((72, 1), (300, 3), (120, 4)):
        out[(M, seed)] = calibrate(M, seed)
    print("BENCH_EV = {")
    for k, v in out.items():
        print(f"    {k}: {v},")
    print("}")
