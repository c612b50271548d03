"""Calibrates the E-value parameters of a workload model
(infernox_amd.models.config1_models, the 72-column cloverleaf;
config2_models, the 300-column 12-hairpin CM; or config3_models, the
120-column three-hairpin CM: argument "1", "2" or "3") with
the CPU oracle, in the manner of
HMMER's p7_Calibrate and of cmcalibrate's exponential tails, with lambda
fixed at ln 2:

* p7 local MSV and Viterbi Gumbel mu by maximum likelihood (known lambda)
  over 200 iid sequences of length 200;
* p7 local Forward exponential tau from the 5.5% tail of 200 iid sequences
  of length 100, glocal Forward tau from the 6.5% tail of 200 of length
  max(100, 2 clen) (the lengths and tail masses of cm_p7_calibrate.rs:273-283);
* CM local CYK and Inside tails: the best root score of 200 iid windows of
  length 2W (non-banded scans; a second argument sets fewer for the large
  model), mu from their 4% tail.

Prints the two dict literals that infernox_amd/models.py holds for it."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from infernox_amd import models  # noqa: E402
from oracle_bind import Oracle  # noqa: E402
from oracle_cm_bind import OracleQdbCM  # noqa: E402
from oracle_env_bind import OracleEnv  # noqa: E402

LAM = math.log(2.0)


def null1(L):
    p1 = L / (L + 1.0)
    return L * math.log(p1) + math.log(1 - p1)


def tail_mu(x, frac=0.04):
    return float(np.quantile(x, 1 - frac)) + math.log(frac) / LAM


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "1"
    ncm = int(sys.argv[2]) if len(sys.argv) > 2 else 200  # CM tail samples (fewer for the large model)
    build = {"1": models.config1_models, "2": models.config2_models, "3": models.config3_models}[which]
    zero = dict(lmmu=0, lmlambda=LAM, lvmu=0, lvlambda=LAM, lftau=0, lflambda=LAM, gfmu=0, gflambda=LAM)
    p7, cm, cp9, _ = build(ev=zero)
    o = Oracle(p7, 200)
    gl = max(100, 2 * int(cm["clen"]))
    ge = OracleEnv(p7, gl)
    r = np.random.default_rng(1001)
    ms, vs, fs, gs = [], [], [], []
    for _ in range(200):
        d = np.concatenate([[255], r.integers(0, 4, 200), [255]]).astype(np.uint8)
        ms.append((o.msv_score(d, 200, 200) - null1(200)) / LAM)
        vs.append((o.vit_scalar(d, 200) - null1(200)) / LAM)
        d2 = np.concatenate([[255], r.integers(0, 4, 100), [255]]).astype(np.uint8)
        fs.append((o.fwd(d2, 100) - null1(100)) / LAM)
        d3 = np.concatenate([[255], r.integers(0, 4, gl), [255]]).astype(np.uint8)
        ge.config(gl)
        gs.append((ge.gforward(d3, gl)[0] - null1(gl)) / LAM)
    ev = dict(lmmu=round(-math.log(np.mean(np.exp(-LAM * np.array(ms)))) / LAM, 4), lmlambda=round(LAM, 6),
              lvmu=round(-math.log(np.mean(np.exp(-LAM * np.array(vs)))) / LAM, 4), lvlambda=round(LAM, 6),
              lftau=round(tail_mu(fs, 0.055), 4), lflambda=round(LAM, 6), gfmu=round(tail_mu(gs, 0.065), 4),
              gflambda=round(LAM, 6))
    q = OracleQdbCM(cm, local=True)
    W = int(cm["W"])
    cyk, ins = [], []
    for _ in range(ncm):
        n = 2 * W
        d = np.concatenate([[255], r.integers(0, 4, n), [255]]).astype(np.uint8)
        _, root = q.scan(d, 1, n, 0, want_root=True)
        cyk.append(float(root.max()))
        _, root = q.scan(d, 1, n, 1, want_root=True)
        ins.append(float(root.max()))
    cm_ev = dict(lc=(round(tail_mu(cyk), 4), round(LAM, 6)), li=(round(tail_mu(ins), 4), round(LAM, 6)))
    print(f"CONFIG{which}_EV =", ev)
    print(f"CONFIG{which}_CM_EV =", cm_ev)


if __name__ == "__main__":
    main()
