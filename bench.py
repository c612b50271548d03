"""Benchmark of the MI355X search path. Prints one JSON line.

Headline: the small plain cmsearch workload -- a 72-column tRNA-like
cloverleaf CM (stems 7/4/5/5 bp, loops 7-9 nt) estimated from a seeded
40-sequence alignment (seed 1), its CP9 and p7 filter HMM, searched against
a 1 Mb iid uniform ACGU target (seed 2) with 20 CM-sampled homologs planted
on random strands, both strands, thresholds of the 2-20 Mb tier (F1 0.35,
F3 = F3b = F4 = F4b = F5 = 0.005, F6 1e-4). One step = one ixg_search call
over the whole target with its sequence resident in HBM, cmsearch's default
pipeline: the STD pass F1 SSV -> F3 Forward -> F3b bias -> F4/F4b glocal
Forward -> glocal Backward and domain definition -> F5 -> CP9 bands (tau
iteration) -> F6 HB CYK with the CYK envelope redefinition -> CP9 bands at
the final tau -> F7 HB Inside with null3 and greedy hit reporting, and the
terminal truncated passes (5' and 3' on the maxW residues at the target's
ends, each with its filter cascade, envelope definition, truncated CP9
bands, TrCYK filter and FTrInside) -> genome coordinates, E-values, overlap
removal and sort. Not in the step: hit alignment.

Secondary sections, each a rate beside the oracle's: the F1-F3b cascade on
a 4.6 Mb genome with a 300-column profile, the HB / truncated HB / QDB
scanners, CP9 banding, full p7 Forward, and the serving loop.
"""
import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from infernox_amd import ixg, models, synth, synth_cm, synth_cp9  # noqa: E402

VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # 256 CUs x 4 SIMD16 x 2.4 GHz, one lane-op per cycle
# HBM bytes per launch of each kernel on this workload, from rocprofv3 FETCH_SIZE +
# WRITE_SIZE passes (KiB per dispatch) scaled by what the same counters read for
# a 256 MiB copy kernel (FETCH x2.0, WRITE x1.0); the measurements are in
# profiles/hbm_traffic_p7_cascade_genome4.6Mb_M300.md. They hold for the
# default workload only and are reported as measured, not recomputed.
TRAFFIC = {"ssv_sweep": 12360715, "ssv_resolve": 6054781, "fwd": 6628113, "bias": 189719}
# The same for the headline search step's two largest kernels, per launch:
# the HB scans (one CYK and one Inside launch per step, their mean: CYK 26940
# KiB FETCH_SIZE / 40299 KiB WRITE_SIZE, Inside 2525 / 3658 KiB) and the glocal
# Forward/Backward (mean of its two launches, 364 / 294 KiB), with the factors
# above; profiles/search_config1_1Mb.md.
SEARCH_TRAFFIC = {"hb_scan": int(((26940 + 2525) * 2.0 + (40299 + 3658) * 1.0) * 1024 / 2),
                  "gfb": int((364 * 2.0 + 294 * 1.0) * 1024)}


def workload():
    M, seed, maxw = 300, 3, 400
    L = 4_600_000
    p = synth.bench_profile(M, seed)
    d = synth.markov2_dsq(L, 3, gc=0.508)
    planted = synth.plant(d, L, p, 100, 3)
    return p, d, L, M, maxw, planted


FP32_PEAK_TFLOPS = 157.3  # MI355X published vector FP32 (packed FMA: 2 flops x 2 lanes per lane-cycle)


def search_workload():
    """The small plain cmsearch workload: (p7, cm, cp9, cm_ev, dsq, L,
    planted)."""
    p7, cm, cp9, cmev = models.config1_models()
    dsq, L, planted = models.target_with_homologs(cm, 1_000_000, 20, seed=2)
    return p7, cm, cp9, cmev, dsq, L, planted


def search_opts(p7, cm, cmev, L):
    """Options of the 2-20 Mb tier (Z = 2L residues, both strands); E-values
    from the calibrated tails, which are per 2W-residue window, so the
    effective database size is Z / 2W."""
    W = int(cm["W"])
    f3 = ixg.F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.005, f3b=0.005, cmw=W, maxw=W,
                    opts=ixg.DO_MSV | ixg.DO_FWD | ixg.DO_FWDBIAS)
    env = ixg.env_opts(f4=0.005, f4b=0.005, f5=0.005, gfmu=p7["gfmu"], gflambda=p7["gflambda"], lftau=p7["lftau"],
                       lflambda=p7["lflambda"], clen=int(cm["clen"]))
    return ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"], 2.0 * L / (2 * W), W=W, trunc=True)


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_quota():
    """CPUs this process may use: (nproc, affinity count, cgroup quota or None)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    q = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            a, b = f.read().split()
            if a != "max":
                q = float(a) / float(b)
    except (OSError, ValueError):
        pass
    return os.cpu_count(), aff, q


def search_cpu_baseline(wl, opts, threads, budget_s):
    """The oracle's search (C restatement of the same pipeline, the STD and
    the truncated passes, not infernox) over the whole workload on `threads`
    threads, repeated while under budget_s: Mres/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_pipe_bind import OracleSearch
    p7, cm, cp9, _, dsq, L, _ = wl
    o = OracleSearch(p7, cm, cp9, int(cm["W"]), trcp9=synth_cp9.build_trunc_cp9s(cm, 1))
    runs, res, t0 = 0, 0, time.time()
    hits = None
    while runs == 0 or time.time() - t0 < budget_s / 2:
        hits, a = o.search(dsq, L, opts, threads=threads)
        res += a.nres
        runs += 1
    dt = time.time() - t0
    nproc, aff, q = cpu_quota()
    return dict(value=res / dt / 1e6, unit="Mres/s", cores=threads, kind="port", nproc=nproc, affinity=aff,
                cgroup_cpus=q, cpu_model=cpu_model(),
                sample=f"the whole 1 Mb x 2 strands workload, {runs} run(s), oracle ixo_search_std_mt (C "
                       f"restatement of the default pipeline, STD and truncated passes, not infernox), "
                       f"{threads} threads, {dt:.1f} s; "
                       f"{len(hits)} hits"), hits


def time_search(ctx, steps, warmup, wl, traffic=None):
    """The headline step: ixg_search over the whole target (the STD and the
    truncated passes). Returns the line's fields and the last step's hits."""
    p7, cm, cp9, cmev, dsq, L, planted = wl
    W = int(cm["W"])
    opts = search_opts(p7, cm, cmev, L)
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    gcm.set_trpenalties(*ixg.cm_trpenalties(cm)[:2])
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    trd = synth_cp9.build_trunc_cp9s(cm, 1)
    tr9 = [ixg.CP9(ctx, h) for h in trd]
    trhb = [ixg.HbMap(cm, h) for h in trd]
    tgt = ixg.Target(ctx, dsq, L)
    for _ in range(warmup):
        ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, opts)
    ctx.kernel_timings(reset=True)
    st = dict.fromkeys(ixg.SEARCH_STAGES, 0.0)
    cyk_cells = ins_cells = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        hits, acct = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, opts)
        ms, cells = ixg.search_stats(ctx)
        for k in st:
            st[k] += ms[k] / steps
        cyk_cells, ins_cells = cells
    dt = (time.perf_counter() - t0) / steps
    kt = ctx.kernel_timings(reset=True)
    f3ms, f3cells = ctx.timings()
    found = sum(any(min(h[0], h[1]) <= e and max(h[0], h[1]) >= s and h[2] == sd for h in hits)
                for s, e, sd in planted)
    M7 = int(p7["M"])
    # algorithmic lane-ops per launch of the kernels with a per-unit figure
    # (the reference loop bodies): SSV 4 u8 ops per (residue, node)
    # (cm_pipeline.rs:598-601); Forward ~16 FP32 ops per (residue, node)
    # (cm_pipeline.rs:1324-1382); glocal Forward / Backward ~20 per (residue,
    # node) (p7_generic.rs:519-640, :939-1060, 3 nats FLogsums of ~4 ops + 8
    # adds/max); HB CYK 8 per cell (the mean of the cloverleaf's 2-3
    # children x add+max, emission, clamp); HB Inside 24 per cell (the same
    # with FLogsum for max)
    # the F4 launch holds every F3b window's glocal Forward and Backward
    # (gfb); the F5 launch rescores the domains (gfwd)
    ops = {"ssv_sweep": 4 * f3cells[0], "fwd": 16 * f3cells[1],
           "gfb": 2 * 20 * acct.f.pos_past_fwdbias * M7, "gfwd": 20 * acct.e.pos_past_edef * M7,
           "hb_scan": 8 * cyk_cells + 24 * ins_cells}
    kern = {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps,
                "ms_per_launch": v[0] / max(1, v[1])} for k, v in kt.items() if v[1] > 0}
    dom = max(kern, key=lambda k: kern[k]["ms_per_step"])
    roof = None
    if dom in ops:
        ms_l = kern[dom]["ms_per_launch"]
        per_launch = ops[dom] / max(1.0, kern[dom]["launches_per_step"])
        a = per_launch / (ms_l * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": dom, "achieved": a, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": a / FP32_PEAK_TFLOPS, "ops_per_launch": per_launch, "ms_per_launch": ms_l,
                "lane_op_ceiling": VALU_PEAK_TOPS, "frac_of_lane_op_ceiling": a / VALU_PEAK_TOPS,
                "why_two_peaks": "the published 157.3 TFLOP/s counts packed FP32 FMA (2 flops x 2 lanes per "
                                 "lane-cycle); the max-plus / log-sum DP ops are single unpacked non-FMA lane "
                                 "ops, built without packed FP32, so 39.3 T lane-op/s is their ceiling",
                "traffic": (SEARCH_TRAFFIC if traffic is None else traffic).get(dom)}
    out = {"value": acct.nres / dt / 1e6, "ms_per_step": dt * 1e3, "residues": int(acct.nres),
           "stage_ms": st, "kernels": kern, "roofline": roof,
           "cyk_gcups": cyk_cells / (st["f6_cyk"] * 1e-3) / 1e9 if st["f6_cyk"] > 0 else None,
           "inside_gcups": ins_cells / (st["f7_inside"] * 1e-3) / 1e9 if st["f7_inside"] > 0 else None,
           "hb_cells": {"cyk": int(cyk_cells), "inside": int(ins_cells)},
           "survivors": {"f1": int(acct.f.n_past_msv), "f3b": int(acct.f.n_past_fwdbias),
                         "f4b": int(acct.e.n_past_gfwdbias), "f5_envelopes": int(acct.e.n_past_edef),
                         "f6": int(acct.n_past_cyk), "hits": int(acct.n_hits)},
           "planted_found": f"{found} of {len(planted)}"}
    out["survivors"]["truncated_passes"] = {
        "residues": [int(x) for x in acct.tr_nres], "f3b": [int(acct.tr_f[q].n_past_fwdbias) for q in range(3)],
        "f5_envelopes": [int(acct.tr_e[q].n_past_edef) for q in range(3)],
        "f6": [int(x) for x in acct.tr_n_past_cyk], "hits": [int(x) for x in acct.tr_hits]}
    for x in [tgt, g9, gcm, g7, hb] + tr9 + trhb:
        x.free()
    return out, opts, hits


def cm_workload(n_env=512, r=8, seed=3):
    """HB scanner workload of the genome config: the 300-column, 12-stem CM
    (about 960 states, W ~ 400), n_env envelopes each holding one emitted
    homolog with 20 flanking residues per side, CP9-shaped bands of
    half-width r (+0..2). Returns (cm, dsq, L, envelopes)."""
    cm = synth_cm.build_cm(synth_cm.multibranch(300, 12, seed), seed)
    rng = np.random.default_rng(seed + 1)
    parts, envs, pos = [], [], 1
    for k in range(n_env):
        h = synth_cm.emit(cm, 1000 + k)
        left, right = rng.integers(0, 4, 20).astype(np.uint8), rng.integers(0, 4, 20).astype(np.uint8)
        parts += [left, h, right]
        i0, j0 = pos, pos + 40 + len(h) - 1
        envs.append((i0, j0, 0, synth_cm.bands(cm, i0, j0, r=r, seed=k)))
        pos = j0 + 1
    seq = np.concatenate(parts)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cm, dsq, L, envs


def time_cm(ctx, steps, warmup, cpu=None, wl=None):
    cm, dsq, L, envl = wl or cm_workload()
    g = ixg.CM(ctx, cm)
    tgt = ixg.Target(ctx, dsq, L)
    envs = ixg.EnvBatch(envl)  # marshalled once, as a caller holding its band arrays would
    work = np.array([hb_work(cm, e[3]) for e in envl], np.int64).sum(0)
    out = {}
    for name, inside in (("cyk", False), ("inside", True)):
        for _ in range(warmup):
            ixg.hb_scan_batch(ctx, g, tgt, envs, inside=inside, report=inside)
        kms, t0 = 0.0, time.perf_counter()
        for _ in range(steps):
            vsc, ei, ej, hits, cells = ixg.hb_scan_batch(ctx, g, tgt, envs, inside=inside, report=inside)
            kms += ctx.timings()[0][2]
        wall = (time.perf_counter() - t0) / steps
        kms /= steps
        c, u, sp = (int(x) for x in work)
        # CYK: add + max per child update and per split, emission + clamp per cell; Inside: ~9 per update
        ops = (9 * (u + sp) + 2 * c) if inside else (2 * u + 2 * c + 2 * sp)
        out[name] = {"gcups": cells / (kms * 1e-3) / 1e9, "gcups_wall": cells / wall / 1e9, "kernel_ms": kms,
                     "wall_ms": wall * 1e3, "cells": int(cells), "updates": u, "b_splits": sp,
                     "envelopes": len(envs), "states": int(cm["M"]), "positive": int((vsc > 0).sum()),
                     "valu_roofline": valu_line(ops, kms)}
        if cpu:
            ob = oracle_cm_bind()

            def run(o, e, inside=inside):
                r = o.scan(e[3], dsq, int(e[0]), int(e[1]), int(inside))
                if inside:
                    o.hits(e[3], dsq, int(e[0]), int(e[1]), r["root"], 0.0)
                return o.ncells(e[3])
            done, w, dt = cpu_rate(lambda: ob.OracleCM(cm), run, envl, *cpu)
            out[name]["cpu"] = cpu_line(done, len(envl), w, dt, cpu[1], "envelopes")
    tgt.free()
    g.free()
    return out


def time_cp9(ctx, steps, cpu=None, wl=None, tau=1e-7):
    """CP9 banding (Forward, Backward, posterior HMM bands) of the HB
    workload's envelopes with a CP9 HMM over the CM's consensus: cells are
    (window position, HMM node) pairs per DP."""
    cm, dsq, L, envl = wl or cm_workload()
    h = synth_cp9.build_cp9(cm, 3)
    g = ixg.CP9(ctx, h)
    tgt = ixg.Target(ctx, dsq, L)
    wins = [(e[0], e[1], 0) for e in envl]
    cells = sum((j0 - i0 + 2) * (h["M"] + 1) for i0, j0, _ in wins)
    ixg.cp9_bands_batch(ctx, g, tgt, wins, tau)
    ctx.kernel_timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ixg.cp9_bands_batch(ctx, g, tgt, wins, tau)
    wall = (time.perf_counter() - t0) / steps
    kms = ctx.kernel_timings(reset=True)["cp9"][0] / steps
    # per cell: Forward ~9 and Backward ~9 ILogsums, the two posterior scans
    # ~6; an ILogsum is ~7 integer ops and one table read
    out = {"gcups": cells / (kms * 1e-3) / 1e9, "kernel_ms": kms, "wall_ms": wall * 1e3, "cells": int(cells),
           "windows": len(wins), "hmm_nodes": int(h["M"]), "tau": tau,
           "valu_roofline": valu_line(24 * 7 * cells, kms)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_cp9_bind import OracleCP9

        def run(o, w):
            o.bands(dsq, int(w[0]), int(w[1]), tau)
            return (w[1] - w[0] + 2) * (h["M"] + 1)
        done, wk, dt = cpu_rate(lambda: OracleCP9(h), run, wins, *cpu)
        out["cpu"] = cpu_line(done, len(wins), wk, dt, cpu[1], "windows")
    g.free()
    tgt.free()
    return out


def trcm_workload(n_env=512, r=8, seed=3, cut=0.3):
    """Truncated-pass workload on the same CM: n_env envelopes, each holding an
    emitted homolog with its 5' (even k) or 3' (odd k) 30% missing at the
    envelope's edge, bands over the consensus columns present, validity of
    the 5'&3' pass (L, R and T planes filled)."""
    cm = synth_cm.build_cm(synth_cm.multibranch(300, 12, seed), seed)
    clen = cm["clen"]
    rng = np.random.default_rng(seed + 2)
    c1, c2 = int(cut * clen), int((1 - cut) * clen)
    parts, envs, pos = [], [], 1
    for k in range(n_env):
        h = synth_cm.emit(cm, 2000 + k)
        cutn = int(len(h) * cut)
        if k % 2 == 0:
            part, cols, vp = h[cutn:], (c1, clen), (1, c1 + 6, clen, clen - 2)
        else:
            part, cols, vp = h[:len(h) - cutn], (1, c2), (1, 3, c2 - 5, c2 + 5)
        gap = rng.integers(0, 4, 10).astype(np.uint8)
        parts += [part, gap]
        i0, j0 = pos, pos + len(part) - 1
        b = synth_cm.tr_bands(cm, i0, j0, r=r, seed=k, c1=cols[0], c2=cols[1])
        val = synth_cm.tr_validity(cm, synth_cm.PASS_5P_AND_3P_FORCE, *vp)
        envs.append((i0, j0, 0, b, val))
        pos = j0 + 1 + len(gap)
    seq = np.concatenate(parts)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cm, dsq, L, envs


def filled_cells(cm, envs):
    """Banded cells summed over the marginal planes each state fills (J, L,
    R; T for bifurcations)."""
    st = cm["sttype"]
    tot = 0
    for (_, _, _, b, val) in envs:
        M = cm["M"]
        w = np.maximum(0, b["hdmax"] - b["hdmin"] + 1)
        per_v = np.add.reduceat(np.concatenate([w, [0]]), b["hd_off"][:M]) if len(w) else np.zeros(M)
        nj = b["hd_off"][1:] - b["hd_off"][:-1]
        per_v = np.where(nj > 0, per_v, 0)
        planes = (val["jvalid"].astype(np.int64) + val["lvalid"] + val["rvalid"] +
                  ((st == synth_cm.B_ST) & (val["tvalid"] > 0)))
        tot += int((per_v * planes)[1:].sum())
    return tot


VALU_PEAK = 39.3216  # T lane-ops/s: 256 CU x 4 SIMD16 x 16 lanes x 2.4 GHz


def hb_work(cm, b, val=None):
    """Algorithmic work of one banded scan (the counts the per-unit op figures
    multiply): cells, child updates (cells x children, per filled plane when
    val is given) and bifurcation split terms (standard scan only), from the
    bands as the reference's loops visit them (cp9.rs:4818-5253)."""
    M = int(cm["M"])
    st, cfirst, cnum = cm["sttype"], cm["cfirst"], cm["cnum"]
    hmn, hmx, off = b["hdmin"], b["hdmax"], b["hd_off"]
    w = np.maximum(0, hmx - hmn + 1)
    per_v = np.add.reduceat(np.concatenate([w, [0]]), off[:M]) if len(w) else np.zeros(M, np.int64)
    per_v = np.where(off[1:M + 1] > off[:M], per_v, 0).astype(np.int64)
    live = (st != synth_cm.E_ST) & (st != synth_cm.B_ST)
    if val is not None:
        planes = (val["jvalid"].astype(np.int64) + val["lvalid"] + val["rvalid"] +
                  ((st == synth_cm.B_ST) & (val["tvalid"] > 0)))
        cells = int((per_v * planes)[1:].sum())
        upd = int((per_v * planes * np.where(live, cnum, 0))[1:].sum())
        return cells, upd, 0
    cells = int(per_v.sum())
    upd = int((per_v * np.where(live, cnum, 0)).sum())
    splits = 0
    for v in np.nonzero(st == synth_cm.B_ST)[0]:
        y, z = int(cfirst[v]), int(cnum[v])
        jn, jx = max(b["jmin"][v], b["jmin"][z]), min(b["jmax"][v], b["jmax"][z])
        if jx < jn:
            continue
        j = np.arange(jn, jx + 1)
        jpv, jpy, jpz = j - b["jmin"][v], j - b["jmin"][y], j - b["jmin"][z]
        dn, dx = hmn[off[v] + jpv], hmx[off[v] + jpv]
        kn = np.maximum(np.maximum(j - b["jmax"][y], hmn[off[z] + jpz]), 0)
        kx = np.minimum(jpy, hmx[off[z] + jpz])
        nd, nk = int(max(0, (dx - dn).max() + 1)), int(max(0, (kx - kn).max() + 1))
        if nd == 0 or nk == 0:
            continue
        d = dn[:, None, None] + np.arange(nd)[None, :, None]
        k = kn[:, None, None] + np.arange(nk)[None, None, :]
        r = np.clip(jpy[:, None, None] - k, 0, off[y + 1] - off[y] - 1)
        ok = ((d <= dx[:, None, None]) & (k <= kx[:, None, None]) & (k >= d - hmx[off[y] + r]) &
              (k <= d - hmn[off[y] + r]) & (jpy[:, None, None] - k >= 0))
        splits += int(ok.sum())
    return cells, upd, splits


def valu_line(ops, ms):
    """VALU roofline entry: algorithmic lane-ops of one launch over its time."""
    t = ops / (ms * 1e-3) / 1e12
    return {"ops": int(ops), "achieved_tops": t, "peak_tops": VALU_PEAK, "frac": t / VALU_PEAK}


def cpu_rate(make, run, items, budget_s, threads):
    """Oracle (C restatement, not infernox) rate over as many items as fit in
    budget_s on `threads` threads (one oracle object per thread; ctypes
    releases the GIL): (items done, work done, seconds). run(o, item) returns
    the item's work units."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    tl = threading.local()

    def one(it):
        if not hasattr(tl, "o"):
            tl.o = make()
        return run(tl.o, it)

    done, work, t0 = 0, 0, time.time()
    with ThreadPoolExecutor(threads) as ex:
        while done < len(items) and time.time() - t0 < budget_s:
            batch = items[done:done + threads]
            work += sum(ex.map(one, batch))
            done += len(batch)
    return done, work, time.time() - t0


def cpu_line(done, total, work, dt, threads, what):
    return {"gcups": work / dt / 1e9, "cores": threads, "kind": "port",
            "sample": f"{done} of {total} {what}, oracle (C restatement, not infernox), {threads} threads, {dt:.1f} s"}


def oracle_cm_bind():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_cm_bind
    return oracle_cm_bind


def time_trcm(ctx, steps, warmup, cpu=None):
    cm, dsq, L, envl = trcm_workload()
    g = ixg.CM(ctx, cm)
    gp, lp = synth_cm.tr_penalties(cm, 3)
    g.set_trpenalties(gp, lp)
    tgt = ixg.Target(ctx, dsq, L)
    cells = filled_cells(cm, envl)
    envs = ixg.EnvBatch(envl, trunc=True)  # marshalled once, as a caller holding its band arrays would
    work = np.array([hb_work(cm, e[3], e[4]) for e in envl], np.int64).sum(0)
    out = {}
    for name, inside in (("trcyk", False), ("trinside", True)):
        for _ in range(warmup):
            ixg.trhb_scan_batch(ctx, g, tgt, envs, synth_cm.PASS_5P_AND_3P_FORCE, inside=inside)
        kms, t0 = 0.0, time.perf_counter()
        for _ in range(steps):
            vsc, vm, ei, ej, hits, _ = ixg.trhb_scan_batch(ctx, g, tgt, envs, synth_cm.PASS_5P_AND_3P_FORCE,
                                                           inside=inside)
            kms += ctx.timings()[0][2]
        wall = (time.perf_counter() - t0) / steps
        kms /= steps
        c, u = int(work[0]), int(work[1])
        ops = (9 * u + 2 * c) if inside else (2 * u + 2 * c)  # bifurcation splits not counted: a lower bound
        out[name] = {"gcups": cells / (kms * 1e-3) / 1e9, "gcups_wall": cells / wall / 1e9, "kernel_ms": kms,
                     "wall_ms": wall * 1e3, "plane_cells": int(cells), "plane_updates": u, "envelopes": len(envs),
                     "positive": int((vsc > 0).sum()), "hits": len(hits), "valu_roofline": valu_line(ops, kms)}
        if cpu:
            ob = oracle_cm_bind()
            pty = lp[ob.PTY_INDEX[synth_cm.PASS_5P_AND_3P_FORCE]]

            def run(o, e, inside=inside):
                o.trscan(e[3], e[4], pty, dsq, int(e[0]), int(e[1]), synth_cm.PASS_5P_AND_3P_FORCE, True, int(inside))
                return filled_cells(cm, [e])
            done, w, dt = cpu_rate(lambda: ob.OracleTrCM(cm), run, envl, *cpu)
            out[name]["cpu"] = cpu_line(done, len(envl), w, dt, cpu[1], "envelopes (plane cells)")
    tgt.free()
    g.free()
    return out


def qdb_workload(n_win=512, seed=4, slack=8):
    """QDB / non-banded scanner workload: the 120-column, 3-stem CM of the
    sequence-database config, n_win windows each holding one emitted homolog
    with 20 random flanking residues per side (both strands alternate), and
    per-state QDB bands of the consensus subtree span +- slack."""
    cm = synth_cm.build_cm(synth_cm.multibranch(120, 3, seed), seed)
    rng = np.random.default_rng(seed + 5)
    parts, wins, pos = [], [], 1
    for k in range(n_win):
        h = synth_cm.emit(cm, 3000 + k)
        left, right = rng.integers(0, 4, 20).astype(np.uint8), rng.integers(0, 4, 20).astype(np.uint8)
        parts += [left, h, right]
        i0, j0 = pos, pos + 40 + len(h) - 1
        wins.append((i0, j0, 0))
        pos = j0 + 1
    seq = np.concatenate(parts)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    lp, rp = synth_cm.state_spans(cm)
    span = np.maximum(0, rp - lp + 1)
    dmin = np.maximum(0, span - slack).astype(np.int32)
    dmax = (span + slack).astype(np.int32)
    return cm, dsq, L, wins, dmin, dmax


def qdb_cells(cm, wins, dmin, dmax, trunc=False):
    """(state, j, d) cells a QDB / non-banded scan fills: per window and j,
    d from the state's lower bound to min(j - i0 + 1, dmax, W)."""
    st = np.asarray(cm["sttype"])
    W = int(cm["W"])
    dx_cap = np.minimum(W, dmax if dmax is not None else W)
    base = np.where(st == synth_cm.MP_ST, 2, 1)
    dn = np.ones_like(base) if trunc else np.minimum(np.maximum(dmin if dmin is not None else 0, base), W)
    live = st != synth_cm.E_ST
    live[0] = False
    n = 0
    for i0, j0, _ in wins:
        for jrow in range(1, j0 - i0 + 2):
            dx = np.minimum(min(jrow, W), dx_cap)
            n += int(np.maximum(0, dx - dn + 1)[live].sum())
    return n


def qdb_work(cm, wins, dmin, dmax):
    """(cells, child updates, bifurcation split terms) of a QDB / non-banded
    scan (cm_nohmm.rs:582-616): per window and j, state v covers d from
    max(dmin, emissions) to min(jrow, dmax, W); a bifurcation cell d splits
    over k in [max(0, dmin_right, d - dmax_left), min(dmax_right, d - dmin_left)]."""
    st, cf, cn = np.asarray(cm["sttype"]), np.asarray(cm["cfirst"]), np.asarray(cm["cnum"])
    W, M = int(cm["W"]), int(cm["M"])
    base = np.where(st == synth_cm.MP_ST, 2, 1)
    dmn = np.zeros(M, np.int64) if dmin is None else np.asarray(dmin, np.int64)
    dxc = np.minimum(W, W if dmax is None else np.asarray(dmax, np.int64)) * np.ones(M, np.int64)
    dn = np.minimum(np.maximum(dmn, base), W)
    dmw = np.minimum(dmn, W)
    live = (st != synth_cm.E_ST)
    live[0] = False
    r = np.arange(W + 1)[:, None]
    cells_r = np.maximum(0, np.minimum(r, dxc[None, :]) - dn[None, :] + 1) * live[None, :]
    F_c = cells_r.sum(1)
    F_u = (cells_r * np.where(st == synth_cm.B_ST, 0, cn)[None, :]).sum(1)
    F_s = np.zeros(W + 1, np.int64)
    for v in np.nonzero(st == synth_cm.B_ST)[0]:
        y, wl = int(cn[v]), int(cf[v])
        d = np.arange(W + 1)
        kmin = np.maximum(0, np.maximum(dmw[y], d - dxc[wl]))
        kmax = np.minimum(dxc[y], d - dmw[wl])
        sv = np.where((d >= dn[v]) & (d <= dxc[v]), np.maximum(0, kmax - kmin + 1), 0)
        F_s += np.cumsum(sv)
    tot = np.zeros(3, np.int64)
    for i0, j0, _ in wins:
        L = j0 - i0 + 1
        jr = np.minimum(np.arange(1, L + 1), min(W, L))
        tot += [F_c[jr].sum(), F_u[jr].sum(), F_s[jr].sum()]
    return tuple(int(x) for x in tot)


def time_qdb(ctx, steps, cpu=None):
    cm, dsq, L, wins, dmin, dmax = qdb_workload(n_win=512)
    tgt = ixg.Target(ctx, dsq, L)
    out = {}
    g = ixg.CM(ctx, cm, local=True)
    for name, inside, band in (("cyk_qdb", False, True), ("inside_qdb", True, True), ("cyk_nonbanded", False, False)):
        dmn, dmx = (dmin, dmax) if band else (None, None)
        cells = qdb_cells(cm, wins, dmn, dmx)
        c, u, sp = qdb_work(cm, wins, dmn, dmx)
        ixg.qdb_scan_batch(ctx, g, tgt, wins, inside=inside, dmin=dmn, dmax=dmx, cutoff=0.0)
        ctx.kernel_timings(reset=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            hits = ixg.qdb_scan_batch(ctx, g, tgt, wins, inside=inside, dmin=dmn, dmax=dmx, cutoff=0.0)
        wall = (time.perf_counter() - t0) / steps
        kms = ctx.kernel_timings(reset=True)["qdb_scan"][0] / steps
        # CYK: add + max per child update and per split, emission + clamp per
        # cell; integer Inside: ~7 int ops + 1 table read per update
        ops = (7 * (u + sp) + 2 * c) if inside else (2 * u + 2 * c + 2 * sp)
        out[name] = {"gcups": cells / (kms * 1e-3) / 1e9, "kernel_ms": kms, "wall_ms": wall * 1e3,
                     "cells": cells, "updates": u, "b_splits": sp, "windows": len(wins), "states": int(cm["M"]),
                     "hits": len(hits), "valu_roofline": valu_line(ops, kms)}
        if cpu:
            ob = oracle_cm_bind()

            def run(o, w, inside=inside, dmn=dmn, dmx=dmx):
                o.scan(dsq, int(w[0]), int(w[1]), int(inside), dmn, dmx, cutoff=0.0)
                return qdb_cells(cm, [w], dmn, dmx)
            done, wk, dt = cpu_rate(lambda: ob.OracleQdbCM(cm, local=True), run, wins, *cpu)
            out[name]["cpu"] = cpu_line(done, len(wins), wk, dt, cpu[1], "windows")
    gp, lp = synth_cm.tr_penalties(cm, 4)
    ig, il = synth_cm.tr_ipenalties(cm, 4)
    g.set_trpenalties(gp, lp)
    g.set_itrpenalties(ig, il)
    # the truncated scanners (a test oracle in the reference) on the first 256
    # windows: their rows for the root's begin states take ~80 MB per window,
    # so the batch runs in launch groups of ~100 windows
    wins = wins[:256]
    cells = qdb_cells(cm, wins, None, None, trunc=True)
    for name, inside in (("trcyk_nonbanded", False), ("trinside_nonbanded", True)):
        ixg.trqdb_scan_batch(ctx, g, tgt, wins, synth_cm.PASS_5P_AND_3P_FORCE, inside=inside, cutoff=0.0)
        ctx.kernel_timings(reset=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            vsc, vm, bf, hits = ixg.trqdb_scan_batch(ctx, g, tgt, wins, synth_cm.PASS_5P_AND_3P_FORCE,
                                                     inside=inside, cutoff=0.0)
        wall = (time.perf_counter() - t0) / steps
        kms = ctx.kernel_timings(reset=True)["trqdb_scan"][0] / steps
        out[name] = {"gcups_per_plane": cells / (kms * 1e-3) / 1e9, "kernel_ms": kms, "wall_ms": wall * 1e3,
                     "plane_cells": cells, "windows": len(wins), "pass": "5'&3'", "hits": len(hits)}
        if cpu:
            ob = oracle_cm_bind()

            def run(o, w, inside=inside):
                o.trscan(dsq, int(w[0]), int(w[1]), int(inside), synth_cm.PASS_5P_AND_3P_FORCE, True, cutoff=0.0)
                return qdb_cells(cm, [w], None, None, trunc=True)
            done, wk, dt = cpu_rate(lambda: ob.OracleTrQdbCM(cm, local=True, seed=4), run, wins, *cpu)
            out[name]["cpu"] = cpu_line(done, len(wins), wk, dt, cpu[1], "windows (per-plane cells)")
    g.free()
    tgt.free()
    return out


def time_p7_forward(ctx, p, d, steps, n=1024, length=400, cpu=None):
    """Full-matrix Forward (retained rows, multihit) over n windows of the
    genome: kernel time and DP cells per second."""
    g = ixg.P7(ctx, p, 400)
    seq = np.ascontiguousarray(d[:n * length + 2], np.uint8)
    regions = [(k * length, length) for k in range(n)]
    g.p7_forward_batch(seq, regions)
    ctx.kernel_timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.p7_forward_batch(seq, regions)
    wall = (time.perf_counter() - t0) / steps
    kms = ctx.kernel_timings(reset=True)["fwd_full"][0] / steps
    cells = n * length * int(p["M"])
    g.free()
    out = {"gcups": cells / (kms * 1e-3) / 1e9, "kernel_ms": kms, "wall_ms": wall * 1e3, "windows": n,
           "length": length, "M": int(p["M"]), "matrix_MB": 3 * n * (length + 1) * (int(p["M"]) + 1) * 4 / 1e6,
           "valu_roofline": valu_line(16 * cells, kms)}  # ~16 FP32 ops per (i, k)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import Oracle

        def run(o, r):
            sub = np.concatenate([[255], seq[r[0] + 1:r[0] + 1 + r[1]], [255]]).astype(np.uint8)
            o.p7_forward(sub, r[1])
            return r[1] * int(p["M"])
        done, wk, dt = cpu_rate(lambda: Oracle(p, 400), run, regions, *cpu)
        out["cpu"] = cpu_line(done, n, wk, dt, cpu[1], "windows")
    return out


def cpu_baseline(p, d, L, M, maxw, tasks, nt, budget_s, threads):
    """Times the oracle (C restatement, not infernox) on a bounded sample of
    the same tasks: as many tasks as fit in budget_s, run on `threads`
    threads (ctypes releases the GIL)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor
    from oracle_bind import Oracle, F3Opts
    o = Oracle(p, maxw)
    oo = F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.005, f3b=0.005, cmw=M, maxw=maxw, opts=7)

    def one(t):
        tk = tasks[t]
        seg = d[tk.start:tk.end + 1]
        if tk.strand:
            seg = np.where(seg < 4, 3 - seg, seg)[::-1]
        sub = np.concatenate([[255], seg, [255]]).astype(np.uint8)
        o.f3_filter_sequence(sub, tk.end - tk.start + 1, oo)
        return tk.end - tk.start + 1 - tk.ctx

    t0 = time.time()
    res = 0
    done = 0
    with ThreadPoolExecutor(threads) as ex:
        while done < nt and time.time() - t0 < budget_s:
            batch = list(range(done, min(nt, done + threads)))
            res += sum(ex.map(one, batch))
            done += len(batch)
    dt = time.time() - t0
    return dict(value=res / dt / 1e6, unit="Mres/s", cores=threads, kind="port",
                sample=f"{done} of {nt} (100 kb chunk x strand) tasks of the same genome, oracle "
                       f"f3_filter_sequence (C restatement, not infernox), {threads} threads, {dt:.1f} s")


def multi_workload(n_cm=32, per_cm=16, r=8, seed=7):
    """cmscan-style HB batch: n_cm CMs with consensus lengths uniform in
    [50, 400] (random nested structures, 40-60% paired), per_cm envelopes
    each, every envelope one emitted homolog with 20 flanking residues per
    side on one query sequence."""
    rng = np.random.default_rng(seed)
    cms = [synth_cm.build_cm(synth_cm.random_nested(int(c), seed + k), seed + k)
           for k, c in enumerate(rng.integers(50, 401, n_cm))]
    parts, envs, env_cm, pos = [], [], [], 1
    for k in range(n_cm * per_cm):
        m = k % n_cm
        h = synth_cm.emit(cms[m], 5000 + k)
        left, right = rng.integers(0, 4, 20).astype(np.uint8), rng.integers(0, 4, 20).astype(np.uint8)
        parts += [left, h, right]
        i0, j0 = pos, pos + 40 + len(h) - 1
        envs.append((i0, j0, 0, synth_cm.bands(cms[m], i0, j0, r=r, seed=k)))
        env_cm.append(m)
        pos = j0 + 1
    seq = np.concatenate(parts)
    L = len(seq)
    dsq = np.concatenate([[255], seq, [255]]).astype(np.uint8)
    return cms, env_cm, dsq, L, envs


def time_multi(ctx, steps=3):
    """Envelopes of 32 CMs: one persistent work-queue launch (ixg_hb_scan_multi)
    against one ixg_hb_scan_batch call per model."""
    cms, env_cm, dsq, L, envl = multi_workload()
    gs = [ixg.CM(ctx, cm) for cm in cms]
    tgt = ixg.Target(ctx, dsq, L)
    batch = ixg.EnvBatch(envl)
    per = [[k for k in range(len(envl)) if env_cm[k] == m] for m in range(len(cms))]
    pb = [ixg.EnvBatch([envl[k] for k in idx]) for idx in per]
    out = {"workload": f"{len(cms)} CMs (consensus 50-400, states {min(c['M'] for c in cms)}-"
                       f"{max(c['M'] for c in cms)}), {len(envl)} envelopes, one query"}
    for name, inside in (("cyk", False), ("inside", True)):
        ixg.hb_scan_multi(ctx, gs, env_cm, tgt, batch, inside=inside)
        t0 = time.perf_counter()
        kms = 0.0
        for _ in range(steps):
            _, _, _, _, cells = ixg.hb_scan_multi(ctx, gs, env_cm, tgt, batch, inside=inside)
            kms += ctx.timings()[0][2]
        wall_multi = (time.perf_counter() - t0) / steps
        kms /= steps
        for m in range(len(cms)):
            ixg.hb_scan_batch(ctx, gs[m], tgt, pb[m], inside=inside)
        t0 = time.perf_counter()
        kper = 0.0
        for _ in range(steps):
            for m in range(len(cms)):
                ixg.hb_scan_batch(ctx, gs[m], tgt, pb[m], inside=inside)
                kper += ctx.timings()[0][2]
        wall_per = (time.perf_counter() - t0) / steps
        kper /= steps
        out[name] = {"cells": int(cells), "one_launch": {"kernel_ms": kms, "wall_ms": wall_multi * 1e3,
                                                         "gcups": cells / (kms * 1e-3) / 1e9,
                                                         "gcups_wall": cells / wall_multi / 1e9},
                     "per_model_launches": {"kernel_ms": kper, "wall_ms": wall_per * 1e3,
                                            "gcups": cells / (kper * 1e-3) / 1e9,
                                            "gcups_wall": cells / wall_per / 1e9}}
    tgt.free()
    for g in gs:
        g.free()
    return out


def serving_workload(n_req=1200, n_models=200, seed=5, median=5000.0, sigma=1.5):
    """Request stream for the serving loop: (model, sequence) pairs with lengths
    log-normal (median 5 kb, sigma 1.5) clipped to [100, 1e6] nt, iid ACGU,
    models drawn uniformly from a pool of 200 p7 profiles with consensus
    lengths uniform in [50, 400]."""
    rng = np.random.default_rng(seed)
    Ms = rng.integers(50, 401, n_models)
    lens = np.clip(np.exp(np.log(median) + sigma * rng.standard_normal(n_req)), 100, 1_000_000).astype(np.int64)
    mids = rng.integers(0, n_models, n_req)
    base = synth.iid_dsq(int(lens.max()), seed + 1)
    # each request's sequence is a window of one long iid sequence (cheap to make)
    starts = [int(rng.integers(0, int(lens.max()) - int(Lq) + 1)) for Lq in lens]
    seqs = []
    for st, Lq in zip(starts, lens):
        dq = np.empty(int(Lq) + 2, np.uint8)
        dq[0] = dq[-1] = 255
        dq[1:-1] = base[1 + st:1 + st + int(Lq)]
        seqs.append(dq)
    return Ms, mids, lens, seqs


def time_serving(ctx, nstreams=4, qps_fracs=(0.5, 1.0, 2.0, 4.0)):
    """Serving loop (continuous batching over nstreams streams) against plain
    per-request calls (ixg_target_create + ixg_f3_filter_tasks, one request at
    a time) on the same requests. Plain latency under an arrival process is a
    single FIFO server with the measured per-request service times."""
    Ms, mids, lens, seqs = serving_workload()
    models = []
    for k, M in enumerate(Ms):
        pk = synth.consensus_profile(int(M), 500 + k)
        pk.update(dict(lmmu=-8.5, lmlambda=0.69, lvmu=-9.0, lvlambda=0.69, lftau=-5.5, lflambda=0.69))
        maxw = int(1.25 * M) + 20
        o = ixg.F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.02, f3b=0.02, cmw=int(M), maxw=maxw, opts=7)
        models.append((ixg.P7(ctx, pk, maxw), o))
    n = len(lens)
    res = 2 * int(lens.sum())  # residues searched: both strands

    def plain_one(k):
        g, o = models[mids[k]]
        tasks, nt = ixg.window_tasks(int(lens[k]), o.maxw)
        tgt = ixg.Target(ctx, seqs[k], int(lens[k]))
        ixg.f3_filter_tasks(ctx, g, tgt, tasks, nt, o)
        tgt.free()

    for k in range(min(20, n)):  # warm-up (scratch growth)
        plain_one(k)
    svc = np.zeros(n)
    for k in range(n):
        t0 = time.perf_counter()
        plain_one(k)
        svc[k] = time.perf_counter() - t0
    plain_rate = res / svc.sum() / 1e6
    cap_qps = n / svc.sum()

    srv = ixg.Server(ctx, nstreams=nstreams, max_inflight=1 << 26, max_batch=1 << 22)
    try:
        for k in range(min(40, n)):  # warm-up
            srv.result(srv.submit(models[mids[k]][0], seqs[k], int(lens[k]), models[mids[k]][1]))
        t0 = time.perf_counter()
        tickets = [srv.submit(models[mids[k]][0], seqs[k], int(lens[k]), models[mids[k]][1]) for k in range(n)]
        for t in tickets:
            srv.result(t)
        burst = time.perf_counter() - t0
        served_peak = res / burst / 1e6
        sweep = []
        rng = np.random.default_rng(55)
        for frac in qps_fracs:
            qps = frac * cap_qps
            arr = np.cumsum(rng.exponential(1.0 / qps, n))
            # plain: one FIFO server with the measured service times
            fin, t = np.zeros(n), 0.0
            for k in range(n):
                t = max(t, arr[k]) + svc[k]
                fin[k] = t
            plat = (fin - arr) * 1e3
            # served: open-loop submission at the arrival times, results collected in a second thread
            lat = np.zeros(n)
            sub_wall = np.zeros(n)
            tick = [0] * n
            def collect():
                for k in range(n):
                    while tick[k] == 0:
                        time.sleep(0.0002)
                    tl, w, acc, tm = srv.result(tick[k])
                    lat[k] = (sub_wall[k] - arr[k]) * 1e3 + tm[2]

            th = threading.Thread(target=collect)
            th.start()
            t0 = time.perf_counter()
            for k in range(n):
                dt = arr[k] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                sub_wall[k] = time.perf_counter() - t0
                tick[k] = srv.submit(models[mids[k]][0], seqs[k], int(lens[k]), models[mids[k]][1])
            th.join()
            span = (time.perf_counter() - t0)
            sweep.append({"qps": qps, "served": {"p50_ms": float(np.percentile(lat, 50)),
                                                 "p99_ms": float(np.percentile(lat, 99)),
                                                 "mres_s": res / span / 1e6},
                          "plain": {"p50_ms": float(np.percentile(plat, 50)), "p99_ms": float(np.percentile(plat, 99)),
                                    "mres_s": res / (fin[-1]) / 1e6}})
        st = srv.stats()
    finally:
        srv.free()
        for g, _ in models:
            g.free()
    return {"workload": f"{n} requests, lengths log-normal median 5 kb sigma 1.5 clipped to [100, 1e6] "
                        f"(seed 5), both strands, models uniform from 200 p7 profiles M in [50, 400], "
                        f"F1 0.35 / F3 0.02 / F3b 0.02",
            "residues": res, "streams": nstreams,
            "plain_mres_s": plain_rate, "plain_capacity_qps": cap_qps,
            "served_peak_mres_s": served_peak, "served_vs_plain": served_peak / plain_rate,
            "sweep": sweep, "server_stats": st}


def search_serving_workload(n_req=300, n_models=8, seed=5, median=5000.0, sigma=1.5):
    """Whole-search requests (sequence, CM): lengths log-normal (median 5 kb,
    sigma 1.5) clipped to [100, 1e6] nt, iid ACGU with a CM-sampled homolog
    inserted into every third; models uniform from a pool of n_models CMs
    with consensus lengths uniform in [50, 400] on random nested structures
    (40-60% paired, 30-sequence alignments), each with its CP9s and p7 filter
    (E-value parameters of the cloverleaf model: not calibrated per model)."""
    from infernox_amd import synth_cm as S
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_models):
        clen = int(rng.integers(50, 401))
        ss = S.random_nested(clen, seed * 100 + k)
        cm = models.cm_from_alignment(ss, models.alignment(ss, 30, seed * 100 + k), seed * 100 + k)
        cp9 = synth_cp9.build_cp9_like(cm, 1, el=True)
        p7 = models.p7_from_cm(cm, models.CONFIG1_EV)
        pool.append((p7, cm, cp9, models.CONFIG1_CM_EV))
    lens = np.clip(np.exp(np.log(median) + sigma * rng.standard_normal(n_req)), 100, 1_000_000).astype(np.int64)
    mids = rng.integers(0, n_models, n_req)
    seqs = []
    for q in range(n_req):
        body = rng.integers(0, 4, int(lens[q])).astype(np.uint8)
        if q % 3 == 0:
            h = S.emit(pool[mids[q]][1], seed * 10000 + q).astype(np.uint8)
            at = int(rng.integers(0, len(body) + 1))
            body = np.concatenate([body[:at], h, body[at:]])
        seqs.append(np.concatenate([[255], body, [255]]).astype(np.uint8))
    return pool, mids, seqs


def time_search_serving(ctx, nstreams=4, qps_fracs=(0.5, 1.0, 2.0, 4.0)):
    """Served whole searches (ixg_server_submit_search: requests of one model
    batched into one ixg_search) against plain per-request calls (one
    ixg_search per request on one context, each sequence already resident),
    both measured under the same open-loop Poisson arrivals: latency from the
    scheduled arrival to the result."""
    import threading
    pool, mids, seqs = search_serving_workload()
    bundles, handles, opts = [], [], []
    for p7, cm, cp9, cmev in pool:
        W = int(cm["W"])
        g7 = ixg.P7(ctx, p7, W)
        gcm = ixg.CM(ctx, cm)
        gcm.set_trpenalties(*ixg.cm_trpenalties(cm)[:2])
        g9 = ixg.CP9(ctx, cp9)
        hb = ixg.HbMap(cm, cp9)
        trd = synth_cp9.build_trunc_cp9s(cm, 1)
        tr9 = [ixg.CP9(ctx, h) for h in trd]
        trhb = [ixg.HbMap(cm, h) for h in trd]
        f3 = ixg.F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.02, f3b=0.02, cmw=W, maxw=W, opts=7)
        env = ixg.env_opts(f4=0.02, f4b=0.02, f5=0.02, gfmu=p7["gfmu"], gflambda=p7["gflambda"],
                           lftau=p7["lftau"], lflambda=p7["lflambda"], clen=int(cm["clen"]))
        opts.append(ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"], 2.0e6 / (2 * W), W=W, trunc=True))
        hs = [g7, gcm, g9, hb] + tr9 + trhb
        handles.append(hs)
        bundles.append(ixg.SearchModel(g7, gcm, g9, hb, tr9, trhb))
    n = len(seqs)
    res = int(sum(2 * (len(d) - 2) for d in seqs))
    tgts = [ixg.Target(ctx, d, len(d) - 2) for d in seqs]

    def plain(q):
        hs = handles[mids[q]]
        return ixg.search(ctx, hs[0], hs[1], hs[2], hs[3], hs[4:7], hs[7:10], tgts[q], opts[mids[q]])[0]

    for q in range(min(8, n)):  # warm-up
        plain(q)
    svc = []
    for q in range(n):
        t0 = time.perf_counter()
        plain(q)
        svc.append(time.perf_counter() - t0)
    cap_qps = n / sum(svc)
    srv = ixg.Server(ctx, nstreams=nstreams, max_inflight=1 << 24, max_batch=1 << 22)
    rng = np.random.default_rng(11)
    sweep = []
    for frac in qps_fracs:
        qps = frac * cap_qps
        arr = np.cumsum(rng.exponential(1.0 / qps, n))
        # plain: one request at a time, in arrival order, from its arrival on
        lat_p, t0 = [], time.perf_counter()
        for q in range(n):
            now = time.perf_counter() - t0
            if now < arr[q]:
                time.sleep(arr[q] - now)
            plain(q)
            lat_p.append(time.perf_counter() - t0 - arr[q])
        t_plain = time.perf_counter() - t0
        # served: a submitter thread on the same schedule, results collected as they finish
        tickets, done_t = [None] * n, [0.0] * n
        t0 = time.perf_counter()

        def submit():
            for q in range(n):
                now = time.perf_counter() - t0
                if now < arr[q]:
                    time.sleep(arr[q] - now)
                tickets[q] = srv.submit_search(bundles[mids[q]], seqs[q], len(seqs[q]) - 2, opts[mids[q]])

        th = threading.Thread(target=submit)
        th.start()
        pending = set(range(n))
        while pending:
            for q in sorted(pending):
                if tickets[q] is not None and srv.done(tickets[q]):
                    srv.search_result(tickets[q])
                    done_t[q] = time.perf_counter() - t0
                    pending.discard(q)
            time.sleep(0.0002)
        th.join()
        t_srv = time.perf_counter() - t0
        lat_s = [done_t[q] - arr[q] for q in range(n)]
        sweep.append({"qps": qps,
                      "plain": {"p50_ms": float(np.percentile(lat_p, 50) * 1e3),
                                "p99_ms": float(np.percentile(lat_p, 99) * 1e3), "mres_s": res / t_plain / 1e6},
                      "served": {"p50_ms": float(np.percentile(lat_s, 50) * 1e3),
                                 "p99_ms": float(np.percentile(lat_s, 99) * 1e3), "mres_s": res / t_srv / 1e6}})
        print(json.dumps({"progress": "search serving", "qps": qps}), file=sys.stderr, flush=True)
    # peak: everything submitted at once
    t0 = time.perf_counter()
    tk = [srv.submit_search(bundles[mids[q]], seqs[q], len(seqs[q]) - 2, opts[mids[q]]) for q in range(n)]
    for t in tk:
        srv.search_result(t)
    t_peak = time.perf_counter() - t0
    st = srv.stats()
    srv.free()
    for t in tgts:
        t.free()
    for hs in handles:
        for x in hs:
            x.free()
    return {"workload": f"{n} whole-search requests (sequence, CM), lengths log-normal median 5 kb sigma 1.5 "
                        f"clipped to [100, 1e6] (seed 5), iid ACGU, a CM-sampled homolog in every third; "
                        f"CMs uniform from {len(pool)} of consensus 50-400 on random nested structures; "
                        f"default pipeline (STD + terminal truncated passes), F1 0.35 / F3-F5 0.02, "
                        f"E-values over a fixed 2 Mb",
            "cm_states": [int(p[1]["M"]) for p in pool], "residues": res, "streams": nstreams,
            "plain_mres_s": res / sum(svc) / 1e6, "plain_capacity_qps": cap_qps,
            "served_peak_mres_s": res / t_peak / 1e6, "served_vs_plain": (res / t_peak) / (res / sum(svc)),
            "sweep": sweep, "server_stats": st}


def time_cascade(ctx, steps, warmup, cpu_s, cpu_threads):
    """The F1-F3b cascade alone on the 4.6 Mb genome with the 300-column
    profile: one ixg_f3_filter_tasks call over every task per step."""
    p, d, L, M, maxw, planted = workload()
    tasks, nt = ixg.window_tasks(L, maxw)
    nres = sum(tasks[t].end - tasks[t].start + 1 - tasks[t].ctx for t in range(nt))
    opts = ixg.F3Opts(f1=0.35, f1b=1.0, f2=1.0, f2b=1.0, f3=0.005, f3b=0.005, cmw=M, maxw=maxw, opts=7)
    g = ixg.P7(ctx, p, maxw)
    tgt = ixg.Target(ctx, d, L)
    for _ in range(warmup):
        ixg.f3_filter_tasks(ctx, g, tgt, tasks, nt, opts)
    stage_ms = np.zeros(5)
    ctx.kernel_timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        wins, accts = ixg.f3_filter_tasks(ctx, g, tgt, tasks, nt, opts)
        ms, cells = ctx.timings()
        stage_ms += np.array(ms)
    dt = (time.perf_counter() - t0) / steps
    stage_ms /= steps
    ms, cells = ctx.timings()
    kt = ctx.kernel_timings(reset=True)
    n_f1 = sum(accts[t].n_past_msv for t in range(nt))
    n_f3b = sum(accts[t].n_past_fwdbias for t in range(nt))
    tgt.free()
    g.free()
    # VALU roofline of the dominant kernel: ops per launch (the last step's
    # cells x ops per cell, from the reference loop bodies) over its mean
    # launch time (HIP events around each launch)
    ops_per_launch = {"ssv_sweep": 4 * cells[0],  # SSV: 4 u8 ops per (residue, k), cm_pipeline.rs:598-601
                      "fwd": 16 * cells[1],        # Forward: ~16 FP32 ops per (i, k), cm_pipeline.rs:1324-1382
                      "bias": 10 * cells[2]}       # bias filter: ~10 FP32 ops per residue, cm_pipeline.rs:1750-1763
    kern = {k: {"ms_per_launch": v[0] / v[1], "launches_per_step": v[1] / steps}
            for k, v in kt.items() if v[1] > 0}
    dom = max(ops_per_launch, key=lambda k: kt[k][0])
    dom_ms = kt[dom][0] / max(1, kt[dom][1])
    achieved = ops_per_launch[dom] / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
    out = {"value": nres / dt / 1e6, "unit": "Mres/s", "ms_per_step": dt * 1e3,
           "workload": "4.6 Mb order-2 Markov 50.8% GC (seed 3), 100 planted homologs, both strands, 100 kb "
                       "chunks + maxW 400 context, p7 M=300, F1 0.35 / F3 0.005 / F3b 0.005; one "
                       "ixg_f3_filter_tasks call per step",
           "tasks": nt, "residues": nres,
           "stage_ms": {"ssv_sweep": stage_ms[0], "ssv_resolve": stage_ms[1], "f3_f3b_windows": stage_ms[2],
                        "host_glue": stage_ms[3], "total_call": stage_ms[4]},
           "survivors": {"f1_windows": int(n_f1), "f3b_windows": int(n_f3b), "merged": len(wins)},
           "kernels": kern,
           "roofline": {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                        "lane_op_ceiling": VALU_PEAK_TOPS, "frac_of_lane_op_ceiling": achieved / VALU_PEAK_TOPS,
                        "ms_per_launch": dom_ms, "ops_per_launch": ops_per_launch[dom],
                        "traffic": TRAFFIC.get(dom)}}
    if cpu_s:
        out["cpu"] = cpu_baseline(p, d, L, M, maxw, tasks, nt, cpu_s, cpu_threads)
    return out


def genome_search_workload():
    """The genome-scale workload: a 4.6 Mb order-2 Markov genome at 50.8% GC
    (seed 3) with 100 CM-sampled homologs inserted on random strands, the
    300-column 12-hairpin CM (959 states, W 398), its CP9s and p7 filter."""
    p7, cm, cp9, cmev = models.config2_models()
    dsq, L, planted = models.markov_genome_with_homologs(cm, 4_600_000, 100, 3)
    return p7, cm, cp9, cmev, dsq, L, planted


def time_genome_search(ctx, steps, warmup, cpu=None):
    """One ixg_search over the genome (both strands, 92 chunk x strand tasks),
    the headline's step on the genome-scale model; the oracle's same search
    on the first 460 kb beside it."""
    wl = genome_search_workload()
    out, opts, hits = time_search(ctx, steps, warmup, wl, traffic={})  # traffic measured for the headline only
    out["workload"] = ("4.6 Mb order-2 Markov genome at 50.8% GC (seed 3) + 100 CM-sampled homologs, both strands; "
                       "300-column 12-hairpin CM (60-sequence alignment, seed 3; 959 states, W 398); tier 2-20 Mb "
                       "(F1 0.35, F3=F3b=F4=F4b=F5 0.005, F6 1e-4); default truncation; one ixg_search per step")
    out["unit"] = "Mres/s"
    if cpu:
        p7, cm, cp9, cmev, dsq, L, planted = wl
        Ls = 460_000
        sub = np.concatenate([dsq[:Ls + 1], [255]]).astype(np.uint8)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_pipe_bind import OracleSearch
        o = OracleSearch(p7, cm, cp9, int(cm["W"]), trcp9=synth_cp9.build_trunc_cp9s(cm, 1))
        so = search_opts(p7, cm, cmev, L)  # the whole genome's E-value scale and tier
        t0 = time.perf_counter()
        _, a = o.search(sub, Ls, so, threads=cpu[1])
        dt = time.perf_counter() - t0
        out["cpu"] = {"value": a.nres / dt / 1e6, "unit": "Mres/s", "cores": cpu[1], "kind": "port",
                      "sample": f"the first {Ls} residues of the genome, both strands, oracle ixo_search_std_mt "
                                f"(C restatement, not infernox), {dt:.1f} s"}
    return out


def batched_workload(n=10000):
    """The batched uneven workload: n iid sequences at 45% GC with log-normal
    lengths (median 2 kb, sigma 1.0, clipped to 200 nt - 200 kb, seed 4), 100
    CM-sampled homologs inserted into random sequences on random strands, the
    120-column three-hairpin CM."""
    p7, cm, cp9, cmev = models.config3_models()
    seqs, placed = models.lognormal_database(n, 2000.0, 1.0, 200, 200000, 0.45, 4, cm=cm, n_hom=100)
    return p7, cm, cp9, cmev, seqs, placed


def batched_opts(p7, cm, cmev, tot):
    """The 20-200 Mb tier (cm_search.rs:1526-1528): F1 0.35, Viterbi F2 and
    its bias filter F2b on at 0.15, F3=F3b=F4=F4b=F5 0.003, F6 1e-4."""
    W = int(cm["W"])
    f3 = ixg.F3Opts(f1=0.35, f1b=1.0, f2=0.15, f2b=0.15, f3=0.003, f3b=0.003, cmw=W, maxw=W,
                    opts=ixg.DO_MSV | ixg.DO_VIT | ixg.DO_VITBIAS | ixg.DO_FWD | ixg.DO_FWDBIAS)
    env = ixg.env_opts(f4=0.003, f4b=0.003, f5=0.003, gfmu=p7["gfmu"], gflambda=p7["gflambda"], lftau=p7["lftau"],
                       lflambda=p7["lflambda"], clen=int(cm["clen"]))
    return ixg.search_opts(f3, env, 1e-4, cmev["lc"], cmev["li"], 2.0 * tot / (2 * W), W=W, trunc=True)


def time_batched(ctx, steps, cpu=None, per_seq_s=4.0):
    """One ixg_search over the whole database (a multi-sequence target: one
    flat task list, every stage one batch over all sequences) against the
    plain form, one ixg_search per sequence (each its own target), timed on
    the leading sequences for about per_seq_s seconds; both at the
    database's E-value scale, so the per-sequence hits are the batched
    call's hits on those sequences."""
    p7, cm, cp9, cmev, seqs, placed = batched_workload()
    tot = sum(len(x) - 2 for x in seqs)
    opts = batched_opts(p7, cm, cmev, tot)
    W = int(cm["W"])
    g7 = ixg.P7(ctx, p7, W)
    gcm = ixg.CM(ctx, cm)
    gcm.set_trpenalties(*ixg.cm_trpenalties(cm)[:2])
    g9 = ixg.CP9(ctx, cp9)
    hb = ixg.HbMap(cm, cp9)
    trd = synth_cp9.build_trunc_cp9s(cm, 1)
    tr9 = [ixg.CP9(ctx, h) for h in trd]
    trhb = [ixg.HbMap(cm, h) for h in trd]
    tgt = ixg.MultiTarget(ctx, seqs)
    hits, acct = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, opts, with_seq=True)  # warm-up
    print(json.dumps({"progress": "batched warm-up done"}), file=sys.stderr, flush=True)
    st = dict.fromkeys(ixg.SEARCH_STAGES, 0.0)
    ctx.kernel_timings(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        hits, acct = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, tgt, opts, with_seq=True)
        ms, cells = ixg.search_stats(ctx)
        for k in st:
            st[k] += ms[k] / steps
    dt = (time.perf_counter() - t0) / steps
    kt = ctx.kernel_timings(reset=True)
    tgt.free()
    found = sum(any(h[7] == s and min(h[0], h[1]) <= e and max(h[0], h[1]) >= b and h[2] == sd for h in hits)
                for s, b, e, sd in placed)
    out = {"autotune": ctx.autotune_choices()}  # (log2 windows class, F3 Forward shape) chosen
    if per_seq_s > 0:
        out["per_sequence_calls"] = per_sequence_calls(ctx, g7, gcm, g9, hb, tr9, trhb, seqs, opts, hits, per_seq_s)
    return finish_batched(out, acct, dt, seqs, cm, W, st, kt, steps, placed, found, cpu, p7, cp9, trd, opts,
                          [g9, gcm, g7, hb] + tr9 + trhb)


def per_sequence_calls(ctx, g7, gcm, g9, hb, tr9, trhb, seqs, opts, hits, per_seq_s):
    """The plain form: one ixg_search per sequence (each its own target) over
    the leading sequences for about per_seq_s seconds; its rate and whether
    its hits are the batched call's hits on those sequences."""
    singles = []
    for k in range(min(len(seqs), 4000)):
        singles.append(ixg.Target(ctx, seqs[k], len(seqs[k]) - 2))
    ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, singles[0], opts)
    per, nres, t0 = [], 0, time.perf_counter()
    k = 0
    while k < len(singles) and (k < 50 or time.perf_counter() - t0 < per_seq_s):
        hs, a = ixg.search(ctx, g7, gcm, g9, hb, tr9, trhb, singles[k], opts)
        per += [h + (k,) for h in hs]
        nres += a.nres
        k += 1
    dts = time.perf_counter() - t0
    for x in singles:
        x.free()
    sub = [h for h in hits if h[7] < k]
    key = lambda h: (h[5], -h[3], h[7], h[0], -h[6], h[1])  # noqa: E731  (the E-value sort's keys)
    same = [h[:3] + h[6:] for h in sub] == [h[:3] + h[6:] for h in sorted(per, key=key)]
    return {"value": nres / dts / 1e6, "unit": "Mres/s", "sequences": k, "ms_per_call": dts / k * 1e3,
            "same_hits_as_batched": bool(same)}


def finish_batched(extra, acct, dt, seqs, cm, W, st, kt, steps, placed, found, cpu, p7, cp9, trd, opts, handles):
    out = {"value": acct.nres / dt / 1e6, "unit": "Mres/s", "ms_per_step": dt * 1e3,
           "workload": "10,000 iid sequences, 45% GC, log-normal lengths (median 2 kb, sigma 1.0, clipped to "
                       "200 nt-200 kb, seed 4), 100 CM-sampled homologs inserted into random sequences; "
                       "120-column three-hairpin CM (stems 8/10/9 bp, 40-seq alignment, seed 4); tier 20-200 Mb: "
                       "F1 0.35, F2 = F2b 0.15 (Viterbi on), F3=F3b=F4=F4b=F5 0.003, F6 1e-4; default "
                       "truncation; one ixg_search over a multi-sequence target per step",
           "sequences": len(seqs), "residues": int(acct.nres), "cm_states": int(cm["M"]), "cm_W": W,
           "stage_ms": st,
           "kernels": {kk: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps}
                       for kk, v in kt.items() if v[1] > 0},
           "survivors": {"f1": int(acct.f.n_past_msv), "f2": int(acct.f.n_past_vit),
                         "f3b": int(acct.f.n_past_fwdbias), "f5_envelopes": int(acct.e.n_past_edef),
                         "f6": int(acct.n_past_cyk), "hits": int(acct.n_hits),
                         "truncated_pass_hits": [int(x) for x in acct.tr_hits]},
           "planted_found": f"{found} of {len(placed)}",
           "autotune_fwd_shapes": extra.pop("autotune", None)}
    out.update(extra)
    if "per_sequence_calls" in out:
        out["batched_over_per_sequence"] = out["value"] / out["per_sequence_calls"]["value"]
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_pipe_bind import OracleSearch
        o = OracleSearch(p7, cm, cp9, W, trcp9=trd)
        budget, threads = cpu
        # a sample sized to the section's CPU budget from a first 500-sequence run
        t0 = time.perf_counter()
        _, oa = o.search_multi(seqs[:500], opts, threads=threads)
        n = int(min(len(seqs), max(500, 500 * budget / max(time.perf_counter() - t0, 1e-3))))
        t0 = time.perf_counter()
        ohits, oa = o.search_multi(seqs[:n], opts, threads=threads)
        dtc = time.perf_counter() - t0
        out["cpu"] = {"value": oa.nres / dtc / 1e6, "unit": "Mres/s", "cores": threads, "kind": "port",
                      "sample": f"the first {n} sequences ({oa.nres} residues searched), oracle "
                                f"ixo_search_multi_mt (C restatement, not infernox), {dtc:.1f} s"}
    for x in handles:
        x.free()
    return out


def default_threads():
    """Host threads for the CPU baselines: the CPUs this process may run on,
    capped by the cgroup CPU quota when one is set."""
    nproc, aff, q = cpu_quota()
    n = aff or nproc or 1
    if q:
        n = min(n, max(1, int(q)))
    return n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="oracle time budget of the headline cpu_baseline")
    ap.add_argument("--cpu-threads", type=int, default=default_threads())
    ap.add_argument("--cpu-section-seconds", type=float, default=2.0,
                    help="oracle time budget per secondary section (CPU rate beside each GPU rate)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cm", action="store_true", help="skip the secondary scanner sections")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="only the headline search step")
    ap.add_argument("--serve-only", action="store_true", help="only run the serving-loop section")
    ap.add_argument("--multi-only", action="store_true", help="only run the several-CM HB launch section")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched uneven-database search section (10,000 sequences)")
    ap.add_argument("--genome-only", action="store_true",
                    help="only run the genome-scale search section (4.6 Mb, 959-state CM)")
    ap.add_argument("--per-seq-seconds", type=float, default=4.0,
                    help="time budget of the batched section's one-call-per-sequence comparison (0: skip it)")
    ap.add_argument("--hb-only", action="store_true",
                    help="only run the HB CYK / Inside section (959-state CM, 512 envelopes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 256 MiB device copy probe (HBM counter calibration) and print its rate")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (the search shards by "
                                   "independent (chunk x strand) tasks, see DESIGN.md)"}))
        sys.exit(2)
    if a.copy_probe:
        with ixg.Context(0) as ctx:
            nb = 256 << 20
            ms = ctx.copy_probe(nb, 2)
        # a profiler's counter pass serialises the launches and the events
        # can read 0: the rate is then unknown, not infinite
        print(json.dumps({"copy_probe_bytes": nb, "ms_per_copy": ms,
                          "GB/s": 2 * nb / (ms * 1e-3) / 1e9 if ms > 0 else None}))
        return
    if a.multi_only:
        with ixg.Context(0) as ctx:
            print(json.dumps({"hb_multi": time_multi(ctx)}))
        return
    if a.hb_only:
        with ixg.Context(0) as ctx:
            cpu = None if a.no_cpu else (a.cpu_section_seconds, a.cpu_threads)
            print(json.dumps({"hb_scanners": time_cm(ctx, a.steps, a.warmup, cpu, cm_workload())}))
        return
    if a.genome_only:
        with ixg.Context(0) as ctx:
            cpu = None if a.no_cpu else (a.cpu_section_seconds, a.cpu_threads)
            print(json.dumps({"genome_search": time_genome_search(ctx, max(1, a.steps // 3), 1, cpu)}))
        return
    if a.batched_only:
        with ixg.Context(0) as ctx:
            cpu = None if a.no_cpu else (a.cpu_section_seconds, a.cpu_threads)
            print(json.dumps({"batched_uneven_search": time_batched(ctx, max(1, a.steps // 3), cpu,
                                                                    a.per_seq_seconds)}))
        return
    if a.serve_only:
        with ixg.Context(0) as ctx:
            print(json.dumps({"serving": time_serving(ctx), "serving_search": time_search_serving(ctx)}))
        return
    wl = search_workload()
    secondary = not (a.headline_only or a.no_cm)
    cpu = None if a.no_cpu else (a.cpu_section_seconds, a.cpu_threads)
    with ixg.Context(0) as ctx:
        sr, opts, gpu_hits = time_search(ctx, a.steps, a.warmup, wl)
        print(json.dumps({"progress": "headline search done", "ms_per_step": sr["ms_per_step"]}),
              file=sys.stderr, flush=True)
        extra = {}
        if secondary:
            extra["filter_cascade_genome"] = time_cascade(ctx, a.steps, a.warmup, None if a.no_cpu else
                                                          a.cpu_section_seconds * 4, a.cpu_threads)
            hwl = cm_workload()
            extra["hb_scanners"] = time_cm(ctx, max(1, a.steps // 2), 1, cpu, hwl)
            extra["cp9_bands"] = time_cp9(ctx, 2, cpu, hwl)
            extra["hb_truncated_scanners"] = time_trcm(ctx, max(1, a.steps // 2), 1, cpu)
            extra["qdb_scanners"] = time_qdb(ctx, 2, cpu)
            p, d, _, _, _, _ = workload()
            extra["p7_forward_full"] = time_p7_forward(ctx, p, d, 2, cpu=cpu)
            extra["hb_multi"] = time_multi(ctx)
            extra["batched_uneven_search"] = time_batched(ctx, max(1, a.steps // 3), cpu, a.per_seq_seconds)
            extra["genome_search"] = time_genome_search(ctx, max(1, a.steps // 3), 1, cpu)
        if secondary and not a.no_serve:
            extra["serving"] = time_serving(ctx)
            extra["serving_search"] = time_search_serving(ctx)
    p7, cm = wl[0], wl[1]
    line = {
        "metric": "Mres/s searched + CYK GCUPS at 1 GPU; p50/p99 latency at fixed QPS",
        "value": sr["value"],
        "unit": "Mres/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sr["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8+i32+f32",
        "data": "synthetic",
        "config": {"workload": "plain cmsearch, small: 72-column tRNA-like cloverleaf CM (stems 7/4/5/5, loops "
                               "7-9; alignment 40 seqs, seed 1), 1 Mb iid ACGU target (seed 2) + 20 planted "
                               "CM samples, both strands, tier thresholds F1 0.35, F3=F3b=F4=F4b=F5 0.005, F6 "
                               "1e-4",
                   "timed_step": "one ixg_search call, cmsearch's default pipeline: STD pass F1 SSV, F3 Forward, "
                                 "F3b bias, F4/F4b glocal Forward, glocal Backward + domain definition, F5, CP9 "
                                 "bands (tau iteration), F6 HB CYK + envelope redefinition, CP9 bands at final "
                                 "tau, F7 HB Inside + null3 + greedy hits; the terminal truncated passes (5' and "
                                 "3' on the maxW residues at each end, both strands: filter cascade, "
                                 "Rgm/Lgm envelope definition, truncated CP9 bands, TrCYK filter, FTrInside); "
                                 "coordinates, E-values, overlap removal, sort; not alignment",
                   "cm_states": int(cm["M"]), "cm_W": int(cm["W"]), "p7_M": int(p7["M"]),
                   "residues": sr["residues"]},
        "cyk_gcups": sr["cyk_gcups"],
        "inside_gcups": sr["inside_gcups"],
        "stage_ms": sr["stage_ms"],
        "survivors": sr["survivors"],
        "planted_found": sr["planted_found"],
        "kernels": sr["kernels"],
        "roofline": sr["roofline"],
    }
    line.update(extra)
    if not a.no_cpu:
        cb, cpu_hits = search_cpu_baseline(wl, opts, a.cpu_threads, a.cpu_seconds)
        cb["same_hits_as_gpu"] = [h[:3] for h in cpu_hits] == [h[:3] for h in gpu_hits]
        line["cpu_baseline"] = cb
    # the ratios the targets are stated in, from the sections above (null when a section did not run)
    bu, ss = line.get("batched_uneven_search", {}), line.get("serving_search", {})
    line["ratios"] = {
        "headline_gpu_over_cpu_baseline": line["value"] / line["cpu_baseline"]["value"] if "cpu_baseline" in line
        else None,
        "batched_over_per_sequence_calls": bu.get("batched_over_per_sequence"),
        "served_search_over_plain_per_request": ss.get("served_vs_plain"),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
