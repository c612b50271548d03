"""CPU replay of the p7 window kernels' table loads (no GPU needed).

Each window kernel (msv_kernel, vit_kernel, ssv_resolve_kernel in
infernox_amd/csrc/p7_kernels.hip) runs one 64-lane wave per window and gives
lane l the model positions k = l*KPL + 1 .. l*KPL + KPL, where KPL is the
launcher's power-of-two ceiling of ceil(M/64). Lanes past M still execute the
loop body, so every table load they issue must either be guarded by k <= M or
fall inside the table. This test replays, for every M the launchers accept
and every residue code (canonical 0..3, degenerate 4..17), the byte offsets
each lane would load, with the guards as written in the kernels, and checks
them against the allocation sizes of ixg_p7_create (api.cpp):
rbv (M+1)*18 bytes, rwv (M+1)*18 int16, the eight Viterbi transition rows of
M+1 int16 each, cost4 ((M+1+3) & ~3) + 4 words.

The unguarded form that msv_kernel had (ssv_cost(p, c4, kb + j, x) for every
lane) is replayed too: it reads past rbv for degenerate codes exactly when
KPL >= 4, where the overshoot crosses into the page after the table (M in
129..255 and 257..), which covers the window-score cases that faulted
(M = 150, 200, 300, 400) and none of those that passed (M <= 120).
"""
import numpy as np
import pytest

KP = 18
PAGE = 4096


def kpl_of(m):
    k = (m + 63) // 64
    for p in (1, 2, 4, 8, 16):
        if k <= p:
            return p
    return None


def lane_positions(m):
    kpl = kpl_of(m)
    lanes = np.arange(64)[:, None]
    j = np.arange(kpl)[None, :]
    return (lanes * kpl + 1 + j).ravel()


def ssv_cost_loads(m, ks, x, guarded_call):
    """rbv byte offsets loaded by ssv_cost(p, c4, k, x) at positions ks.

    guarded_call: the call site tests k <= m before calling (resolve kernel,
    sweep kernel, msv_kernel)."""
    if x < 4:
        return np.empty(0, dtype=np.int64)  # canonical codes come from c4 (a register)
    ks = ks[ks <= m] if guarded_call else ks
    return ks * KP + x


def cost4_loads(m, ks):
    # c4[j] = (kb + j <= m) ? p.cost4[kb + j] : 0xFFFFFFFF
    return ks[ks <= m]


def test_window_kernel_table_loads_in_bounds():
    for m in range(1, 1025):
        ks = lane_positions(m)
        rbv_bytes = (m + 1) * KP
        cost4_words = ((m + 1 + 3) & ~3) + 4
        for x in range(KP):
            # msv_kernel / ssv_resolve_kernel / sweep: call site guarded and ssv_cost guarded
            loads = ssv_cost_loads(m, ks, x, guarded_call=True)
            assert loads.size == 0 or loads.max() < rbv_bytes, (m, x)
            # ssv_cost on its own (any caller): guarded inside
            inner = ks[ks <= m] * KP + x if x >= 4 else np.empty(0, dtype=np.int64)
            assert inner.size == 0 or inner.max() < rbv_bytes
            # vit_kernel: e = k <= m ? p.rwv[k * KP + x] : NEG
            vl = ks[ks <= m] * KP + x
            assert vl.max() < (m + 1) * KP
        c4 = cost4_loads(m, ks)
        assert c4.max() < cost4_words
        # vit_kernel transition rows: t*[k-1] and vbm/tmi/tii[k], only for k <= m
        kk = ks[ks <= m]
        assert (kk - 1).min() >= 0 and kk.max() <= m


def test_unguarded_msv_read_overshoots_rbv_from_kpl4():
    """The pre-fix msv_kernel loaded rbv[(kb+j)*18 + x] on every lane for
    degenerate x. Replay that form and show where it leaves the allocation and
    where it crosses into the next page past the allocation's end."""
    crosses = []
    for m in (2, 10, 50, 64, 65, 100, 101, 120, 128, 129, 150, 200, 256, 257, 300, 400):
        ks = lane_positions(m)
        rbv_bytes = (m + 1) * KP
        worst = max(ssv_cost_loads(m, ks, x, guarded_call=False).max() for x in range(4, KP))
        if worst >= rbv_bytes:
            # the allocation starts page-aligned (device allocations are), so
            # the read crosses a page boundary past the end when the last
            # byte read is on a later page than the last byte allocated
            if worst // PAGE > (rbv_bytes - 1) // PAGE:
                crosses.append(m)
    # every M the faulting GPU cases used (150, 200, 300, 400) crosses, as
    # does any M in 129..255 or 257..; none of the M <= 128 cases that passed
    # does, and M = 256 (all lanes in range) reads nothing past the table
    assert crosses == [129, 150, 200, 257, 300, 400]
