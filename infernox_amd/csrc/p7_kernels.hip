// gfx950 kernels for the p7 filter stages: F1 SSV (speculative sweep +
// exact per-chunk resolve), F2 Viterbi, F3 Forward (striped order), F3b
// composition bias. All integer paths are exact; float paths reproduce the
// reference's evaluation order (built with -ffp-contract=off).
#include "ixg_internal.h"

namespace ixg {

#define DPP_WAVE_SHR1 0x138
#define DPP_QUAD_RSZ 0x90   // quad_perm [0,0,1,2]
#define DPP_QUAD_X1 0xB1    // quad_perm [1,0,3,2]
#define DPP_QUAD_X2 0x4E    // quad_perm [2,3,0,1]

// SSV cost of model position k for residue code x. Canonical codes come from
// the packed word c4 (0xFFFFFFFF past m); degenerate codes read rbv, whose
// rows exist for k = 0..m only, so positions past m cost 255 without a load.
__device__ __forceinline__ uint32_t ssv_cost(const SsvDev &p, uint32_t c4, int k, int x) {
  if (x < 4) return (c4 >> (8 * x)) & 0xFFu;
  return (x < KP && k <= p.m) ? (uint32_t)p.rbv[k * KP + x] : 255u;
}

// One SSV cell: subs(adds(max(prev, xb), bias), cost) on u8 values held in u32.
__device__ __forceinline__ uint32_t ssv_cell(uint32_t prev, uint32_t xb, uint32_t bias, uint32_t c) {
  uint32_t t = prev > xb ? prev : xb;
  t = t + bias;
  t = t < 255u ? t : 255u;
  return t > c ? t - c : 0u;
}

// ---------------------------------------------------------------------------
// F1 sweep. Each 64-lane wave owns SWEEP_SEG rows of one task and computes,
// for every row, the SSV row maximum WITHOUT the reference's reset-and-skip
// (cm_pipeline.rs:703-712). A cell (i,k) depends only on residues
// i-k+1..i (the diagonal enters at k=1 from the zero column), so starting the
// sweep M-1 rows early with a zero row reproduces every row of the segment.
// Resetting only lowers cells (max/adds/subs are monotone and the reset value
// is the minimum), so rows whose reset-free maximum stays below the threshold
// can never trigger; the others are flagged in a bitmask for the resolve
// kernel.
//
// Layout: lane = row within a 64-row tile, and a wave carries SWEEP_TILES
// tiles (SWEEP_TILES*64 consecutive rows) in registers, all advancing one
// model position k per step. The diagonal dependency (i-1,k-1) -> (i,k) is a
// one-lane wave shift inside a tile; lane 0 of tile t takes lane 63 of tile
// t-1 (still holding step k-1, as tiles are updated from the last to the
// first) through a wave rotate, and lane 0 of the first tile takes the
// previous tile group's last row from a small LDS carry row. The row maximum
// and its leftmost position are tracked as one packed key
// (value << 16 | 0xFFFF - k), whose maximum is the largest value at the
// smallest k. Costs come from wave-uniform scalar loads, four positions at a
// time.
// ---------------------------------------------------------------------------
#define DPP_WAVE_ROR1 0x13C
typedef __attribute__((address_space(4))) const uint32_t CU32;

__device__ __forceinline__ uint32_t ssv_in(uint32_t prev_tile, uint32_t cur_tile) {
  // lanes 1..63: cur_tile of lane-1; lane 0: prev_tile of lane 63
  const int rot = __builtin_amdgcn_update_dpp(0, (int)prev_tile, DPP_WAVE_ROR1, 0xF, 0xF, false);
  return (uint32_t)__builtin_amdgcn_update_dpp(rot, (int)cur_tile, DPP_WAVE_SHR1, 0xF, 0xF, false);
}

template <int R, bool DEGEN>
__device__ __forceinline__ void ssv_group(const SsvDev &p, const CU32 *c4g, const uint8_t *cprev, uint8_t *cnext,
                                          const int (&x)[R], uint32_t (&key)[R]) {
  const int m = p.m, lane = threadIdx.x;
  const uint32_t xb = p.xb, bias = p.bias;
  uint32_t a[R];
#pragma unroll
  for (int t = 0; t < R; t++) a[t] = 0;
  const uint32_t *cp32 = (const uint32_t *)cprev;
  uint32_t *cn32 = (uint32_t *)cnext;
  // Carry byte b holds the last row's value at position b (byte 0: the zero
  // column). Word j = bytes 4j..4j+3 is read for steps 4j+1..4j+4 and written
  // after them, holding position 4j from the step before (lastb).
  uint32_t lastb = 0;
  for (int k0 = 1; k0 <= m; k0 += 4) {
    const uint32_t carry4 = cp32[(k0 - 1) >> 2];  // carry[k0-1 .. k0+2]
    uint32_t out4 = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k = k0 + u;
      const uint32_t c4 = c4g[k];
      const uint32_t cin = (carry4 >> (8 * u)) & 0xFFu;
#pragma unroll
      for (int t = R - 1; t >= 0; t--) {
        const uint32_t in = t ? ssv_in(a[t - 1], a[t])
                              : (uint32_t)__builtin_amdgcn_update_dpp((int)cin, (int)a[0], DPP_WAVE_SHR1, 0xF, 0xF,
                                                                      false);
        uint32_t c;
        if (DEGEN) c = k <= m ? ssv_cost(p, c4, k, x[t]) : 255u;  // positions past m cost 255
        else c = __builtin_amdgcn_ubfe(c4, 8u * (uint32_t)x[t], 8);
        a[t] = ssv_cell(in, xb, bias, c);
        key[t] = max(key[t], (a[t] << 16) | (uint32_t)(0xFFFF - k));
      }
      out4 |= (a[R - 1] & 0xFFu) << (8 * u);
    }
    if (lane == 63) cn32[(k0 - 1) >> 2] = lastb | (out4 << 8);
    lastb = out4 >> 24;
  }
}

__global__ __launch_bounds__(64) void ssv_sweep_kernel(SsvDev p, const TaskDev *tasks,
                                                       const SweepItem *items, uint64_t *mask,
                                                       uint32_t *cinfo, int thresh) {
  constexpr int R = SWEEP_TILES;
  extern __shared__ uint32_t sh[];  // carry rows 2 x cstride bytes
  const SweepItem it = items[blockIdx.x];
  const TaskDev t = tasks[it.task];
  const int m = p.m;
  const int lane = threadIdx.x;
  const int cstride = ((m + 8) + 3) & ~3;
  uint8_t *carry = (uint8_t *)sh;
  for (int k = lane; k < 2 * cstride; k += 64) carry[k] = 0;
  __syncthreads();
  const CU32 *c4g = (const CU32 *)p.cost4;
  const int64_t s0 = 1 + (int64_t)it.seg * SWEEP_SEG;
  const int64_t s1 = min(s0 + (int64_t)SWEEP_SEG, t.L + 1);
  const int64_t warm = (int64_t)((m - 1 + 63) / 64) * 64;
  int64_t w0 = s0 - warm;
  if (w0 < 1) w0 = 1;
  int cb = 0;
  for (int64_t g = w0; g < s1; g += 64 * R) {
    int x[R];
    bool degenerate = false;
#pragma unroll
    for (int u = 0; u < R; u++) {
      const int64_t i = g + 64 * u + lane;
      x[u] = i < s1 ? (int)t.dsq[i] : 0;
      degenerate |= x[u] >= 4;
    }
    uint32_t key[R];
#pragma unroll
    for (int u = 0; u < R; u++) key[u] = 0;
    const uint8_t *cprev = carry + cb * cstride;
    uint8_t *cnext = carry + (cb ^ 1) * cstride;
    if (__any(degenerate)) ssv_group<R, true>(p, c4g, cprev, cnext, x, key);
    else ssv_group<R, false>(p, c4g, cprev, cnext, x, key);
    __syncthreads();
    cb ^= 1;
#pragma unroll
    for (int u = 0; u < R; u++) {
      const int64_t b = g + 64 * u;
      if (b >= s0 && b < s1) {
        const int64_t i = b + lane;
        const uint32_t rowmax = key[u] >> 16;
        const bool cand = i < s1 && (int)rowmax >= thresh;
        const uint64_t bits = __ballot(cand);
        if (lane == 0) mask[t.mask_off + (b - 1) / 64] = bits;
        // the reset-free row maximum and its leftmost position, for the resolve
        if (cand) cinfo[t.mask_off * 64 + (i - 1)] = (rowmax << 16) | (0xFFFFu - (key[u] & 0xFFFFu));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// F1 resolve: replays the reference's sequential reset-and-skip logic
// (cm_pipeline.rs:521-717) exactly, one wave per SWEEP_SEG-row block of a
// task. A trigger at row c' ends its forward extension at most M-2 rows
// later, and after a restart at r only rows below r+M-1 differ from the
// reset-free sweep. Hence a flagged row c with no flagged row in
// [c-2M, c) sees the same state as at the start of the task: every earlier
// restart r satisfies c >= r+M-1, so c triggers with the sweep's values and
// everything after it depends only on c ("fresh" row). The wave of a block
// starts at the block's first fresh row and stops at the first fresh row
// past the block's end, where the next block's wave takes over; flagged rows
// before the first fresh row belong to the wave of an earlier block.
// Within a wave, a flagged row at least M-1 rows past the last restart has
// its reset-free values, so it triggers with the maximum and leftmost argmax
// the sweep recorded; rows closer to a restart are recomputed from the
// restart (lanes hold KPL consecutive model positions each). The diagonal
// is then recovered backward and extended forward on lane 0.
// ---------------------------------------------------------------------------
// SSV cost of model position k (1..m) for residue code x.
__device__ __forceinline__ int ssv_cost_at(const SsvDev &p, int k, int x) {
  return (x < KP && k >= 0 && k <= p.m) ? (int)p.rbv[k * KP + x] : 255;
}
// Wave-wide inclusive prefix sum / maximum over lanes 0..63, exclusive maximum.
__device__ __forceinline__ int wave_incl_sum(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}
__device__ __forceinline__ int wave_incl_max(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off) v = max(v, t);
  }
  return v;
}
__device__ __forceinline__ int wave_excl_max(int v) {
  const int lane = threadIdx.x & 63;
  const int in = wave_incl_max(v);
  const int t = __shfl_up(in, 1);
  return lane ? t : INT_MIN;
}

template <int KPL>
__global__ __launch_bounds__(64) void ssv_resolve_kernel(SsvDev p, const TaskDev *tasks, const SweepItem *items,
                                                         const uint64_t *mask, const uint32_t *cinfo, int thresh,
                                                         RawWinDev *out, int64_t cap,
                                                         unsigned long long *counter) {
  const SweepItem it = items[blockIdx.x];
  const TaskDev t = tasks[it.task];
  const int m = p.m;
  const int lane = threadIdx.x;
  const int kb = lane * KPL + 1;
  uint32_t c4[KPL], d[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    c4[j] = (kb + j <= m) ? p.cost4[kb + j] : 0xFFFFFFFFu;
    d[j] = 0;
  }
  const uint32_t xb = p.xb, bias = p.bias;
  const int64_t L = t.L;
  const int64_t nwords = (L + 63) / 64;
  const int64_t bs = 1 + (int64_t)it.seg * SWEEP_SEG, be = bs + SWEEP_SEG;
  const int64_t gap = 2 * (int64_t)m;
  const int64_t s0 = bs - gap > 1 ? bs - gap : 1;
  int64_t lastf = -1;  // last flagged row seen (-1: none since s0)
  bool active = false;
  int64_t r = 1, pos = 1;
  // mask words are fetched 64 at a time (one per lane) and read back with
  // readlane, so the walk over a block waits on one load per 4096 rows
  uint64_t wbuf = 0;
  for (int64_t w = (s0 - 1) / 64; w < nwords; w++) {
    const int wl = (int)((w - (s0 - 1) / 64) & 63);
    if (wl == 0) wbuf = w + lane < nwords ? mask[t.mask_off + w + lane] : 0ull;
    uint64_t bits = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(wbuf >> 32), wl) << 32) |
                    (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wbuf, wl);
    if (w == (s0 - 1) / 64) bits &= ~0ull << ((s0 - 1) % 64);
    while (bits) {
      const int bi = __builtin_ctzll(bits);
      bits &= bits - 1;
      const int64_t c = 64 * w + 1 + bi;
      const bool fresh = lastf < 0 || c - lastf > gap;
      if (!active) {
        if (c >= be) return;  // no fresh row in this block
        if (c < bs || !fresh) {
          lastf = c;
          continue;
        }
        active = true;
      } else if (c >= be && fresh) {
        return;
      }
      lastf = c;
      if (c < r) continue;
      uint32_t bv;
      int bk;
      // Before the wave's first trigger the scan is reset-free, so the
      // sweep's values hold as well.
      if (c >= r + m - 1 || r == 1) {
        const uint32_t ci = cinfo[t.mask_off * 64 + (c - 1)];
        bv = ci >> 16;
        bk = (int)(ci & 0xFFFF);
      } else {
        // rows pos..c from the restart state, residues fetched 64 at a time
        for (int64_t base = pos; base <= c; base += 64) {
          const int64_t ri = base + lane;
          const int xr = ri <= c ? (int)t.dsq[ri] : 0;
          const int nrow = (int)min((int64_t)64, c - base + 1);
          for (int rr = 0; rr < nrow; rr++) {
            const int x = __builtin_amdgcn_readlane(xr, rr);
            const uint32_t prev =
                (uint32_t)__builtin_amdgcn_update_dpp(0, (int)d[KPL - 1], DPP_WAVE_SHR1, 0xF, 0xF, false);
            if (x < 4) {
              // one uniform branch per row; positions past m cost 255
              // (c4 = 0xFFFFFFFF there), which makes their cells 0
              const uint32_t sh8 = 8u * (uint32_t)x;
#pragma unroll
              for (int j = KPL - 1; j >= 0; j--) {
                const uint32_t in = j ? d[j - 1] : prev;
                d[j] = ssv_cell(in, xb, bias, __builtin_amdgcn_ubfe(c4[j], sh8, 8));
              }
            } else {
#pragma unroll
              for (int j = KPL - 1; j >= 0; j--) {
                const uint32_t in = j ? d[j - 1] : prev;
                const uint32_t cc = kb + j <= m ? ssv_cost(p, c4[j], kb + j, x) : 255u;
                d[j] = (kb + j <= m) ? ssv_cell(in, xb, bias, cc) : 0u;
              }
            }
          }
        }
        pos = c + 1;
        // The row triggers only if some cell reaches the threshold; a ballot
        // of the lanes' maxima decides that, and only then is the row maximum
        // and its leftmost position reduced (one packed key per lane:
        // value << 16 | 0xFFFF - k, whose maximum is the largest value at the
        // smallest k).
        uint32_t lmax = 0;
#pragma unroll
        for (int j = 0; j < KPL; j++)
          if (kb + j <= m) lmax = max(lmax, d[j]);
        if (!__any((int)lmax >= thresh)) continue;
        uint32_t key = 0;
#pragma unroll
        for (int j = 0; j < KPL; j++)
          if (kb + j <= m) key = max(key, (d[j] << 16) | (uint32_t)(0xFFFF - (kb + j)));
        for (int off = 32; off >= 1; off >>= 1) key = max(key, (uint32_t)__shfl_xor((int)key, off));
        bv = key >> 16;
        bk = bv ? (int)(0xFFFFu - (key & 0xFFFFu)) : 0;
      }
      if ((int)bv >= thresh && bk > 0) {
        const int end_k = bk;
        // backward diagonal recovery (cm_pipeline.rs:657-672), 64 steps per
        // chunk: step j reads (end_k - j, c - j) and runs while the remaining
        // score stays above the baseline and both indices are >= 1
        const int baseline = (int)p.base_b - (int)p.tjb_b - (int)p.tbm_b;
        int64_t nit = 0;
        {
          int rem0 = (int)bv;
          for (int64_t j0 = 0;; j0 += 64) {
            const int64_t j = j0 + lane;
            const bool inb = end_k - j >= 1 && c - j >= 1;
            const int term = inb ? (int)p.bias - ssv_cost_at(p, end_k - j, t.dsq[c - j]) : 0;
            const int incl = wave_incl_sum(term);
            const bool ok = inb && rem0 - (incl - term) > baseline;
            const uint64_t fail = __ballot(!ok);
            if (fail) {
              nit = j0 + __builtin_ctzll(fail);
              break;
            }
            rem0 -= __shfl(incl, 63);
          }
        }
        const int64_t start = end_k - nit + 1, tstart = c - nit + 1;
        // forward extension (cm_pipeline.rs:674-690), 64 steps per chunk:
        // step j scores (end_k+1+j, c+1+j) while k < m and n <= L, and stops
        // after the fifth consecutive step below the running maximum
        int max_sc = (int)bv, since = 0;
        int64_t max_end = c;
        {
          int sc0 = (int)bv;
          for (int64_t j0 = 0;; j0 += 64) {
            const int64_t j = j0 + lane, k = end_k + 1 + j, n = c + 1 + j;
            const bool valid = k < m && n <= L;
            const int term = valid ? (int)p.bias - ssv_cost_at(p, (int)k, t.dsq[n]) : 0;
            const int sc2 = sc0 + wave_incl_sum(term);
            const int before = max(max_sc, wave_excl_max(sc2));  // running maximum before step j
            const bool imp = valid && sc2 >= before;
            const int last = wave_incl_max(imp ? lane : -1);
            const int sinc = last >= 0 ? lane - last : since + lane + 1;
            const uint64_t stop = __ballot(!valid || sinc == 5);
            // steps taken in this chunk: up to the first invalid one
            // (excluded) or the one that reaches five (included)
            int taken = 64;
            if (stop) {
              const int s = __builtin_ctzll(stop);
              taken = __shfl((int)valid, s) ? s + 1 : s;
            }
            if (taken > 0) {
              const int lt = __shfl(last, taken - 1);
              if (lt >= 0) {
                max_end = c + 1 + j0 + lt;
                max_sc = __shfl(sc2, lt);
              }
              since = __shfl(sinc, taken - 1);
              sc0 = __shfl(sc2, taken - 1);
            }
            if (stop) break;
          }
        }
        if (lane == 0) {
          const int64_t end_model = end_k + (max_end - c);
          // Records of one block are appended in trigger order; the host
          // orders them by (task, trigger row).
          const unsigned long long slot = atomicAdd(counter, 1ull);
          if ((int64_t)slot < cap) {
            RawWinDev rw;
            rw.n = tstart;
            rw.length = end_model - start + 1;
            rw.k = end_model;
            rw.row = c;
            rw.score = ((float)(max_sc - (int)p.tjb_b) - (float)p.base_b) / p.scale_b - 3.0f;
            rw.task = it.task;
            out[slot] = rw;
          }
        }
        r = max_end + 1;
        pos = r;
#pragma unroll
        for (int j = 0; j < KPL; j++) d[j] = 0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// F3 Forward, striped order. Four lanes emulate the reference's four SSE
// lanes for one window (cm_pipeline.rs:1268-1418): lane z owns model
// positions q+1+z*nq and runs the q loop in the same order, so every add and
// multiply happens in the same sequence as on the CPU. A wave serves 16
// windows. The three DP columns live in LDS, one slot per lane.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float quad_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// ---------------------------------------------------------------------------
// F3b composition bias: the reference's 2-state filter-null Forward with
// per-row max rescaling (cm_pipeline.rs:1722-1774), one lane per window. The
// length-only tail (L*ln p1 + ln(1-p1)) arrives precomputed in (a, b).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void bias_kernel(BiasDev p, const uint8_t *seq, const WinDev *wins,
                                                  int64_t nwin, float *sc) {
  const int64_t wi = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (wi >= nwin) return;
  const WinDev w = wins[wi];
  const int64_t L = w.len;
  if (L == 0) { sc[w.idx] = 0.f; return; }
  const float p1 = (float)L / ((float)L + 1.0f);
  const float t00 = p1, t01 = 1.0f - p1;
  const float t10 = p.t1[0], t11 = p.t1[1], t12 = p.t1[2];
  const uint8_t *s = seq + w.off;
  float logsc = 0.0f;
  int x = s[1];
  float d0 = p.eo[x][0] * p.pi[0], d1 = p.eo[x][1] * p.pi[1];
  float mx = d0 > d1 ? d0 : d1;
  d0 /= mx;
  d1 /= mx;
  logsc += (float)log((double)mx);
  for (int64_t i = 2; i <= L; i++) {
    x = s[i];
    float s0 = 0.0f;
    s0 += d0 * t00;
    s0 += d1 * t10;
    float s1 = 0.0f;
    s1 += d0 * t01;
    s1 += d1 * t11;
    const float n0 = s0 * p.eo[x][0], n1 = s1 * p.eo[x][1];
    mx = n0 > n1 ? n0 : n1;
    d0 = n0 / mx;
    d1 = n1 / mx;
    logsc += (float)log((double)mx);
  }
  const float end = d0 * 1.0f + d1 * t12;
  logsc += (float)log((double)end);
  sc[w.idx] = (logsc + w.a) + w.b;
}

// ---------------------------------------------------------------------------
// F2 Viterbi (p7_vitfilter.rs:305-386), one wave per window, KPL model
// positions per lane. M and I cells are independent within a row; the D row
// is a max-plus prefix over k whose added terms are all <= 0, so saturating
// at -32768 after every step equals saturating once at the end; the prefix
// is computed exactly in 32-bit with a wave scan.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int adds16(int a, int b) {
  int r = a + b;
  r = r > 32767 ? 32767 : r;
  return r < -32768 ? -32768 : r;
}

template <int KPL>
__global__ __launch_bounds__(64) void vit_kernel(VitDev p, const uint8_t *seq, const WinDev *wins,
                                                 int64_t nwin, float *sc, uint8_t *ovf) {
  const WinDev w = wins[blockIdx.x];
  const int m = p.m, lane = threadIdx.x, kb = lane * KPL + 1;
  const int16_t *tmm = p.tr, *tim = p.tr + (m + 1), *tdm = p.tr + 2 * (m + 1), *vbm = p.tr + 3 * (m + 1);
  const int16_t *tmi = p.tr + 4 * (m + 1), *tii = p.tr + 5 * (m + 1), *tmd = p.tr + 6 * (m + 1),
                *tdd = p.tr + 7 * (m + 1);
  int rmm[KPL], rim[KPL], rdm[KPL], rbm[KPL], rmi[KPL], rii[KPL], rmd[KPL], rdd[KPL];
  int mp[KPL], ip[KPL], dp[KPL];
  const int NEG = -32768;
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    const int k = kb + j;
    const bool ok = k <= m;
    rmm[j] = ok ? tmm[k - 1] : NEG;
    rim[j] = ok ? tim[k - 1] : NEG;
    rdm[j] = ok ? tdm[k - 1] : NEG;
    rbm[j] = ok ? vbm[k] : NEG;
    rmi[j] = ok ? tmi[k] : NEG;
    rii[j] = ok ? tii[k] : NEG;
    rmd[j] = ok ? tmd[k - 1] : NEG;
    rdd[j] = ok ? tdd[k - 1] : NEG;
    mp[j] = ip[j] = dp[j] = NEG;
  }
  const int xw_move = (int)w.a;
  int xn = p.base_w, xb = adds16(xn, xw_move), xj = NEG, xc = NEG;
  bool overflow = false;
  const uint8_t *s = seq + w.off;
  for (int64_t i = 1; i <= w.len; i++) {
    const int x = s[i];
    int mpl = __shfl_up(mp[KPL - 1], 1), ipl = __shfl_up(ip[KPL - 1], 1), dpl = __shfl_up(dp[KPL - 1], 1);
    if (lane == 0) mpl = ipl = dpl = NEG;
    int mc[KPL], ic[KPL];
    int lmax = NEG;
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int pm = j ? mp[j - 1] : mpl, pi = j ? ip[j - 1] : ipl, pd = j ? dp[j - 1] : dpl;
      int v = adds16(xb, rbm[j]);
      v = max(v, adds16(pm, rmm[j]));
      v = max(v, adds16(pi, rim[j]));
      v = max(v, adds16(pd, rdm[j]));
      const int k = kb + j;
      const int e = k <= m ? (int)p.rwv[k * KP + x] : NEG;
      mc[j] = k <= m ? adds16(v, e) : NEG;
      ic[j] = k <= m ? max(adds16(mp[j], rmi[j]), adds16(ip[j], rii[j])) : NEG;
      lmax = max(lmax, mc[j]);
    }
    for (int off = 32; off >= 1; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off));
    // D row: dc[k] = max(mc[k-1] + tmd[k-1], dc[k-1] + tdd[k-1]), clamp once.
    int mcl = __shfl_up(mc[KPL - 1], 1);
    if (lane == 0) mcl = NEG;
    int A[KPL];
    int fa = -(1 << 30), fb = 0;
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      const int pmc = j ? mc[j - 1] : mcl;
      A[j] = pmc + rmd[j];
      fa = max(A[j], fa + rdd[j]);
      fb = fb + rdd[j];
    }
    for (int off = 1; off < 64; off <<= 1) {
      const int oa = __shfl_up(fa, off), ob = __shfl_up(fb, off);
      if (lane >= off) {
        fa = max(fa, oa + fb);
        fb = ob + fb;
      }
    }
    int cin = __shfl_up(max(fa, NEG + fb), 1);
    if (lane == 0) cin = NEG;
    int dcur = cin;
#pragma unroll
    for (int j = 0; j < KPL; j++) {
      dcur = max(A[j], dcur + rdd[j]);
      dp[j] = (kb + j <= m) ? max(dcur, NEG) : NEG;
      dcur = max(dcur, NEG);
      mp[j] = mc[j];
      ip[j] = ic[j];
    }
    const int xe = lmax;
    if (xe >= 32767) { overflow = true; break; }
    xc = max(xc, adds16(xe, p.xw_e_move));
    xj = max(xj, adds16(xe, p.xw_e_loop));
    xb = max(adds16(xj, xw_move), adds16(xn, xw_move));
  }
  if (lane == 0) {
    ovf[w.idx] = overflow ? 1 : 0;
    float r = -INFINITY;
    if (!overflow && xc > NEG) {
      float t = (float)xc + (float)xw_move - (float)p.base_w;
      t /= p.scale_w;
      r = t - 3.0f;
    }
    sc[w.idx] = overflow ? 0.f : r;
  }
}

// ---------------------------------------------------------------------------
// F1b full MSV with the J state (cm_pipeline.rs:426-452), one wave per
// window, KPL model positions per lane; overflow (the 255 ceiling) -> ovf.
// ---------------------------------------------------------------------------
template <int KPL>
__global__ __launch_bounds__(64) void msv_kernel(SsvDev p, const uint8_t *seq, const WinDev *wins,
                                                 int64_t nwin, float *sc, uint8_t *ovf) {
  const WinDev w = wins[blockIdx.x];
  const int m = p.m, lane = threadIdx.x, kb = lane * KPL + 1;
  const uint32_t tjb_b = (uint32_t)w.a;               // reconfigured J->B cost
  const uint32_t tjbm = (tjb_b + p.tbm_b) & 0xFFu;    // wrapping u8 add
  uint32_t c4[KPL], d[KPL];
#pragma unroll
  for (int j = 0; j < KPL; j++) {
    c4[j] = (kb + j <= m) ? p.cost4[kb + j] : 0xFFFFFFFFu;
    d[j] = 0;
  }
  const uint32_t bias = p.bias, base = p.base_b, tec = (uint32_t)w.b;
  uint32_t xj = 0;
  uint32_t xb = base > tjbm ? base - tjbm : 0;
  bool overflow = false;
  const uint8_t *s = seq + w.off;
  for (int64_t i = 1; i <= w.len; i++) {
    const int x = s[i];
    uint32_t prev = (uint32_t)__shfl_up((int)d[KPL - 1], 1);
    if (lane == 0) prev = 0;
    uint32_t lmax = 0;
#pragma unroll
    for (int j = KPL - 1; j >= 0; j--) {
      const uint32_t in = j ? d[j - 1] : prev;
      const uint32_t cc = kb + j <= m ? ssv_cost(p, c4[j], kb + j, x) : 255u;
      d[j] = (kb + j <= m) ? ssv_cell(in, xb, bias, cc) : 0u;
      lmax = d[j] > lmax ? d[j] : lmax;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)lmax, off);
      lmax = o > lmax ? o : lmax;
    }
    uint32_t xe = lmax;
    if (min(xe + bias, 255u) == 255u) { overflow = true; break; }
    xe = xe > tec ? xe - tec : 0;
    xj = xj > xe ? xj : xe;
    const uint32_t bm = base > xj ? base : xj;
    xb = bm > tjbm ? bm - tjbm : 0;
  }
  if (lane == 0) {
    ovf[w.idx] = overflow ? 1 : 0;
    sc[w.idx] = overflow ? 0.f : ((float)((int)xj - (int)tjb_b) - (float)base) / p.scale_b - 3.0f;
  }
}

__global__ void revcomp_kernel(const uint8_t *fwd, uint8_t *rc, int64_t L) {
  // rc[j] = comp(fwd[L+1-j]) for j in 1..L (cm_search.rs:1822-1827); sentinels kept.
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > L + 1) return;
  if (j == 0 || j == L + 1) { rc[j] = 255; return; }
  const uint8_t d = fwd[L + 1 - j];
  rc[j] = d < 4 ? (uint8_t)(3 - d) : d;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_ssv_sweep(const SsvDev &p, const TaskDev *tasks, const SweepItem *items, int64_t nitems,
                            uint64_t *mask, uint32_t *cinfo, int thresh, hipStream_t s) {
  if (nitems == 0) return hipSuccess;
  const size_t lds = 2 * (size_t)(((p.m + 8) + 3) & ~3);
  hipLaunchKernelGGL(ssv_sweep_kernel, dim3((unsigned)nitems), dim3(64), lds, s, p, tasks, items, mask, cinfo,
                     thresh);
  return hipGetLastError();
}

template <int KPL>
static void resolve_k(const SsvDev &p, const TaskDev *tasks, const SweepItem *items, int64_t nitems,
                      const uint64_t *mask, const uint32_t *cinfo, int thresh, RawWinDev *out, int64_t cap,
                      unsigned long long *counter, hipStream_t s) {
  hipLaunchKernelGGL(ssv_resolve_kernel<KPL>, dim3((unsigned)nitems), dim3(64), 0, s, p, tasks, items, mask, cinfo,
                     thresh, out, cap, counter);
}
hipError_t launch_ssv_resolve(const SsvDev &p, const TaskDev *tasks, const SweepItem *items, int64_t nitems,
                              const uint64_t *mask, const uint32_t *cinfo, int thresh, RawWinDev *out, int64_t cap,
                              unsigned long long *counter, hipStream_t s) {
  if (nitems == 0) return hipSuccess;
  const int kpl = (p.m + 63) / 64;
#define IXG_RES(K) resolve_k<K>(p, tasks, items, nitems, mask, cinfo, thresh, out, cap, counter, s)
  if (kpl <= 1) IXG_RES(1);
  else if (kpl <= 2) IXG_RES(2);
  else if (kpl <= 4) IXG_RES(4);
  else if (kpl <= 8) IXG_RES(8);
  else if (kpl <= 16) IXG_RES(16);
#undef IXG_RES
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_bias(const BiasDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                       hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  hipLaunchKernelGGL(bias_kernel, dim3((unsigned)((nwin + 63) / 64)), dim3(64), 0, s, p, seq, wins, nwin, sc);
  return hipGetLastError();
}

template <int KPL>
static void vit_k(const VitDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc, uint8_t *ovf,
                  hipStream_t s) {
  hipLaunchKernelGGL(vit_kernel<KPL>, dim3((unsigned)nwin), dim3(64), 0, s, p, seq, wins, nwin, sc, ovf);
}
hipError_t launch_vit(const VitDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                      uint8_t *ovf, hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  const int kpl = (p.m + 63) / 64;
  if (kpl <= 1) vit_k<1>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 2) vit_k<2>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 4) vit_k<4>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 8) vit_k<8>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 16) vit_k<16>(p, seq, wins, nwin, sc, ovf, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int KPL>
static void msv_k(const SsvDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc, uint8_t *ovf,
                  hipStream_t s) {
  hipLaunchKernelGGL(msv_kernel<KPL>, dim3((unsigned)nwin), dim3(64), 0, s, p, seq, wins, nwin, sc, ovf);
}
hipError_t launch_msv(const SsvDev &p, const uint8_t *seq, const WinDev *wins, int64_t nwin, float *sc,
                      uint8_t *ovf, hipStream_t s) {
  if (nwin == 0) return hipSuccess;
  const int kpl = (p.m + 63) / 64;
  if (kpl <= 1) msv_k<1>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 2) msv_k<2>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 4) msv_k<4>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 8) msv_k<8>(p, seq, wins, nwin, sc, ovf, s);
  else if (kpl <= 16) msv_k<16>(p, seq, wins, nwin, sc, ovf, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// Plain device-to-device copy, 16 bytes per lane per step (HBM calibration).
__global__ void copy_probe_kernel(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

hipError_t launch_copy_probe(const void *a, void *b, int64_t n16, hipStream_t s) {
  hipLaunchKernelGGL(copy_probe_kernel, dim3(8192), dim3(256), 0, s, (const float4 *)a, (float4 *)b, n16);
  return hipGetLastError();
}

hipError_t launch_revcomp(const uint8_t *fwd, uint8_t *rc, int64_t L, hipStream_t s) {
  const int64_t n = L + 2;
  hipLaunchKernelGGL(revcomp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fwd, rc, L);
  return hipGetLastError();
}

}  // namespace ixg
