// gfx950 HMM-banded CYK / Inside scanners (F6/F7), one workgroup of 1024
// lanes per envelope, as a wavefront over the subsequence length d:
//
// * Deck: each state's cells sit in a rectangle laid out diagonal-major,
//   (d, x) with x = j - jlo or x = i - ilo, whichever of the state's j and
//   i band widths is smaller, d over the state's band's d range. The cells
//   of one d are contiguous, which is what a wavefront step reads and writes;
//   a child cell's address is arithmetic on the child's rectangle (12 ints
//   per state, in LDS), never a per-row lookup. Rectangle cells outside the
//   true band hold -inf, the identity of max and of FLogsum, so reading one
//   is the same as the reference skipping the out-of-band child; the cells
//   whose j is outside the state's j band are never read and never written.
// * Schedule: at step t state v fills its cells of length d = t - level(v)
//   (all j at once), where a state's level is one more than its children's
//   highest. Every child cell is then at a smaller d (ML, MR, MP, IL, IR,
//   including the inserts' self transitions, and a bifurcation's splits
//   with 0 < k < d) or at the same d in a lower level (S, D, and the k = 0 /
//   k = d splits), so it was finished at an earlier step. One barrier per
//   step, no serial chain inside a step except a bifurcation cell's splits.
// * Lanes: each step the states' cell counts at their d are prefix-summed
//   in LDS and the step's cells are dealt round-robin over all lanes.
//
// Per cell the children are combined in increasing child order (max for
// CYK, the FLogsum table for Inside) and a bifurcation's splits in
// increasing k, the order of the reference's loops (cp9.rs:4788-5253,
// 5514-5885), so CYK is bit-identical and Inside applies the same float
// operations in the same order. Local begins are folded into the root by
// MAX after the wavefront (cp9.rs:5264-5304, 5908-5937); root cells also go
// to the envelope's compact root output.
#include "cm_internal.h"

namespace ixg {

namespace {

__device__ __forceinline__ int sdelta(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }
__device__ __forceinline__ int srdelta(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

template <bool INSIDE>
__device__ __forceinline__ float combine(float cur, float cand, const float *__restrict__ ftbl) {
  if (!INSIDE) return cand > cur ? cand : cur;
  // FLogsum (cp9.rs:5405-5413), branch-free: the table is read at index 0
  // when the smaller term is -inf or 23 or more below the larger
  const float mx = cur > cand ? cur : cand, mn = cur < cand ? cur : cand;
  const float diff = mx - mn;
  const bool near = !(mn == -INFINITY) && diff < 23.0f;
  const float tv = ftbl[near ? (int)(diff * 1000.0f) : 0];
  return near ? mx + tv : mx;
}

// The EL start value of a cell (branch-free: nothing waits on the endsc load)
__device__ __forceinline__ float cell_init(const CmDev &cm, int v, int st, int d) {
  const float es = gmem(cm.endsc)[v];
  const int sd = sdelta(st);
  const int dp = d - sd > 0 ? d - sd : 0;
  const float el = cm.el_selfsc * (float)dp + es;
  return es > -9.999e35f ? el : IMPOSSIBLE;
}

// A state's rectangle: band rows jlo .. jlo+nj-1, d = dlo .. dlo+nd-1,
// column x = j - jlo (mode 1) or i - xlo (mode 0) in 0 .. nx-1, cell (d, x)
// at alpha + base + (d - dlo) * nx + x. Words 0-7 are what a cell address
// needs (two 16-byte reads); word 9 packs the state type, child count and
// first child (a bifurcation's children are read from the model instead).
// box: the band is the (i, j) box i = xlo .. xlo+ni-1 clipped to the
// rectangle's d range, so a rectangle cell is a band cell exactly when its
// i is in that range (always true in mode 0).
struct Rect {
  int jlo, nj, xlo, nx, base, lvl, mode, box, dlo, nd, rowbase, st, cn, cf, ni;
};
__device__ __forceinline__ Rect rect_addr(const int32_t *rt, int v) {
  const int4 a = *reinterpret_cast<const int4 *>(rt + HB_RECT * v);
  const int4 b = *reinterpret_cast<const int4 *>(rt + HB_RECT * v + 4);
  Rect r;
  r.jlo = a.x, r.nj = a.y, r.xlo = a.z, r.nx = a.w;
  r.base = b.x, r.lvl = b.y >> 2, r.box = (b.y >> 1) & 1, r.mode = b.y & 1, r.dlo = b.z, r.nd = b.w;
  r.rowbase = 0, r.st = 0, r.cn = 0, r.cf = 0, r.ni = 0;
  return r;
}
__device__ __forceinline__ Rect rect_of(const int32_t *rt, int v) {
  Rect r = rect_addr(rt, v);
  const int4 c = *reinterpret_cast<const int4 *>(rt + HB_RECT * v + 8);
  r.rowbase = c.x;
  r.st = c.y & 15, r.cn = (c.y >> 4) & 15, r.cf = (int)((unsigned)c.y >> 8);
  r.ni = c.z;
  return r;
}
// Deck index of cell (j, d) of a state, or -1 outside its rectangle
// (branch-free).
__device__ __forceinline__ int rect_cell(const Rect &r, int j, int d) {
  const int jp = j - r.jlo, dp = d - r.dlo;
  const int x = r.mode ? jp : (j - d + 1) - r.xlo;
  const bool in = ((unsigned)jp < (unsigned)r.nj) & ((unsigned)dp < (unsigned)r.nd) & ((unsigned)x < (unsigned)r.nx);
  return in ? r.base + dp * r.nx + x : -1;
}

// One cell of a non-bifurcation state (S, D, ML, MR, MP, IL, IR, E; an
// insert's self transition reads its own cell at d-1): every operand is
// fetched first (the children's rectangles from LDS, then their cells, the
// transition, end and emission scores from memory), then the children are
// combined in child order. Called for every rectangle cell; the caller keeps
// the value only for band cells (every load stays in bounds). sq: the
// envelope's residues i0..j0.
template <bool INSIDE>
__device__ __forceinline__ float cell_generic(const CmDev &cm, const EnvDev &e, const int32_t *rt, const Rect &r,
                                              int v, int j, int d, const uint8_t *sq, const float *__restrict__ ftbl) {
  const int st = r.st;
  if (st == E_ST) return 0.0f;  // an end state's band is d = 0
  const int sd = sdelta(st), sdr = srdelta(st);
  const int cf = r.cf, cn = r.cn, last = cm.M - 1;
  // LDS first: the children's rectangles (a slot past cn reads a valid
  // record and is masked) and the residues; i and j are clamped into the
  // envelope (a cell outside the band is computed for nothing, its value
  // discarded, and may reach past the envelope)
  Rect ry[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) ry[yo] = rect_addr(rt, min(cf + yo, last));
  const int i = min(max(j - d + 1, e.i0), e.j0), jc = min(max(j, e.i0), e.j0);
  const int xl = sq[(st == MR_ST || st == IR_ST ? jc : i) - e.i0], xr = sq[jc - e.i0];
  __builtin_amdgcn_sched_barrier(0);
  // then every memory operand, all in flight together
  int c[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    const int cc = rect_cell(ry[yo], j - sdr, d - sd);
    c[yo] = yo < cn ? cc : -1;
  }
  const IXG_GLOBAL float *al = gmem(e.alpha);
  const IXG_GLOBAL float *tsc = gmem(cm.tsc) + v * MAXC;
  float a[MAXC], t[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    a[yo] = al[c[yo] >= 0 ? c[yo] : 0];
    t[yo] = tsc[yo];
  }
  const float es = gmem(cm.endsc)[v];
  const float em = gmem(cm.oesc)[(size_t)v * (KP * KP) + (st == MP_ST ? xl * KP + xr : xl)];
  __builtin_amdgcn_sched_barrier(0);
  // EL start (cell_init, on the loaded end score), then the children in order
  const int dpe = d - sd > 0 ? d - sd : 0;
  const float el = cm.el_selfsc * (float)dpe + es;
  float cur = es > -9.999e35f ? el : IMPOSSIBLE;
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (c[yo] >= 0) cur = combine<INSIDE>(cur, a[yo] + t[yo], ftbl);
  if (st == ML_ST || st == MR_ST || st == IL_ST || st == IR_ST || st == MP_ST) cur += em;
  return cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
}

// One cell of a bifurcation: left child y at (j-k, d-k), right child z at
// (j, k), k increasing over z's d range (terms outside either band read
// -inf, which max and FLogsum ignore exactly as the reference skips them).
template <bool INSIDE>
__device__ __forceinline__ float cell_bif(const CmDev &cm, const EnvDev &e, const int32_t *rt, int v, int j, int d,
                                          const float *__restrict__ ftbl) {
  float cur = cell_init(cm, v, B_ST, d);
  const int y = gmem(cm.cfirst)[v], z = gmem(cm.cnum)[v];
  const Rect ry = rect_addr(rt, y), rz = rect_addr(rt, z);
  if ((unsigned)(j - rz.jlo) >= (unsigned)rz.nj || rz.nd == 0 || ry.nj == 0) return cur;
  const int kn = max(rz.dlo, max(j - (ry.jlo + ry.nj - 1), 0));
  const int kx = min(min(rz.dlo + rz.nd - 1, j - ry.jlo), d);  // k > d: the left child below d = 0
  const IXG_GLOBAL float *al = gmem(e.alpha);
  constexpr int KB = 8;
  for (int k0 = kn; k0 <= kx; k0 += KB) {
    // a term with either cell outside its rectangle is -inf as a whole: the
    // placeholder load (cell 0, which may not be written yet in this scan)
    // never enters the sum
    // the batch's 2 KB loads all issued before any is used
    float av[KB], bv[KB];
    bool ok[KB];
#pragma unroll
    for (int u = 0; u < KB; u++) {
      const int k = k0 + u;
      const int cy = rect_cell(ry, j - k, d - k), cz = rect_cell(rz, j, k);
      ok[u] = (k <= kx) & (cy >= 0) & (cz >= 0);
      av[u] = al[ok[u] ? cy : 0];
      bv[u] = al[ok[u] ? cz : 0];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < KB; u++) cur = combine<INSIDE>(cur, ok[u] ? av[u] + bv[u] : -INFINITY, ftbl);
  }
  return cur;
}

// One rectangle cell of state v at (j, d): its band cell or -inf, and its
// deck index.
template <bool INSIDE>
__device__ __forceinline__ float hb_cell(const CmDev &cm, const EnvDev &e, const int32_t *rt, const Rect &r, int v,
                                         int j, int d, const uint8_t *sq, const float *ftbl, int &idx) {
  const int jp = j - r.jlo;
  const int x = r.mode ? jp : (j - d + 1) - r.xlo;
  bool in;
  if (r.box) {
    in = r.mode ? (unsigned)((j - d + 1) - r.xlo) < (unsigned)r.ni : true;
  } else {
    const IXG_GLOBAL int32_t *rb = gmem(e.rowx) + 2 * (r.rowbase + jp);
    in = d >= rb[0] && d <= rb[1];  // the row's d band
  }
  idx = r.base + (d - r.dlo) * r.nx + x;
  if (r.st == B_ST) return in ? cell_bif<INSIDE>(cm, e, rt, v, j, d, ftbl) : -INFINITY;
  const float cur = cell_generic<INSIDE>(cm, e, rt, r, v, j, d, sq, ftbl);  // loads overlap the band's
  return in ? cur : -INFINITY;
}

// One envelope, scanned by the whole workgroup: a wavefront over d (step t:
// state v fills its rectangle cells of length d = t - level(v)). Each step
// the states' cell counts are prefix-summed in LDS (pre[v], first row
// jst[v]) and every lane takes every nt-th cell of the step (one cell per
// lane at a time keeps the most cells in flight). rt / bl / bs: the
// rectangle table (HB_RECT ints per state), the begin states and their begin
// scores; ftbl: the FLogsum table.
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_body(const CmDev &cm, const EnvDev &e, const int32_t *rt, const int32_t *bl,
                                             const float *bs, int nb, const float *ftbl, int32_t *pre, int32_t *jst,
                                             int32_t *wsum, const uint8_t *sq) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  const int M = cm.M;
  const int per = (M + nt - 1) / nt, v0 = tid * per;
  IXG_GLOBAL float *al = gmem(e.alpha);
  for (int t = 0; t < e.nsteps; t++) {
    // this thread's states: first row and cell count at this step
    int my = 0;
    for (int q = 0; q < per; q++) {
      const int v = v0 + q;
      if (v >= M) break;
      const Rect r = rect_addr(rt, v);
      const int d = t - r.lvl;
      int n = 0, ja = 0;
      if ((unsigned)(d - r.dlo) < (unsigned)r.nd) {
        if (r.mode) {
          ja = r.jlo, n = r.nj;
        } else {
          ja = max(r.jlo, d - 1 + r.xlo);
          n = max(0, min(r.jlo + r.nj - 1, d - 2 + r.xlo + r.nx) - ja + 1);
        }
      }
      jst[v] = ja;
      pre[v] = my;
      my += n;
    }
    // block exclusive scan of the threads' counts
    int inc = my;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int off = inc - my, total = 0;
    for (int w = 0; w < nw; w++) {
      const int s = wsum[w];
      if (w < wv) off += s;
      total += s;
    }
    for (int q = 0; q < per; q++) {
      const int v = v0 + q;
      if (v >= M) break;
      pre[v] += off;
    }
    __syncthreads();
    // the step's cells, every nt-th to a lane
    for (int c = tid; c < total; c += nt) {
      int lo = 0, hi = M - 1;  // the last state whose first cell is at or before c
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= c) lo = mid;
        else hi = mid - 1;
      }
      const Rect r = rect_of(rt, lo);
      int ix;
      const float val = hb_cell<INSIDE>(cm, e, rt, r, lo, jst[lo] + (c - pre[lo]), t - r.lvl, sq, ftbl, ix);
      al[ix] = val;
    }
    __syncthreads();
  }
  // local begins into the root by MAX (cp9.rs:5264-5304, 5908-5937), begin
  // state by begin state in list order (the reference's strict comparison
  // per cell, so ties resolve as there): all lanes stream through the
  // state's rectangle, a cell outside either band contributing nothing (the
  // root's own out-of-band cells are reset below)
  const Rect r0 = rect_of(rt, 0);
  for (int b = 0; b < nb; b++) {
    const int y = bl[b];
    const float sc = bs[b];
    const Rect ry = rect_addr(rt, y);
    const int n = ry.nd * ry.nx;
    for (int c = tid; c < n; c += nt) {
      const int dp = c / ry.nx, x = c - dp * ry.nx, d = ry.dlo + dp;
      const int j = ry.mode ? ry.jlo + x : x + ry.xlo + d - 1;
      if ((unsigned)(j - ry.jlo) >= (unsigned)ry.nj) continue;  // never written
      const int c0 = rect_cell(r0, j, d);
      if (c0 < 0) continue;
      const float sv = al[ry.base + c] + sc;
      if (sv > al[c0]) al[c0] = sv;
    }
    __syncthreads();
  }
  // the root rows out (increasing d per row), and -inf back into the root's
  // out-of-band cells
  const int n0 = r0.nd * r0.nx;
  for (int c = tid; c < n0; c += nt) {
    const int dp = c / r0.nx, x = c - dp * r0.nx, d = r0.dlo + dp;
    const int j = r0.mode ? r0.jlo + x : x + r0.xlo + d - 1, jp = j - r0.jlo;
    if ((unsigned)jp >= (unsigned)r0.nj) continue;
    const IXG_GLOBAL int32_t *rb = gmem(e.rowx) + 2 * (r0.rowbase + jp);
    const int lo = rb[0], hi = rb[1];
    if (d >= lo && d <= hi) gmem(e.rout)[e.rootoff[jp] + (d - lo)] = al[r0.base + c];
    else if (nb) al[r0.base + c] = -INFINITY;
  }
}

// Scans one envelope with its rectangle table and begin list in LDS (RL) or
// read from memory, and the Inside FLogsum table in LDS (TL) or memory.
// LDS: [FLogsum table (Inside, TL)] [pre (M), jst (M), wsum (32), padded]
// [residues i0..j0, padded] [rectangle table (HB_RECT * M), begin states
// (nb), begin scores (nb)] (RL); the choice is a template parameter so that
// every access compiles to LDS or global instructions.
template <bool INSIDE, bool RL, bool TL>
__device__ __forceinline__ void hb_scan_staged(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem) {
  const int M = cm.M, nb = e.nbegin;
  int32_t *pre = smem + (TL ? FLOGSUM_TBL : 0), *jst = pre + M, *wsum = jst + M;
  int32_t *sqi = pre + hb_scan_ints(M);
  const int32_t *base = RL ? sqi + hb_seq_ints(e.j0 - e.i0 + 1) : (const int32_t *)gmem(e.rect);
  const int32_t *bl = base + HB_RECT * M;
  const float *bs = reinterpret_cast<const float *>(bl + nb);
  const float *tb = TL ? reinterpret_cast<const float *>(smem) : (const float *)gmem(ftbl);
  hb_scan_body<INSIDE>(cm, e, base, bl, bs, nb, tb, pre, jst, wsum, reinterpret_cast<const uint8_t *>(sqi));
}

// Stages what fits in LDS and scans the envelope; the host sizes lds_ints
// for the step scan and the residues at least (hb_scan_ints(M) +
// hb_seq_ints(j0 - i0 + 1)).
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_env(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem,
                                            int lds_ints) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int M = cm.M, nb = e.nbegin, len = e.j0 - e.i0 + 1;
  const int scan = hb_scan_ints(M) + hb_seq_ints(len), need = HB_RECT * M + 2 * nb;
  const bool rl = scan + need <= lds_ints, tl = INSIDE && rl && FLOGSUM_TBL + scan + need <= lds_ints;
  if (tl)
    for (int q = tid; q < FLOGSUM_TBL; q += nt) smem[q] = __float_as_int(ftbl[q]);
  int32_t *sqi = smem + (tl ? FLOGSUM_TBL : 0) + hb_scan_ints(M);
  uint8_t *sq = reinterpret_cast<uint8_t *>(sqi);
  const IXG_GLOBAL uint8_t *dsq = gmem(e.dsq);
  for (int q = tid; q < len; q += nt) sq[q] = dsq[e.i0 + q];
  if (rl) {
    int32_t *at = smem + (tl ? FLOGSUM_TBL : 0) + scan;
    for (int q = tid; q < need; q += nt) at[q] = e.rect[q];
  }
  __syncthreads();
  if (tl) hb_scan_staged<INSIDE, true, INSIDE>(cm, e, ftbl, smem);
  else if (rl) hb_scan_staged<INSIDE, true, false>(cm, e, ftbl, smem);
  else hb_scan_staged<INSIDE, false, false>(cm, e, ftbl, smem);
}

// amdgpu_waves_per_eu: a workgroup of HB_THREADS lanes fills a CU at 4 waves
// per SIMD; telling the compiler so lets its scheduler keep loads in flight
// instead of trading them for an occupancy the launch cannot use.
template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void hb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl,
                                                             int lds_ints) {
  extern __shared__ int32_t smem[];
  const EnvDev e = envs[blockIdx.x];  // by value: its fields stay in registers
  hb_scan_env<INSIDE>(cm, e, ftbl, smem, lds_ints);
}

// Persistent form over envelopes of several models: each workgroup claims
// the next envelope of `order` (largest first) from a device counter until
// the list is exhausted; an envelope's model is cms[e.cmi]. Every workgroup
// leaves once a claim passes n, so the grid drains whatever its size.
template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void hb_scan_queue_kernel(const CmDev *__restrict__ cms,
                                                                   const EnvDev *__restrict__ envs,
                                                                   const int32_t *__restrict__ order, int32_t n,
                                                                   uint32_t *counter, const float *ftbl,
                                                                   int lds_ints) {
  extern __shared__ int32_t smem[];
  __shared__ int32_t claim;
  for (;;) {
    if (threadIdx.x == 0) claim = (int32_t)atomicAdd(counter, 1u);
    __syncthreads();
    const int32_t k = __builtin_amdgcn_readfirstlane(claim);
    __syncthreads();  // every lane has read the claim before the next one is written
    if (k >= n) break;
    const int32_t ek = order[k];
    const EnvDev e = envs[ek];
    const CmDev cm = cms[e.cmi];
    hb_scan_env<INSIDE>(cm, e, ftbl, smem, lds_ints);
    __syncthreads();  // the next envelope's LDS staging waits for this one's last reads
  }
}

}  // namespace

hipError_t launch_hb_scan_queue(const CmDev *cms, const EnvDev *envs, const int32_t *order, int64_t n,
                                uint32_t *counter, int grid, const float *ftbl, int inside, size_t lds_ints,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t lds = lds_ints * sizeof(int32_t);
  hipError_t r = hipMemsetAsync(counter, 0, sizeof(uint32_t), s);
  if (r != hipSuccess) return r;
  const unsigned g = (unsigned)std::min<int64_t>(n, grid);
  if (inside)
    hipLaunchKernelGGL(hb_scan_queue_kernel<true>, dim3(g), dim3(HB_THREADS), lds, s, cms, envs, order, (int32_t)n,
                       counter, ftbl, (int)lds_ints);
  else
    hipLaunchKernelGGL(hb_scan_queue_kernel<false>, dim3(g), dim3(HB_THREADS), lds, s, cms, envs, order,
                       (int32_t)n, counter, ftbl, (int)lds_ints);
  return hipGetLastError();
}

hipError_t launch_hb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, int inside,
                          size_t lds_ints, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t lds = lds_ints * sizeof(int32_t);
  if (inside)
    hipLaunchKernelGGL(hb_scan_kernel<true>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl,
                       (int)lds_ints);
  else
    hipLaunchKernelGGL(hb_scan_kernel<false>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl,
                       (int)lds_ints);
  return hipGetLastError();
}

}  // namespace ixg
