// gfx950 HMM-banded CYK / Inside scanners (F6/F7), one workgroup of 1024
// lanes per envelope, as a wavefront over end positions j:
//
// * Deck: each state's band cells sit in a dense rectangle over (j, x),
//   x = i - xlo (or d - xlo for states whose band is narrower in d, such as
//   end states), so a child cell's address needs only the child's rectangle
//   (in LDS), never a per-row lookup. Rectangle cells outside the true band
//   hold -inf, the identity of max and of FLogsum, so reading one is the
//   same as the reference skipping the out-of-band child. Every rectangle
//   cell with 0 <= d <= the envelope's largest d is written; no read asks
//   for another d.
// * Schedule: a wavefront over the subsequence length d. At step t state v
//   fills its cells of length d = t - level(v) (all j at once), where a
//   state's level is one more than its children's highest. Every child cell
//   is then at a smaller d (ML, MR, MP, IL, IR, including the inserts' self
//   transitions, and a bifurcation's splits with 0 < k < d) or at the same d
//   in a lower level (S, D, and the k = 0 / k = d splits), so it was
//   finished at an earlier step. One barrier per step, no serial chain
//   inside a step except a bifurcation cell's splits.
// * Lanes: each step the states' cell counts at their d are prefix-summed
//   in LDS and the step's cells are dealt round-robin over all lanes.
//
// Per cell the children are combined in increasing child order (max for
// CYK, the FLogsum table for Inside) and a bifurcation's splits in
// increasing k, the order of the reference's loops (cp9.rs:4788-5253,
// 5514-5885), so CYK is bit-identical and Inside applies the same float
// operations in the same order. Local begins are folded into each root cell
// by MAX as it is computed (cp9.rs:5264-5304, 5908-5937); root cells also go
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

__device__ __forceinline__ float cell_init(const CmDev &cm, int v, int st, int d) {
  const float es = gmem(cm.endsc)[v];
  if (es > -9.999e35f) {
    const int sd = sdelta(st);
    const int dp = d - sd > 0 ? d - sd : 0;
    return cm.el_selfsc * (float)dp + es;
  }
  return IMPOSSIBLE;
}

// A state's rectangle: rows jlo .. jlo+nj-1, columns x = 0 .. nx-1 (x = i -
// xlo, or d - xlo in d mode), row-major at alpha + base.
// The last word packs the state type, child count and first child (a
// bifurcation's children are read from the model instead).
struct Rect {
  int jlo, nj, xlo, nx, base, lvl, mode, rowbase, st, cn, cf;
};
__device__ __forceinline__ Rect rect_of(const int32_t *rt, int v) {
  const int4 a = *reinterpret_cast<const int4 *>(rt + 8 * v);
  const int4 b = *reinterpret_cast<const int4 *>(rt + 8 * v + 4);
  Rect r;
  r.jlo = a.x, r.nj = a.y, r.xlo = a.z, r.nx = a.w;
  r.base = b.x, r.lvl = b.y >> 1, r.mode = b.y & 1, r.rowbase = b.z;
  r.st = b.w & 15, r.cn = (b.w >> 4) & 15, r.cf = (int)((unsigned)b.w >> 8);
  return r;
}
// Deck index of cell (j, d) of a state, or -1 outside its rectangle.
__device__ __forceinline__ int rect_cell(const Rect &r, int j, int d) {
  const int jp = j - r.jlo;
  const int x = r.mode ? d - r.xlo : (j - d + 1) - r.xlo;
  return ((unsigned)jp < (unsigned)r.nj && (unsigned)x < (unsigned)r.nx) ? r.base + jp * r.nx + x : -1;
}
__device__ __forceinline__ int rect_cell(const int32_t *rt, int y, int j, int d) { return rect_cell(rect_of(rt, y), j, d); }

// One cell of a non-bifurcation state (S, D, ML, MR, MP, IL, IR, E; an
// insert's self transition reads its own cell at d-1): children fetched
// first, then combined in child order. Called for every rectangle cell; the
// caller keeps the value only for band cells (every load stays in bounds).
template <bool INSIDE>
__device__ __forceinline__ float cell_generic(const CmDev &cm, const EnvDev &e, const int32_t *rt, const Rect &r,
                                              int v, int j, int d, const float *__restrict__ ftbl) {
  const int st = r.st;
  if (st == E_ST) return 0.0f;  // an end state's band is d = 0
  float cur = cell_init(cm, v, st, d);
  const int sd = sdelta(st), sdr = srdelta(st);
  const int cf = r.cf, cn = r.cn;
  const IXG_GLOBAL float *tsc = gmem(cm.tsc) + v * MAXC;
  const IXG_GLOBAL float *al = gmem(e.alpha);
  float a[MAXC];
  bool ok[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    const int c = yo < cn ? rect_cell(rt, cf + yo, j - sdr, d - sd) : -1;
    ok[yo] = c >= 0;
    a[yo] = al[ok[yo] ? c : 0];
  }
  const IXG_GLOBAL uint8_t *dsq = gmem(e.dsq);
  // i clamped into the envelope: a cell outside the band is computed for
  // nothing (its value is discarded) and may have i < 1
  const int i = max(j - d + 1, e.i0);
  const int xl = dsq[st == MR_ST || st == IR_ST ? j : i], xr = dsq[j];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (ok[yo]) cur = combine<INSIDE>(cur, a[yo] + tsc[yo], ftbl);
  const IXG_GLOBAL float *oe = gmem(cm.oesc) + (size_t)v * (KP * KP);
  if (st == ML_ST || st == MR_ST || st == IR_ST) cur += oe[xl];
  else if (st == MP_ST) cur += oe[xl * KP + xr];
  return cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
}

// One cell of a bifurcation: left child y at (j-k, d-k), right child z at
// (j, k), k increasing over z's rectangle at row j (terms outside either
// band read -inf).
template <bool INSIDE>
__device__ __forceinline__ float cell_bif(const CmDev &cm, const EnvDev &e, const int32_t *rt, int v, int j, int d,
                                          const float *__restrict__ ftbl) {
  float cur = cell_init(cm, v, B_ST, d);
  const int y = gmem(cm.cfirst)[v], z = gmem(cm.cnum)[v];
  const Rect ry = rect_of(rt, y), rz = rect_of(rt, z);
  const int jpz = j - rz.jlo;
  if ((unsigned)jpz >= (unsigned)rz.nj || rz.nx == 0 || ry.nj == 0) return cur;
  // z's d range at row j, and j - k within y's rows
  int kn = rz.mode ? rz.xlo : j - (rz.xlo + rz.nx - 1) + 1;
  int kx = rz.mode ? rz.xlo + rz.nx - 1 : j - rz.xlo + 1;
  kn = max(kn, max(j - (ry.jlo + ry.nj - 1), 0));
  kx = min(min(kx, j - ry.jlo), d);  // k > d puts the left child below d = 0, outside every band
  const IXG_GLOBAL float *al = gmem(e.alpha);
  const int zrow = rz.base + jpz * rz.nx;
  constexpr int KB = 8;
  for (int k0 = kn; k0 <= kx; k0 += KB) {
    // a term whose left cell is outside y's rectangle (or past kx) is -inf,
    // which max and FLogsum ignore exactly as the reference skips it
    float a[KB], b[KB];
#pragma unroll
    for (int u = 0; u < KB; u++) {
      const int k = k0 + u, kc = k <= kx ? k : kx;
      const int cy = k <= kx ? rect_cell(ry, j - k, d - k) : -1;
      const int xz = rz.mode ? kc - rz.xlo : (j - kc + 1) - rz.xlo;
      const float av = al[cy >= 0 ? cy : 0];
      a[u] = cy >= 0 ? av : -INFINITY;
      b[u] = al[zrow + xz];
    }
#pragma unroll
    for (int u = 0; u < KB; u++) cur = combine<INSIDE>(cur, a[u] + b[u], ftbl);
  }
  return cur;
}

// One rectangle cell of state v at (j, d): its band cell or -inf, and its
// deck index.
template <bool INSIDE>
__device__ __forceinline__ float hb_cell(const CmDev &cm, const EnvDev &e, const int32_t *rt, const Rect &r, int v,
                                         int j, int d, const float *ftbl, int &idx) {
  const int jp = j - r.jlo;
  const int x = r.mode ? d - r.xlo : (j - d + 1) - r.xlo;
  const IXG_GLOBAL int32_t *rx = gmem(e.rowx) + 2 * (r.rowbase + jp);
  const int xa = rx[0], xb = rx[1];
  const bool in = x >= xa && x <= xb;
  idx = r.base + jp * r.nx + x;
  if (r.st == B_ST) return in ? cell_bif<INSIDE>(cm, e, rt, v, j, d, ftbl) : -INFINITY;
  const float cur = cell_generic<INSIDE>(cm, e, rt, r, v, j, d, ftbl);  // loads overlap the band range's
  return in ? cur : -INFINITY;
}

// One envelope, scanned by the whole workgroup: a wavefront over d (step t:
// state v fills its rectangle cells of length d = t - level(v)). Each step
// the states' cell counts are prefix-summed in LDS (pre[v], first row
// jst[v]) and every lane takes every nt-th cell of the step (one cell per
// lane at a time keeps the most cells in flight). rt / bl / bs:
// the rectangle table (8 ints per state), the begin states and their begin
// scores; ftbl: the FLogsum table.
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_body(const CmDev &cm, const EnvDev &e, const int32_t *rt, const int32_t *bl,
                                             const float *bs, int nb, const float *ftbl, int32_t *pre, int32_t *jst,
                                             int32_t *wsum) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  const int M = cm.M;
  const int per = (M + nt - 1) / nt, v0 = tid * per;
  for (int t = 0; t < e.nsteps; t++) {
    // this thread's states: first row and cell count at this step
    int my = 0;
    for (int q = 0; q < per; q++) {
      const int v = v0 + q;
      if (v >= M) break;
      const Rect r = rect_of(rt, v);
      const int d = t - r.lvl;
      int n = 0, ja = 0;
      if (d >= 0 && r.nx > 0) {
        if (r.mode) {
          if ((unsigned)(d - r.xlo) < (unsigned)r.nx) ja = r.jlo, n = r.nj;
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
    IXG_GLOBAL float *al = gmem(e.alpha);
    for (int c = tid; c < total; c += nt) {
      int lo = 0, hi = M - 1;  // the last state whose first cell is at or before c
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= c) lo = mid;
        else hi = mid - 1;
      }
      const Rect r = rect_of(rt, lo);
      int ix;
      const float val = hb_cell<INSIDE>(cm, e, rt, r, lo, jst[lo] + (c - pre[lo]), t - r.lvl, ftbl, ix);
      al[ix] = val;
    }
    __syncthreads();
  }
  // local begins into the root by MAX (cp9.rs:5264-5304, 5908-5937), begin
  // state by begin state in list order (the reference's strict comparison
  // per cell, so ties resolve as there): all lanes stream through the
  // state's rectangle rows that overlap the root's rows; band and rectangle
  // checks on both sides (a cell outside either band contributes nothing)
  IXG_GLOBAL float *al = gmem(e.alpha);
  const Rect r0 = rect_of(rt, 0);
  for (int b = 0; b < nb; b++) {
    const int y = bl[b];
    const float sc = bs[b];
    const Rect ry = rect_of(rt, y);
    const int ja = max(ry.jlo, r0.jlo), jb = min(ry.jlo + ry.nj, r0.jlo + r0.nj) - 1;
    const int n = jb >= ja ? (jb - ja + 1) * ry.nx : 0;
    for (int c = tid; c < n; c += nt) {
      const int jr = c / ry.nx, xy = c - jr * ry.nx, j = ja + jr;
      const int d = ry.mode ? ry.xlo + xy : j - ry.xlo - xy + 1;
      const int c0 = rect_cell(r0, j, d);
      if (c0 < 0) continue;
      const float a = al[ry.base + (j - ry.jlo) * ry.nx + xy];  // -inf outside y's band
      const float sv = a + sc;
      if (sv > al[c0]) al[c0] = sv;  // -inf root cells (outside its band) are rewritten below
    }
    __syncthreads();
  }
  // the root rows out, and -inf back into the root's out-of-band cells
  const int n0 = r0.nj * r0.nx;
  for (int c = tid; c < n0; c += nt) {
    const int jp = c / r0.nx, x = c - jp * r0.nx;
    const IXG_GLOBAL int32_t *rx = gmem(e.rowx) + 2 * (r0.rowbase + jp);
    const int xa = rx[0], xb = rx[1];
    if (x >= xa && x <= xb) gmem(e.rout)[e.rootoff[jp] + (r0.mode ? x - xa : xb - x)] = al[r0.base + c];
    else if (nb) al[r0.base + c] = -INFINITY;
  }
}

// Scans one envelope with its rectangle table and begin list in LDS (RL) or
// read from memory, and the Inside FLogsum table in LDS (TL) or memory.
// LDS: [FLogsum table (Inside, TL)] [pre (M), jst (M), wsum (32)]
// [rectangle table (8M), begin states (nb), begin scores (nb)] (RL); the
// choice is a template parameter so that every access compiles to LDS or
// global instructions.
template <bool INSIDE, bool RL, bool TL>
__device__ __forceinline__ void hb_scan_staged(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem) {
  const int M = cm.M, nb = e.nbegin;
  int32_t *pre = smem + (TL ? FLOGSUM_TBL : 0), *jst = pre + M, *wsum = jst + M;
  const int32_t *base = RL ? wsum + 32 : (const int32_t *)gmem(e.rect);
  const int32_t *bl = base + 8 * M;
  const float *bs = reinterpret_cast<const float *>(bl + nb);
  const float *tb = TL ? reinterpret_cast<const float *>(smem) : (const float *)gmem(ftbl);
  hb_scan_body<INSIDE>(cm, e, base, bl, bs, nb, tb, pre, jst, wsum);
}

// Stages what fits in LDS and scans the envelope; the host sizes lds_ints
// for the step scan at least (2M + 32 ints).
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_env(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem,
                                            int lds_ints) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int M = cm.M, nb = e.nbegin;
  const int scan = 2 * M + 32, need = 8 * M + 2 * nb;
  const bool rl = scan + need <= lds_ints, tl = INSIDE && rl && FLOGSUM_TBL + scan + need <= lds_ints;
  if (tl)
    for (int q = tid; q < FLOGSUM_TBL; q += nt) smem[q] = __float_as_int(ftbl[q]);
  if (rl) {
    int32_t *at = smem + (tl ? FLOGSUM_TBL : 0) + scan;
    for (int q = tid; q < need; q += nt) at[q] = e.rect[q];
  }
  __syncthreads();
  if (tl) hb_scan_staged<INSIDE, true, INSIDE>(cm, e, ftbl, smem);
  else if (rl) hb_scan_staged<INSIDE, true, false>(cm, e, ftbl, smem);
  else hb_scan_staged<INSIDE, false, false>(cm, e, ftbl, smem);
}

template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) void hb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl,
                                                             int lds_ints) {
  extern __shared__ int32_t smem[];
  hb_scan_env<INSIDE>(cm, envs[blockIdx.x], ftbl, smem, lds_ints);
}

// Persistent form over envelopes of several models: each workgroup claims
// the next envelope of `order` (largest first) from a device counter until
// the list is exhausted; an envelope's model is cms[e.cmi]. Every workgroup
// leaves once a claim passes n, so the grid drains whatever its size.
template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) void hb_scan_queue_kernel(const CmDev *__restrict__ cms,
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
