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
// * Lanes: the host deals each state one or more lane slots for the
//   envelope in proportion to its cells per step (a bifurcation weighted by
//   its splits), grouped by kind so that a wave runs one code path. A slot
//   with nsub lanes takes every nsub-th cell of its state's d diagonal.
//
// Per cell the children are combined in increasing child order (max for
// CYK, the FLogsum table for Inside) and a bifurcation's splits in
// increasing k, the order of the reference's loops (cp9.rs:4788-5253,
// 5514-5885), so CYK is bit-identical and Inside applies the same float
// operations in the same order. Local begins are folded into the root by
// MAX as each root row is finished (cp9.rs:5264-5304, 5908-5937); root rows
// also go to the envelope's compact root output.
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
struct Rect {
  int jlo, nj, xlo, nx, base, lvl, mode, rowbase, st;
};
__device__ __forceinline__ Rect rect_of(const int32_t *rt, int v) {
  const int4 a = *reinterpret_cast<const int4 *>(rt + 8 * v);
  const int4 b = *reinterpret_cast<const int4 *>(rt + 8 * v + 4);
  Rect r;
  r.jlo = a.x, r.nj = a.y, r.xlo = a.z, r.nx = a.w;
  r.base = b.x, r.lvl = b.y >> 1, r.mode = b.y & 1, r.rowbase = b.z, r.st = b.w;
  return r;
}
// Deck index of cell (j, d) of a state, or -1 outside its rectangle.
__device__ __forceinline__ int rect_cell(const Rect &r, int j, int d) {
  const int jp = j - r.jlo;
  const int x = r.mode ? d - r.xlo : (j - d + 1) - r.xlo;
  return ((unsigned)jp < (unsigned)r.nj && (unsigned)x < (unsigned)r.nx) ? r.base + jp * r.nx + x : -1;
}
__device__ __forceinline__ int rect_cell(const int32_t *rt, int y, int j, int d) { return rect_cell(rect_of(rt, y), j, d); }

// d of column x of row j.
__device__ __forceinline__ int d_of(const Rect &r, int j, int x) { return r.mode ? r.xlo + x : j - r.xlo - x + 1; }

// One cell of a non-insert, non-bifurcation state (S, D, ML, MR, MP, E) or
// of an IR state (its self transition reads its own row j-1): children
// fetched first, then combined in child order.
template <bool INSIDE>
__device__ __forceinline__ float cell_generic(const CmDev &cm, const EnvDev &e, const int32_t *rt, int v, int st,
                                              int j, int d, const float *__restrict__ ftbl) {
  if (st == E_ST) return 0.0f;  // an end state's band is d = 0
  float cur = cell_init(cm, v, st, d);
  const int sd = sdelta(st), sdr = srdelta(st);
  const int cf = gmem(cm.cfirst)[v], cn = gmem(cm.cnum)[v];
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
  const int i = j - d + 1;
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
    float a[KB], b[KB];
    bool ok[KB];
#pragma unroll
    for (int u = 0; u < KB; u++) {
      const int k = k0 + u;
      const int cy = k <= kx ? rect_cell(ry, j - k, d - k) : -1;
      const int xz = rz.mode ? k - rz.xlo : (j - k + 1) - rz.xlo;
      ok[u] = cy >= 0;
      a[u] = al[ok[u] ? cy : 0];
      b[u] = al[ok[u] ? zrow + xz : 0];
    }
#pragma unroll
    for (int u = 0; u < KB; u++)
      if (ok[u]) cur = combine<INSIDE>(cur, a[u] + b[u], ftbl);
  }
  return cur;
}

__device__ __forceinline__ void slot_decode(int32_t code, int &v, int &sub, int &nsub, int &kind) {
  // state (bits 0-15), sub (16-21), nsub - 1 (22-27), kind (28-29)
  v = code & 0xFFFF, sub = (code >> 16) & 63, nsub = ((code >> 22) & 63) + 1, kind = (code >> 28) & 3;
}

// The slot's share of the cells of state v at subsequence length d: every
// rectangle cell with that d, band cells computed, the others set to -inf.
template <bool INSIDE>
__device__ __forceinline__ void hb_diag(const CmDev &cm, const EnvDev &e, const int32_t *rt, const Rect &r, int v,
                                        int kind, int sub, int nsub, int d, const float *ftbl) {
  // rows whose rectangle holds a cell with this d
  int ja = r.jlo, jb = r.jlo + r.nj - 1;
  if (r.mode) {
    if ((unsigned)(d - r.xlo) >= (unsigned)r.nx) return;
  } else {
    ja = max(ja, d - 1 + r.xlo);
    jb = min(jb, d - 2 + r.xlo + r.nx);
  }
  IXG_GLOBAL float *al = gmem(e.alpha);
  const IXG_GLOBAL int32_t *rowx = gmem(e.rowx) + 2 * r.rowbase;
  for (int j = ja + sub; j <= jb; j += nsub) {
    const int jp = j - r.jlo;
    const int x = r.mode ? d - r.xlo : (j - d + 1) - r.xlo;
    const int xa = rowx[2 * jp], xb = rowx[2 * jp + 1];
    float cur = -INFINITY;
    if (x >= xa && x <= xb)
      cur = kind == HB_SLOT_BIF ? cell_bif<INSIDE>(cm, e, rt, v, j, d, ftbl)
                                : cell_generic<INSIDE>(cm, e, rt, v, r.st, j, d, ftbl);
    al[r.base + jp * r.nx + x] = cur;
  }
}

// One envelope, scanned by the whole workgroup. rt / sl / bl / bs: the
// rectangle table (8 ints per state), the lane slots, the begin states and
// their begin scores; ftbl: the FLogsum table. At step t state v fills its
// cells of length d = t - level(v).
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_body(const CmDev &cm, const EnvDev &e, const int32_t *rt, const int32_t *sl,
                                             const int32_t *bl, const float *bs, int nb, const float *ftbl) {
  const int tid = threadIdx.x, nt = blockDim.x;
  IXG_GLOBAL float *al = gmem(e.alpha);
  auto row_range = [&](const Rect &r, int jp, int &xa, int &xb) {
    const IXG_GLOBAL int32_t *rx = gmem(e.rowx) + 2 * (r.rowbase + jp);
    xa = rx[0];
    xb = rx[1];
  };
  if (e.nslots <= nt) {
    int v = 0, sub = 0, nsub = 1, kind = 0;
    Rect r{};
    const bool mine = tid < e.nslots;
    if (mine) {
      slot_decode(sl[tid], v, sub, nsub, kind);
      r = rect_of(rt, v);
    }
    for (int t = 0; t < e.nsteps; t++) {
      const int d = t - r.lvl;
      if (mine && d >= 0) hb_diag<INSIDE>(cm, e, rt, r, v, kind, sub, nsub, d, ftbl);
      __syncthreads();
    }
  } else {
    for (int t = 0; t < e.nsteps; t++) {
      for (int s = tid; s < e.nslots; s += nt) {
        int v, sub, nsub, kind;
        slot_decode(sl[s], v, sub, nsub, kind);
        const Rect r = rect_of(rt, v);
        const int d = t - r.lvl;
        if (d >= 0) hb_diag<INSIDE>(cm, e, rt, r, v, kind, sub, nsub, d, ftbl);
      }
      __syncthreads();
    }
  }
  // root: local begins by MAX in begin-list order (cp9.rs:5264-5304,
  // 5908-5937), then the compact root rows (increasing d per row). Each lane
  // owns RB root cells at a time; their begin terms are fetched together.
  const Rect r0 = rect_of(rt, 0);
  const int n0 = r0.nj * r0.nx;
  constexpr int RB = 4;
  for (int c0 = 0; c0 < n0; c0 += RB * nt) {
    float cur[RB];
    int jq[RB], dq[RB], oq[RB];
#pragma unroll
    for (int q = 0; q < RB; q++) {
      const int c = c0 + q * nt + tid;
      cur[q] = -INFINITY;
      oq[q] = -1;
      if (c < n0) {
        const int jp = c / r0.nx, x = c - jp * r0.nx;
        int xa, xb;
        row_range(r0, jp, xa, xb);
        jq[q] = r0.jlo + jp;
        dq[q] = d_of(r0, jq[q], x);
        if (x >= xa && x <= xb) {
          cur[q] = al[r0.base + c];
          oq[q] = e.rootoff[jp] + (r0.mode ? x - xa : xb - x);
        }
      }
    }
    for (int b = 0; b < nb; b++) {
      const int y = bl[b];
      const float sc = bs[b];
      const Rect ry = rect_of(rt, y);
      float a[RB];
      int ci[RB];
#pragma unroll
      for (int q = 0; q < RB; q++) {
        ci[q] = oq[q] >= 0 ? rect_cell(ry, jq[q], dq[q]) : -1;
        a[q] = al[ci[q] >= 0 ? ci[q] : 0];
      }
#pragma unroll
      for (int q = 0; q < RB; q++)
        if (ci[q] >= 0) {
          const float sv = a[q] + sc;
          if (sv > cur[q]) cur[q] = sv;
        }
    }
#pragma unroll
    for (int q = 0; q < RB; q++)
      if (oq[q] >= 0) {
        al[r0.base + c0 + q * nt + tid] = cur[q];
        gmem(e.rout)[oq[q]] = cur[q];
      }
  }
}

// Scans one envelope with its rectangle table, slots and begin list in LDS
// (RL) or read from memory, and the Inside FLogsum table in LDS (TL) or
// memory; the choice is a template parameter so that every access compiles
// to LDS or global instructions.
template <bool INSIDE, bool RL, bool TL>
__device__ __forceinline__ void hb_scan_staged(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem) {
  const int M = cm.M, nb = e.nbegin;
  const int32_t *base = RL ? smem + (TL ? FLOGSUM_TBL : 0) : (const int32_t *)gmem(e.rect);
  const int32_t *sl = base + 8 * M, *bl = sl + e.nslots;
  const float *bs = reinterpret_cast<const float *>(bl + nb);
  const float *tb = TL ? reinterpret_cast<const float *>(smem) : (const float *)gmem(ftbl);
  hb_scan_body<INSIDE>(cm, e, base, sl, bl, bs, nb, tb);
}

// Stages what fits in LDS (the FLogsum table for Inside, then the envelope's
// rectangle table, slots, begin list and begin scores, which the host packs
// contiguously at e.rect) and scans the envelope.
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_env(const CmDev &cm, const EnvDev &e, const float *ftbl, int32_t *smem,
                                            int lds_ints) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int M = cm.M, nb = e.nbegin;
  const int need = 8 * M + e.nslots + 2 * nb;
  const bool rl = need <= lds_ints, tl = INSIDE && FLOGSUM_TBL + need <= lds_ints;
  int32_t *at = smem + (tl ? FLOGSUM_TBL : 0);
  if (tl)
    for (int q = tid; q < FLOGSUM_TBL; q += nt) smem[q] = __float_as_int(ftbl[q]);
  if (rl)
    for (int q = tid; q < need; q += nt) at[q] = e.rect[q];
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
