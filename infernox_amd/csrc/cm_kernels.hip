// gfx950 HMM-banded CYK / Inside scanners (F6/F7): one 256-lane workgroup per
// envelope walks the model's dependency levels (a state's level is one more
// than its children's highest, so the two subtrees of a bifurcation and the
// match states of a node proceed together), filling the banded decks of all
// states of a level in parallel before the next level starts. Per cell the children are combined in increasing child order
// (max for CYK, the FLogsum table for Inside), exactly the order of the
// reference's scalar loops (cp9.rs:4788-5253, 5514-5885), so CYK is
// bit-identical and Inside applies the same float operations in the same
// order. The deck lives in HBM; every cell is written once, by the lane
// that computes it.
#include "hb_sched.h"

namespace ixg {

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

// Value of child y's cell (jy, dy) if it lies in y's bands.
__device__ __forceinline__ bool child_cell(const EnvDev &e, int y, int jy, int dy, float &val) {
  const int jmn = e.jmin[y];
  if (jy < jmn || jy > e.jmax[y]) return false;
  const int h = e.hd_off[y] + (jy - jmn);
  const int lo = e.hdmin[h];
  if (dy < lo || dy > e.hdmax[h]) return false;
  val = gmem(e.alpha)[e.roff[h] + (dy - lo)];
  return true;
}

template <bool INSIDE>
__device__ __forceinline__ float cell_init(const CmDev &cm, int v, int st, int d) {
  const float es = cm.endsc[v];
  if (es > -9.999e35f) {
    const int sd = sdelta(st);
    const int dp = d - sd > 0 ? d - sd : 0;
    return cm.el_selfsc * (float)dp + es;
  }
  return IMPOSSIBLE;
}

// Non-insert, non-bifurcation states (S, D, ML, MR, MP) and E: all children's
// cells are fetched first, then combined in child order.
template <bool INSIDE>
__device__ float cell_generic(const CmDev &cm, const EnvDev &e, int v, int st, int j, int d, int dp,
                              const float *__restrict__ ftbl) {
  float cur = cell_init<INSIDE>(cm, v, st, d);
  if (st == E_ST) return dp == 0 ? 0.0f : cur;
  const int sd = sdelta(st), sdr = srdelta(st);
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  float a[MAXC];
  bool ok[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    const int y = min(cf + yo, cm.M - 1);
    const RowRef r = row_ref(e, y, j - sdr);
    const int dy = d - sd;
    ok[yo] = yo < cn && dy >= r.lo && dy <= r.hi;
    a[yo] = gmem(e.alpha)[ok[yo] ? r.base + dy : 0];
  }
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (ok[yo]) cur = combine<INSIDE>(cur, a[yo] + tsc[yo], ftbl);
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) cur += oe[gmem(e.dsq)[i]];
  else if (st == MR_ST) cur += oe[gmem(e.dsq)[j]];
  else if (st == MP_ST) cur += oe[gmem(e.dsq)[i] * KP + gmem(e.dsq)[j]];
  return cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
}

template <bool INSIDE>
__device__ float cell_bif(const CmDev &cm, const EnvDev &e, int v, int j, int d, const float *__restrict__ ftbl) {
  float cur = cell_init<INSIDE>(cm, v, B_ST, d);
  const int y = cm.cfirst[v], z = cm.cnum[v];
  const int jminy = e.jmin[y], jminz = e.jmin[z];
  if (j < jminz || j > e.jmax[z]) return cur;
  const int jpy = j - jminy, hz = e.hd_off[z] + (j - jminz);
  const int hz0 = e.hdmin[hz];
  int kn = j - e.jmax[y] > hz0 ? j - e.jmax[y] : hz0;
  kn = kn > 0 ? kn : 0;
  const int kx = jpy < e.hdmax[hz] ? jpy : e.hdmax[hz];
  const IXG_GLOBAL float *az = gmem(e.alpha) + e.roff[hz] - hz0;
  const int hy0 = e.hd_off[y];
#pragma unroll 4
  for (int k = kn; k <= kx; k++) {
    const int hy = hy0 + (jpy - k);
    const int lo = e.hdmin[hy];
    const bool ok = k >= d - e.hdmax[hy] && k <= d - lo;
    const float a = gmem(e.alpha)[ok ? e.roff[hy] + (d - lo - k) : 0];
    const float b = az[k];
    if (ok) cur = combine<INSIDE>(cur, a + b, ftbl);
  }
  return cur;
}

constexpr int ICH = 4;  // cells of an insert chain fetched ahead of their serial combine

// One j row of an IL state: the self transition reads (j, d-1), the previous
// cell of the same row, so the row is a serial chain over d. The other
// children's rows are looked up once per row; their cells are fetched ICH
// cells ahead; the chain itself carries the previous cell in a register.
template <bool INSIDE>
__device__ void il_row(const CmDev &cm, const EnvDev &e, int v, int u, const float *__restrict__ ftbl) {
  const int j = e.jmin[v] + u, h = e.hd_off[v] + u;
  const int lo = e.hdmin[h], hi = e.hdmax[h];
  if (hi < lo) return;
  IXG_GLOBAL float *row = gmem(e.alpha) + e.roff[h] - lo;
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  RowRef rr[MAXC];
  float t[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    rr[yo] = row_ref(e, min(cf + yo, cm.M - 1), j);
    t[yo] = tsc[yo];
  }
  float prev = 0.f;
  for (int d0 = lo; d0 <= hi; d0 += ICH) {
    float a[ICH][MAXC];
    bool ok[ICH][MAXC];
    int x[ICH];
#pragma unroll
    for (int c = 0; c < ICH; c++) {
      const int d = d0 + c;
      x[c] = gmem(e.dsq)[d <= hi ? j - d + 1 : j];
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        const int y = cf + yo;
        ok[c][yo] = d <= hi && yo < cn && y != v && d - 1 >= rr[yo].lo && d - 1 <= rr[yo].hi;
        a[c][yo] = gmem(e.alpha)[ok[c][yo] ? rr[yo].base + d - 1 : 0];
      }
    }
#pragma unroll
    for (int c = 0; c < ICH; c++) {
      const int d = d0 + c;
      if (d > hi) break;
      float cur = cell_init<INSIDE>(cm, v, IL_ST, d);
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (yo >= cn) break;
        if (cf + yo == v) {
          if (d > lo) cur = combine<INSIDE>(cur, prev + t[yo], ftbl);
        } else if (ok[c][yo]) {
          cur = combine<INSIDE>(cur, a[c][yo] + t[yo], ftbl);
        }
      }
      cur += oe[x[c]];
      cur = cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
      row[d] = cur;
      prev = cur;
    }
  }
}

// One i diagonal of an IR state: the self transition reads (j-1, d-1), the
// previous cell of the same diagonal, so the diagonal is a serial chain over
// j. Cells ahead are located and their other children fetched ICH rows ahead.
template <bool INSIDE>
__device__ void ir_diag(const CmDev &cm, const EnvDev &e, int v, int u, const float *__restrict__ ftbl) {
  const int i = e.ibase[v] + u;
  const int jmn = e.jmin[v], nj = e.jmax[v] - jmn + 1, ho = e.hd_off[v];
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  float t[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) t[yo] = tsc[yo];
  float prev = 0.f;
  bool have_prev = false;
  for (int jp0 = 0; jp0 < nj; jp0 += ICH) {
    float a[ICH][MAXC];
    bool ok[ICH][MAXC], on[ICH];
    int cell[ICH], x[ICH];
#pragma unroll
    for (int c = 0; c < ICH; c++) {
      const int jp = jp0 + c, j = jmn + jp, d = j - i + 1;
      const int h = ho + (jp < nj ? jp : 0);
      const int lo = e.hdmin[h], hi = e.hdmax[h];
      on[c] = jp < nj && d >= lo && d <= hi;
      cell[c] = e.roff[h] + (d - lo);
      x[c] = gmem(e.dsq)[on[c] ? j : i];
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        const int y = cf + yo;
        const RowRef r = row_ref(e, min(y, cm.M - 1), j - 1);
        ok[c][yo] = on[c] && yo < cn && y != v && d - 1 >= r.lo && d - 1 <= r.hi;
        a[c][yo] = gmem(e.alpha)[ok[c][yo] ? r.base + d - 1 : 0];
      }
    }
#pragma unroll
    for (int c = 0; c < ICH; c++) {
      if (!on[c]) {
        have_prev = false;
        continue;
      }
      const int d = jmn + jp0 + c - i + 1;
      float cur = cell_init<INSIDE>(cm, v, IR_ST, d);
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (yo >= cn) break;
        if (cf + yo == v) {
          if (have_prev) cur = combine<INSIDE>(cur, prev + t[yo], ftbl);
        } else if (ok[c][yo]) {
          cur = combine<INSIDE>(cur, a[c][yo] + t[yo], ftbl);
        }
      }
      cur += oe[x[c]];
      cur = cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
      gmem(e.alpha)[cell[c]] = cur;
      prev = cur;
      have_prev = true;
    }
  }
}

// One envelope, scanned by the whole workgroup.
template <bool INSIDE>
__device__ __forceinline__ void hb_scan_env(const CmDev &cm, EnvDev e, const float *ftbl, int32_t *smem) {
  const int tid = threadIdx.x, nt = blockDim.x;
  hb_stage_scalars(e, cm.M, smem);
  auto unit = [&](int v, int u) {
    const int st = cm.sttype[v];
    const int jmn = e.jmin[v], nj = e.jmax[v] - jmn + 1;
    const int ho = e.hd_off[v];
    const int32_t *hmn = e.hdmin + ho, *ro = e.roff + ho;
    if (st == IL_ST) {
      il_row<INSIDE>(cm, e, v, u, ftbl);
    } else if (st == IR_ST) {
      ir_diag<INSIDE>(cm, e, v, u, ftbl);
    } else {
      const int tgt = ro[0] + u;
      const int r = hb_row_of(ro, nj - 1, tgt);
      const int dp = tgt - ro[r];
      const int d = hmn[r] + dp, j = jmn + r;
      gmem(e.alpha)[tgt] = st == B_ST ? cell_bif<INSIDE>(cm, e, v, j, d, ftbl)
                                : cell_generic<INSIDE>(cm, e, v, st, j, d, dp, ftbl);
    }
  };
  // states level by level: the states of one level read only lower levels
  for (int l = 0; l < cm.nlev; l++) {
    hb_level_units(cm, e, l, unit);
    __syncthreads();
  }
  // local begins into the root deck by MAX (cp9.rs:5264-5304, 5908-5937);
  // the root rows also go to the envelope's compact root output
  const int jmn = e.jmin[0], jmx = e.jmax[0];
  const int nj = jmx >= jmn ? jmx - jmn + 1 : 0;
  if (nj == 0) return;
  const int32_t *hmn = e.hdmin + e.hd_off[0], *hmx = e.hdmax + e.hd_off[0], *ro = e.roff + e.hd_off[0];
  const int last = nj - 1;
  const int wl = hmx[last] - hmn[last] + 1;
  const int ncv = ro[last] + (wl > 0 ? wl : 0) - ro[0];
  for (int c = tid; c < ncv; c += nt) {
    const int tgt = ro[0] + c;
    const int lo = hb_row_of(ro, last, tgt);
    const int d = hmn[lo] + (tgt - ro[lo]), j = jmn + lo;
    float cur = gmem(e.alpha)[tgt];
    if (cm.local_begin)
      for (int b = 0; b < cm.nbegin; b++) {
        const int y = cm.begin_list[b];
        float a;
        if (child_cell(e, y, j, d, a)) {
          const float s = a + cm.beginsc[y];
          if (s > cur) cur = s;
        }
      }
    gmem(e.alpha)[tgt] = cur;
    gmem(e.rout)[c] = cur;
  }
}

template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) void hb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl) {
  extern __shared__ int32_t smem[];
  hb_scan_env<INSIDE>(cm, envs[blockIdx.x], ftbl, smem);
}

// Persistent form over envelopes of several models: each workgroup claims
// the next envelope of `order` (largest first) from a device counter until
// the list is exhausted; an envelope's model is cms[e.cmi]. Every workgroup
// leaves once a claim passes n, so the grid drains whatever its size.
template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) void hb_scan_queue_kernel(const CmDev *__restrict__ cms,
                                                                   const EnvDev *__restrict__ envs,
                                                                   const int32_t *__restrict__ order, int32_t n,
                                                                   uint32_t *counter, const float *ftbl) {
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
    hb_scan_env<INSIDE>(cm, e, ftbl, smem);
    __syncthreads();  // the next envelope's LDS staging waits for this one's last reads
  }
}

hipError_t launch_hb_scan_queue(const CmDev *cms, int max_m, const EnvDev *envs, const int32_t *order, int64_t n,
                                uint32_t *counter, int grid, const float *ftbl, int inside, hipStream_t s) {
  if (n == 0) return hipSuccess;
  // models with more states than the LDS holds read their band scalars from
  // memory; the others stage 20 bytes per state
  const size_t lds = (size_t)20 * std::min(max_m, HB_LDS_STATES);
  hipError_t r = hipMemsetAsync(counter, 0, sizeof(uint32_t), s);
  if (r != hipSuccess) return r;
  const unsigned g = (unsigned)std::min<int64_t>(n, grid);
  if (inside)
    hipLaunchKernelGGL(hb_scan_queue_kernel<true>, dim3(g), dim3(HB_THREADS), lds, s, cms, envs, order, (int32_t)n,
                       counter, ftbl);
  else
    hipLaunchKernelGGL(hb_scan_queue_kernel<false>, dim3(g), dim3(HB_THREADS), lds, s, cms, envs, order,
                       (int32_t)n, counter, ftbl);
  return hipGetLastError();
}

hipError_t launch_hb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, int inside,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t lds = hb_lds_bytes(cm.M);
  if (inside)
    hipLaunchKernelGGL(hb_scan_kernel<true>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl);
  else
    hipLaunchKernelGGL(hb_scan_kernel<false>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl);
  return hipGetLastError();
}

}  // namespace ixg
