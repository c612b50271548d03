// gfx950 HMM-banded CYK / Inside scanners (F6/F7): one 256-lane workgroup per
// envelope walks the states from M-1 down to 0 (children have larger
// indices), filling each state's banded deck in parallel before the next
// state starts. Per cell the children are combined in increasing child order
// (max for CYK, the FLogsum table for Inside), exactly the order of the
// reference's scalar loops (cp9.rs:4788-5253, 5514-5885), so CYK is
// bit-identical and Inside applies the same float operations in the same
// order. The deck lives in HBM; every cell is written once, by the lane
// that computes it.
#include "cm_internal.h"

namespace ixg {

__device__ __forceinline__ int sdelta(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }
__device__ __forceinline__ int srdelta(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

template <bool INSIDE>
__device__ __forceinline__ float combine(float cur, float cand, const float *__restrict__ ftbl) {
  if (!INSIDE) return cand > cur ? cand : cur;
  // FLogsum (cp9.rs:5405-5413)
  const float mx = cur > cand ? cur : cand, mn = cur < cand ? cur : cand;
  if (mn == -INFINITY || (mx - mn) >= 23.0f) return mx;
  return mx + ftbl[(int)((mx - mn) * 1000.0f)];
}

// Value of child y's cell (j - sdr, d - sd) if it lies in y's bands.
__device__ __forceinline__ bool child_cell(const EnvDev &e, int y, int jy, int dy, float &val) {
  const int jmn = e.jmin[y];
  if (jy < jmn || jy > e.jmax[y]) return false;
  const int h = e.hd_off[y] + (jy - jmn);
  const int lo = e.hdmin[h];
  if (dy < lo || dy > e.hdmax[h]) return false;
  val = e.alpha[e.roff[h] + (dy - lo)];
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

// Non-insert, non-bifurcation states (S, D, ML, MR, MP) and E.
template <bool INSIDE>
__device__ float cell_generic(const CmDev &cm, const EnvDev &e, int v, int st, int j, int d, int dp,
                              const float *__restrict__ ftbl) {
  float cur = cell_init<INSIDE>(cm, v, st, d);
  if (st == E_ST) return dp == 0 ? 0.0f : cur;
  const int sd = sdelta(st), sdr = srdelta(st);
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  for (int yo = 0; yo < cn; yo++) {
    float a;
    if (child_cell(e, cf + yo, j - sdr, d - sd, a)) cur = combine<INSIDE>(cur, a + tsc[yo], ftbl);
  }
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) cur += oe[e.dsq[i]];
  else if (st == MR_ST) cur += oe[e.dsq[j]];
  else if (st == MP_ST) cur += oe[e.dsq[i] * KP + e.dsq[j]];
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
  const float *az = e.alpha + e.roff[hz] - hz0;
  const int hy0 = e.hd_off[y];
  for (int k = kn; k <= kx; k++) {
    const int hy = hy0 + (jpy - k);
    const int lo = e.hdmin[hy];
    if (k >= d - e.hdmax[hy] && k <= d - lo)
      cur = combine<INSIDE>(cur, e.alpha[e.roff[hy] + (d - lo - k)] + az[k], ftbl);
  }
  return cur;
}

// IL/IR cell: children in increasing order (the self transition included),
// then the single emission and the clamp.
template <bool INSIDE>
__device__ __forceinline__ float cell_insert(const CmDev &cm, const EnvDev &e, int v, int st, int j, int d,
                                             const float *__restrict__ ftbl) {
  float cur = cell_init<INSIDE>(cm, v, st, d);
  const int sdr = st == IR_ST ? 1 : 0;
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  for (int yo = 0; yo < cn; yo++) {
    float a;
    if (child_cell(e, cf + yo, j - sdr, d - 1, a)) cur = combine<INSIDE>(cur, a + tsc[yo], ftbl);
  }
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  cur += oe[st == IL_ST ? e.dsq[j - d + 1] : e.dsq[j]];
  return cur < IMPOSSIBLE ? IMPOSSIBLE : cur;
}

template <bool INSIDE>
__global__ __launch_bounds__(256) void hb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl) {
  const EnvDev e = envs[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int v = cm.M - 1; v >= 0; v--) {
    const int st = cm.sttype[v];
    const int jmn = e.jmin[v], jmx = e.jmax[v];
    const int nj = jmx >= jmn ? jmx - jmn + 1 : 0;
    if (nj > 0) {
      const int ho = e.hd_off[v];
      const int32_t *hmn = e.hdmin + ho, *hmx = e.hdmax + ho, *ro = e.roff + ho;
      if (st == IL_ST) {
        // self transition reads (j, d-1): serial over d within a row
        for (int jp = tid; jp < nj; jp += nt) {
          const int j = jmn + jp;
          float *row = e.alpha + ro[jp] - hmn[jp];
          for (int d = hmn[jp]; d <= hmx[jp]; d++) row[d] = cell_insert<INSIDE>(cm, e, v, st, j, d, ftbl);
        }
      } else if (st == IR_ST) {
        // self transition reads (j-1, d-1), the same i: serial along each i
        int imin = 0x7FFFFFFF, imax = -0x7FFFFFFF;
        for (int jp = 0; jp < nj; jp++)
          if (hmx[jp] >= hmn[jp]) {
            const int j = jmn + jp;
            imin = min(imin, j - hmx[jp] + 1);
            imax = max(imax, j - hmn[jp] + 1);
          }
        for (int i = imin + tid; i <= imax; i += nt)
          for (int jp = 0; jp < nj; jp++) {
            const int j = jmn + jp, d = j - i + 1;
            if (d < hmn[jp] || d > hmx[jp]) continue;
            e.alpha[ro[jp] + (d - hmn[jp])] = cell_insert<INSIDE>(cm, e, v, st, j, d, ftbl);
          }
      } else {
        const int last = nj - 1;
        const int wl = hmx[last] - hmn[last] + 1;
        const int base = ro[0];
        const int ncv = ro[last] + (wl > 0 ? wl : 0) - base;
        for (int c = tid; c < ncv; c += nt) {
          const int tgt = base + c;
          int lo = 0, hi = last;  // last row whose first cell is at or before tgt
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ro[mid] <= tgt) lo = mid;
            else hi = mid - 1;
          }
          const int dp = tgt - ro[lo];
          const int d = hmn[lo] + dp, j = jmn + lo;
          e.alpha[tgt] = st == B_ST ? cell_bif<INSIDE>(cm, e, v, j, d, ftbl)
                                    : cell_generic<INSIDE>(cm, e, v, st, j, d, dp, ftbl);
        }
      }
    }
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
    int lo = 0, hi = last;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ro[mid] <= tgt) lo = mid;
      else hi = mid - 1;
    }
    const int d = hmn[lo] + (tgt - ro[lo]), j = jmn + lo;
    float cur = e.alpha[tgt];
    if (cm.local_begin)
      for (int b = 0; b < cm.nbegin; b++) {
        const int y = cm.begin_list[b];
        float a;
        if (child_cell(e, y, j, d, a)) {
          const float s = a + cm.beginsc[y];
          if (s > cur) cur = s;
        }
      }
    e.alpha[tgt] = cur;
    e.rout[c] = cur;
  }
}

hipError_t launch_hb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, int inside,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (inside) hipLaunchKernelGGL(hb_scan_kernel<true>, dim3((unsigned)n), dim3(256), 0, s, cm, envs, ftbl);
  else hipLaunchKernelGGL(hb_scan_kernel<false>, dim3((unsigned)n), dim3(256), 0, s, cm, envs, ftbl);
  return hipGetLastError();
}

}  // namespace ixg
