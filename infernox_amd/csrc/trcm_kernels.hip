// gfx950 truncated HMM-banded scanners, TrCYK and FTrInside: four marginal
// planes (J, L, R and, for bifurcations and the root, T) over the same banded
// deck layout as the standard scanners. One 256-lane workgroup per envelope
// walks the model's dependency levels (the states of a level read only lower
// levels); every cell of a level's states is computed by one lane, all of its
// valid planes at once. Per cell and plane the
// children are combined in increasing child order, and split points in
// increasing k, which is the order in which the reference's state loops
// reach that cell (cm_dpsearch_trunc.rs:354-880, 1140-1640), so TrCYK is
// bit-identical and FTrInside applies the same FLogsum sequence. The root
// takes truncated begins by MAX with the per-cell best entry state and mode
// (cm_dpsearch_trunc.rs:882-987); reporting runs on the host.
#include "hb_sched.h"

namespace ixg {

namespace {

enum { BJ = 1, BL = 2, BR = 4, BT = 8 };                          // mode-mask bits
enum { MODE_T = 0, MODE_R = 1, MODE_L = 2, MODE_J = 3, MODE_UNKNOWN = 4 };  // cm_trunc.rs:34-38

__device__ __forceinline__ int t_sd(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }
__device__ __forceinline__ int t_sdl(int st) { return (st == MP_ST || st == ML_ST || st == IL_ST) ? 1 : 0; }
__device__ __forceinline__ int t_sdr(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

template <bool INSIDE>
__device__ __forceinline__ float tcomb(float cur, float cand, const float *__restrict__ ftbl) {
  if (!INSIDE) return cand > cur ? cand : cur;
  const float mx = cur > cand ? cur : cand, mn = cur < cand ? cur : cand;  // FLogsum, cp9.rs:5405-5413
  if (mn == -INFINITY || (mx - mn) >= 23.0f) return mx;
  return mx + ftbl[(int)((mx - mn) * 1000.0f)];
}

__device__ __forceinline__ float clampi(float x) { return x < IMPOSSIBLE ? IMPOSSIBLE : x; }

// Plane-0 index of state y's cell (jy, dy), or -1 outside y's bands.
__device__ __forceinline__ int cell_at(const EnvDev &e, int y, int jy, int dy) {
  const int jmn = e.jmin[y];
  if (jy < jmn || jy > e.jmax[y]) return -1;
  const int h = e.hd_off[y] + (jy - jmn);
  const int lo = e.hdmin[h];
  if (dy < lo || dy > e.hdmax[h]) return -1;
  return e.roff[h] + (dy - lo);
}

// Local-end initialisation of a plane whose emission count is sdx
// (cm_dpsearch_trunc.rs:357-409): cells with d < sdx start IMPOSSIBLE.
__device__ __forceinline__ float tr_init(const CmDev &cm, int v, int d, int sdx) {
  const float es = cm.endsc[v];
  if (es > IMPOSSIBLE + 1.0f && d >= sdx) return cm.el_selfsc * (float)(d - sdx) + es;
  return IMPOSSIBLE;
}

// Cell (j, d) of a non-bifurcation state v: J, L and R planes.
template <bool INSIDE>
__device__ void tr_cell(const CmDev &cm, const EnvDev &e, int v, int st, int j, int d, int dp, int c,
                        const float *__restrict__ ftbl) {
  const int mv = e.mask[v];
  const bool dJ = mv & BJ, dL = mv & BL, dR = mv & BR;
  const int nc = e.nc;
  float *A = e.alpha;
  float J = tr_init(cm, v, d, t_sd(st)), L = tr_init(cm, v, d, t_sdl(st)), R = tr_init(cm, v, d, t_sdr(st));
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  if (st == E_ST) {
    if (dp == 0) J = L = R = 0.0f;
  } else if (st == ML_ST || st == IL_ST) {
    if (dJ || dL) {
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo, my = e.mask[y];
        if (!(my & (BJ | BL))) continue;
        const int cy = cell_at(e, y, j, d - 1);
        if (cy < 0) continue;
        if (dJ && (my & BJ)) J = tcomb<INSIDE>(J, A[cy] + tsc[yo], ftbl);
        if (dL && (my & BL)) L = tcomb<INSIDE>(L, A[nc + cy] + tsc[yo], ftbl);
      }
      const float em = oe[e.dsq[j - d + 1]];
      J = clampi(J + em);
      L = clampi(d >= 2 ? L + em : em);
    }
    if (dR) {
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo, my = e.mask[y];
        if (y == v || !(my & (BJ | BR))) continue;
        const int cy = cell_at(e, y, j, d);
        if (cy < 0) continue;
        if (my & BJ) R = tcomb<INSIDE>(R, A[cy] + tsc[yo], ftbl);
        if (my & BR) R = tcomb<INSIDE>(R, A[2 * nc + cy] + tsc[yo], ftbl);
      }
    }
  } else if (st == MR_ST || st == IR_ST) {
    if (dJ || dR) {
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo, my = e.mask[y];
        if (!(my & (BJ | BR))) continue;
        const int cy = cell_at(e, y, j - 1, d - 1);
        if (cy < 0) continue;
        if (dJ && (my & BJ)) J = tcomb<INSIDE>(J, A[cy] + tsc[yo], ftbl);
        if (dR && (my & BR)) R = tcomb<INSIDE>(R, A[2 * nc + cy] + tsc[yo], ftbl);
      }
      const float em = oe[e.dsq[j]];
      J = clampi(J + em);
      R = clampi(d >= 2 ? R + em : em);
    }
    if (dL) {
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo, my = e.mask[y];
        if (y == v || !(my & (BJ | BL))) continue;
        const int cy = cell_at(e, y, j, d);
        if (cy < 0) continue;
        if (my & BJ) L = tcomb<INSIDE>(L, A[cy] + tsc[yo], ftbl);
        if (my & BL) L = tcomb<INSIDE>(L, A[nc + cy] + tsc[yo], ftbl);
      }
    }
  } else if (st == MP_ST) {
    for (int yo = 0; yo < cn; yo++) {
      const int y = cf + yo, my = e.mask[y];
      const float t = tsc[yo];
      if (dJ && (my & BJ)) {
        const int cy = cell_at(e, y, j - 1, d - 2);
        if (cy >= 0) J = tcomb<INSIDE>(J, A[cy] + t, ftbl);
      }
      if (dR && (my & (BJ | BR))) {
        const int cy = cell_at(e, y, j - 1, d - 1);
        if (cy >= 0) {
          if (my & BJ) R = tcomb<INSIDE>(R, A[cy] + t, ftbl);
          if (my & BR) R = tcomb<INSIDE>(R, A[2 * nc + cy] + t, ftbl);
        }
      }
      if (dL && (my & (BJ | BL))) {
        const int cy = cell_at(e, y, j, d - 1);
        if (cy >= 0) {
          if (my & BJ) L = tcomb<INSIDE>(L, A[cy] + t, ftbl);
          if (my & BL) L = tcomb<INSIDE>(L, A[nc + cy] + t, ftbl);
        }
      }
    }
    const int xi = e.dsq[j - d + 1], xj = e.dsq[j];
    if (d >= 2) {
      J += oe[xi * KP + xj];
      L += cm.lmesc[v * KP + xi];
      R += cm.rmesc[v * KP + xj];
    } else {
      J = IMPOSSIBLE;
      L = cm.lmesc[v * KP + xi];
      R = cm.rmesc[v * KP + xj];
    }
    J = clampi(J);
    L = clampi(L);
    R = clampi(R);
  } else {  // D, S: no emission; a zero-length L or R cell reached from a child is IMPOSSIBLE
    for (int yo = 0; yo < cn; yo++) {
      const int y = cf + yo, my = e.mask[y];
      const bool uj = dJ && (my & BJ), ul = dL && (my & BL), ur = dR && (my & BR);
      if (!(uj || ul || ur)) continue;
      const int cy = cell_at(e, y, j, d);
      if (cy < 0) continue;
      const float t = tsc[yo];
      if (uj) J = tcomb<INSIDE>(J, A[cy] + t, ftbl);
      if (ul) L = tcomb<INSIDE>(L, A[nc + cy] + t, ftbl);
      if (ur) R = tcomb<INSIDE>(R, A[2 * nc + cy] + t, ftbl);
      if (d == 0) L = R = IMPOSSIBLE;
    }
  }
  if (dJ) A[c] = J;
  if (dL) A[nc + c] = L;
  if (dR) A[2 * nc + c] = R;
}

// Cell (j, d) of a bifurcation: split point k = length of the right child's
// part (J·J, J·L, R·J, R·L), then L from the left child alone and R from the
// right child alone (cm_dpsearch_trunc.rs:773-879).
template <bool INSIDE>
__device__ void tr_cell_bif(const CmDev &cm, const EnvDev &e, int v, int j, int d, int c,
                            const float *__restrict__ ftbl) {
  const int mv = e.mask[v];
  const bool dJ = mv & BJ, dL = mv & BL, dR = mv & BR, dT = mv & BT;
  const int nc = e.nc;
  float *A = e.alpha;
  float J = tr_init(cm, v, d, 0), L = J, R = J, T = IMPOSSIBLE;
  const int y = cm.cfirst[v], z = cm.cnum[v];
  const int my = e.mask[y], mz = e.mask[z];
  const bool kJ = dJ && (my & BJ) && (mz & BJ), kL = dL && (my & BJ) && (mz & BL);
  const bool kR = dR && (my & BR) && (mz & BJ), kT = dT && (my & BR) && (mz & BL);
  if ((kJ || kL || kR || kT) && j >= e.jmin[z] && j <= e.jmax[z]) {
    const int jpy = j - e.jmin[y], hz = e.hd_off[z] + (j - e.jmin[z]);
    const int hz0 = e.hdmin[hz];
    int kn = j - e.jmax[y] > hz0 ? j - e.jmax[y] : hz0;
    kn = kn > 0 ? kn : 0;
    const int kx = jpy < e.hdmax[hz] ? jpy : e.hdmax[hz];
    const int zoff = e.roff[hz] - hz0;
    const int hy0 = e.hd_off[y];
    for (int k = kn; k <= kx; k++) {
      const int hy = hy0 + (jpy - k);
      const int lo = e.hdmin[hy];
      if (k < d - e.hdmax[hy] || k > d - lo) continue;
      const int cy = e.roff[hy] + (d - lo - k), cz = zoff + k;
      if (kJ) J = tcomb<INSIDE>(J, A[cy] + A[cz], ftbl);
      if (kL) L = tcomb<INSIDE>(L, A[cy] + A[nc + cz], ftbl);
      if (kR) R = tcomb<INSIDE>(R, A[2 * nc + cy] + A[cz], ftbl);
      if (kT && k != 0 && k != d) T = tcomb<INSIDE>(T, A[2 * nc + cy] + A[nc + cz], ftbl);
    }
  }
  if (dL && (my & (BJ | BL))) {
    const int cy = cell_at(e, y, j, d);
    if (cy >= 0) {
      if (my & BJ) L = tcomb<INSIDE>(L, A[cy], ftbl);
      if (my & BL) L = tcomb<INSIDE>(L, A[nc + cy], ftbl);
    }
  }
  if (dR && (mz & (BJ | BR))) {
    const int cz = cell_at(e, z, j, d);
    if (cz >= 0) {
      if (mz & BJ) R = tcomb<INSIDE>(R, A[cz], ftbl);
      if (mz & BR) R = tcomb<INSIDE>(R, A[2 * nc + cz], ftbl);
    }
  }
  if (dJ) A[c] = J;
  if (dL) A[nc + c] = L;
  if (dR) A[2 * nc + c] = R;
  if (dT) A[3 * nc + c] = T;
}

template <bool INSIDE>
__global__ __launch_bounds__(HB_THREADS) void trhb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl,
                                                               const float *pty, const int32_t *ylist, int ny) {
  extern __shared__ int32_t smem[];
  EnvDev e = envs[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  hb_stage_scalars(e, cm.M, smem);
  auto unit = [&](int v, int u) {
    const int st = cm.sttype[v];
    const int jmn = e.jmin[v], nj = e.jmax[v] - jmn + 1;
    const int ho = e.hd_off[v];
    const int32_t *hmn = e.hdmin + ho, *hmx = e.hdmax + ho, *ro = e.roff + ho;
    if (st == IL_ST) {
      // J and L read their own (j, d-1): serial over d within row u
      for (int d = hmn[u]; d <= hmx[u]; d++)
        tr_cell<INSIDE>(cm, e, v, st, jmn + u, d, d - hmn[u], ro[u] + (d - hmn[u]), ftbl);
    } else if (st == IR_ST) {
      // J and R read their own (j-1, d-1), the same i: serial along diagonal i
      const int i = e.ibase[v] + u;
      for (int jp = 0; jp < nj; jp++) {
        const int j = jmn + jp, d = j - i + 1;
        if (d < hmn[jp] || d > hmx[jp]) continue;
        tr_cell<INSIDE>(cm, e, v, st, j, d, d - hmn[jp], ro[jp] + (d - hmn[jp]), ftbl);
      }
    } else {
      const int tgt = ro[0] + u;
      const int r = hb_row_of(ro, nj - 1, tgt);
      const int dp = tgt - ro[r];
      const int d = hmn[r] + dp, j = jmn + r;
      if (st == B_ST) tr_cell_bif<INSIDE>(cm, e, v, j, d, tgt, ftbl);
      else tr_cell<INSIDE>(cm, e, v, st, j, d, dp, tgt, ftbl);
    }
  };
  // states level by level (the root, with no units, takes truncated begins below)
  for (int l = 0; l < cm.nlev; l++) {
    hb_level_units(cm, e, l, unit);
    __syncthreads();
  }
  // root: truncated begins by MAX, per cell best (entry state, mode)
  const int jmn = e.jmin[0], jmx = e.jmax[0];
  const int nj = jmx >= jmn ? jmx - jmn + 1 : 0;
  if (nj == 0) return;
  const int32_t *hmn = e.hdmin + e.hd_off[0], *ro = e.roff + e.hd_off[0], *hmx = e.hdmax + e.hd_off[0];
  const int last = nj - 1;
  const int wl = hmx[last] - hmn[last] + 1;
  const int ncv = ro[last] + (wl > 0 ? wl : 0) - ro[0];
  const int m0 = e.mask[0], nc = e.nc;
  const float *A = e.alpha;
  for (int c = tid; c < ncv; c += nt) {
    const int tgt = ro[0] + c;
    const int r = hb_row_of(ro, last, tgt);
    const int d = hmn[r] + (tgt - ro[r]), j = jmn + r;
    float J = IMPOSSIBLE, L = IMPOSSIBLE, R = IMPOSSIBLE, T = IMPOSSIBLE, best = IMPOSSIBLE;
    int bm = MODE_UNKNOWN, br = 0;
    for (int q = 0; q < ny; q++) {
      const int y = ylist[q];
      const int cy = cell_at(e, y, j, d);
      if (cy < 0) continue;
      const int my = e.mask[y];
      const float p = pty[y];
      if ((m0 & BJ) && (my & BJ)) {
        const float s = A[cy] + p;
        if (s > J) {
          J = s;
          if (s > best) best = s, bm = MODE_J, br = y;
        }
      }
      if ((m0 & BL) && (my & BL)) {
        const float s = A[nc + cy] + p;
        if (s > L) {
          L = s;
          if (s > best) best = s, bm = MODE_L, br = y;
        }
      }
      if ((m0 & BR) && (my & BR)) {
        const float s = A[2 * nc + cy] + p;
        if (s > R) {
          R = s;
          if (s > best) best = s, bm = MODE_R, br = y;
        }
      }
      if ((m0 & BT) && (my & BT) && cm.sttype[y] == B_ST) {
        const float s = A[3 * nc + cy] + p;
        if (s > T) {
          T = s;
          if (s > best) best = s, bm = MODE_T, br = y;
        }
      }
    }
    e.rout[c] = J;
    e.rout[ncv + c] = L;
    e.rout[2 * ncv + c] = R;
    e.rout[3 * ncv + c] = T;
    e.rout[4 * ncv + c] = __int_as_float(br * 8 + bm);
  }
}

}  // namespace

hipError_t launch_trhb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, const float *pty,
                            const int32_t *ylist, int ny, int inside, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t lds = hb_lds_bytes(cm.M);
  if (inside)
    hipLaunchKernelGGL(trhb_scan_kernel<true>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl, pty,
                       ylist, ny);
  else
    hipLaunchKernelGGL(trhb_scan_kernel<false>, dim3((unsigned)n), dim3(HB_THREADS), lds, s, cm, envs, ftbl, pty,
                       ylist, ny);
  return hipGetLastError();
}

}  // namespace ixg
