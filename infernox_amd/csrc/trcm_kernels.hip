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
  // FLogsum (cp9.rs:5405-5413), branch-free: the table is read at index 0
  // when the smaller term is -inf or 23 or more below the larger
  const float mx = cur > cand ? cur : cand, mn = cur < cand ? cur : cand;
  const float diff = mx - mn;
  const bool near = !(mn == -INFINITY) && diff < 23.0f;
  const float tv = ftbl[near ? (int)(diff * 1000.0f) : 0];
  return near ? mx + tv : mx;
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

__device__ __forceinline__ bool in_row(const RowRef &r, int d) { return d >= r.lo && d <= r.hi; }

// Cell (j, d) of a non-insert, non-bifurcation state v (E, ML, MR, MP, D,
// S): J, L and R planes. Every child's cells are fetched first (rows j and,
// for right emitters, j-1), then combined per plane in child order.
template <bool INSIDE>
__device__ void tr_cell(const CmDev &cm, const EnvDev &e, int v, int st, int j, int d, int dp, int c,
                        const float *__restrict__ ftbl) {
  const int mv = e.mask[v];
  const bool dJ = mv & BJ, dL = mv & BL, dR = mv & BR;
  const int nc = e.nc;
  IXG_GLOBAL float *A = gmem(e.alpha);
  float J = tr_init(cm, v, d, t_sd(st)), L = tr_init(cm, v, d, t_sdl(st)), R = tr_init(cm, v, d, t_sdr(st));
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  if (st == E_ST) {
    if (dp == 0) J = L = R = 0.0f;
  } else if (st == ML_ST) {
    // J, L from (j, d-1); R from (j, d)
    float a1[MAXC], l1[MAXC], a0[MAXC], r0[MAXC];
    int my[MAXC];
    bool k1[MAXC], k0[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const int y = min(cf + yo, cm.M - 1);
      my[yo] = yo < cn ? e.mask[y] : 0;
      const RowRef r = row_ref(e, y, j);
      k1[yo] = in_row(r, d - 1);
      k0[yo] = in_row(r, d);
      const int i1 = k1[yo] ? r.base + d - 1 : 0, i0 = k0[yo] ? r.base + d : 0;
      a1[yo] = A[i1];
      l1[yo] = (dL && (my[yo] & BL)) ? A[nc + i1] : 0.f;
      a0[yo] = A[i0];
      r0[yo] = (dR && (my[yo] & BR)) ? A[2 * nc + i0] : 0.f;
    }
    const float em = oe[gmem(e.dsq)[j - d + 1]];
    if (dJ || dL) {
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (!(my[yo] & (BJ | BL)) || !k1[yo]) continue;
        if (dJ && (my[yo] & BJ)) J = tcomb<INSIDE>(J, a1[yo] + tsc[yo], ftbl);
        if (dL && (my[yo] & BL)) L = tcomb<INSIDE>(L, l1[yo] + tsc[yo], ftbl);
      }
      J = clampi(J + em);
      L = clampi(d >= 2 ? L + em : em);
    }
    if (dR) {
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (!(my[yo] & (BJ | BR)) || !k0[yo]) continue;
        if (my[yo] & BJ) R = tcomb<INSIDE>(R, a0[yo] + tsc[yo], ftbl);
        if (my[yo] & BR) R = tcomb<INSIDE>(R, r0[yo] + tsc[yo], ftbl);
      }
    }
  } else if (st == MR_ST) {
    // J, R from (j-1, d-1); L from (j, d)
    float a1[MAXC], r1[MAXC], a0[MAXC], l0[MAXC];
    int my[MAXC];
    bool k1[MAXC], k0[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const int y = min(cf + yo, cm.M - 1);
      my[yo] = yo < cn ? e.mask[y] : 0;
      const RowRef rp = row_ref(e, y, j - 1), rq = row_ref(e, y, j);
      k1[yo] = in_row(rp, d - 1);
      k0[yo] = in_row(rq, d);
      const int i1 = k1[yo] ? rp.base + d - 1 : 0, i0 = k0[yo] ? rq.base + d : 0;
      a1[yo] = A[i1];
      r1[yo] = (dR && (my[yo] & BR)) ? A[2 * nc + i1] : 0.f;
      a0[yo] = A[i0];
      l0[yo] = (dL && (my[yo] & BL)) ? A[nc + i0] : 0.f;
    }
    const float em = oe[gmem(e.dsq)[j]];
    if (dJ || dR) {
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (!(my[yo] & (BJ | BR)) || !k1[yo]) continue;
        if (dJ && (my[yo] & BJ)) J = tcomb<INSIDE>(J, a1[yo] + tsc[yo], ftbl);
        if (dR && (my[yo] & BR)) R = tcomb<INSIDE>(R, r1[yo] + tsc[yo], ftbl);
      }
      J = clampi(J + em);
      R = clampi(d >= 2 ? R + em : em);
    }
    if (dL) {
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        if (!(my[yo] & (BJ | BL)) || !k0[yo]) continue;
        if (my[yo] & BJ) L = tcomb<INSIDE>(L, a0[yo] + tsc[yo], ftbl);
        if (my[yo] & BL) L = tcomb<INSIDE>(L, l0[yo] + tsc[yo], ftbl);
      }
    }
  } else if (st == MP_ST) {
    // J from (j-1, d-2); R from (j-1, d-1); L from (j, d-1)
    float a2[MAXC], a1[MAXC], r1[MAXC], b1[MAXC], l1[MAXC];
    int my[MAXC];
    bool k2[MAXC], k1[MAXC], kq[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const int y = min(cf + yo, cm.M - 1);
      my[yo] = yo < cn ? e.mask[y] : 0;
      const RowRef rp = row_ref(e, y, j - 1), rq = row_ref(e, y, j);
      k2[yo] = in_row(rp, d - 2);
      k1[yo] = in_row(rp, d - 1);
      kq[yo] = in_row(rq, d - 1);
      const int i2 = k2[yo] ? rp.base + d - 2 : 0, i1 = k1[yo] ? rp.base + d - 1 : 0;
      const int iq = kq[yo] ? rq.base + d - 1 : 0;
      a2[yo] = A[i2];
      a1[yo] = A[i1];
      r1[yo] = (dR && (my[yo] & BR)) ? A[2 * nc + i1] : 0.f;
      b1[yo] = A[iq];
      l1[yo] = (dL && (my[yo] & BL)) ? A[nc + iq] : 0.f;
    }
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const int my_ = my[yo];
      const float t = tsc[yo];
      if (dJ && (my_ & BJ) && k2[yo]) J = tcomb<INSIDE>(J, a2[yo] + t, ftbl);
      if (dR && (my_ & (BJ | BR)) && k1[yo]) {
        if (my_ & BJ) R = tcomb<INSIDE>(R, a1[yo] + t, ftbl);
        if (my_ & BR) R = tcomb<INSIDE>(R, r1[yo] + t, ftbl);
      }
      if (dL && (my_ & (BJ | BL)) && kq[yo]) {
        if (my_ & BJ) L = tcomb<INSIDE>(L, b1[yo] + t, ftbl);
        if (my_ & BL) L = tcomb<INSIDE>(L, l1[yo] + t, ftbl);
      }
    }
    const int xi = gmem(e.dsq)[j - d + 1], xj = gmem(e.dsq)[j];
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
    float a0[MAXC], l0[MAXC], r0[MAXC];
    int my[MAXC];
    bool k0[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const int y = min(cf + yo, cm.M - 1);
      my[yo] = yo < cn ? e.mask[y] : 0;
      const RowRef r = row_ref(e, y, j);
      k0[yo] = in_row(r, d);
      const int i0 = k0[yo] ? r.base + d : 0;
      a0[yo] = A[i0];
      l0[yo] = (dL && (my[yo] & BL)) ? A[nc + i0] : 0.f;
      r0[yo] = (dR && (my[yo] & BR)) ? A[2 * nc + i0] : 0.f;
    }
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++) {
      const bool uj = dJ && (my[yo] & BJ), ul = dL && (my[yo] & BL), ur = dR && (my[yo] & BR);
      if (!(uj || ul || ur) || !k0[yo]) continue;
      const float t = tsc[yo];
      if (uj) J = tcomb<INSIDE>(J, a0[yo] + t, ftbl);
      if (ul) L = tcomb<INSIDE>(L, l0[yo] + t, ftbl);
      if (ur) R = tcomb<INSIDE>(R, r0[yo] + t, ftbl);
      if (d == 0) L = R = IMPOSSIBLE;
    }
  }
  if (dJ) A[c] = J;
  if (dL) A[nc + c] = L;
  if (dR) A[2 * nc + c] = R;
}

// cells of an insert chain fetched ahead of their serial combine: two for
// TrCYK, one for FTrInside (whose table lookups leave no registers for a
// second cell's operands without spilling)
template <bool INSIDE>
struct Tch {
  static constexpr int n = INSIDE ? 1 : 2;
};

// One j row of an IL state: J and L read their own (j, d-1), the previous
// cell of the row, carried in registers; the other children's cells (rows
// looked up once per row) are fetched TCH cells ahead. R has no self term.
template <bool INSIDE>
__device__ void tr_il_row(const CmDev &cm, const EnvDev &e, int v, int u, const float *__restrict__ ftbl) {
  const int j = e.jmin[v] + u, h = e.hd_off[v] + u;
  const int lo = e.hdmin[h], hi = e.hdmax[h];
  if (hi < lo) return;
  const int cbase = e.roff[h] - lo;
  const int mv = e.mask[v];
  const bool dJ = mv & BJ, dL = mv & BL, dR = mv & BR;
  const int nc = e.nc;
  IXG_GLOBAL float *A = gmem(e.alpha);
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  RowRef rr[MAXC];
  int my[MAXC];
  float t[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    const int y = min(cf + yo, cm.M - 1);
    rr[yo] = row_ref(e, y, j);
    my[yo] = (yo < cn && cf + yo != v) ? e.mask[y] : 0;  // the self child is handled by the chain
    t[yo] = tsc[yo];
  }
  int selfo = -1;
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn && cf + yo == v) selfo = yo;
  float pJ = 0.f, pL = 0.f;
  constexpr int TCH = Tch<INSIDE>::n;
  for (int d0 = lo; d0 <= hi; d0 += TCH) {
    float a1[TCH][MAXC], l1[TCH][MAXC], a0[TCH][MAXC], r0[TCH][MAXC];
    int x[TCH];
#pragma unroll
    for (int q = 0; q < TCH; q++) {
      const int d = d0 + q;
      x[q] = gmem(e.dsq)[d <= hi ? j - d + 1 : j];
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        const bool k1 = d <= hi && my[yo] && in_row(rr[yo], d - 1), k0 = d <= hi && my[yo] && in_row(rr[yo], d);
        const int i1 = k1 ? rr[yo].base + d - 1 : 0, i0 = k0 ? rr[yo].base + d : 0;
        a1[q][yo] = k1 ? A[i1] : IMPOSSIBLE;
        l1[q][yo] = (k1 && dL && (my[yo] & BL)) ? A[nc + i1] : IMPOSSIBLE;
        a0[q][yo] = k0 ? A[i0] : IMPOSSIBLE;
        r0[q][yo] = (k0 && dR && (my[yo] & BR)) ? A[2 * nc + i0] : IMPOSSIBLE;
      }
    }
#pragma unroll
    for (int q = 0; q < TCH; q++) {
      const int d = d0 + q;
      if (d > hi) break;
      float J = tr_init(cm, v, d, 1), L = tr_init(cm, v, d, 1), R = tr_init(cm, v, d, 0);
      if (dJ || dL) {
#pragma unroll
        for (int yo = 0; yo < MAXC; yo++) {
          if (yo == selfo) {
            if (d > lo) {  // (j, d-1) of this state, valid in every plane this state fills
              if (dJ) J = tcomb<INSIDE>(J, pJ + t[yo], ftbl);
              if (dL) L = tcomb<INSIDE>(L, pL + t[yo], ftbl);
            }
            continue;
          }
          if (!(my[yo] & (BJ | BL)) || !in_row(rr[yo], d - 1)) continue;
          if (dJ && (my[yo] & BJ)) J = tcomb<INSIDE>(J, a1[q][yo] + t[yo], ftbl);
          if (dL && (my[yo] & BL)) L = tcomb<INSIDE>(L, l1[q][yo] + t[yo], ftbl);
        }
        const float em = oe[x[q]];
        J = clampi(J + em);
        L = clampi(d >= 2 ? L + em : em);
      }
      if (dR) {
#pragma unroll
        for (int yo = 0; yo < MAXC; yo++) {
          if (!(my[yo] & (BJ | BR)) || !in_row(rr[yo], d)) continue;
          if (my[yo] & BJ) R = tcomb<INSIDE>(R, a0[q][yo] + t[yo], ftbl);
          if (my[yo] & BR) R = tcomb<INSIDE>(R, r0[q][yo] + t[yo], ftbl);
        }
      }
      const int c = cbase + d;
      if (dJ) A[c] = J;
      if (dL) A[nc + c] = L;
      if (dR) A[2 * nc + c] = R;
      pJ = J;
      pL = L;
    }
  }
}

// One i diagonal of an IR state: J and R read their own (j-1, d-1), the
// previous cell of the diagonal, carried in registers; the other children's
// cells are fetched TCH cells ahead. L has no self term.
template <bool INSIDE>
__device__ void tr_ir_diag(const CmDev &cm, const EnvDev &e, int v, int u, const float *__restrict__ ftbl) {
  const int i = e.ibase[v] + u;
  const int jmn = e.jmin[v], nj = e.jmax[v] - jmn + 1, ho = e.hd_off[v];
  const int mv = e.mask[v];
  const bool dJ = mv & BJ, dL = mv & BL, dR = mv & BR;
  const int nc = e.nc;
  IXG_GLOBAL float *A = gmem(e.alpha);
  const int cf = cm.cfirst[v], cn = cm.cnum[v];
  const float *tsc = cm.tsc + v * MAXC;
  const float *oe = cm.oesc + (size_t)v * (KP * KP);
  int my[MAXC], ys[MAXC];
  float t[MAXC];
  int selfo = -1;
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++) {
    ys[yo] = min(cf + yo, cm.M - 1);
    my[yo] = (yo < cn && cf + yo != v) ? e.mask[ys[yo]] : 0;
    t[yo] = tsc[yo];
    if (yo < cn && cf + yo == v) selfo = yo;
  }
  float pJ = 0.f, pR = 0.f;
  bool have = false;
  constexpr int TCH = Tch<INSIDE>::n;
  for (int jp0 = 0; jp0 < nj; jp0 += TCH) {
    float a1[TCH][MAXC], r1[TCH][MAXC], a0[TCH][MAXC], l0[TCH][MAXC];
    bool k1[TCH][MAXC], k0[TCH][MAXC], on[TCH];
    int cell[TCH], x[TCH];
#pragma unroll
    for (int q = 0; q < TCH; q++) {
      const int jp = jp0 + q, j = jmn + jp, d = j - i + 1;
      const int h = ho + (jp < nj ? jp : 0);
      const int lo = e.hdmin[h], hi = e.hdmax[h];
      on[q] = jp < nj && d >= lo && d <= hi;
      cell[q] = e.roff[h] + (d - lo);
      x[q] = gmem(e.dsq)[on[q] ? j : i];
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++) {
        const RowRef rp = row_ref(e, ys[yo], j - 1), rq = row_ref(e, ys[yo], j);
        k1[q][yo] = on[q] && my[yo] && in_row(rp, d - 1);
        k0[q][yo] = on[q] && my[yo] && in_row(rq, d);
        const int i1 = k1[q][yo] ? rp.base + d - 1 : 0, i0 = k0[q][yo] ? rq.base + d : 0;
        a1[q][yo] = k1[q][yo] ? A[i1] : IMPOSSIBLE;
        r1[q][yo] = (k1[q][yo] && dR && (my[yo] & BR)) ? A[2 * nc + i1] : IMPOSSIBLE;
        a0[q][yo] = k0[q][yo] ? A[i0] : IMPOSSIBLE;
        l0[q][yo] = (k0[q][yo] && dL && (my[yo] & BL)) ? A[nc + i0] : IMPOSSIBLE;
      }
    }
#pragma unroll
    for (int q = 0; q < TCH; q++) {
      if (!on[q]) {
        have = false;
        continue;
      }
      const int d = jmn + jp0 + q - i + 1;
      float J = tr_init(cm, v, d, 1), L = tr_init(cm, v, d, 0), R = tr_init(cm, v, d, 1);
      if (dJ || dR) {
#pragma unroll
        for (int yo = 0; yo < MAXC; yo++) {
          if (yo == selfo) {
            if (have) {  // (j-1, d-1) of this state
              if (dJ) J = tcomb<INSIDE>(J, pJ + t[yo], ftbl);
              if (dR) R = tcomb<INSIDE>(R, pR + t[yo], ftbl);
            }
            continue;
          }
          if (!(my[yo] & (BJ | BR)) || !k1[q][yo]) continue;
          if (dJ && (my[yo] & BJ)) J = tcomb<INSIDE>(J, a1[q][yo] + t[yo], ftbl);
          if (dR && (my[yo] & BR)) R = tcomb<INSIDE>(R, r1[q][yo] + t[yo], ftbl);
        }
        const float em = oe[x[q]];
        J = clampi(J + em);
        R = clampi(d >= 2 ? R + em : em);
      }
      if (dL) {
#pragma unroll
        for (int yo = 0; yo < MAXC; yo++) {
          if (!(my[yo] & (BJ | BL)) || !k0[q][yo]) continue;
          if (my[yo] & BJ) L = tcomb<INSIDE>(L, a0[q][yo] + t[yo], ftbl);
          if (my[yo] & BL) L = tcomb<INSIDE>(L, l0[q][yo] + t[yo], ftbl);
        }
      }
      if (dJ) A[cell[q]] = J;
      if (dL) A[nc + cell[q]] = L;
      if (dR) A[2 * nc + cell[q]] = R;
      pJ = J;
      pR = R;
      have = true;
    }
  }
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
  IXG_GLOBAL float *A = gmem(e.alpha);
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
__global__ __launch_bounds__(TRHB_THREADS) void trhb_scan_kernel(CmDev cm, const EnvDev *envs, const float *ftbl,
                                                               const float *pty, const int32_t *ylist, int ny) {
  extern __shared__ int32_t smem[];
  EnvDev e = envs[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  hb_stage_scalars(e, cm.M, smem);
  auto unit = [&](int v, int u) {
    const int st = cm.sttype[v];
    const int jmn = e.jmin[v], nj = e.jmax[v] - jmn + 1;
    const int ho = e.hd_off[v];
    const int32_t *hmn = e.hdmin + ho, *ro = e.roff + ho;
    if (st == IL_ST) {
      tr_il_row<INSIDE>(cm, e, v, u, ftbl);
    } else if (st == IR_ST) {
      tr_ir_diag<INSIDE>(cm, e, v, u, ftbl);
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
  const IXG_GLOBAL float *A = gmem(e.alpha);
  // root cells in runs of RCH consecutive d of one row: each begin state's row
  // is looked up once per run; per cell the states go in list order with the
  // reference's strict comparisons, as cell by cell
  constexpr int RCH = 8;
  const int cpr = (e.j0 - e.i0 + 2 + RCH - 1) / RCH;  // runs per row: d <= j0 - i0 + 1
  (void)ncv;
  for (int id = tid; id < nj * cpr; id += nt) {
    const int r = id / cpr, k0 = (id - r * cpr) * RCH;
    const int lo = hmn[r], hi = hmx[r], d0 = lo + k0;
    if (d0 > hi) continue;
    const int n = min(RCH, hi - d0 + 1);
    const int j = jmn + r, c0 = ro[r] - ro[0] + k0;
    float J[RCH], L[RCH], R[RCH], T[RCH], best[RCH];
    int bm[RCH], br[RCH];
#pragma unroll
    for (int t = 0; t < RCH; t++) {
      J[t] = L[t] = R[t] = T[t] = best[t] = IMPOSSIBLE;
      bm[t] = MODE_UNKNOWN;
      br[t] = 0;
    }
    for (int q = 0; q < ny; q++) {
      const int y = ylist[q];
      const int jmy = e.jmin[y];
      if (j < jmy || j > e.jmax[y]) continue;
      const int h = e.hd_off[y] + (j - jmy);
      const int ylo = e.hdmin[h], yhi = e.hdmax[h], yro = e.roff[h];
      const int my = e.mask[y];
      const float p = pty[y];
      const bool bt = (m0 & BT) && (my & BT) && cm.sttype[y] == B_ST;
#pragma unroll
      for (int t = 0; t < RCH; t++) {
        const int d = d0 + t;
        if (t >= n || d < ylo || d > yhi) continue;
        const int cy = yro + (d - ylo);
        if ((m0 & BJ) && (my & BJ)) {
          const float s = A[cy] + p;
          if (s > J[t]) {
            J[t] = s;
            if (s > best[t]) best[t] = s, bm[t] = MODE_J, br[t] = y;
          }
        }
        if ((m0 & BL) && (my & BL)) {
          const float s = A[nc + cy] + p;
          if (s > L[t]) {
            L[t] = s;
            if (s > best[t]) best[t] = s, bm[t] = MODE_L, br[t] = y;
          }
        }
        if ((m0 & BR) && (my & BR)) {
          const float s = A[2 * nc + cy] + p;
          if (s > R[t]) {
            R[t] = s;
            if (s > best[t]) best[t] = s, bm[t] = MODE_R, br[t] = y;
          }
        }
        if (bt) {
          const float s = A[3 * nc + cy] + p;
          if (s > T[t]) {
            T[t] = s;
            if (s > best[t]) best[t] = s, bm[t] = MODE_T, br[t] = y;
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < RCH; t++) {
      if (t >= n) continue;
      const int c = c0 + t;
      gmem(e.rout)[c] = J[t];
      gmem(e.rout)[ncv + c] = L[t];
      gmem(e.rout)[2 * ncv + c] = R[t];
      gmem(e.rout)[3 * ncv + c] = T[t];
      gmem(e.rout)[4 * ncv + c] = __int_as_float(br[t] * 8 + bm[t]);
    }
  }
}

}  // namespace

hipError_t launch_trhb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, const float *pty,
                            const int32_t *ylist, int ny, int inside, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t lds = hb_lds_bytes(cm.M);
  if (inside)
    hipLaunchKernelGGL(trhb_scan_kernel<true>, dim3((unsigned)n), dim3(TRHB_THREADS), lds, s, cm, envs, ftbl, pty,
                       ylist, ny);
  else
    hipLaunchKernelGGL(trhb_scan_kernel<false>, dim3((unsigned)n), dim3(TRHB_THREADS), lds, s, cm, envs, ftbl, pty,
                       ylist, ny);
  return hipGetLastError();
}

}  // namespace ixg
