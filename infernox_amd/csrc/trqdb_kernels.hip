// gfx950 QDB / non-banded truncated scanners: truncated CYK on float
// scores and truncated Inside on integer scores with the ILogsum table
// (cm_dpsearch_trunc.rs:1800-2316 generic_tr_scan). One workgroup per
// window runs the same wavefront as the standard scanners: at step t each
// state v fills the four marginal planes J/L/R/T of its cells at end position
// j = i0 + t - lvl[v] (d = 1..min(jrow, dmax, W)); every read goes to lower
// levels at j or to j - 1, written at earlier steps, so one barrier per step
// suffices. Rows live in per-state rings (four planes per row) deep enough
// for their last reader; the states the pass gives a truncated begin keep
// their rows until the root has read them at the same j. BEGL_S decks keep
// J/L/R over w+1 rows plus the steps between writer and bifurcation. An IL
// state's row is one lane's serial walk along its self transition. Children
// are folded in increasing order, the bifurcation and pair-state marginal
// cells take the maximum of their J and L (or R) terms exactly where the
// reference does, and the root's two parity rows are never reset between
// j's: a truncated begin must beat what the same-parity row holds at d. The
// root scans the begin states in increasing order, as the reference does,
// and keeps per cell the best begin (score, state, mode); cells at or above
// the reporting cutoff go to a candidate list, and the window's best root
// score, its mode and the full-window best to a summary. Greedy reporting,
// truncation legality, null3 and overlap removal run on the host.
#include <type_traits>
#include "cm_internal.h"

namespace ixg {

namespace {

constexpr int MODE_T = 0, MODE_R = 1, MODE_L = 2, MODE_J = 3, MODE_UNKNOWN = 4;

template <typename T>
struct Tq;
template <>
struct Tq<float> {
  static __device__ __forceinline__ float zero() { return IMPOSSIBLE; }
  static __device__ __forceinline__ float comb(float a, float b, const uint16_t *) { return a >= b ? a : b; }
  static __device__ __forceinline__ float max2(float a, float b) { return a >= b ? a : b; }
  static __device__ __forceinline__ float add(float a, float b) { return a + b; }
  static __device__ __forceinline__ float bits(float a) { return a; }
  static __device__ __forceinline__ bool possible(float a) { return a > IMPOSSIBLE + 1.0f; }
};
template <>
struct Tq<int32_t> {
  static __device__ __forceinline__ int32_t zero() { return -INFTY_I; }
  // ILogsum (cp9.rs:62-74), branch-free: the table (in LDS, its entries
  // 0..1000 held as 16-bit) is read at 0 when unused
  static __device__ __forceinline__ int32_t comb(int32_t a, int32_t b, const uint16_t *tbl) {
    const int32_t mx = max(-INFTY_I, max(a, b)), mn = min(a, b);
    const int32_t diff = mx - mn;
    const bool use = mn > -INFTY_I && diff < FLOGSUM_TBL;
    const int32_t t = tbl[use ? diff : 0];
    return use ? mx + t : mx;
  }
  static __device__ __forceinline__ int32_t max2(int32_t a, int32_t b) { return a >= b ? a : b; }
  static __device__ __forceinline__ int32_t add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
  static __device__ __forceinline__ float bits(int32_t a) {  // Scorify (cm_nohmm.rs:37)
    return a == -INFTY_I ? IMPOSSIBLE : (float)((double)a / 1000.0);
  }
  static __device__ __forceinline__ bool possible(int32_t a) { return a != -INFTY_I; }
};

// first cell of plane 0 of the row of end position j in a packed ring (four planes per row)
__device__ __forceinline__ size_t tring_at(int32_t ring, int j, int W1) {
  const int base = ring & ((1 << RING_SHIFT) - 1), mask = (1 << (ring >> RING_SHIFT)) - 1;
  return (size_t)(base + (j & mask)) * 4 * W1;
}

// The window's residues: staged in LDS when the window fits, else read from
// the device sequence.
constexpr int TRQ_SEQ_LDS = 32768;
// Plane rows and decks, read from a window descriptor, in global memory.
template <class T>
using Gp = IXG_GLOBAL T;

template <bool LDS>
struct TSeq {
  const Gp<uint8_t> *g;
  const uint8_t *l;
  int i0;
  __device__ __forceinline__ int operator[](int i) const { return LDS ? (int)l[i - i0] : (int)g[i]; }
};

__device__ __forceinline__ QdbUnit load_unit(const QdbUnit *u) {
  const int4 *p = (const int4 *)u;
  QdbUnit r;
  int4 *o = (int4 *)&r;
  o[0] = p[0];
  o[1] = p[1];
  o[2] = p[2];
  o[3] = p[3];
  return r;
}

template <typename T>
__device__ __forceinline__ T tinit(const TrqTabs<T> &tb, int v, int dp) {
  const T es = tb.endsc[v];
  return Tq<T>::possible(es) ? Tq<T>::add(tb.el_sca[dp], es) : Tq<T>::zero();
}

__device__ __forceinline__ int t_sd(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }

// One cell (j, d) of a state other than IL and the root, every filled plane
// (cm_dpsearch_trunc.rs:1978-2217). P: plane stride (W1); cj / cp: each
// child's row of j and of j - 1.
template <typename T, bool LS>
__device__ void trq_cell(const TrqDev &q, const TrqTabs<T> &tb, const uint16_t *tbl, Gp<T> *__restrict__ rows,
                         Gp<T> *__restrict__ dk, const TSeq<LS> &dsq, int W1, const QdbUnit &u, int st,
                         int cn, int j, int d) {
  using R = Tq<T>;
  const bool fl = q.fill & 2, fr = q.fill & 4, ft = q.fill & 8;
  const int v = u.code & 0x0FFFFFFF;
  const int sd = t_sd(st);
  const T z = R::zero();
  const size_t P = W1;
  Gp<T> *out = rows + tring_at(u.ring, j, W1) + d;
  if (st == B_ST) {
    const int kmin = u.k0, kmax = u.k1, DR = W1 + u.cring[2];
    const Gp<T> *db = dk + ((size_t)u.deck * W1 + u.cring[1]) * 3 * P;
    const size_t rs = 3 * P;  // deck row stride
    const Gp<T> *yj = rows + tring_at(u.cring[0], j, W1), *yl = yj + P;
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z, tsc = ft ? jsc : z;
    int jr = (j - kmin) % DR;  // deck row of j - k (j - kmin >= i0 - 1 >= 0)
#pragma unroll 4
    for (int k = kmin; k <= kmax; k++) {
      const Gp<T> *dr = db + jr * rs + (d - k);
      const T lft = dr[0];
      jsc = R::comb(jsc, R::add(lft, yj[k]), tbl);
      if (fl) lsc = R::comb(lsc, R::add(lft, yl[k]), tbl);
      if (fr) rsc = R::comb(rsc, R::add(dr[2 * P], yj[k]), tbl);
      jr = jr == 0 ? DR - 1 : jr - 1;
    }
    if (ft) {
      const int k0 = max(1, kmin);
      jr = (j - k0) % DR;
      for (int k = k0; k <= min(d - 1, kmax); k++) {
        tsc = R::comb(tsc, R::add(db[jr * rs + 2 * P + (d - k)], yl[k]), tbl);
        jr = jr == 0 ? DR - 1 : jr - 1;
      }
    }
    out[0] = jsc;
    if (ft) out[3 * P] = tsc;
    const Gp<T> *dj = db + (size_t)(j % DR) * rs + d;
    if (fl) out[P] = kmin == 0 ? R::comb(lsc, R::max2(dj[0], dj[P]), tbl) : lsc;
    if (fr) out[2 * P] = kmax == d ? R::comb(rsc, R::max2(yj[d], yj[2 * P + d]), tbl) : rsc;
    return;
  }
  const T *ts = tb.tsc + v * MAXC;
  const Gp<T> *cj[MAXC], *cp[MAXC];  // children's rows of j and j - 1
  T tsr[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) {
      cj[yo] = rows + tring_at(u.cring[yo], j, W1);
      cp[yo] = rows + tring_at(u.cring[yo], j - 1, W1);
      tsr[yo] = ts[yo];
    }
  if (u.deck >= 0) {  // BEGL_S: into its J/L/R decks (:2046-2067)
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z;
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) {
        jsc = R::comb(jsc, R::add(cj[yo][d - sd], tsr[yo]), tbl);
        if (fl) lsc = R::comb(lsc, R::add(cj[yo][P + d - sd], tsr[yo]), tbl);
        if (fr) rsc = R::comb(rsc, R::add(cj[yo][2 * P + d - sd], tsr[yo]), tbl);
      }
    Gp<T> *dr = dk + ((size_t)u.deck * W1 + u.k0 + j % (W1 + u.k1)) * 3 * P + d;
    dr[0] = jsc;
    if (fl) dr[P] = lsc;
    if (fr) dr[2 * P] = rsc;
    return;
  }
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) {  // EMITLEFT, non-self (:2068-2106)
    const T e = oe[dsq[i]];
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z;
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) {
        jsc = R::comb(jsc, R::add(cj[yo][d - sd], tsr[yo]), tbl);
        if (fl) lsc = R::comb(lsc, R::add(cj[yo][P + d - sd], tsr[yo]), tbl);
      }
    out[0] = R::add(jsc, e);
    if (fl) out[P] = d >= 2 ? R::add(lsc, e) : e;
    if (fr) {
      T rsc = tinit(tb, v, d);
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++)
        if (yo < cn) rsc = R::comb(rsc, R::max2(R::add(cj[yo][d], tsr[yo]), R::add(cj[yo][2 * P + d], tsr[yo])), tbl);
      out[2 * P] = rsc;
    }
  } else if (st == MR_ST || st == IR_ST) {  // EMITRIGHT (:2107-2141)
    const T esc_j = oe[dsq[j]];
    T jsc = tinit(tb, v, d - sd);
    T rsc = fr ? jsc : z;
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) {
        jsc = R::comb(jsc, R::add(cp[yo][d - sd], tsr[yo]), tbl);
        if (fr) rsc = R::comb(rsc, R::add(cp[yo][2 * P + d - sd], tsr[yo]), tbl);
      }
    out[0] = R::add(jsc, esc_j);
    if (fr) out[2 * P] = d >= 2 ? R::add(rsc, esc_j) : esc_j;
    if (fl) {
      T lsc = tinit(tb, v, d);
      const int y1 = st == IR_ST ? 1 : 0;
#pragma unroll
      for (int yo = 0; yo < MAXC; yo++)
        if (yo >= y1 && yo < cn)
          lsc = R::comb(lsc, R::max2(R::add(cj[yo][d], tsr[yo]), R::add(cj[yo][P + d], tsr[yo])), tbl);
      out[P] = lsc;
    }
  } else if (st == MP_ST) {  // EMITPAIR (:2142-2190)
    T jsc = d >= 2 ? tinit(tb, v, d - sd) : z;
    T lsc = fl ? tinit(tb, v, d - 1) : z, rsc = fr ? tinit(tb, v, d - 1) : z;
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) {
        if (d >= 2) jsc = R::comb(jsc, R::add(cp[yo][d - 2], tsr[yo]), tbl);
        if (fl) lsc = R::comb(lsc, R::max2(R::add(cj[yo][d - 1], tsr[yo]), R::add(cj[yo][P + d - 1], tsr[yo])), tbl);
        if (fr) rsc = R::comb(rsc, R::max2(R::add(cp[yo][d - 1], tsr[yo]), R::add(cp[yo][2 * P + d - 1], tsr[yo])), tbl);
      }
    out[0] = d >= 2 ? R::add(jsc, oe[dsq[i] * KP + dsq[j]]) : z;
    const T lm = tb.lmesc[v * KP + dsq[i]], rm = tb.rmesc[v * KP + dsq[j]];
    if (fl) out[P] = d >= 2 ? R::add(lsc, lm) : lm;
    if (fr) out[2 * P] = d >= 2 ? R::add(rsc, rm) : rm;
  } else {  // EMITNONE: D, S (:2191-2217)
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z;
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) {
        jsc = R::comb(jsc, R::add(cj[yo][d - sd], tsr[yo]), tbl);
        if (fl) lsc = R::comb(lsc, R::add(cj[yo][P + d - sd], tsr[yo]), tbl);
        if (fr) rsc = R::comb(rsc, R::add(cj[yo][2 * P + d - sd], tsr[yo]), tbl);
      }
    out[0] = jsc;
    if (fl) out[P] = lsc;
    if (fr) out[2 * P] = rsc;
  }
}

// The row d = 1..dx of an IL state at j: its J and L cells read the same
// state at (j, d-1), carried in registers along the row (:2068-2106). The R
// plane does not read the state itself and is filled first, with no
// dependence between d's; the J/L walk then fetches four d's of its other
// inputs at a time.
template <typename T, bool LS>
__device__ void trq_il_row(const TrqDev &q, const TrqTabs<T> &tb, const uint16_t *tbl, Gp<T> *__restrict__ rows,
                           const TSeq<LS> &dsq, int W1, const QdbUnit &u, int cn, int j, int dx) {
  using R = Tq<T>;
  constexpr int B = 4;
  const bool fl = q.fill & 2, fr = q.fill & 4;
  const T z = R::zero();
  const int v = u.code & 0x0FFFFFFF;
  const size_t P = W1;
  const T *ts = tb.tsc + v * MAXC;
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  Gp<T> *out = rows + tring_at(u.ring, j, W1);
  const Gp<T> *cj[MAXC];
  T tsr[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) {
      cj[yo] = rows + tring_at(u.cring[yo], j, W1);
      tsr[yo] = ts[yo];
    }
  if (fr)
    for (int d0 = 1; d0 <= dx; d0 += B) {
      T a[MAXC][B], b[MAXC][B], in[B];
#pragma unroll
      for (int k = 0; k < B; k++) {
        const int d = min(d0 + k, dx);
#pragma unroll
        for (int yo = 1; yo < MAXC; yo++)
          if (yo < cn) {
            a[yo][k] = cj[yo][d];
            b[yo][k] = cj[yo][2 * P + d];
          }
        in[k] = tinit(tb, v, d);
      }
#pragma unroll
      for (int k = 0; k < B; k++)
        if (d0 + k <= dx) {
          T rsc = in[k];
#pragma unroll
          for (int yo = 1; yo < MAXC; yo++)
            if (yo < cn) rsc = R::comb(rsc, R::max2(R::add(a[yo][k], tsr[yo]), R::add(b[yo][k], tsr[yo])), tbl);
          out[2 * P + d0 + k] = rsc;
        }
    }
  T pj = out[0], pl = out[P];  // the state's own J and L at (j, d-1)
  for (int d0 = 1; d0 <= dx; d0 += B) {
    T cJ[MAXC][B], cL[MAXC][B], in[B], e[B];
    int x[B];
#pragma unroll
    for (int k = 0; k < B; k++) {
      const int d = min(d0 + k, dx);
#pragma unroll
      for (int yo = 1; yo < MAXC; yo++)
        if (yo < cn) {
          cJ[yo][k] = cj[yo][d - 1];
          if (fl) cL[yo][k] = cj[yo][P + d - 1];
        }
      in[k] = tinit(tb, v, d - 1);
      x[k] = dsq[j - d + 1];
    }
#pragma unroll
    for (int k = 0; k < B; k++) e[k] = oe[x[k]];
#pragma unroll
    for (int k = 0; k < B; k++)
      if (d0 + k <= dx) {
        const int d = d0 + k;
        T jsc = in[k];
        T lsc = fl ? jsc : z;
        jsc = R::comb(jsc, R::add(pj, tsr[0]), tbl);
        if (fl) lsc = R::comb(lsc, R::add(pl, tsr[0]), tbl);
#pragma unroll
        for (int yo = 1; yo < MAXC; yo++)
          if (yo < cn) {
            jsc = R::comb(jsc, R::add(cJ[yo][k], tsr[yo]), tbl);
            if (fl) lsc = R::comb(lsc, R::add(cL[yo][k], tsr[yo]), tbl);
          }
        pj = R::add(jsc, e[k]);
        out[d] = pj;
        if (fl) {
          pl = d >= 2 ? R::add(lsc, e[k]) : e[k];
          out[P + d] = pl;
        }
      }
  }
}

template <typename T, bool LS>
__global__ __launch_bounds__(QDB_MAX_THREADS) void trqdb_kernel(TrqDev q, TrqTabs<T> tb, const QdbWinDev *wins,
                                                                float cutoff, TrqCand *cand, int64_t cap,
                                                                unsigned long long *counter, TrqSummary *summ) {
  using R = Tq<T>;
  __shared__ float s_sc[QDB_MAX_THREADS];
  __shared__ int s_j[QDB_MAX_THREADS], s_d[QDB_MAX_THREADS], s_m[QDB_MAX_THREADS];
  __shared__ float s_full;
  const QdbWinDev wd = wins[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const int W1 = wd.w + 1;
  const size_t P = W1;
  Gp<T> *rows = gmem((T *)wd.rows), *dk = gmem((T *)wd.begl);
  TSeq<LS> dsq;
  dsq.g = gmem(wd.dsq);
  dsq.i0 = wd.i0;
  dsq.l = nullptr;
  if constexpr (LS) {
    __shared__ uint8_t s_seq[TRQ_SEQ_LDS];
    for (int x = tid; x < wd.j0 - wd.i0 + 1; x += nt) s_seq[x] = wd.dsq[wd.i0 + x];
    dsq.l = s_seq;
  }
  const uint16_t *tbl = nullptr;
  if constexpr (!std::is_same<T, float>::value) {
    __shared__ uint16_t s_ilog[FLOGSUM_TBL];
    for (int x = tid; x < FLOGSUM_TBL; x += nt) s_ilog[x] = (uint16_t)q.ilog[x];
    tbl = s_ilog;
  }
  // every cell starts at the semiring's zero; d = 0 at the base cases, set
  // once (:1873-1922): J of non-BEGL states, L and R of E states, the J decks
  for (int r = wave; r < 4 * q.nrows; r += nt >> 6) {
    const int v = q.row_state[r >> 2], pl = r & 3;
    T b0 = R::zero();
    if (q.begl[v] < 0) {
      if (pl == 0) b0 = tb.base0[v];
      else if ((pl == 1 || pl == 2) && q.sttype[v] == E_ST) b0 = (T)0;
    }
    for (int d = lane; d < W1; d += 64) rows[(size_t)r * W1 + d] = d == 0 ? b0 : R::zero();
  }
  for (int b = 0; b < q.nbegl; b++) {
    const int2 k = q.deck[b];
    Gp<T> *db = dk + ((size_t)b * W1 + k.x) * 3 * P;
    for (int r = wave; r < 3 * (W1 + k.y); r += nt >> 6)
      for (int d = lane; d < W1; d += 64) db[(size_t)r * W1 + d] = (r % 3 == 0 && d == 0) ? tb.bbase0[b] : R::zero();
  }
  if (tid == 0) s_full = IMPOSSIBLE;
  __syncthreads();
  const bool fl = q.fill & 2, fr = q.fill & 4, ft = q.fill & 8;
  float best_sc = IMPOSSIBLE;
  int best_j = 0, best_d = 0, best_m = MODE_UNKNOWN;
  const int L = wd.j0 - wd.i0 + 1, nsteps = L + q.lroot;
  const int ib = wave < q.nw ? q.wave_item[wave] : 0, ie = wave < q.nw ? q.wave_item[wave + 1] : 0;
  for (int t = 0; t < nsteps; t++) {
    // items are in increasing lowest level: stop at the first one whose
    // lowest level is above t
    QdbUnit nxt;
    if (ib < ie) nxt = load_unit(q.unit + (size_t)ib * 64 + lane);
    for (int it = ib; it < ie; it++) {
      const int4 im = q.item[it];
      if (im.x > t) break;
      const QdbUnit u = nxt;
      if (it + 1 < ie) nxt = load_unit(q.unit + (size_t)(it + 1) * 64 + lane);
      if (im.y < t - (L - 1) || im.z > min(t - im.x + 1, wd.w)) continue;
      const int kind = u.code >> 28, d = u.d;
      if (kind == QU_NONE) continue;
      const int j = wd.i0 + t - u.lvl;
      if (j < wd.i0 || j > wd.j0) continue;
      const int jrow = min(j - wd.i0 + 1, wd.w);
      const int st = u.info & 15, cn = (u.info >> 4) & 15;
      if (kind == QU_IL) {
        const int dx = min(jrow, u.y1);
        if (dx >= 1) trq_il_row<T, LS>(q, tb, tbl, rows, dsq, W1, u, cn, j, dx);
        continue;
      }
      if (d > jrow) continue;
      if (kind == QU_CELL) {
        trq_cell<T, LS>(q, tb, tbl, rows, dk, dsq, W1, u, st, cn, j, d);
        continue;
      }
      // the root: truncated begins into the persistent root rows (:2220-2289),
      // the begin states in increasing order, the row held in registers
      float bsc = IMPOSSIBLE;
      int br = 0, bm = MODE_UNKNOWN;
      Gp<T> *root = rows + tring_at(u.ring, j, W1) + d;
      T rt[4] = {root[0], root[P], root[2 * P], root[3 * P]};
      constexpr int TB = 8;  // begin states whose cells are fetched together
      for (int e0 = 0; e0 < q.ntb; e0 += TB) {
        T val[TB][4];
        int4 tev[TB];
#pragma unroll
        for (int k = 0; k < TB; k++) {
          tev[k] = q.tblist[min(e0 + k, q.ntb - 1)];
          const Gp<T> *yc = rows + tring_at(tev[k].x, j, W1) + d;
#pragma unroll
          for (int pl = 0; pl < 4; pl++) val[k][pl] = yc[pl * P];
        }
#pragma unroll
        for (int k = 0; k < TB; k++) {
          const int4 te = tev[k];
          if (e0 + k >= q.ntb || d > min(jrow, te.z)) continue;
          T tp;
          memcpy(&tp, &te.y, 4);
          const int ny = (ft && (te.w >> 30)) ? 4 : 3;
#pragma unroll
          for (int pl = 0; pl < 4; pl++) {
            if (pl >= ny || (pl == 1 && !fl) || (pl == 2 && !fr)) continue;
            const T sc = R::add(val[k][pl], tp);
            if (sc > rt[pl]) {
              rt[pl] = sc;
              const float fsc = R::bits(sc);
              if (fsc > bsc) {
                bsc = fsc;
                bm = pl == 0 ? MODE_J : pl == 1 ? MODE_L : pl == 2 ? MODE_R : MODE_T;
                br = te.w & 0x3FFFFFFF;
              }
            }
          }
        }
      }
      root[0] = rt[0];
      root[P] = rt[1];
      root[2 * P] = rt[2];
      root[3 * P] = rt[3];
      if (bsc > best_sc || (bsc == best_sc && bsc > IMPOSSIBLE && (j < best_j || (j == best_j && d < best_d)))) {
        best_sc = bsc;
        best_j = j;
        best_d = d;
        best_m = bm;
      }
      if (j == wd.j0 && d == j && j <= wd.w) s_full = bsc;
      if (bsc >= cutoff && bsc > IMPOSSIBLE + 1.0f) {
        const unsigned long long k = atomicAdd(counter, 1ull);
        if ((int64_t)k < cap) {
          TrqCand c;
          c.win = wd.win;
          c.j = j;
          c.d = d;
          c.sc = bsc;
          c.root = br;
          c.mode = bm;
          cand[k] = c;
        }
      }
    }
    __syncthreads();
  }
  // the window's best root score: the first maximum in (j, d) order
  s_sc[tid] = best_sc;
  s_j[tid] = best_j;
  s_d[tid] = best_d;
  s_m[tid] = best_m;
  __syncthreads();
  if (tid == 0) {
    float b = IMPOSSIBLE;
    int bj = 0, bd = 0, bm = MODE_UNKNOWN;
    for (int t = 0; t < nt; t++) {
      const bool better = s_sc[t] > b || (s_sc[t] == b && b > IMPOSSIBLE && (s_j[t] < bj || (s_j[t] == bj && s_d[t] < bd)));
      if (better) {
        b = s_sc[t];
        bj = s_j[t];
        bd = s_d[t];
        bm = s_m[t];
      }
    }
    TrqSummary sm;
    sm.vsc = b;
    sm.vmode = b > IMPOSSIBLE ? bm : MODE_UNKNOWN;
    sm.bsc_full = s_full;
    summ[wd.win] = sm;
  }
}

}  // namespace

hipError_t launch_trqdb_scan(const TrqDev &q, const void *tabs, int inside, const QdbWinDev *wins, int64_t n,
                             float cutoff, TrqCand *cand, int64_t cap, unsigned long long *counter, TrqSummary *summ,
                             int maxlen, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)n), block(64 * q.nw);
  const bool ls = maxlen <= TRQ_SEQ_LDS;
  if (inside) {
    const TrqTabs<int32_t> &t = *(const TrqTabs<int32_t> *)tabs;
    if (ls)
      hipLaunchKernelGGL((trqdb_kernel<int32_t, true>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter, summ);
    else
      hipLaunchKernelGGL((trqdb_kernel<int32_t, false>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter, summ);
  } else {
    const TrqTabs<float> &t = *(const TrqTabs<float> *)tabs;
    if (ls)
      hipLaunchKernelGGL((trqdb_kernel<float, true>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter, summ);
    else
      hipLaunchKernelGGL((trqdb_kernel<float, false>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter, summ);
  }
  return hipGetLastError();
}

}  // namespace ixg
