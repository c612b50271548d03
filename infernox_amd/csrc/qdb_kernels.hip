// gfx950 query-dependent-banded (or non-banded) scanners: CYK on float
// scores and Inside on integer scores with the ILogsum table
// (cm_nohmm.rs:406-767 generic_scan). One workgroup per window runs a
// wavefront over the model's dependency levels: at step t each state v fills
// its cells of end position j = i0 + t - lvl[v]. A state reads lower-level
// states at the same j, or itself and lower levels at j - 1, so everything a
// step reads was written at an earlier step and one barrier per step is the
// only synchronisation; per step every level is busy at once instead of one
// level per barrier. The host deals the step's (state, d) cells as 64-lane
// items to the waves. Rows live in per-state rings deep enough that a cell is
// read before its slot comes round again; cells outside a state's band are
// never written, so they keep their initial values exactly as in the
// reference's rolling rows. BEGL_S decks keep w+1 rows plus the steps between
// the deck's writer and its bifurcation. Local begins are combined by max
// into per-(j, slot, d) rows, one slot per begin state of a level, which the
// root reads and clears; max is order-free, so the result equals the
// reference's scan over the begin list. Children are combined in the
// reference's unrolled order (reduce_c_order), so both semirings are
// bit-identical to it. Root cells at or above the reporting cutoff are
// appended to a candidate list; greedy reporting, null3 and overlap removal
// run on the host.
#include <type_traits>
#include "cm_internal.h"

namespace ixg {

namespace {


__device__ __forceinline__ int q_sd(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }
__device__ __forceinline__ int q_sdr(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

// semiring operations (cm_nohmm.rs:88-185)
template <typename T>
struct Sr;
template <>
struct Sr<float> {
  static __device__ __forceinline__ float zero() { return IMPOSSIBLE; }
  static __device__ __forceinline__ float comb(float a, float b, const uint16_t *) { return a >= b ? a : b; }
  static __device__ __forceinline__ float add(float a, float b) { return a + b; }
  static __device__ __forceinline__ float bits(float a) { return a; }
  static __device__ __forceinline__ bool possible(float a) { return a > IMPOSSIBLE + 1.0f; }
};
template <>
struct Sr<int32_t> {
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
  static __device__ __forceinline__ int32_t add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
  static __device__ __forceinline__ float bits(int32_t a) {  // Scorify (cm_nohmm.rs:37)
    return a == -INFTY_I ? IMPOSSIBLE : (float)((double)a / 1000.0);
  }
  static __device__ __forceinline__ bool possible(int32_t a) { return a != -INFTY_I; }
};

// cm_nohmm.rs:187-244 reduce_c_order: init and the child terms folded in C's
// hand-unrolled order
template <typename T>
__device__ __forceinline__ T reduce_c(int cn, const T *t, T init, const uint16_t *tbl) {
  using R = Sr<T>;
  T s;
  switch (cn) {
    case 1: return R::comb(init, t[0], tbl);
    case 2: s = R::comb(t[1], init, tbl); return R::comb(s, t[0], tbl);
    case 3:
      s = R::comb(t[2], t[1], tbl);
      s = R::comb(s, init, tbl);
      return R::comb(s, t[0], tbl);
    case 4:
      s = R::comb(t[3], t[2], tbl);
      s = R::comb(s, init, tbl);
      s = R::comb(s, t[1], tbl);
      return R::comb(s, t[0], tbl);
    case 5:
      s = R::comb(t[4], t[3], tbl);
      s = R::comb(s, init, tbl);
      s = R::comb(s, t[1], tbl);
      s = R::comb(s, t[2], tbl);
      return R::comb(s, t[0], tbl);
    case 6:
      s = R::comb(t[5], init, tbl);
      s = R::comb(s, t[4], tbl);
      s = R::comb(s, t[3], tbl);
      s = R::comb(s, t[2], tbl);
      s = R::comb(s, t[1], tbl);
      return R::comb(s, t[0], tbl);
    default:  // not reached: a state has 1..MAXC children
      s = init;
#pragma unroll
      for (int k = 0; k < MAXC; k++)
        if (k < cn) s = R::comb(s, t[k], tbl);
      return s;
  }
}

template <typename T>
__device__ __forceinline__ T init_sc(const QdbTabs<T> &tb, int v, int dp) {
  const T es = tb.endsc[v];
  return Sr<T>::possible(es) ? Sr<T>::add(tb.el_sca[dp], es) : Sr<T>::zero();
}

// first cell of the row of end position j in a packed ring
__device__ __forceinline__ size_t ring_at(int32_t ring, int j, int W1) {
  const int base = ring & ((1 << RING_SHIFT) - 1), mask = (1 << (ring >> RING_SHIFT)) - 1;
  return (size_t)(base + (j & mask)) * W1;
}

// The window's residues: staged in LDS when the window fits, else read from
// the device sequence.
constexpr int QDB_SEQ_LDS = 32768;
// Ring, deck and local-begin rows, read from a window descriptor, in global
// memory.
template <class T>
using Gp = IXG_GLOBAL T;

template <bool LDS>
struct Seq {
  const Gp<uint8_t> *g;
  const uint8_t *l;
  int i0;
  __device__ __forceinline__ int operator[](int i) const { return LDS ? (int)l[i - i0] : (int)g[i]; }
};

// an IL state's scratch row: the row just past its ring
__device__ __forceinline__ size_t ring_scratch(int32_t ring) {
  return (size_t)((ring & ((1 << RING_SHIFT) - 1)) + (1 << (ring >> RING_SHIFT)));
}

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

// One cell (j, d) of a state other than IL and the root; returns its value
// (for the local-begin rows).
template <typename T, bool LS>
__device__ __forceinline__ T qdb_cell(const QdbTabs<T> &tb, const uint16_t *tbl, Gp<T> *__restrict__ rows,
                                      Gp<T> *__restrict__ deck, const Seq<LS> &dsq, int W1,
                                      const QdbUnit &u, int st, int cn, int j, int d) {
  using R = Sr<T>;
  const int v = u.code & 0x0FFFFFFF;
  const int sd = q_sd(st);
  if (st == B_ST) {  // cm_nohmm.rs:589-606
    const int kmin = u.k0, kmax = u.k1, DR = W1 + u.cring[2];
    const Gp<T> *db = deck + ((size_t)u.deck * W1 + u.cring[1]) * W1;
    const Gp<T> *yr = rows + ring_at(u.cring[0], j, W1);
    T sc = init_sc(tb, v, d);
    int jr = (j - kmin) % DR;  // deck row of j - k (j - kmin >= i0 - 1 >= 0)
#pragma unroll 4
    for (int k = kmin; k <= kmax; k++) {
      sc = R::comb(sc, R::add(db[(size_t)jr * W1 + d - k], yr[k]), tbl);
      jr = jr == 0 ? DR - 1 : jr - 1;
    }
    rows[ring_at(u.ring, j, W1) + d] = sc;
    return sc;
  }
  const int jy = q_sdr(st) ? j - 1 : j;
  const T *ts = tb.tsc + v * MAXC;
  T terms[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) terms[yo] = R::add(rows[ring_at(u.cring[yo], jy, W1) + d - sd], ts[yo]);
  const T init = init_sc(tb, v, d - sd);
  if (u.deck >= 0) {  // BEGL_S: natural child order, into its deck (cm_nohmm.rs:607-616)
    T sc = init;
    for (int yo = 0; yo < cn; yo++) sc = R::comb(sc, terms[yo], tbl);
    deck[((size_t)u.deck * W1 + u.k0 + j % (W1 + u.k1)) * W1 + d] = sc;
    return sc;
  }
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  const int ei = st == ML_ST ? dsq[i] : st == MP_ST ? dsq[i] * KP + dsq[j] : dsq[j];
  T sc = reduce_c<T>(cn, terms, init, tbl);
  if (st == ML_ST || st == MR_ST || st == IR_ST || st == MP_ST) sc = R::add(sc, oe[ei]);
  rows[ring_at(u.ring, j, W1) + d] = sc;
  return sc;
}

// reduce_c_order with the first child left out: every unrolled order above
// combines t[0] last, so reduce_c(cn, t, init) == comb(reduce_c_rest(...), t[0]).
template <typename T>
__device__ __forceinline__ T reduce_c_rest(int cn, const T *t, T init, const uint16_t *tbl) {
  using R = Sr<T>;
  T s;
  switch (cn) {
    case 1: return init;
    case 2: return R::comb(t[1], init, tbl);
    case 3:
      s = R::comb(t[2], t[1], tbl);
      return R::comb(s, init, tbl);
    case 4:
      s = R::comb(t[3], t[2], tbl);
      s = R::comb(s, init, tbl);
      return R::comb(s, t[1], tbl);
    case 5:
      s = R::comb(t[4], t[3], tbl);
      s = R::comb(s, init, tbl);
      s = R::comb(s, t[1], tbl);
      return R::comb(s, t[2], tbl);
    default:  // 6
      s = R::comb(t[5], init, tbl);
      s = R::comb(s, t[4], tbl);
      s = R::comb(s, t[3], tbl);
      s = R::comb(s, t[2], tbl);
      return R::comb(s, t[1], tbl);
  }
}

// The d row dn..dx of an IL state at j (cm_nohmm.rs:617-647). Its first
// child is itself, read at (j, d-1), and reduce_c_order combines that term
// last; so a first pass, with no dependence between d's, folds the other
// children, the local end and the emission into the row and a scratch row,
// and a second pass adds the self term along d with nothing but that chain.
template <typename T, bool LS>
__device__ void qdb_il_row(const QdbTabs<T> &tb, const uint16_t *tbl, Gp<T> *__restrict__ rows, const Seq<LS> &dsq,
                           int W1, const QdbUnit &u, int cn, int j, int dn, int dx) {
  using R = Sr<T>;
  constexpr int B1 = 4, B2 = 8;  // d's whose loads are issued together
  const int v = u.code & 0x0FFFFFFF;
  const T *ts = tb.tsc + v * MAXC;
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  Gp<T> *out = rows + ring_at(u.ring, j, W1);
  Gp<T> *scr = rows + ring_scratch(u.ring) * W1;
  const Gp<T> *cr[MAXC];
  T tsr[MAXC];
#pragma unroll
  for (int yo = 1; yo < MAXC; yo++)
    if (yo < cn) {
      cr[yo] = rows + ring_at(u.cring[yo], j, W1);
      tsr[yo] = ts[yo];
    }
  const T es = tb.endsc[v], tself = ts[0];
  for (int d0 = dn; d0 <= dx; d0 += B1) {
    T c[MAXC][B1], el[B1], e[B1];
    int x[B1];
#pragma unroll
    for (int k = 0; k < B1; k++) {
      const int d = min(d0 + k, dx);
#pragma unroll
      for (int yo = 1; yo < MAXC; yo++)
        if (yo < cn) c[yo][k] = cr[yo][d - 1];
      el[k] = tb.el_sca[d - 1];
      x[k] = dsq[j - d + 1];
    }
#pragma unroll
    for (int k = 0; k < B1; k++) e[k] = oe[x[k]];
#pragma unroll
    for (int k = 0; k < B1; k++)
      if (d0 + k <= dx) {
        T terms[MAXC];
        terms[0] = R::zero();
#pragma unroll
        for (int yo = 1; yo < MAXC; yo++)
          if (yo < cn) terms[yo] = R::add(c[yo][k], tsr[yo]);
        const T init = R::possible(es) ? R::add(el[k], es) : R::zero();
        out[d0 + k] = reduce_c_rest<T>(cn, terms, init, tbl);
        scr[d0 + k] = e[k];
      }
  }
  T prev = out[dn - 1];
  for (int d0 = dn; d0 <= dx; d0 += B2) {
    T a[B2], e[B2];
#pragma unroll
    for (int k = 0; k < B2; k++) {
      const int d = min(d0 + k, dx);
      a[k] = out[d];
      e[k] = scr[d];
    }
#pragma unroll
    for (int k = 0; k < B2; k++)
      if (d0 + k <= dx) {
        const T sc = R::add(R::comb(a[k], R::add(prev, tself), tbl), e[k]);
        out[d0 + k] = sc;
        prev = sc;
      }
  }
}

template <typename T, bool LS>
__global__ __launch_bounds__(QDB_MAX_THREADS) void qdb_kernel(QdbDev q, QdbTabs<T> tb, const QdbWinDev *wins,
                                                              float cutoff, QdbCand *cand, int64_t cap,
                                                              unsigned long long *counter) {
  using R = Sr<T>;
  const QdbWinDev wd = wins[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const int W1 = wd.w + 1;
  Gp<T> *rows = gmem((T *)wd.rows);
  Gp<T> *deck = rows + (size_t)q.nrows * W1;
  Gp<T> *acc = deck + ((size_t)q.nbegl * W1 + q.deck_extra) * W1;
  Seq<LS> dsq;
  dsq.g = gmem(wd.dsq);
  dsq.i0 = wd.i0;
  dsq.l = nullptr;
  if constexpr (LS) {
    __shared__ uint8_t s_seq[QDB_SEQ_LDS];
    for (int x = tid; x < wd.j0 - wd.i0 + 1; x += nt) s_seq[x] = wd.dsq[wd.i0 + x];
    dsq.l = s_seq;
  }
  const uint16_t *tbl = nullptr;
  if constexpr (sizeof(T) == 4 && !std::is_same<T, float>::value) {
    __shared__ uint16_t s_ilog[FLOGSUM_TBL];
    for (int x = tid; x < FLOGSUM_TBL; x += nt) s_ilog[x] = (uint16_t)q.ilog[x];
    tbl = s_ilog;
  }
  // every cell starts at the semiring's zero, d = 0 at its base case
  // (cm_nohmm.rs:489-532)
  for (int r = wave; r < q.nrows; r += nt >> 6) {
    const int v = q.row_state[r];
    const T b0 = q.begl[v] < 0 ? tb.base0[v] : R::zero();
    for (int d = lane; d < W1; d += 64) rows[(size_t)r * W1 + d] = d == 0 ? b0 : R::zero();
  }
  for (int b = 0; b < q.nbegl; b++) {
    const int2 dk = q.deck[b];
    Gp<T> *db = deck + ((size_t)b * W1 + dk.x) * W1;
    for (int r = wave; r < W1 + dk.y; r += nt >> 6)
      for (int d = lane; d < W1; d += 64) db[(size_t)r * W1 + d] = d == 0 ? tb.bbase0[b] : R::zero();
  }
  if (q.local)
    for (int x = tid; x < (q.amask + 1) * q.nslot * W1; x += nt) acc[x] = R::zero();
  __syncthreads();
  const int L = wd.j0 - wd.i0 + 1, nsteps = L + q.lroot;
  const int ib = wave < q.nw ? q.wave_item[wave] : 0, ie = wave < q.nw ? q.wave_item[wave + 1] : 0;
  for (int t = 0; t < nsteps; t++) {
    // items are in increasing lowest level: the live ones at this step start
    // where the highest level reaches t - (L - 1) and end before lowest > t
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
        if (dx >= d) qdb_il_row<T, LS>(tb, tbl, rows, dsq, W1, u, cn, j, d, dx);
        continue;
      }
      if (d > jrow) continue;
      if (kind == QU_CELL) {
        const T sc = qdb_cell<T, LS>(tb, tbl, rows, deck, dsq, W1, u, st, cn, j, d);
        const int slot = (u.info >> 8) - 1;
        if (q.local && slot >= 0) {  // this begin's candidate for the root at (j, d)
          Gp<T> &a = acc[((size_t)(j & q.amask) * q.nslot + slot) * W1 + d];
          const T cnd = R::add(sc, tb.beginsc[u.code & 0x0FFFFFFF]);
          if (a < cnd) a = cnd;
        }
        continue;
      }
      // the root: its own transitions, then local begins by MAX (cm_nohmm.rs:686-746)
      T a = R::comb(R::zero(), R::add(rows[ring_at(u.cring[0], j, W1) + d], tb.tsc[0]), tbl);
#pragma unroll
      for (int yo = 1; yo < MAXC; yo++)
        if (yo < cn) a = R::comb(a, R::add(rows[ring_at(u.cring[yo], j, W1) + d], tb.tsc[yo]), tbl);
      if (q.local)
        for (int s = 0; s < q.nslot; s++) {
          Gp<T> &b = acc[((size_t)(j & q.amask) * q.nslot + s) * W1 + d];
          if (a < b) a = b;
          b = R::zero();
        }
      const float sc = R::bits(a);
      if (sc >= cutoff && sc > IMPOSSIBLE + 1.0f) {
        const unsigned long long k = atomicAdd(counter, 1ull);
        if ((int64_t)k < cap) {
          QdbCand c;
          c.win = wd.win;
          c.j = j;
          c.d = d;
          c.sc = sc;
          cand[k] = c;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_qdb_scan(const QdbDev &q, const void *tabs, int inside, const QdbWinDev *wins, int64_t n,
                           float cutoff, QdbCand *cand, int64_t cap, unsigned long long *counter, int maxlen,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)n), block(64 * q.nw);
  const bool ls = maxlen <= QDB_SEQ_LDS;
  if (inside) {
    const QdbTabs<int32_t> &t = *(const QdbTabs<int32_t> *)tabs;
    if (ls) hipLaunchKernelGGL((qdb_kernel<int32_t, true>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter);
    else hipLaunchKernelGGL((qdb_kernel<int32_t, false>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter);
  } else {
    const QdbTabs<float> &t = *(const QdbTabs<float> *)tabs;
    if (ls) hipLaunchKernelGGL((qdb_kernel<float, true>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter);
    else hipLaunchKernelGGL((qdb_kernel<float, false>), grid, block, 0, s, q, t, wins, cutoff, cand, cap, counter);
  }
  return hipGetLastError();
}

}  // namespace ixg
