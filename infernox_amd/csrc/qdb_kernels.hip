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
  static __device__ __forceinline__ float comb(float a, float b, const int32_t *) { return a >= b ? a : b; }
  static __device__ __forceinline__ float add(float a, float b) { return a + b; }
  static __device__ __forceinline__ float bits(float a) { return a; }
  static __device__ __forceinline__ bool possible(float a) { return a > IMPOSSIBLE + 1.0f; }
};
template <>
struct Sr<int32_t> {
  static __device__ __forceinline__ int32_t zero() { return -INFTY_I; }
  // ILogsum (cp9.rs:62-74), branch-free: the table is read at 0 when unused
  static __device__ __forceinline__ int32_t comb(int32_t a, int32_t b, const int32_t *tbl) {
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
__device__ __forceinline__ T reduce_c(int cn, const T *t, T init, const int32_t *tbl) {
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
    default:
      s = init;
      for (int k = 0; k < cn; k++) s = R::comb(s, t[k], tbl);
      return s;
  }
}

template <typename T>
__device__ __forceinline__ T init_sc(const QdbTabs<T> &tb, int v, int dp) {
  const T es = tb.endsc[v];
  return Sr<T>::possible(es) ? Sr<T>::add(tb.el_sca[dp], es) : Sr<T>::zero();
}

// the row of end position j in state v's ring
template <typename T>
__device__ __forceinline__ T *ring_row(T *rows, int2 ri, int j, int W1) {
  return rows + (size_t)(ri.x + (j & ri.y)) * W1;
}

struct QdbGeo {  // one window's arena
  int i0, j0, w, W1;
};

// One cell (j, d) of a state v other than IL and the root; returns its value
// (for the local-begin rows).
template <typename T>
__device__ __forceinline__ T qdb_cell(const QdbDev &q, const QdbTabs<T> &tb, T *rows, T *deck, const uint8_t *dsq,
                                      const QdbGeo &g, int v, int st, int y0, int cn, int j, int d) {
  using R = Sr<T>;
  const int32_t *tbl = q.ilog;
  const int W1 = g.W1;
  const int sd = q_sd(st);
  if (st == B_ST) {  // cm_nohmm.rs:589-606; cn holds the right child
    const int wl = y0, y = cn, bi = q.begl[wl];
    const int dn_y = q.dminw[y], dx_y = q.dxcap[y], dn_w = q.dminw[wl], dx_w = q.dxcap[wl];
    const int kmin = max(0, max(dn_y, d - dx_w)), kmax = min(dx_y, d - dn_w);
    const int2 dk = q.deck[bi];
    const int DR = W1 + dk.y;
    const T *db = deck + ((size_t)bi * W1 + dk.x) * W1;
    const T *yr = ring_row(rows, q.ring[y], j, W1);
    T sc = init_sc(tb, v, d);
    int jr = (j - kmin) % DR;  // deck row of j - k (j - kmin >= i0 - 1 >= 0)
    for (int k = kmin; k <= kmax; k++) {
      sc = R::comb(sc, R::add(db[(size_t)jr * W1 + d - k], yr[k]), tbl);
      jr = jr == 0 ? DR - 1 : jr - 1;
    }
    ring_row(rows, q.ring[v], j, W1)[d] = sc;
    return sc;
  }
  const int jy = q_sdr(st) ? j - 1 : j;
  const T *ts = tb.tsc + v * MAXC;
  const int bv = q.begl[v];
  if (bv >= 0) {  // BEGL_S: natural child order, into its deck (cm_nohmm.rs:607-616)
    T sc = init_sc(tb, v, d - sd);
    for (int yo = 0; yo < cn; yo++)
      sc = R::comb(sc, R::add(ring_row(rows, q.ring[y0 + yo], jy, W1)[d - sd], ts[yo]), tbl);
    const int2 dk = q.deck[bv];
    deck[((size_t)bv * W1 + dk.x + j % (W1 + dk.y)) * W1 + d] = sc;
    return sc;
  }
  T terms[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) terms[yo] = R::add(ring_row(rows, q.ring[y0 + yo], jy, W1)[d - sd], ts[yo]);
  T sc = reduce_c<T>(cn, terms, init_sc(tb, v, d - sd), tbl);
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) sc = R::add(sc, oe[dsq[i]]);
  else if (st == MR_ST || st == IR_ST) sc = R::add(sc, oe[dsq[j]]);
  else if (st == MP_ST) sc = R::add(sc, oe[dsq[i] * KP + dsq[j]]);
  ring_row(rows, q.ring[v], j, W1)[d] = sc;
  return sc;
}

// The d row dn..dx of an IL state at j: the self transition reads (j, d-1),
// carried in a register along the row (cm_nohmm.rs:617-647).
template <typename T>
__device__ void qdb_il_row(const QdbDev &q, const QdbTabs<T> &tb, T *rows, const uint8_t *dsq, int W1, int v,
                           int y0, int cn, int j, int dn, int dx) {
  using R = Sr<T>;
  const int32_t *tbl = q.ilog;
  const T *ts = tb.tsc + v * MAXC;
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  T *out = ring_row(rows, q.ring[v], j, W1);
  const T *cr[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) cr[yo] = ring_row(rows, q.ring[y0 + yo], j, W1);
  T prev = out[dn - 1];
  for (int d = dn; d <= dx; d++) {
    T terms[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) terms[yo] = R::add(y0 + yo == v ? prev : cr[yo][d - 1], ts[yo]);
    T sc = reduce_c<T>(cn, terms, init_sc(tb, v, d - 1), tbl);
    sc = R::add(sc, oe[dsq[j - d + 1]]);
    out[d] = sc;
    prev = sc;
  }
}

template <typename T>
__global__ __launch_bounds__(QDB_MAX_THREADS) void qdb_kernel(QdbDev q, QdbTabs<T> tb, const QdbWinDev *wins,
                                                              float cutoff, QdbCand *cand, int64_t cap,
                                                              unsigned long long *counter) {
  using R = Sr<T>;
  const QdbWinDev wd = wins[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  QdbGeo g;
  g.i0 = wd.i0;
  g.j0 = wd.j0;
  g.w = wd.w;
  g.W1 = wd.w + 1;
  const int W1 = g.W1;
  T *rows = (T *)wd.rows;
  T *deck = rows + (size_t)q.nrows * W1;
  T *acc = deck + ((size_t)q.nbegl * W1 + q.deck_extra) * W1;
  const uint8_t *dsq = wd.dsq;
  // every cell starts at the semiring's zero, d = 0 at its base case
  // (cm_nohmm.rs:489-532)
  for (int64_t x = tid; x < (int64_t)q.nrows * W1; x += nt) {
    const int d = (int)(x % W1), v = q.row_state[x / W1];
    rows[x] = (d == 0 && q.begl[v] < 0) ? tb.base0[v] : R::zero();
  }
  for (int b = 0; b < q.nbegl; b++) {
    const int2 dk = q.deck[b];
    T *db = deck + ((size_t)b * W1 + dk.x) * W1;
    for (int64_t x = tid; x < (int64_t)(W1 + dk.y) * W1; x += nt) db[x] = (x % W1) == 0 ? tb.bbase0[b] : R::zero();
  }
  if (q.local)
    for (int64_t x = tid; x < (int64_t)(q.amask + 1) * q.nslot * W1; x += nt) acc[x] = R::zero();
  __syncthreads();
  const int L = g.j0 - g.i0 + 1, nsteps = L + q.lroot;
  const int ib = wave < q.nw ? q.wave_item[wave] : 0, ie = wave < q.nw ? q.wave_item[wave + 1] : 0;
  const int r0 = q.cfirst[0], cn0 = q.cnum[0];
  for (int t = 0; t < nsteps; t++) {
    for (int it = ib; it < ie; it++) {
      const int4 im = q.item[it];  // wave-uniform: skip items with no live cell at this step
      if (im.y < t - (L - 1) || im.x > t || im.z > min(t - im.x + 1, g.w)) continue;
      const int2 u = q.unit[(size_t)it * 64 + lane];
      const int kind = u.x >> 28, v = u.x & 0x0FFFFFFF, d = u.y;
      if (kind == QU_NONE) continue;
      const int j = g.i0 + t - q.lvl[v];
      if (j < g.i0 || j > g.j0) continue;
      const int jrow = min(j - g.i0 + 1, g.w);
      if (kind == QU_IL) {
        const int dx = min(jrow, q.dxcap[v]);
        if (dx >= d) qdb_il_row<T>(q, tb, rows, dsq, W1, v, q.cfirst[v], q.cnum[v], j, d, dx);
        continue;
      }
      if (d > jrow) continue;
      if (kind == QU_CELL) {
        const T sc = qdb_cell<T>(q, tb, rows, deck, dsq, g, v, q.sttype[v], q.cfirst[v], q.cnum[v], j, d);
        const int slot = q.acc[v];
        if (q.local && slot >= 0) {  // this begin's candidate for the root at (j, d)
          T &a = acc[((size_t)(j & q.amask) * q.nslot + slot) * W1 + d];
          const T cnd = R::add(sc, tb.beginsc[v]);
          if (a < cnd) a = cnd;
        }
        continue;
      }
      // the root: its own transitions, then local begins by MAX (cm_nohmm.rs:686-746)
      T a = R::comb(R::zero(), R::add(ring_row(rows, q.ring[r0], j, W1)[d], tb.tsc[0]), q.ilog);
      for (int yo = 1; yo < cn0; yo++)
        a = R::comb(a, R::add(ring_row(rows, q.ring[r0 + yo], j, W1)[d], tb.tsc[yo]), q.ilog);
      if (q.local)
        for (int s = 0; s < q.nslot; s++) {
          T &b = acc[((size_t)(j & q.amask) * q.nslot + s) * W1 + d];
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
                           float cutoff, QdbCand *cand, int64_t cap, unsigned long long *counter, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (inside)
    hipLaunchKernelGGL(qdb_kernel<int32_t>, dim3((unsigned)n), dim3(64 * q.nw), 0, s, q,
                       *(const QdbTabs<int32_t> *)tabs, wins, cutoff, cand, cap, counter);
  else
    hipLaunchKernelGGL(qdb_kernel<float>, dim3((unsigned)n), dim3(64 * q.nw), 0, s, q,
                       *(const QdbTabs<float> *)tabs, wins, cutoff, cand, cap, counter);
  return hipGetLastError();
}

}  // namespace ixg
