// gfx950 query-dependent-banded (or non-banded) scanners: CYK on float
// scores and Inside on integer scores with the ILogsum table
// (cm_nohmm.rs:406-767 generic_scan). One 256-lane workgroup per window
// walks the end positions j in order; for each j it fills the rows of the
// model's dependency levels one level at a time (a level's states read only
// lower levels at the same j, or the previous j), each lane computing whole
// cells of one state. Rows roll over two j's ([2][M][w+1]); BEGL_S decks keep
// w+1 rows for the bifurcations' splits, as the reference's matrices do.
// Children are combined in the reference's unrolled order (reduce_c_order),
// so both semirings are bit-identical to it. Root cells at or above the
// reporting cutoff are appended to a candidate list; greedy reporting, null3
// and overlap removal run on the host.
#include "cm_internal.h"

namespace ixg {

namespace {

constexpr int QDB_THREADS = 256;

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
struct Win {
  T *rows, *bg;
  int M, W1, nbegl;
  __device__ __forceinline__ T &r(int par, int v, int d) const { return rows[((size_t)par * M + v) * W1 + d]; }
  __device__ __forceinline__ T &b(int jr, int bi, int d) const { return bg[((size_t)jr * nbegl + bi) * W1 + d]; }
};

template <typename T>
__device__ __forceinline__ T init_sc(const QdbTabs<T> &tb, int v, int dp) {
  const T es = tb.endsc[v];
  return Sr<T>::possible(es) ? Sr<T>::add(tb.el_sca[dp], es) : Sr<T>::zero();
}

// One cell (j, d) of a state v other than IL and the root.
template <typename T>
__device__ void qdb_cell(const QdbDev &q, const QdbTabs<T> &tb, const Win<T> &w, const uint8_t *dsq, int v, int st,
                         int j, int d, int cur, int prv) {
  using R = Sr<T>;
  const int32_t *tbl = q.ilog;
  const int sd = q_sd(st);
  if (st == B_ST) {  // cm_nohmm.rs:589-606
    const int wl = q.cfirst[v], y = q.cnum[v], bi = q.begl[wl];
    const int dn_y = q.dminw[y], dx_y = q.dxcap[y], dn_w = q.dminw[wl], dx_w = q.dxcap[wl];
    const int kmin = max(0, max(dn_y, d - dx_w)), kmax = min(dx_y, d - dn_w);
    T sc = init_sc(tb, v, d - sd);
    for (int k = kmin; k <= kmax; k++) {
      const int jr = ((j - k) % w.W1 + w.W1) % w.W1;
      sc = R::comb(sc, R::add(w.b(jr, bi, d - k), w.r(cur, y, k)), tbl);
    }
    w.r(cur, v, d) = sc;
    return;
  }
  const int y0 = q.cfirst[v], cn = q.cnum[v];
  const int jp_y = q_sdr(st) ? prv : cur;
  const T *ts = tb.tsc + v * MAXC;
  if (q.begl[v] >= 0) {  // BEGL_S: natural child order, into its deck (cm_nohmm.rs:607-616)
    T sc = init_sc(tb, v, d - sd);
    for (int yo = 0; yo < cn; yo++) sc = R::comb(sc, R::add(w.r(jp_y, y0 + yo, d - sd), ts[yo]), tbl);
    w.b(j % w.W1, q.begl[v], d) = sc;
    return;
  }
  T terms[MAXC];
#pragma unroll
  for (int yo = 0; yo < MAXC; yo++)
    if (yo < cn) terms[yo] = R::add(w.r(jp_y, y0 + yo, d - sd), ts[yo]);
  T sc = reduce_c<T>(cn, terms, init_sc(tb, v, d - sd), tbl);
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) sc = R::add(sc, oe[dsq[i]]);
  else if (st == MR_ST || st == IR_ST) sc = R::add(sc, oe[dsq[j]]);
  else if (st == MP_ST) sc = R::add(sc, oe[dsq[i] * KP + dsq[j]]);
  w.r(cur, v, d) = sc;
}

// The d row dn..dx of an IL state at j: the self transition reads (j, d-1),
// carried in a register along the row (cm_nohmm.rs:617-647).
template <typename T>
__device__ void qdb_il_row(const QdbDev &q, const QdbTabs<T> &tb, const Win<T> &w, const uint8_t *dsq, int v, int j,
                           int dn, int dx, int cur) {
  using R = Sr<T>;
  const int32_t *tbl = q.ilog;
  const int y0 = q.cfirst[v], cn = q.cnum[v];
  const T *ts = tb.tsc + v * MAXC;
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  T prev = w.r(cur, v, dn - 1);
  for (int d = dn; d <= dx; d++) {
    T terms[MAXC];
#pragma unroll
    for (int yo = 0; yo < MAXC; yo++)
      if (yo < cn) terms[yo] = R::add(y0 + yo == v ? prev : w.r(cur, y0 + yo, d - 1), ts[yo]);
    T sc = reduce_c<T>(cn, terms, init_sc(tb, v, d - 1), tbl);
    sc = R::add(sc, oe[dsq[j - d + 1]]);
    w.r(cur, v, d) = sc;
    prev = sc;
  }
}

template <typename T>
__global__ __launch_bounds__(QDB_THREADS) void qdb_kernel(QdbDev q, QdbTabs<T> tb, const QdbWinDev *wins,
                                                          float cutoff, QdbCand *cand, int64_t cap,
                                                          unsigned long long *counter) {
  using R = Sr<T>;
  const QdbWinDev wd = wins[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x, M = q.M;
  Win<T> w;
  w.rows = (T *)wd.rows;
  w.bg = (T *)wd.begl;
  w.M = M;
  w.W1 = wd.w + 1;
  w.nbegl = q.nbegl;
  const uint8_t *dsq = wd.dsq;
  // every cell starts at the semiring's zero, d = 0 at its base case
  // (cm_nohmm.rs:489-532)
  for (int64_t x = tid; x < 2 * (int64_t)M * w.W1; x += nt) {
    const int d = (int)(x % w.W1), v = (int)((x / w.W1) % M);
    w.rows[x] = (d == 0 && q.begl[v] < 0) ? tb.base0[v] : R::zero();
  }
  for (int64_t x = tid; x < (int64_t)w.W1 * q.nbegl * w.W1; x += nt) {
    const int d = (int)(x % w.W1), bi = (int)((x / w.W1) % q.nbegl);
    w.bg[x] = d == 0 ? tb.bbase0[bi] : R::zero();
  }
  __syncthreads();
  const int r0 = q.cfirst[0], cn0 = q.cnum[0];
  for (int j = wd.i0; j <= wd.j0; j++) {
    const int jrow = min(j - wd.i0 + 1, wd.w), cur = j & 1, prv = (j - 1) & 1;
    for (int l = 0; l < q.nlev; l++) {
      int off = 0;  // units dealt so far in this level, modulo nt
      for (int s = q.lvl_off[l]; s < q.lvl_off[l + 1]; s++) {
        const int v = q.lvl_states[s];
        const int st = q.sttype[v];
        if (v == 0 || st == E_ST) continue;
        const int dn = q.dn[v], dx = min(jrow, q.dxcap[v]);
        if (dx < dn) continue;
        const int n = st == IL_ST ? 1 : dx - dn + 1;
        int r = (tid - off) % nt;
        if (r < 0) r += nt;
        for (int u = r; u < n; u += nt) {
          if (st == IL_ST) qdb_il_row<T>(q, tb, w, dsq, v, j, dn, dx, cur);
          else qdb_cell<T>(q, tb, w, dsq, v, st, j, dn + u, cur, prv);
        }
        off = (off + n) % nt;
      }
      __syncthreads();
    }
    // the root: its own transitions, then local begins by MAX (cm_nohmm.rs:686-746)
    const int dn0 = q.dn[0], dx0 = min(jrow, q.dxcap[0]);
    for (int d = dn0 + tid; d <= dx0; d += nt) {
      T a = R::comb(R::zero(), R::add(w.r(cur, r0, d), tb.tsc[0]), q.ilog);
      for (int yo = 1; yo < cn0; yo++) a = R::comb(a, R::add(w.r(cur, r0 + yo, d), tb.tsc[yo]), q.ilog);
      if (q.local)
        for (int b = 0; b < q.nb; b++) {
          const int y = q.blist[b];
          if (d < q.dn[y] || d > min(jrow, q.dxcap[y])) continue;
          const T cnd = R::add(q.begl[y] >= 0 ? w.b(j % w.W1, q.begl[y], d) : w.r(cur, y, d), tb.beginsc[y]);
          if (a < cnd) a = cnd;
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
  }
}

}  // namespace

hipError_t launch_qdb_scan(const QdbDev &q, const void *tabs, int inside, const QdbWinDev *wins, int64_t n,
                           float cutoff, QdbCand *cand, int64_t cap, unsigned long long *counter, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (inside)
    hipLaunchKernelGGL(qdb_kernel<int32_t>, dim3((unsigned)n), dim3(QDB_THREADS), 0, s, q,
                       *(const QdbTabs<int32_t> *)tabs, wins, cutoff, cand, cap, counter);
  else
    hipLaunchKernelGGL(qdb_kernel<float>, dim3((unsigned)n), dim3(QDB_THREADS), 0, s, q,
                       *(const QdbTabs<float> *)tabs, wins, cutoff, cand, cap, counter);
  return hipGetLastError();
}

}  // namespace ixg
