// gfx950 QDB / non-banded truncated scanners: truncated CYK on float
// scores and truncated Inside on integer scores with the ILogsum table
// (cm_dpsearch_trunc.rs:1800-2316 generic_tr_scan). One 256-lane workgroup
// per window walks the end positions j in order; for each j it fills the
// four marginal planes J/L/R/T (rows rolling over two j's, [2][M][w+1] per
// plane) and the BEGL_S decks of J/L/R ([w+1][nbegl][w+1]) one dependency
// level at a time, each lane computing whole cells (d = 1..min(jrow, dmax, W)
// for every state); an IL state's row is one lane's serial walk along its
// self transition. Children are folded in increasing order, the bifurcation
// and pair-state marginal cells take the maximum of their J and L (or R)
// terms exactly where the reference does, and the root rows are never reset
// between j's: a truncated begin must beat what the same-parity row holds at
// d. Per root cell the best begin (score, state, mode) is kept; cells at or
// above the reporting cutoff go to a candidate list, and the window's best
// root score, its mode and the full-window best to a summary. Greedy
// reporting, truncation legality, null3 and overlap removal run on the host.
#include "cm_internal.h"

namespace ixg {

namespace {

constexpr int TRQ_THREADS = 256;
constexpr int MODE_T = 0, MODE_R = 1, MODE_L = 2, MODE_J = 3, MODE_UNKNOWN = 4;

template <typename T>
struct Tq;
template <>
struct Tq<float> {
  static __device__ __forceinline__ float zero() { return IMPOSSIBLE; }
  static __device__ __forceinline__ float comb(float a, float b, const int32_t *) { return a >= b ? a : b; }
  static __device__ __forceinline__ float max2(float a, float b) { return a >= b ? a : b; }
  static __device__ __forceinline__ float add(float a, float b) { return a + b; }
  static __device__ __forceinline__ float bits(float a) { return a; }
  static __device__ __forceinline__ bool possible(float a) { return a > IMPOSSIBLE + 1.0f; }
};
template <>
struct Tq<int32_t> {
  static __device__ __forceinline__ int32_t zero() { return -INFTY_I; }
  // ILogsum (cp9.rs:62-74), branch-free: the table is read at 0 when unused
  static __device__ __forceinline__ int32_t comb(int32_t a, int32_t b, const int32_t *tbl) {
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

template <typename T>
struct Mx {
  T *rows, *dk;  // planes J|L|R|T of [2][M][W1]; decks J|L|R of [W1][nbegl][W1]
  int M, W1, nbegl;
  __device__ __forceinline__ T &p(int pl, int par, int v, int d) const {
    return rows[(((size_t)pl * 2 + par) * M + v) * W1 + d];
  }
  __device__ __forceinline__ T &b(int pl, int jr, int bi, int d) const {
    return dk[(((size_t)pl * W1 + jr) * nbegl + bi) * W1 + d];
  }
};

template <typename T>
__device__ __forceinline__ T tinit(const TrqTabs<T> &tb, int v, int dp) {
  const T es = tb.endsc[v];
  return Tq<T>::possible(es) ? Tq<T>::add(tb.el_sca[dp], es) : Tq<T>::zero();
}

__device__ __forceinline__ int t_sd(int st) { return st == MP_ST ? 2 : ((st >= ML_ST && st <= IR_ST) ? 1 : 0); }
__device__ __forceinline__ int t_sdr(int st) { return (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0; }

// One cell (j, d) of a state other than IL and the root, every filled plane
// (cm_dpsearch_trunc.rs:1978-2217).
template <typename T>
__device__ void trq_cell(const TrqDev &q, const TrqTabs<T> &tb, const Mx<T> &x, const uint8_t *dsq, int v, int st,
                         int j, int d, int cur, int prv) {
  using R = Tq<T>;
  const int32_t *tbl = q.ilog;
  const bool fl = q.fill & 2, fr = q.fill & 4, ft = q.fill & 8;
  const int sd = t_sd(st);
  const T z = R::zero();
  if (st == B_ST) {
    const int wl = q.cfirst[v], y = q.cnum[v], bi = q.begl[wl];
    int kmin = 0, kmax = d;
    if (!q.nonbanded) {
      kmin = max(0, d - q.dxcap[wl]);
      kmax = min(q.dxcap[y], d);
    }
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z, tsc = ft ? jsc : z;
    for (int k = kmin; k <= kmax; k++) {
      const int jr = ((j - k) % x.W1 + x.W1) % x.W1;
      const T lft = x.b(0, jr, bi, d - k);
      jsc = R::comb(jsc, R::add(lft, x.p(0, cur, y, k)), tbl);
      if (fl) lsc = R::comb(lsc, R::add(lft, x.p(1, cur, y, k)), tbl);
      if (fr) rsc = R::comb(rsc, R::add(x.b(2, jr, bi, d - k), x.p(0, cur, y, k)), tbl);
    }
    if (ft)
      for (int k = max(1, kmin); k <= min(d - 1, kmax); k++) {
        const int jr = ((j - k) % x.W1 + x.W1) % x.W1;
        tsc = R::comb(tsc, R::add(x.b(2, jr, bi, d - k), x.p(1, cur, y, k)), tbl);
      }
    x.p(0, cur, v, d) = jsc;
    if (ft) x.p(3, cur, v, d) = tsc;
    const int j0r = j % x.W1;
    if (fl) x.p(1, cur, v, d) = kmin == 0 ? R::comb(lsc, R::max2(x.b(0, j0r, bi, d), x.b(1, j0r, bi, d)), tbl) : lsc;
    if (fr) x.p(2, cur, v, d) = kmax == d ? R::comb(rsc, R::max2(x.p(0, cur, y, d), x.p(2, cur, y, d)), tbl) : rsc;
    return;
  }
  const int y0 = q.cfirst[v], cn = q.cnum[v];
  const int jp_y = t_sdr(st) ? prv : cur, jq_y = t_sdr(st) ? cur : prv;
  const T *ts = tb.tsc + v * MAXC;
  if (q.begl[v] >= 0) {  // BEGL_S: into its J/L/R decks (:2046-2067)
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z;
    for (int yo = 0; yo < cn; yo++) {
      jsc = R::comb(jsc, R::add(x.p(0, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fl) lsc = R::comb(lsc, R::add(x.p(1, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fr) rsc = R::comb(rsc, R::add(x.p(2, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
    }
    const int jr = j % x.W1, bi = q.begl[v];
    x.b(0, jr, bi, d) = jsc;
    if (fl) x.b(1, jr, bi, d) = lsc;
    if (fr) x.b(2, jr, bi, d) = rsc;
    return;
  }
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  const int i = j - d + 1;
  if (st == ML_ST) {  // EMITLEFT, non-self (:2068-2106)
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z;
    for (int yo = 0; yo < cn; yo++) {
      jsc = R::comb(jsc, R::add(x.p(0, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fl) lsc = R::comb(lsc, R::add(x.p(1, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
    }
    const T e = oe[dsq[i]];
    x.p(0, cur, v, d) = R::add(jsc, e);
    if (fl) x.p(1, cur, v, d) = d >= 2 ? R::add(lsc, e) : e;
    if (fr) {
      T rsc = tinit(tb, v, d);
      for (int yo = 0; yo < cn; yo++)
        rsc = R::comb(rsc, R::max2(R::add(x.p(0, jp_y, y0 + yo, d), ts[yo]), R::add(x.p(2, jp_y, y0 + yo, d), ts[yo])),
                      tbl);
      x.p(2, cur, v, d) = rsc;
    }
  } else if (st == MR_ST || st == IR_ST) {  // EMITRIGHT (:2107-2141)
    const T esc_j = oe[dsq[j]];
    T jsc = tinit(tb, v, d - sd);
    T rsc = fr ? jsc : z;
    for (int yo = 0; yo < cn; yo++) {
      jsc = R::comb(jsc, R::add(x.p(0, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fr) rsc = R::comb(rsc, R::add(x.p(2, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
    }
    x.p(0, cur, v, d) = R::add(jsc, esc_j);
    if (fr) x.p(2, cur, v, d) = d >= 2 ? R::add(rsc, esc_j) : esc_j;
    if (fl) {
      T lsc = tinit(tb, v, d);
      for (int yo = st == IR_ST ? 1 : 0; yo < cn; yo++)
        lsc = R::comb(lsc, R::max2(R::add(x.p(0, jq_y, y0 + yo, d), ts[yo]), R::add(x.p(1, jq_y, y0 + yo, d), ts[yo])),
                      tbl);
      x.p(1, cur, v, d) = lsc;
    }
  } else if (st == MP_ST) {  // EMITPAIR (:2142-2190)
    T jsc = d >= 2 ? tinit(tb, v, d - sd) : z;
    T lsc = fl ? tinit(tb, v, d - 1) : z, rsc = fr ? tinit(tb, v, d - 1) : z;
    for (int yo = 0; yo < cn; yo++) {
      if (d >= 2) jsc = R::comb(jsc, R::add(x.p(0, jp_y, y0 + yo, d - 2), ts[yo]), tbl);
      if (fl)
        lsc = R::comb(lsc, R::max2(R::add(x.p(0, jq_y, y0 + yo, d - 1), ts[yo]), R::add(x.p(1, jq_y, y0 + yo, d - 1), ts[yo])),
                      tbl);
      if (fr)
        rsc = R::comb(rsc, R::max2(R::add(x.p(0, jp_y, y0 + yo, d - 1), ts[yo]), R::add(x.p(2, jp_y, y0 + yo, d - 1), ts[yo])),
                      tbl);
    }
    x.p(0, cur, v, d) = d >= 2 ? R::add(jsc, oe[dsq[i] * KP + dsq[j]]) : z;
    const T lm = tb.lmesc[v * KP + dsq[i]], rm = tb.rmesc[v * KP + dsq[j]];
    if (fl) x.p(1, cur, v, d) = d >= 2 ? R::add(lsc, lm) : lm;
    if (fr) x.p(2, cur, v, d) = d >= 2 ? R::add(rsc, rm) : rm;
  } else {  // EMITNONE: D, S (:2191-2217)
    T jsc = tinit(tb, v, d - sd);
    T lsc = fl ? jsc : z, rsc = fr ? jsc : z;
    for (int yo = 0; yo < cn; yo++) {
      jsc = R::comb(jsc, R::add(x.p(0, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fl) lsc = R::comb(lsc, R::add(x.p(1, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
      if (fr) rsc = R::comb(rsc, R::add(x.p(2, jp_y, y0 + yo, d - sd), ts[yo]), tbl);
    }
    x.p(0, cur, v, d) = jsc;
    if (fl) x.p(1, cur, v, d) = lsc;
    if (fr) x.p(2, cur, v, d) = rsc;
  }
}

// The row d = 1..dx of an IL state at j: its J and L cells read the same
// state at (j, d-1), so one lane walks the row (:2068-2106).
template <typename T>
__device__ void trq_il_row(const TrqDev &q, const TrqTabs<T> &tb, const Mx<T> &x, const uint8_t *dsq, int v, int j,
                           int dx, int cur) {
  using R = Tq<T>;
  const int32_t *tbl = q.ilog;
  const bool fl = q.fill & 2, fr = q.fill & 4;
  const T z = R::zero();
  const int y0 = q.cfirst[v], cn = q.cnum[v];
  const T *ts = tb.tsc + v * MAXC;
  const T *oe = tb.oesc + (size_t)v * (KP * KP);
  for (int d = 1; d <= dx; d++) {
    T jsc = tinit(tb, v, d - 1);
    T lsc = fl ? jsc : z;
    for (int yo = 0; yo < cn; yo++) {
      jsc = R::comb(jsc, R::add(x.p(0, cur, y0 + yo, d - 1), ts[yo]), tbl);
      if (fl) lsc = R::comb(lsc, R::add(x.p(1, cur, y0 + yo, d - 1), ts[yo]), tbl);
    }
    const T e = oe[dsq[j - d + 1]];
    x.p(0, cur, v, d) = R::add(jsc, e);
    if (fl) x.p(1, cur, v, d) = d >= 2 ? R::add(lsc, e) : e;
    if (fr) {
      T rsc = tinit(tb, v, d);
      for (int yo = 1; yo < cn; yo++)
        rsc = R::comb(rsc, R::max2(R::add(x.p(0, cur, y0 + yo, d), ts[yo]), R::add(x.p(2, cur, y0 + yo, d), ts[yo])),
                      tbl);
      x.p(2, cur, v, d) = rsc;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(TRQ_THREADS) void trqdb_kernel(TrqDev q, TrqTabs<T> tb, const QdbWinDev *wins,
                                                             float cutoff, TrqCand *cand, int64_t cap,
                                                             unsigned long long *counter, TrqSummary *summ) {
  using R = Tq<T>;
  __shared__ float s_sc[TRQ_THREADS];
  __shared__ int s_j[TRQ_THREADS], s_d[TRQ_THREADS], s_m[TRQ_THREADS];
  __shared__ float s_full;
  const QdbWinDev wd = wins[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x, M = q.M;
  Mx<T> x;
  x.rows = (T *)wd.rows;
  x.dk = (T *)wd.begl;
  x.M = M;
  x.W1 = wd.w + 1;
  x.nbegl = q.nbegl;
  const uint8_t *dsq = wd.dsq;
  // every cell starts at the semiring's zero; d = 0 at the base cases, set
  // once (:1873-1922): J of non-BEGL states, L and R of E states, the J decks
  const int64_t plane = 2 * (int64_t)M * x.W1;
  for (int64_t k = tid; k < 4 * plane; k += nt) {
    const int pl = (int)(k / plane);
    const int64_t r = k % plane;
    const int d = (int)(r % x.W1), v = (int)((r / x.W1) % M);
    T val = R::zero();
    if (d == 0 && q.begl[v] < 0) {
      if (pl == 0) val = tb.base0[v];
      else if ((pl == 1 || pl == 2) && q.sttype[v] == E_ST) val = (T)0;
    }
    x.rows[k] = val;
  }
  const int64_t deck = (int64_t)x.W1 * q.nbegl * x.W1;
  for (int64_t k = tid; k < 3 * deck; k += nt) {
    const int64_t r = k % deck;
    const int d = (int)(r % x.W1), bi = (int)((r / x.W1) % q.nbegl);
    x.dk[k] = (k < deck && d == 0) ? tb.bbase0[bi] : R::zero();
  }
  if (tid == 0) s_full = IMPOSSIBLE;
  __syncthreads();
  const bool fl = q.fill & 2, fr = q.fill & 4, ft = q.fill & 8;
  float best_sc = IMPOSSIBLE;
  int best_j = 0, best_d = 0, best_m = MODE_UNKNOWN;
  for (int j = wd.i0; j <= wd.j0; j++) {
    const int jrow = min(j - wd.i0 + 1, wd.w), cur = j & 1, prv = (j - 1) & 1;
    for (int l = 0; l < q.nlev; l++) {
      int off = 0;  // units dealt so far in this level, modulo nt
      for (int s = q.lvl_off[l]; s < q.lvl_off[l + 1]; s++) {
        const int v = q.lvl_states[s];
        const int st = q.sttype[v];
        if (v == 0 || st == E_ST) continue;
        // an emitter followed by an END node's E state is detached (:88)
        if ((st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST) && v + 1 < M && q.sttype[v + 1] == E_ST)
          continue;
        const int dx = min(jrow, q.dxcap[v]);
        if (dx < 1) continue;
        const int n = st == IL_ST ? 1 : dx;
        int r = (tid - off) % nt;
        if (r < 0) r += nt;
        for (int u = r; u < n; u += nt) {
          if (st == IL_ST) trq_il_row<T>(q, tb, x, dsq, v, j, dx, cur);
          else trq_cell<T>(q, tb, x, dsq, v, st, j, 1 + u, cur, prv);
        }
        off = (off + n) % nt;
      }
      __syncthreads();
    }
    // the root: truncated begins into the persistent root rows (:2220-2289)
    const int dx0 = min(jrow, q.dxcap[0]);
    for (int d = 1 + tid; d <= dx0; d += nt) {
      float bsc = IMPOSSIBLE;
      int br = 0, bm = MODE_UNKNOWN;
      for (int y = 1; y < M; y++) {
        const T tp = tb.pty[y];
        if (!R::possible(tp) || d > min(jrow, q.dxcap[y])) continue;
        const int ny = (ft && q.sttype[y] == B_ST) ? 4 : 3;
        for (int pl = 0; pl < ny; pl++) {
          if ((pl == 1 && !fl) || (pl == 2 && !fr)) continue;
          const T sc = R::add(x.p(pl, cur, y, d), tp);
          T &root = x.p(pl, cur, 0, d);
          if (sc > root) {
            root = sc;
            const float fsc = R::bits(sc);
            if (fsc > bsc) {
              bsc = fsc;
              bm = pl == 0 ? MODE_J : pl == 1 ? MODE_L : pl == 2 ? MODE_R : MODE_T;
              br = y;
            }
          }
        }
      }
      if (bsc > best_sc) {
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
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (inside)
    hipLaunchKernelGGL(trqdb_kernel<int32_t>, dim3((unsigned)n), dim3(TRQ_THREADS), 0, s, q,
                       *(const TrqTabs<int32_t> *)tabs, wins, cutoff, cand, cap, counter, summ);
  else
    hipLaunchKernelGGL(trqdb_kernel<float>, dim3((unsigned)n), dim3(TRQ_THREADS), 0, s, q,
                       *(const TrqTabs<float> *)tabs, wins, cutoff, cand, cap, counter, summ);
  return hipGetLastError();
}

}  // namespace ixg
