// HMM -> CM band conversion of the HMM-banded scanners (cp9.rs:2893-3879):
// the CP9 posterior bands of an envelope (pn_* from the GPU CP9 kernels)
// become, per CM state, i and j bands (cp9_hmm2ijbands, non-truncated path,
// with hmm_bands_enforce_valid_parse) and then the d bands the scanners read
// (ij2d_bands); cm_hb_mx_size_needed gives the matrix size that drives the
// tau iteration of cp9_iterate_seq2bands (cp9.rs:4075).
//
// The CM's guide tree is walked once per model, when the map is built: the
// walk's order and the consensus positions each step reads depend only on
// the tree, so a per-envelope conversion is a linear pass over a precomputed
// step list. Envelopes are converted on host worker threads.
#include <algorithm>
#include <climits>
#include <cstring>
#include "ctx.h"

using namespace ixg;

namespace {
enum { D_ST = 0, MP_ST = 1, ML_ST = 2, MR_ST = 3, IL_ST = 4, IR_ST = 5, S_ST = 6, E_ST = 7, B_ST = 8 };
enum { BIF_ND = 0, MATP_ND = 1, MATL_ND = 2, MATR_ND = 3, BEGL_ND = 4, BEGR_ND = 5, ROOT_ND = 6, END_ND = 7 };
constexpr int32_t NINF_I = -987654321;
}  // namespace

struct ixg_hbmap {
  int M = 0, nodes = 0, hmm_m = 0;
  bool cm_local = false, cp9_lb = false, cp9_le = false, cp9_el = false;
  std::vector<int8_t> sttype, ndtype;
  std::vector<int32_t> nodemap, cfirst, cnum;
  std::vector<int32_t> esc, tdm, tim, el_ct, el_off, el_idx;  // CP9 terms, [hmm_m+1] / EL lists
  struct Step {
    int32_t nd;
    int8_t right, type;
    int32_t lpos, rpos;  // the consensus positions this step reads
  };
  std::vector<Step> walk;
  std::vector<int32_t> slp, srp;  // per state: consensus span (emit map with the per-node-type offsets)
};

namespace ixg {

// cp9.rs:144-207 create_emit_map: consensus left/right position of every node
static void emit_positions(const ixg_hbmap &m, const std::vector<int32_t> &ndidx, std::vector<int32_t> &lpos,
                           std::vector<int32_t> &rpos) {
  lpos.assign(m.nodes, -1);
  rpos.assign(m.nodes, -1);
  std::vector<std::pair<int, int>> st{{0, 0}};
  int cpos = 0;
  while (!st.empty()) {
    const auto [nd, right] = st.back();
    st.pop_back();
    const int t = m.ndtype[nd];
    if (right) {
      rpos[nd] = cpos + 1;
      if (t == MATP_ND || t == MATR_ND) cpos++;
      continue;
    }
    if (t == MATP_ND || t == MATL_ND) cpos++;
    lpos[nd] = cpos;
    st.push_back({nd, 1});
    if (t == BIF_ND) {
      const int v = m.nodemap[nd];
      st.push_back({ndidx[m.cnum[v]], 0});
      st.push_back({ndidx[m.cfirst[v]], 0});
    } else if (t != END_ND) {
      st.push_back({nd + 1, 0});
    }
  }
}

// The walk of cp9_hmm2ijbands (cp9.rs:3416-3621) with the consensus
// positions each step reads: left visits of MATP/MATL set i bands, right
// visits set j bands (and the BIF / BEG / END / ROOT combinations).
static void build_walk(ixg_hbmap &m) {
  std::vector<int32_t> ndidx(m.M);
  for (int nd = 0; nd < m.nodes; nd++) {
    const int end = nd + 1 < m.nodes ? m.nodemap[nd + 1] : m.M;
    for (int v = m.nodemap[nd]; v < end; v++) ndidx[v] = nd;
  }
  std::vector<int32_t> lp, rp;
  emit_positions(m, ndidx, lp, rp);
  auto nd2l = [&](int nd) { return (m.ndtype[nd] == MATP_ND || m.ndtype[nd] == MATL_ND) ? lp[nd] : -1; };
  auto nd2r = [&](int nd) { return (m.ndtype[nd] == MATP_ND || m.ndtype[nd] == MATR_ND) ? rp[nd] : -1; };
  std::vector<std::pair<int, int>> st{{0, 0}};
  std::vector<int> lstack;
  int lpos = 0, rpos = 0;
  m.walk.clear();
  while (!st.empty()) {
    const auto [nd, right] = st.back();
    st.pop_back();
    const int t = m.ndtype[nd];
    ixg_hbmap::Step s{nd, (int8_t)right, (int8_t)t, 0, 0};
    if (right) {
      switch (t) {
        case MATP_ND: lpos = nd2l(nd); rpos = nd2r(nd); break;
        case MATL_ND: lpos = nd2l(nd); break;
        case MATR_ND: rpos = nd2r(nd); break;
        default: break;
      }
      s.lpos = lpos;
      s.rpos = rpos;
      if (t == BEGR_ND) {
        lpos = lstack.back();
        lstack.pop_back();
      } else if (t == BEGL_ND) {
        lstack.push_back(lpos);
        lpos = rpos + 1;
      } else if (t == END_ND) {
        rpos = lpos;
      }
      m.walk.push_back(s);
      continue;
    }
    if (t == MATP_ND || t == MATL_ND) lpos = nd2l(nd);
    s.lpos = lpos;
    s.rpos = rpos;
    m.walk.push_back(s);
    st.push_back({nd, 1});
    if (t == BIF_ND) {
      const int v = m.nodemap[nd];
      st.push_back({ndidx[m.cnum[v]], 0});
      st.push_back({ndidx[m.cfirst[v]], 0});
    } else if (t != END_ND) {
      st.push_back({nd + 1, 0});
    }
  }
}

namespace {

struct Reach {  // residues from which each HMM state (and node) can be reached
  std::vector<int> mn, mx, in, ix, dn, dx, nn_i, nx_i, nn_j, nx_j;
};

// cp9.rs:2965-3376 hmm_bands_enforce_valid_parse. pn [6][hmm_m+1]:
// pn_min_m, pn_max_m, pn_min_i, pn_max_i, pn_min_d, pn_max_d (doctored in
// place when local begins/ends are off).
void reach(const ixg_hbmap &m, int32_t *pn, int i0, int j0, bool search, Reach &r) {
  const int H = m.hmm_m, sz = H + 1;
  int32_t *pmn[3] = {pn, pn + 2 * sz, pn + 4 * sz}, *pmx[3] = {pn + sz, pn + 3 * sz, pn + 5 * sz};
  enum { SM = 0, SI = 1, SD = 2 };
  const bool lbe = m.cp9_lb && m.cp9_le;
  std::vector<int> *lo[3] = {&r.mn, &r.in, &r.dn}, *hi[3] = {&r.mx, &r.ix, &r.dx};
  for (auto *v : {&r.mn, &r.in, &r.dn, &r.nn_i, &r.nn_j}) v->assign(sz, INT_MAX);
  for (auto *v : {&r.mx, &r.ix, &r.dx, &r.nx_i, &r.nx_j}) v->assign(sz, INT_MIN);
  std::vector<int> hn(sz, INT_MAX), hx(sz, INT_MIN);
  int begn = INT_MAX, begx = INT_MIN, endn = INT_MAX, endx = INT_MIN;
  if (pmn[SM][0] != -1) r.mn[0] = pmn[SM][0], r.mx[0] = pmx[SM][0];
  auto prv = [&](const std::vector<int> &a, int k) { return k > 0 ? a[k - 1] : INT_MAX; };
  // cp9.rs:2940-2963: widen I_k so that a gap in the target's reach closes
  auto fill_gap = [&](int k, int n, int tmin) {
    const int right_min = n <= tmin ? tmin : n;
    int in = INT_MAX;
    if (prv(r.mn, k) != INT_MAX) in = std::min(in, prv(r.mn, k) + 1);
    if (prv(r.dn, k) != INT_MAX) in = std::min(in, prv(r.dn, k));
    int32_t &a = pmn[SI][k], &b = pmx[SI][k];
    a = a != -1 ? std::min(a, in) : in;
    b = b != -1 ? std::max(b, right_min - 1) : right_min - 1;
  };
  // cp9.rs:2893-2938: make node k reachable
  auto fix_unreachable = [&](int k, int pmin) {
    int nm = -1, nd = -1;
    if (pmn[SM][k] != -1 && pmx[SM][k] != -1 && pmx[SM][k] - 1 > pmin) nm = std::max(pmn[SM][k] - 1, pmin);
    if (pmn[SD][k] != -1 && pmx[SD][k] != -1 && pmx[SD][k] > pmin) nd = std::max(pmn[SD][k], pmin);
    if (nm != -1 || nd != -1) {
      const int nn = nm == -1 ? nd : nd == -1 ? nm : std::min(nm, nd);
      int32_t &a = pmn[SI][k - 1], &b = pmx[SI][k - 1];
      a = a != -1 ? std::min(a, pmin + 1) : pmin + 1;
      b = b != -1 ? std::max(b, nn) : nn;
    } else {
      for (int kp = k; kp <= H && pmx[SM][kp] < pmin + 1 && pmx[SD][kp] < pmin; kp++) pmn[SD][kp] = pmx[SD][kp] = pmin;
    }
  };
  // one transition from a source reach band [n0, x0] into state (t, k1);
  // gap: whether a gap check applies (guarded: only when the target is
  // reached; unguarded: the I->M check, whose INT_MIN - INT_MAX wraps)
  auto into = [&](int t, int k1, int n0, int x0, int k, int gap_mode) -> bool {
    if (pmn[t][k1] == -1) return false;
    if (std::min(x0, pmx[t][k1]) - std::max(n0, pmn[t][k1]) < 0) return false;
    const int n = std::min(std::max(n0, pmn[t][k1]), pmx[t][k1]), x = std::min(x0, pmx[t][k1]);
    std::vector<int> &L = *lo[t], &X = *hi[t];
    bool filled = false;
    if (gap_mode && !lbe) {
      const int a = std::min(x, X[k1]), b = std::max(n, L[k1]);
      const bool gap = gap_mode == 1 ? (L[k1] != INT_MAX && a - b < -1) : ((int)((unsigned)a - (unsigned)b) < -1);
      if (gap) {
        fill_gap(k, n, L[k1]);
        filled = true;
      }
    }
    L[k1] = std::min(L[k1], n);
    X[k1] = std::max(X[k1], x);
    return filled;
  };
  int k = 0;
  while (k <= H) {
    bool filled = false;
    if (r.mn[k] <= r.mx[k]) into(SI, k, r.mn[k] + 1, r.mx[k] + 1, k, 0);
    if (r.dn[k] <= r.dx[k]) into(SI, k, r.dn[k] + 1, r.dx[k] + 1, k, 0);
    if (r.in[k] <= r.ix[k] && pmn[SI][k] != -1) {  // the insert's self loop
      if (r.in[k] <= pmx[SI][k]) r.ix[k] = pmx[SI][k];
      else r.in[k] = INT_MAX, r.ix[k] = INT_MIN;
    }
    if (k < H) {
      const int k1 = k + 1;
      if (r.mn[k] <= r.mx[k]) filled |= into(SM, k1, r.mn[k] + 1, r.mx[k] + 1, k, 1);
      if (r.dn[k] <= r.dx[k]) filled |= into(SM, k1, r.dn[k] + 1, r.dx[k] + 1, k, 1);
      if (r.in[k] <= r.ix[k]) filled |= into(SM, k1, r.in[k] + 1, r.ix[k] + 1, k, 2);
      if (m.cp9_el)
        for (int c = 0; c < m.el_ct[k1]; c++) {
          const int kk = m.el_idx[m.el_off[k1] + c];
          if (r.mn[kk] <= r.mx[kk]) filled |= into(SM, k1, r.mn[kk], j0, k, 1);
        }
      if (lbe && pmn[SM][k1] != -1) {  // local begin into M_k+1
        if (search) {
          r.mn[k1] = std::min(r.mn[k1], pmn[SM][k1]);
          r.mx[k1] = std::max(r.mx[k1], pmx[SM][k1]);
        } else if (pmn[SM][k1] == r.mn[0] + 1) {
          r.mn[k1] = std::min(r.mn[k1], r.mn[0] + 1);
          r.mx[k1] = std::max(r.mx[k1], r.mn[0] + 1);
        }
      }
    }
    if ((k == H || lbe) && r.mn[k] <= r.mx[k] && m.esc[k] != NINF_I) endn = std::min(endn, r.mn[k]), endx = std::max(endx, r.mx[k]);
    if (k == H) {
      if (r.dn[k] <= r.dx[k] && m.tdm[k] != NINF_I) endn = std::min(endn, r.dn[k]), endx = std::max(endx, r.dx[k]);
      if (r.in[k] <= r.ix[k] && m.tim[k] != NINF_I) endn = std::min(endn, r.in[k]), endx = std::max(endx, r.in[k]);
    }
    if (k < H) {
      const int k1 = k + 1;
      if (r.mn[k] <= r.mx[k]) filled |= into(SD, k1, r.mn[k], r.mx[k], k, 1);
      if (r.in[k] <= r.ix[k]) filled |= into(SD, k1, r.in[k], r.ix[k], k, 1);
      if (r.dn[k] <= r.dx[k]) filled |= into(SD, k1, r.dn[k], r.dx[k], k, 1);
    }
    // residues that reach node k from the left (nn_i) and right (nn_j)
    auto low = [](std::vector<int> &v, int q, int x) { v[q] = std::min(v[q], x); };
    auto high = [](std::vector<int> &v, int q, int x) { v[q] = std::max(v[q], x); };
    if (r.mn[k] <= r.mx[k]) {
      low(hn, k, r.mn[k]), high(hx, k, r.mx[k]);
      if (k != H) low(r.nn_i, k + 1, r.mn[k] + 1), high(r.nx_i, k + 1, r.mx[k] + 1);
      if (k != 0) low(r.nn_j, k - 1, r.mn[k] - 1), high(r.nx_j, k - 1, r.mx[k] - 1);
      if ((lbe && k > 0) || k == H) {
        if (search) low(r.nn_j, k, r.mn[k]), high(r.nx_j, k, r.mx[k]);
        else if (r.mx[k] == j0) low(r.nn_j, k, j0), high(r.nx_j, k, j0);
      }
      if ((lbe && k > 0) || k == 1) {
        if (search) {
          low(r.nn_i, k, r.mn[k]), high(r.nx_i, k, r.mx[k]);
          begn = std::min(begn, r.mn[k]), begx = std::max(begx, r.mx[k]);
        } else if (r.mn[k] == i0) {
          low(r.nn_i, k, i0), high(r.nx_i, k, i0);
        }
      }
    }
    if (r.in[k] <= r.ix[k]) {
      low(hn, k, r.in[k]), high(hx, k, r.ix[k]);
      if (k != H) low(r.nn_i, k + 1, r.in[k] + 1), high(r.nx_i, k + 1, r.ix[k] + 1);
      low(r.nn_j, k, r.in[k] - 1), high(r.nx_j, k, r.ix[k] - 1);
      if (k == 0) begn = std::min(begn, r.in[k]), begx = std::max(begx, r.ix[k]);
    }
    if (r.dn[k] <= r.dx[k]) {
      low(hn, k, r.dn[k]), high(hx, k, r.dx[k]);
      if (k != H) low(r.nn_i, k + 1, r.dn[k] + 1), high(r.nx_i, k + 1, r.dx[k] + 1);
      if (k != 0) low(r.nn_j, k - 1, r.dn[k]), high(r.nx_j, k - 1, r.dx[k]);
      if (k == 1) begn = std::min(begn, r.dn[k] + 1), begx = std::max(begx, r.dx[k] + 1);
    }
    if (!lbe && r.mn[k] > r.mx[k] && r.dn[k] > r.dx[k]) {
      fix_unreachable(k, prv(hn, k));
      k -= 2;
    } else if (filled) {
      k -= 1;
    }
    k++;
  }
  if (!search) begn = begx = i0, endn = endx = j0;
  r.nn_i[1] = std::min(r.nn_i[1], begn);
  r.nx_i[1] = std::max(r.nx_i[1], begx);
  r.nn_j[H] = std::min(r.nn_j[H], endn);
  r.nx_j[H] = std::max(r.nx_j[H], endx);
  auto empty = [](std::vector<int> &a, std::vector<int> &b) {
    for (size_t q = 0; q < a.size(); q++) {
      if (a[q] == INT_MAX) a[q] = -1;
      if (b[q] == INT_MIN) b[q] = -2;
    }
  };
  empty(r.mn, r.mx);
  empty(r.in, r.ix);
  empty(r.dn, r.dx);
  if (lbe) empty(r.nn_i, r.nx_i), empty(r.nn_j, r.nx_j);
}

int sdelta_of(int st) { return st == MP_ST ? 2 : (st == ML_ST || st == MR_ST || st == IL_ST || st == IR_ST); }
int ldelta_of(int st) { return st == MP_ST || st == ML_ST || st == IL_ST; }
int rdelta_of(int st) { return st == MP_ST || st == MR_ST || st == IR_ST; }

}  // namespace

// cp9.rs:3380-3794 cp9_hmm2ijbands; tb != null: do_trunc, with the L/R/T
// band widening and the truncated special cases (tb's validity updated)
void hmm2ij(const ixg_hbmap &m, int32_t *pn, int i0, int j0, bool search, int32_t *imin, int32_t *imax,
            int32_t *jmin, int32_t *jmax, TrBand *tb) {
  const int M = m.M, H = m.hmm_m;
  const bool trunc = tb != nullptr;
  for (int v = 0; v < M; v++) imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
  Reach r;
  reach(m, pn, i0, j0, search, r);
  auto set = [&](int v, int i_lo, int i_hi, int j_lo, int j_hi) {
    imin[v] = i_lo, imax[v] = i_hi, jmin[v] = j_lo, jmax[v] = j_hi;
  };
  auto clear = [&](int v) { set(v, -1, -2, -1, -2); };
  auto shifted = [](int a, int empty) { return a == empty ? empty : a + 1; };
  for (const ixg_hbmap::Step &s : m.walk) {
    const int v0 = m.nodemap[s.nd], L = s.lpos, R = s.rpos;
    if (!s.right) {  // left visit: i bands of MATP / MATL states
      if (s.type == MATP_ND) {
        imin[v0] = r.mn[L], imax[v0] = r.mx[L];                                             // MP
        imin[v0 + 1] = r.mn[L], imax[v0 + 1] = r.mx[L];                                     // ML
        imin[v0 + 2] = shifted(r.dn[L], -1), imax[v0 + 2] = shifted(r.dx[L], -2);           // MR
        imin[v0 + 3] = shifted(r.dn[L], -1), imax[v0 + 3] = shifted(r.dx[L], -2);           // D
        imin[v0 + 4] = r.in[L], imax[v0 + 4] = r.ix[L];                                     // IL
      } else if (s.type == MATL_ND) {
        imin[v0] = r.mn[L], imax[v0] = r.mx[L];                                             // ML
        imin[v0 + 1] = shifted(r.dn[L], -1), imax[v0 + 1] = shifted(r.dx[L], -2);           // D
        imin[v0 + 2] = r.in[L], imax[v0 + 2] = r.ix[L];                                     // IL
      }
      continue;
    }
    switch (s.type) {
      case BIF_ND: {
        const int w = m.cfirst[v0], y = m.cnum[v0];
        imin[v0] = imin[w] != -1 ? imin[w] : imin[y];
        imax[v0] = imax[w] != -2 ? imax[w] : imax[y];
        jmin[v0] = jmin[y] != -1 ? jmin[y] : jmin[w];
        jmax[v0] = jmax[y] != -2 ? jmax[y] : jmax[w];
        if (!trunc &&
            (imin[w] == -1 || jmin[w] == -1 || imin[y] == -1 || jmin[y] == -1 || imax[w] == -2 || jmax[y] == -2)) {
          clear(v0), clear(w), clear(y), clear(y + 1);
        }
        break;
      }
      case MATP_ND:
        jmin[v0] = r.mn[R], jmax[v0] = r.mx[R];                    // MP
        jmin[v0 + 1] = r.dn[R], jmax[v0 + 1] = r.dx[R];            // ML
        jmin[v0 + 2] = r.mn[R], jmax[v0 + 2] = r.mx[R];            // MR
        jmin[v0 + 3] = r.dn[R], jmax[v0 + 3] = r.dx[R];            // D
        jmin[v0 + 4] = r.nn_j[R - 1], jmax[v0 + 4] = r.nx_j[R - 1]; // IL
        set(v0 + 5, r.nn_i[L + 1], r.nx_i[L + 1], r.in[R - 1], r.ix[R - 1]);  // IR
        break;
      case MATL_ND:
        for (int q = 0; q < 3; q++) jmin[v0 + q] = r.nn_j[R], jmax[v0 + q] = r.nx_j[R];
        break;
      case MATR_ND:
        set(v0, r.nn_i[L], r.nx_i[L], r.mn[R], r.mx[R]);            // MR
        set(v0 + 1, r.nn_i[L], r.nx_i[L], r.dn[R], r.dx[R]);        // D
        set(v0 + 2, r.nn_i[L], r.nx_i[L], r.in[R - 1], r.ix[R - 1]); // IR
        break;
      case BEGL_ND:
      case BEGR_ND: {
        int a = INT_MAX, b = INT_MIN, c = INT_MAX, d = INT_MIN;
        for (int y = m.cfirst[v0]; y < m.cfirst[v0] + m.cnum[v0]; y++) {
          if (imin[y] != -1) a = std::min(a, imin[y]), b = std::max(b, imax[y]);
          if (jmin[y] != -1) c = std::min(c, jmin[y]), d = std::max(d, jmax[y]);
        }
        if (a == INT_MAX) a = -1, b = -2;
        if (c == INT_MAX) c = -1, d = -2;
        set(v0, a, b, c, d);
        if (s.type == BEGR_ND) {
          const int vil = v0 + 1;
          imin[vil] = r.in[L - 1], imax[vil] = r.ix[L - 1];
          if (imin[v0] != -1 && imin[vil] != -1) {
            imin[v0] = std::min(imin[v0], imin[vil]);
            jmin[vil] = jmin[v0] == -1 ? -1 : std::max(jmin[v0], i0);
            jmax[vil] = jmax[v0];
          } else {
            clear(vil);
          }
        }
        break;
      }
      case END_ND: {
        int a = r.nn_i[L], b = r.nx_i[L] == -2 ? -2 : std::min(r.nx_i[L] + 1, j0 + 1);
        if (r.in[L] != -1) {
          a = std::min(a, std::max(r.in[L] - 1, i0));
          b = std::max(b, std::max(r.ix[L] - 1, i0));
        }
        imin[v0] = a, imax[v0] = b;
        if (a != -1) jmin[v0] = a - 1, jmax[v0] = b - 1;
        break;
      }
      case ROOT_ND: {
        set(v0, r.nn_i[1], r.nx_i[1], r.nn_j[H], r.nx_j[H]);  // S
        int jl = r.nn_j[H] == -1 ? -1 : std::max(r.nn_j[H], i0), jh = r.nx_j[H];
        if (r.in[H] != -1) jl = std::min(jl, r.in[H]), jh = std::min(jh, r.ix[H]);
        set(v0 + 1, r.in[0], r.ix[0], jl, jh);  // IL
        if (r.in[H] != -1) {                   // IR
          int a = r.nn_i[1], b = r.nx_i[1];
          if (imin[v0 + 1] != -1) a = std::min(a, imin[v0 + 1] + 1), b = std::max(b, imax[v0 + 1] + 1);
          set(v0 + 2, a, b, r.in[H], r.ix[H]);
        }
        break;
      }
      default: break;
    }
  }
  if (trunc)  // cp9.rs:3646-3661: L / R / T alignments widen the j / i bands
    for (int v = 0; v < M; v++) {
      if (tb->lv[v] || tb->tv[v]) {
        jmin[v] = jmin[v] == -1 ? tb->lmarg_jmin : std::min(jmin[v], tb->lmarg_jmin);
        jmax[v] = jmax[v] == -2 ? tb->lmarg_jmax : std::max(jmax[v], tb->lmarg_jmax);
      }
      if (tb->rv[v] || tb->tv[v]) {
        imin[v] = imin[v] == -1 ? tb->rmarg_imin : std::min(imin[v], tb->rmarg_imin);
        imax[v] = imax[v] == -2 ? tb->rmarg_imax : std::max(imax[v], tb->rmarg_imax);
      }
    }
  if (!search) {
    imin[0] = i0;
    if (imin[1] != -1) imin[1] = i0;
    jmax[0] = j0;
    if (jmin[1] != -1) jmax[1] = j0;
    if (jmin[2] != -1) jmax[2] = j0;
  }
  for (int v = 0; v < M; v++) {
    if (imin[v] == -1 || jmin[v] == -1) clear(v);
    if (v + 1 < M && m.sttype[v + 1] == E_ST) clear(v);  // detached insert
    if (trunc) continue;
    const int st = m.sttype[v];
    if (st == MP_ST) {
      if (jmax[v] == i0) clear(v);
      else if (jmin[v] == i0) jmin[v]++;
    }
    if (ldelta_of(st) && imin[v] != -1) {
      if (jmax[v] == i0 - 1) clear(v);
      else if (jmin[v] == i0 - 1) jmin[v] = i0;
    }
  }
  // the localized case keeps at least one parse (cp9.rs:3655-3791)
  if ((m.cp9_lb || m.cp9_le || m.cp9_el) && m.cm_local) {
    if (trunc) tb->jv[0] = 1;
    if (imin[0] == -1) set(0, i0, i0, j0, j0);
    int nd1 = 1;
    if (i0 == j0)
      while (nd1 < m.nodes && m.ndtype[nd1] == MATP_ND) nd1++;
    const int v = m.nodemap[nd1];
    if (trunc) tb->jv[v] = 1;  // cp9.rs:3710-3714 (the B node's S children too) and :3756
    if (m.ndtype[nd1] == BIF_ND) {
      const int w = m.cfirst[v], y = m.cnum[v];
      if (trunc) tb->jv[w] = tb->jv[y] = 1;
      if (imin[v] != -1 && imin[w] != -1 && imin[y] != -1) {
        if (!search) {
          imin[v] = std::min(imin[v], i0), imax[v] = std::max(imax[v], i0);
          jmin[v] = std::min(jmin[v], j0), jmax[v] = std::max(jmax[v], j0);
          imin[w] = imin[v], imax[w] = imax[v];
          jmax[w] = std::max(jmax[w], std::min(j0, imax[w]));
          jmin[y] = jmin[v], jmax[y] = jmax[v];
          imin[y] = std::min(imin[y], std::max(i0, jmin[y]));
        }
        imin[y] = std::min(imin[y], std::max(i0, jmax[w] + 1));
        imax[y] = std::max(imin[y], imax[y]);
      } else if (!search) {
        set(v, i0, i0, j0, j0), set(w, i0, i0, j0 - 1, j0 - 1), set(y, j0, j0, j0, j0);
      } else {
        const int a = imin[0], b = jmax[0];
        set(v, a, a, b, b), set(w, a, a, b - 1, b - 1), set(y, b, b, b, b);
      }
    } else if (!search) {
      if (imin[v] == -1) set(v, i0, i0, j0, j0);
      else set(v, std::min(imin[v], i0), std::max(imax[v], i0), std::min(jmin[v], j0), std::max(jmax[v], j0));
    } else if (imin[v] == -1) {
      set(v, imin[0], imax[0], jmin[0], jmax[0]);
    } else {
      set(0, std::min(imin[0], imin[v]), std::max(imax[0], imax[v]), std::min(jmin[0], jmin[v]),
          std::max(jmax[0], jmax[v]));
    }
  }
}

// cp9.rs:1755-1916 cp9_predict_start_and_end_positions: the first nodes
// (from the left) whose posterior occupancy passes thresh1 / thresh2 start
// the alignment, the first from the right end it; the R marginal i band
// comes from the start nodes' posterior bands, the L marginal j band from
// the end nodes' (a delete's j is one less). pn as the CP9 DPs return it.
void tr_predict(const int32_t *pn, const int32_t *occ, int H, int i0, int j0, TrBand &tb) {
  const int H1 = H + 1;
  const int32_t *mn = pn, *mx = pn + H1, *in = pn + 2 * H1, *ix = pn + 3 * H1, *dn = pn + 4 * H1, *dx = pn + 5 * H1;
  auto prob = [](int32_t sc) {  // Score2Prob(sc, 1.0), cp9.rs:1741
    return sc == NINF_I ? 0.0f : (float)std::exp((double)((float)sc / 1000.0f) * 0.69314718);
  };
  auto banded = [&](int k) { return !(mn[k] == -1 && in[k] == -1 && dn[k] == -1); };
  int sp1 = -1, sp2 = -1, ep1, ep2, k = 1;
  for (; k <= H && (sp1 == -1 || sp2 == -1); k++)
    if (banded(k)) {
      const float p = prob(occ[k]);
      if (sp1 == -1 && p > tb.thresh1) sp1 = k;
      if (sp2 == -1 && p > tb.thresh2) sp2 = k;
    }
  if (k == H + 1) {
    if (sp1 == -1) sp1 = H + 1;
    if (sp2 == -1) sp2 = H + 1;
  }
  if (sp1 == H + 1 && sp2 == H + 1) {
    ep1 = ep2 = 0;
  } else {
    ep1 = ep2 = -1;
    for (k = H; k >= 1 && (ep1 == -1 || ep2 == -1); k--)
      if (banded(k)) {
        const float p = prob(occ[k]);
        if (ep1 == -1 && p > tb.thresh1) ep1 = k;
        if (ep2 == -1 && p > tb.thresh2) ep2 = k;
      }
    if (k == 0) {
      if (ep1 == -1) ep1 = 0;
      if (ep2 == -1) ep2 = 0;
    }
  }
  tb.sp1 = sp1, tb.sp2 = sp2, tb.ep1 = ep1, tb.ep2 = ep2;
  if (sp1 == H + 1) {
    tb.rmarg_imin = i0, tb.rmarg_imax = j0;
  } else {
    int lo = INT_MAX, hi = INT_MIN;
    for (int s : {sp1, sp2}) {
      if (s == H + 1) continue;
      for (const int32_t *a : {mn, in, dn})
        if (a[s] >= 0) lo = std::min(lo, a[s]);
      for (const int32_t *a : {mx, ix, dx})
        if (a[s] >= 0) hi = std::max(hi, a[s]);
    }
    if (lo == INT_MAX || sp2 == H + 1) lo = i0;
    if (hi == INT_MIN || sp2 == H + 1) hi = j0 + 1;
    tb.rmarg_imin = std::min(j0 + 1, std::max(i0, lo));
    tb.rmarg_imax = std::min(j0 + 1, std::max(i0, hi));
  }
  if (ep1 == 0) {
    tb.lmarg_jmin = i0 - 1, tb.lmarg_jmax = j0;
  } else {
    int lo = INT_MAX, hi = INT_MIN;
    for (int e : {ep1, ep2}) {
      if (e == 0) continue;
      if (mn[e] >= 0) lo = std::min(lo, mn[e]);
      if (in[e] >= 0) lo = std::min(lo, in[e]);
      if (dn[e] >= 0) lo = std::min(lo, dn[e] - 1);
      if (mx[e] >= 0) hi = std::max(hi, mx[e]);
      if (ix[e] >= 0) hi = std::max(hi, ix[e]);
      if (dx[e] >= 0) hi = std::max(hi, dx[e] - 1);
    }
    if (lo == INT_MAX || ep2 == 0) lo = i0 - 1;
    if (hi == INT_MIN || ep2 == 0) hi = j0;
    tb.lmarg_jmin = std::min(j0, std::max(i0 - 1, lo));
    tb.lmarg_jmax = std::min(j0, std::max(i0 - 1, hi));
  }
}

// cp9.rs:1918-1990 cp9_marginal_candidates_from_start_end_positions: which
// marginal modes may pass through each state given the predicted start and
// end nodes and the pass's fill flags; the root takes every mode a child
// of its kind allows.
void tr_candidates(const ixg_hbmap &m, int pass, TrBand &tb) {
  const bool al = pass == IXG_PASS_3P_ONLY_FORCE || pass == IXG_PASS_5P_AND_3P_FORCE || pass == IXG_PASS_5P_AND_3P_ANY;
  const bool ar = pass == IXG_PASS_5P_ONLY_FORCE || pass == IXG_PASS_5P_AND_3P_FORCE || pass == IXG_PASS_5P_AND_3P_ANY;
  const bool at = pass == IXG_PASS_5P_AND_3P_FORCE || pass == IXG_PASS_5P_AND_3P_ANY;
  const int M = m.M;
  tb.jv.assign(M, 0), tb.lv.assign(M, 0), tb.rv.assign(M, 0), tb.tv.assign(M, 0);
  for (int v = 0; v < M; v++) {
    const int lp = m.slp[v], rp = m.srp[v];
    tb.jv[v] = lp >= tb.sp1 && rp <= tb.ep1;
    tb.lv[v] = al && lp >= tb.sp1 && lp <= tb.ep1 && rp > tb.ep2;
    tb.rv[v] = ar && rp <= tb.ep1 && rp >= tb.sp1 && lp < tb.sp2;
    tb.tv[v] = m.sttype[v] == B_ST && at && lp < tb.sp2 && rp > tb.ep2;
  }
  for (int v = 0; v < M; v++) {
    const int st = m.sttype[v];
    if ((st == B_ST || st == MP_ST || st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST) && tb.jv[v]) tb.jv[0] = 1;
    if ((st == B_ST || st == MP_ST || st == ML_ST || st == IL_ST) && tb.lv[v]) tb.lv[0] = 1;
    if ((st == B_ST || st == MP_ST || st == MR_ST || st == IR_ST) && tb.rv[v]) tb.rv[0] = 1;
    if (st == B_ST && tb.tv[v]) tb.tv[0] = 1;
    if (tb.jv[0] && tb.lv[0] && tb.rv[0] && tb.tv[0]) break;
  }
}

// cp9.rs:3815-3850 ij2d_bands into the scanners' layout: rows of state v
// start at hd_off[v]; returns the row count
int64_t ij2d(const ixg_hbmap &m, const int32_t *imin, const int32_t *imax, const int32_t *jmin, const int32_t *jmax,
             bool trunc, int32_t *hd_off, std::vector<int32_t> &hdmin, std::vector<int32_t> &hdmax) {
  int64_t rows = 0;
  for (int v = 0; v < m.M; v++) rows += jmax[v] >= jmin[v] ? jmax[v] - jmin[v] + 1 : 0;
  hdmin.resize(rows);
  hdmax.resize(rows);
  int64_t at = 0;
  for (int v = 0; v < m.M; v++) {
    hd_off[v] = (int32_t)at;
    const int st = m.sttype[v];
    const int dn = trunc ? std::max(ldelta_of(st), rdelta_of(st)) : sdelta_of(st);
    for (int j = jmin[v]; j <= jmax[v]; j++, at++) {
      if (st == E_ST) {
        hdmin[at] = hdmax[at] = 0;
        continue;
      }
      const int lo = j - imax[v] + 1, hi = j - imin[v] + 1;
      if (hi < dn) hdmin[at] = -1, hdmax[at] = -2;
      else hdmin[at] = std::max(lo, dn), hdmax[at] = hi;
    }
  }
  hd_off[m.M] = (int32_t)at;
  return rows;
}

// cp9.rs:3854-3879 cm_hb_mx_size_needed (Mb), the EL deck counted when the CM
// has local ends
float hb_mx_mb(const ixg_hbmap &m, const int32_t *jmin, const int32_t *jmax, const int32_t *hd_off,
               const int32_t *hdmin, const int32_t *hdmax, int L, int64_t *ncells) {
  int64_t cells = 0;
  float mb = (float)(64 + (m.M + 1) * 8 + (m.M + 1) * 4);
  for (int v = 0; v < m.M; v++) {
    const int jbw = jmax[v] - jmin[v];
    mb += (float)(8 * (jbw + 1));
    for (int jp = 0; jp <= jbw; jp++) cells += hdmax[hd_off[v] + jp] - hdmin[hd_off[v] + jp] + 1;
  }
  if (m.cm_local) cells += (int64_t)((double)(L + 2) * (double)(L + 1) * 0.5);
  mb += 4.0f * (float)cells;
  mb *= 0.000001f;
  if (ncells) *ncells = cells;
  return mb;
}

const ixg_hbmap *hbmap_check(const ixg_hbmap *m) { return m; }
int hbmap_M(const ixg_hbmap *m) { return m->M; }
int hbmap_hmm_m(const ixg_hbmap *m) { return m->hmm_m; }

}  // namespace ixg

extern "C" {

int ixg_hbmap_create(const ixg_cm_params *cm, int cm_local, const ixg_cp9_params *cp9, int cp9_local,
                     ixg_hbmap **out) {
  if (!cm || !cp9 || !out || cm->M < 1 || cm->nodes < 2 || !cm->sttype || !cm->ndtype || !cm->nodemap ||
      !cm->cfirst || !cm->cnum || !cp9->otsc || !cp9->el_from_ct || !cp9->el_from_off ||
      (cp9->el && !cp9->el_from_idx)) {
    g_err = "ixg_hbmap_create: bad arguments";
    return IXG_EINVAL;
  }
  auto m = std::make_unique<ixg_hbmap>();
  m->M = cm->M;
  m->nodes = cm->nodes;
  m->hmm_m = cp9->M;
  m->cm_local = cm_local != 0;
  m->cp9_lb = m->cp9_le = cp9_local != 0;
  m->cp9_el = cp9->el != 0;
  m->sttype.assign(cm->sttype, cm->sttype + cm->M);
  m->ndtype.assign(cm->ndtype, cm->ndtype + cm->nodes);
  m->nodemap.assign(cm->nodemap, cm->nodemap + cm->nodes);
  m->cfirst.assign(cm->cfirst, cm->cfirst + cm->M);
  m->cnum.assign(cm->cnum, cm->cnum + cm->M);
  const int H = cp9->M;
  enum { O_IM = 1, O_DM = 2, O_ME = 10 };
  for (int k = 0; k <= H; k++) {
    m->esc.push_back(cp9->otsc[k * 12 + O_ME]);
    m->tdm.push_back(cp9->otsc[k * 12 + O_DM]);
    m->tim.push_back(cp9->otsc[k * 12 + O_IM]);
  }
  m->el_ct.assign(cp9->el_from_ct, cp9->el_from_ct + H + 2);
  m->el_off.assign(cp9->el_from_off, cp9->el_from_off + H + 3);
  if (cp9->el) m->el_idx.assign(cp9->el_from_idx, cp9->el_from_idx + std::max(m->el_off[H + 2], 1));
  // the tree and its consensus length must agree with the HMM
  for (int nd = 0; nd < m->nodes; nd++)
    if (m->ndtype[nd] < BIF_ND || m->ndtype[nd] > END_ND || m->nodemap[nd] < 0 || m->nodemap[nd] >= m->M) {
      g_err = "ixg_hbmap_create: bad node table";
      return IXG_EINVAL;
    }
  build_walk(*m);
  {
    std::vector<int32_t> ndidx(m->M), lp, rp;
    for (int nd = 0, v = 0; v < m->M; v++) {
      while (nd + 1 < m->nodes && m->nodemap[nd + 1] <= v) nd++;
      ndidx[v] = nd;
    }
    emit_positions(*m, ndidx, lp, rp);
    m->slp.resize(m->M), m->srp.resize(m->M);
    for (int v = 0; v < m->M; v++) {
      const int nd = ndidx[v], t = m->ndtype[nd];
      m->slp[v] = (t == MATP_ND || t == MATL_ND) ? lp[nd] : lp[nd] + 1;
      m->srp[v] = (t == MATP_ND || t == MATR_ND) ? rp[nd] : rp[nd] - 1;
    }
  }
  int clen = 0;
  for (int nd = 0; nd < m->nodes; nd++) clen += m->ndtype[nd] == MATP_ND ? 2 : (m->ndtype[nd] == MATL_ND || m->ndtype[nd] == MATR_ND);
  if (clen != H) {
    g_err = "ixg_hbmap_create: CP9 node count differs from the CM's consensus length";
    return IXG_EINVAL;
  }
  *out = m.release();
  return IXG_OK;
}

int ixg_hbmap_free(ixg_hbmap *m) {
  delete m;
  return IXG_OK;
}

int ixg_hmm2ij_bands(const ixg_hbmap *m, int32_t *pn, int32_t i0, int32_t j0, int doing_search, int32_t *imin,
                     int32_t *imax, int32_t *jmin, int32_t *jmax) {
  if (!m || !pn || !imin || !imax || !jmin || !jmax || i0 < 1 || j0 < i0) {
    g_err = "ixg_hmm2ij_bands: bad arguments";
    return IXG_EINVAL;
  }
  hmm2ij(*m, pn, i0, j0, doing_search != 0, imin, imax, jmin, jmax, nullptr);
  return IXG_OK;
}

int ixg_hmm2ij_bands_trunc(const ixg_hbmap *m, int32_t *pn, const int32_t *occ, int32_t i0, int32_t j0, int pass_idx,
                           int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax, uint8_t *jvalid,
                           uint8_t *lvalid, uint8_t *rvalid, uint8_t *tvalid, int32_t *spep) {
  if (!m || !pn || !occ || !imin || !imax || !jmin || !jmax || !jvalid || !lvalid || !rvalid || !tvalid || i0 < 1 ||
      j0 < i0 || pass_idx < IXG_PASS_5P_ONLY_FORCE || pass_idx > IXG_PASS_5P_AND_3P_ANY) {
    g_err = "ixg_hmm2ij_bands_trunc: bad arguments";
    return IXG_EINVAL;
  }
  TrBand tb;
  tr_predict(pn, occ, m->hmm_m, i0, j0, tb);
  tr_candidates(*m, pass_idx, tb);
  hmm2ij(*m, pn, i0, j0, true, imin, imax, jmin, jmax, &tb);
  std::copy(tb.jv.begin(), tb.jv.end(), jvalid);
  std::copy(tb.lv.begin(), tb.lv.end(), lvalid);
  std::copy(tb.rv.begin(), tb.rv.end(), rvalid);
  std::copy(tb.tv.begin(), tb.tv.end(), tvalid);
  if (spep) {
    const int32_t v[8] = {tb.sp1, tb.sp2, tb.ep1, tb.ep2, tb.rmarg_imin, tb.rmarg_imax, tb.lmarg_jmin, tb.lmarg_jmax};
    std::copy(v, v + 8, spep);
  }
  return IXG_OK;
}

int ixg_ij2d_bands(const ixg_hbmap *m, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                   const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax, int64_t cap,
                   int64_t *nrows) {
  if (!m || !imin || !imax || !jmin || !jmax || !hd_off || !nrows || (cap && (!hdmin || !hdmax))) {
    g_err = "ixg_ij2d_bands: bad arguments";
    return IXG_EINVAL;
  }
  std::vector<int32_t> a, b;
  *nrows = ij2d(*m, imin, imax, jmin, jmax, do_trunc != 0, hd_off, a, b);
  if (*nrows > cap) return IXG_ERANGE;
  std::copy(a.begin(), a.end(), hdmin);
  std::copy(b.begin(), b.end(), hdmax);
  return IXG_OK;
}

}  // extern "C"
