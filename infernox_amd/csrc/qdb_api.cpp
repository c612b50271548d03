// C ABI of the query-dependent-banded (or non-banded) scanners
// (cm_nohmm.rs generic_scan, cyk_seq_filter, final_stage_inside_opt): the
// resident integer tables, per-call band vectors, windows launched in groups
// that fit the row budget, and the host side of the reporting (greedy per j
// with null3, cm_nohmm.rs:769-813; overlap removal, cp9.rs:6544; the
// envelope padding and merge of cm_nohmm.rs:855-894).
#include <algorithm>
#include <cmath>
#include <cstring>
#include "cm_handle.h"

using namespace ixg;

int qdb_upload(ixg_ctx *c, ixg_cm *cm) {
  const CmHost &h = cm->h;
  std::vector<int32_t> q(h.itsc);
  auto put = [&](const std::vector<int32_t> &x) { q.insert(q.end(), x.begin(), x.end()); };
  auto putf = [&](const std::vector<float> &x) {
    for (float f : x) {
      int32_t b;
      memcpy(&b, &f, 4);
      q.push_back(b);
    }
  };
  put(h.ioesc);
  put(h.ibeginsc);
  put(h.iendsc);
  put(h.ilog);
  put(h.qbegl);
  put(h.qibase0);
  put(h.qiel);
  put(h.qibbase0);
  putf(h.qfbase0);
  putf(h.qfel);
  putf(h.qfbbase0);
  put(h.qblist);
  q.push_back(0);
  cm->b_q = q.size() * 4;
  RCHK(dmalloc(c, (void **)&cm->d_q, cm->b_q));
  HIPCHK(hipMemcpy(cm->d_q, q.data(), cm->b_q, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  // truncated QDB: integer marginal emissions, then integer penalties
  // (-INFTY, i.e. no truncated begin, until ixg_cm_set_itrpenalties)
  std::vector<int32_t> t(h.ilmesc);
  t.insert(t.end(), h.irmesc.begin(), h.irmesc.end());
  t.resize(t.size() + 6 * (size_t)h.M, -INFTY_I);
  cm->b_tq = t.size() * 4;
  RCHK(dmalloc(c, (void **)&cm->d_tq, cm->b_tq));
  HIPCHK(hipMemcpy(cm->d_tq, t.data(), cm->b_tq, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  return IXG_OK;
}

namespace {

struct QdbView {  // device pointers into cm->d_q
  const int32_t *itsc, *ioesc, *ibeginsc, *iendsc, *ilog, *begl, *ibase0, *iel, *ibbase0, *blist;
  const float *fbase0, *fel, *fbbase0;
};
QdbView qview(const ixg_cm *cm) {
  const CmHost &h = cm->h;
  const int M = h.M, nb = (int)h.qibbase0.size();
  QdbView v;
  const int32_t *p = cm->d_q;
  v.itsc = p;
  p += (size_t)M * MAXC;
  v.ioesc = p;
  p += (size_t)M * KP * KP;
  v.ibeginsc = p;
  p += M;
  v.iendsc = p;
  p += M;
  v.ilog = p;
  p += FLOGSUM_TBL;
  v.begl = p;
  p += M;
  v.ibase0 = p;
  p += M;
  v.iel = p;
  p += h.W + 1;
  v.ibbase0 = p;
  p += nb;
  v.fbase0 = (const float *)p;
  p += M;
  v.fel = (const float *)p;
  p += h.W + 1;
  v.fbbase0 = (const float *)p;
  p += nb;
  v.blist = p;
  return v;
}

struct HostCand {
  int32_t j, d;
  float sc;
};

// Greedy reporting per j over one window's candidates, null3, then overlap
// removal (cm_nohmm.rs:769-813, cp9.rs:6544-6574); appends to out.
void qdb_report(const ixg_cm &cm, const uint8_t *dsq, int32_t i0, int32_t j0, std::vector<HostCand> &cand,
                float cutoff, int do_null3, int32_t env, std::vector<ixg_hit> &out) {
  std::sort(cand.begin(), cand.end(), [](const HostCand &a, const HostCand &b) {
    return a.j != b.j ? a.j < b.j : a.d < b.d;
  });
  std::vector<double> act;
  if (do_null3) {
    const int L = j0 - i0 + 1;
    act.assign((size_t)(L + 1) * 4, 0.0);
    for (int jp = 1; jp <= L; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      if (dsq[i0 + jp - 1] < 4) act[jp * 4 + dsq[i0 + jp - 1]] += 1.0;
    }
  }
  std::vector<ixg_hit> tmp;
  size_t k = 0;
  while (k < cand.size()) {
    const int j = cand[k].j;
    float maxrep = IMPOSSIBLE;
    for (; k < cand.size() && cand[k].j == j; k++) {
      const int d = cand[k].d, i = j - d + 1;
      if (d < 1) continue;
      float hs = cand[k].sc, bias = 0.f;
      if (!(hs > maxrep && hs >= cutoff && hs > IMPOSSIBLE + 1.0f)) continue;
      bool rep = true;
      if (do_null3) {
        float comp[4];
        const int ip = i - i0 + 1, jj = j - i0 + 1;
        for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
        float s = 0.f;
        for (int a = 0; a < 4; a++) s += comp[a];
        if (s > 0.f)
          for (int a = 0; a < 4; a++) comp[a] /= s;
        else
          for (int a = 0; a < 4; a++) comp[a] = 0.25f;
        const float corr = null3_score(cm.h, cm.ftbl_host, comp, d);
        hs -= corr;
        bias = corr;
        rep = hs > maxrep && hs >= cutoff;
      }
      if (rep) {
        tmp.push_back({i, j, hs, bias, env});
        maxrep = hs;
      }
    }
  }
  std::stable_sort(tmp.begin(), tmp.end(), [](const ixg_hit &a, const ixg_hit &b) {
    if (a.score != b.score) return a.score > b.score;
    return a.i < b.i;
  });
  std::vector<char> removed(tmp.size(), 0);
  for (size_t a = 0; a < tmp.size(); a++) {
    if (removed[a]) continue;
    for (size_t b = a + 1; b < tmp.size(); b++)
      if (!removed[b] && !(tmp[b].j < tmp[a].i) && !(tmp[a].j < tmp[b].i)) removed[b] = 1;
  }
  for (size_t a = 0; a < tmp.size(); a++)
    if (!removed[a]) out.push_back(tmp[a]);
}

// The wavefront schedule of one call (the per-call bands decide which cells
// exist): per-state levels, ring depths and local-begin slots, the deck
// gaps, and the (state, d) cells dealt as 64-lane items to the waves.
struct QdbSched {
  std::vector<int32_t> blob;  // lvl | acc | row_state | wave_item | ring | deck | item | unit
  size_t o_lvl, o_acc, o_row, o_wave, o_ring, o_deck, o_item, o_unit, o_tb;
  int nw, nrows, deck_extra, nslot, amask, lroot, nitems, ntb;
};

static int pow2_at_least(int n) {
  int r = 1;
  while (r < n) r <<= 1;
  return r;
}

// tbegin (truncated scanners): per state, whether the pass gives it a
// truncated begin (tpen: the penalties' bits); the root then reads those
// states at its own j, keeps two parity rows of its own, and has no
// transitions of its own.
void qdb_schedule(const CmHost &h, const std::vector<int32_t> &bands, int local, const std::vector<char> *tbegin,
                  const int32_t *tpen, int nonbanded, QdbSched &s) {
  const int M = h.M, W = h.W;
  const int32_t *dn = bands.data(), *dx = dn + M, *dminw = dn + 2 * M;
  const int nbegl = (int)h.qibbase0.size();
  std::vector<int32_t> lvl(M, 0);
  for (size_t l = 0; l + 1 < h.lvl_off.size(); l++)
    for (int k = h.lvl_off[l]; k < h.lvl_off[l + 1]; k++) lvl[h.lvl_states[k]] = (int32_t)l;
  s.lroot = lvl[0];
  // a cell of state y written at step j + lvl[y] is read at latest by a
  // parent p at step j + delta + lvl[p] (delta = 1 for right emitters), so
  // its ring must not come round before then
  std::vector<int32_t> need(M, 1), gap(std::max(nbegl, 1), 0);
  if (tbegin) {
    need[0] = 2;
    for (int y = 1; y < M; y++)
      if ((*tbegin)[y]) need[y] = std::max(need[y], s.lroot - lvl[y] + 1);
  }
  for (int p = tbegin ? 1 : 0; p < M; p++) {
    const int st = h.sttype[p];
    if (st == E_ST) continue;
    if (st == B_ST) {
      const int y = h.cnum[p], wl = h.cfirst[p];
      need[y] = std::max(need[y], lvl[p] - lvl[y] + 1);
      gap[h.qbegl[wl]] = std::max(gap[h.qbegl[wl]], lvl[p] - lvl[wl]);
      continue;
    }
    const int dl = (st == MP_ST || st == MR_ST || st == IR_ST) ? 1 : 0;
    for (int k = 0; k < h.cnum[p]; k++) {
      const int y = h.cfirst[p] + k;
      need[y] = std::max(need[y], lvl[p] - lvl[y] + dl + 1);
    }
  }
  // (the standard scanners' IL states keep one scratch row past the ring)
  std::vector<int32_t> ring(2 * (size_t)M), row_state;
  for (int v = 0; v < M; v++) {
    const int r = pow2_at_least(need[v]);
    ring[2 * v] = (int32_t)row_state.size();
    ring[2 * v + 1] = r - 1;
    row_state.insert(row_state.end(), r + (!tbegin && h.sttype[v] == IL_ST ? 1 : 0), v);
  }
  s.nrows = (int)row_state.size();
  std::vector<int32_t> deck(2 * (size_t)std::max(nbegl, 1), 0);
  s.deck_extra = 0;
  for (int b = 0; b < nbegl; b++) {
    deck[2 * b] = s.deck_extra;
    deck[2 * b + 1] = gap[b];
    s.deck_extra += gap[b];
  }
  // local begins: one slot per begin state of a level; a (j, slot) row is
  // first written at step j + (lowest begin level) and cleared by the root at
  // step j + lroot
  std::vector<int32_t> acc(M, -1);
  s.nslot = 0;
  s.amask = 0;
  if (!tbegin && local && !h.qblist.empty()) {
    std::vector<int32_t> per(s.lroot + 1, 0);
    int amin = s.lroot;
    for (int y : h.qblist) {
      acc[y] = per[lvl[y]]++;
      s.nslot = std::max(s.nslot, per[lvl[y]]);
      amin = std::min(amin, (int)lvl[y]);
    }
    s.amask = pow2_at_least(s.lroot - amin + 1) - 1;
  }
  // the cells: class 0 ordinary cells and the root, 1 bifurcations, 2 IL rows
  struct U {
    int cls, lvl, v, d, cost, code;
  };
  std::vector<U> us;
  const int pw = tbegin ? 3 : 1;  // cost weight: planes per cell
  int ntb = 0;
  if (tbegin)
    for (int y = 1; y < M; y++) ntb += (*tbegin)[y] ? 1 : 0;
  for (int v = 1; v < M; v++) {
    const int st = h.sttype[v], cn = h.cnum[v];
    if (st == E_ST) continue;
    // truncated scanners: an emitter followed by an END node's E state is
    // detached (cm_dpsearch_trunc.rs:88)
    if (tbegin && (st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST) && v + 1 < M && h.sttype[v + 1] == E_ST)
      continue;
    const int a = dn[v], b = std::min(W, (int)dx[v]);
    if (b < a) continue;
    if (st == IL_ST) {
      us.push_back({2, lvl[v], v, a, (b - a + 1) * (cn + 4) * pw, QU_IL << 28 | v});
    } else if (st == B_ST) {
      const int wl = h.cfirst[v], y = cn;
      for (int d = a; d <= b; d++) {
        int kmin, kmax;
        if (!tbegin) {
          kmin = std::max(0, std::max((int)dminw[y], d - (int)dx[wl]));
          kmax = std::min((int)dx[y], d - (int)dminw[wl]);
        } else {
          kmin = nonbanded ? 0 : std::max(0, d - (int)dx[wl]);
          kmax = nonbanded ? d : std::min((int)dx[y], d);
        }
        us.push_back({1, lvl[v], v, d, (std::max(0, kmax - kmin + 1) * 3 + 4) * pw, QU_CELL << 28 | v});
      }
    } else {
      for (int d = a; d <= b; d++)
        us.push_back({0, lvl[v], v, d, (cn + 4 + (acc[v] >= 0 ? 2 : 0)) * pw, QU_CELL << 28 | v});
    }
  }
  for (int d = dn[0], b = std::min(W, (int)dx[0]); d <= b; d++)
    us.push_back({0, s.lroot, 0, d, tbegin ? ntb * 4 * 12 : h.cnum[0] + 4 + s.nslot, QU_ROOT << 28});
  std::sort(us.begin(), us.end(), [](const U &x, const U &y) {
    if (x.cls != y.cls) return x.cls < y.cls;
    if (x.lvl != y.lvl) return x.lvl < y.lvl;
    if (x.v != y.v) return x.v < y.v;
    return x.d < y.d;
  });
  struct It {
    int lmin, lmax, dmin, cost;
    size_t u0, n;
  };
  std::vector<It> its;
  for (size_t k = 0; k < us.size(); k++) {
    if (its.empty() || its.back().n == 64 || us[its.back().u0].cls != us[k].cls)
      its.push_back({us[k].lvl, us[k].lvl, us[k].d, 0, k, 0});
    It &t = its.back();
    t.lmin = std::min(t.lmin, us[k].lvl);
    t.lmax = std::max(t.lmax, us[k].lvl);
    t.dmin = std::min(t.dmin, us[k].d);
    t.cost = std::max(t.cost, us[k].cost);
    t.n++;
  }
  s.nitems = (int)its.size();
  // waves: enough that the heaviest item is about one wave's share per step
  int64_t total = 0, heavy = 1;
  for (const It &t : its) {
    total += t.cost;
    heavy = std::max<int64_t>(heavy, t.cost);
  }
  s.nw = (int)std::min<int64_t>(QDB_MAX_WAVES, std::max<int64_t>({4, (total + heavy - 1) / heavy, s.nitems / 2}));
  // deal items in level order to the least-loaded wave, which keeps every
  // level (and so the steps where only some levels are live) spread out
  std::vector<size_t> order(its.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return its[a].lmin < its[b].lmin; });
  std::vector<int64_t> load(s.nw, 0);
  std::vector<std::vector<size_t>> per_wave(s.nw);
  for (size_t k : order) {
    const int wv = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[wv] += its[k].cost;
    per_wave[wv].push_back(k);
  }
  auto pack = [&](int v) {
    int lg = 0;
    while ((1 << lg) <= ring[2 * v + 1]) lg++;
    return ring[2 * v] | lg << RING_SHIFT;
  };
  // one lane's record (QdbUnit, sixteen words)
  auto record = [&](const U &x, std::vector<int32_t> &out) {
    QdbUnit r;
    memset(&r, 0, sizeof r);
    r.deck = -1;
    r.code = x.code;
    r.d = x.d;
    const int v = x.v, kind = x.code >> 28;
    if (kind != QU_NONE) {
      const int st = h.sttype[v], cn = st == B_ST ? 0 : h.cnum[v];
      r.lvl = lvl[v];
      r.info = st | (cn & 15) << 4 | (acc[v] + 1) << 8;
      r.y0 = h.cfirst[v];
      r.ring = pack(v);
      if (st == B_ST) {
        const int wl = h.cfirst[v], y = h.cnum[v], bi = h.qbegl[wl];
        r.y1 = y;
        r.deck = bi;
        r.cring[0] = pack(y);
        r.cring[1] = deck[2 * bi];
        r.cring[2] = deck[2 * bi + 1];
        if (!tbegin) {
          r.k0 = std::max(0, std::max((int)dminw[y], x.d - (int)dx[wl]));
          r.k1 = std::min((int)dx[y], x.d - (int)dminw[wl]);
        } else {
          r.k0 = nonbanded ? 0 : std::max(0, x.d - (int)dx[wl]);
          r.k1 = nonbanded ? x.d : std::min((int)dx[y], x.d);
        }
      } else {
        if (st != E_ST && !(tbegin && v == 0))
          for (int k = 0; k < cn && k < MAXC; k++) r.cring[k] = pack(h.cfirst[v] + k);
        if (h.qbegl[v] >= 0) {
          r.deck = h.qbegl[v];
          r.k0 = deck[2 * r.deck];
          r.k1 = deck[2 * r.deck + 1];
        }
        if (st == IL_ST) r.y1 = std::min(W, (int)dx[v]);
      }
    }
    const int32_t *w = (const int32_t *)&r;
    out.insert(out.end(), w, w + 16);
  };
  std::vector<int32_t> wave_item(s.nw + 1, 0), item, unit;
  for (int wv = 0; wv < s.nw; wv++) {
    for (size_t k : per_wave[wv]) {
      const It &t = its[k];
      item.insert(item.end(), {t.lmin, t.lmax, t.dmin, 0});
      for (size_t l = 0; l < 64; l++) record(l < t.n ? us[t.u0 + l] : U{0, 0, 0, 0, 0, QU_NONE << 28}, unit);
    }
    wave_item[wv + 1] = wave_item[wv] + (int32_t)per_wave[wv].size();
  }
  // one blob, 16-byte aligned parts
  s.blob.clear();
  auto part = [&](const std::vector<int32_t> &x) {
    while (s.blob.size() % 4) s.blob.push_back(0);
    const size_t o = s.blob.size();
    s.blob.insert(s.blob.end(), x.begin(), x.end());
    return o;
  };
  s.o_lvl = part(lvl);
  s.o_acc = part(acc);
  s.o_row = part(row_state);
  s.o_wave = part(wave_item);
  s.o_ring = part(ring);
  s.o_deck = part(deck);
  s.o_item = part(item);
  s.o_unit = part(unit);
  // the truncated root's begin states in increasing order:
  // {ring, penalty bits, dxcap, state | bifurcation << 30}
  std::vector<int32_t> tbl;
  if (tbegin)
    for (int y = 1; y < M; y++)
      if ((*tbegin)[y]) tbl.insert(tbl.end(), {pack(y), tpen[y], dx[y], y | (h.sttype[y] == B_ST ? 1 << 30 : 0)});
  s.ntb = (int)(tbl.size() / 4);
  s.o_tb = part(tbl);
  s.blob.push_back(0);
}

// The kernels hold the ILogsum table in LDS as 16-bit entries.
int check_ilog16(const CmHost &h) {
  for (int32_t x : h.ilog)
    if (x < 0 || x > 65535) {
      g_err = "ILogsum table entry outside 0..65535";
      return IXG_EINVAL;
    }
  return IXG_OK;
}

struct QWin {
  int32_t i0, j0;
  int64_t seq_off;
  const uint8_t *host_dsq;
};

// Runs the windows (over the device sequence base + seq_off) and appends
// every window's hits (env = window index) to out.
int qdb_run(ixg_ctx *c, const ixg_cm *cm, const uint8_t *d_base, const std::vector<QWin> &wins,
            const int32_t *dmin, const int32_t *dmax, int inside, float cutoff, int do_null3,
            std::vector<ixg_hit> &out) {
  const CmHost &h = cm->h;
  const int M = h.M, W = h.W, nbegl = (int)std::count_if(h.qbegl.begin(), h.qbegl.end(), [](int b) { return b >= 0; });
  if (inside) RCHK(check_ilog16(h));
  // per-state bands of this call (cm_nohmm.rs:444-455)
  std::vector<int32_t> bands(3 * (size_t)M);
  for (int v = 0; v < M; v++) {
    const int mn = dmin ? dmin[v] : 0, mx = dmax ? dmax[v] : W;
    if (mn < 0 || mx < 0) {
      g_err = "QDB scan: negative band";
      return IXG_EINVAL;
    }
    const int base = h.sttype[v] == MP_ST ? 2 : 1;
    bands[v] = std::min(std::max(mn, base), W);
    bands[M + v] = std::min(mx, W);
    bands[2 * M + v] = std::min(mn, W);
  }
  int32_t *d_bands;
  RCHK(scratch(c, S_CM_BANDS, bands.size() * 4, (void **)&d_bands));
  HIPCHK(hipMemcpyAsync(d_bands, bands.data(), bands.size() * 4, hipMemcpyHostToDevice, c->stream));
  const QdbView qv = qview(cm);
  const CmDev cd = cm->dev();
  QdbSched sc;
  qdb_schedule(h, bands, h.local_begin, nullptr, nullptr, 0, sc);
  int32_t *d_sched;
  RCHK(scratch(c, S_CM_SCHED, sc.blob.size() * 4, (void **)&d_sched));
  HIPCHK(hipMemcpyAsync(d_sched, sc.blob.data(), sc.blob.size() * 4, hipMemcpyHostToDevice, c->stream));
  QdbDev q;
  q.sttype = cd.sttype;
  q.cfirst = cd.cfirst;
  q.cnum = cd.cnum;
  q.begl = qv.begl;
  q.dn = d_bands;
  q.dxcap = d_bands + M;
  q.dminw = d_bands + 2 * M;
  q.ilog = qv.ilog;
  q.lvl = d_sched + sc.o_lvl;
  q.acc = d_sched + sc.o_acc;
  q.ring = (const int2 *)(d_sched + sc.o_ring);
  q.row_state = d_sched + sc.o_row;
  q.deck = (const int2 *)(d_sched + sc.o_deck);
  q.item = (const int4 *)(d_sched + sc.o_item);
  q.unit = (const QdbUnit *)(d_sched + sc.o_unit);
  q.wave_item = d_sched + sc.o_wave;
  q.M = M;
  q.W = W;
  q.nbegl = nbegl;
  q.local = h.local_begin;
  q.nw = sc.nw;
  q.nrows = sc.nrows;
  q.deck_extra = sc.deck_extra;
  q.nslot = sc.nslot;
  q.amask = sc.amask;
  q.lroot = sc.lroot;
  QdbTabs<float> tf{cd.tsc, cd.oesc, cd.endsc, cd.beginsc, qv.fbase0, qv.fel, qv.fbbase0};
  QdbTabs<int32_t> ti{qv.itsc, qv.ioesc, qv.iendsc, qv.ibeginsc, qv.ibase0, qv.iel, qv.ibbase0};
  const void *tabs = inside ? (const void *)&ti : (const void *)&tf;
  const int64_t n = (int64_t)wins.size();
  const int64_t budget = arena_elems(c, S_CM_ALPHA);  // row and deck elements per launch group
  int64_t k0 = 0;
  while (k0 < n) {
    int64_t k1 = k0, elems = 0, res = 0;
    std::vector<int64_t> at;
    while (k1 < n) {
      const int L = wins[k1].j0 - wins[k1].i0 + 1, w = std::min(W, L), W1 = w + 1;
      const int64_t need =
          ((int64_t)sc.nrows + (int64_t)nbegl * W1 + sc.deck_extra + (int64_t)(sc.amask + 1) * sc.nslot) * W1;
      if (k1 > k0 && elems + need > budget) break;
      at.push_back(elems);
      elems += need;
      res += L;
      k1++;
    }
    void *d_mx;
    QdbWinDev *d_w;
    RCHK(scratch(c, S_CM_ALPHA, (size_t)elems * 4, &d_mx));
    RCHK(scratch(c, S_CM_ENV, (size_t)(k1 - k0) * sizeof(QdbWinDev), (void **)&d_w));
    std::vector<QdbWinDev> wd(k1 - k0);
    for (int64_t k = k0; k < k1; k++) {
      QdbWinDev &x = wd[k - k0];
      const int L = wins[k].j0 - wins[k].i0 + 1, w = std::min(W, L);
      x.dsq = d_base + wins[k].seq_off;
      x.i0 = wins[k].i0;
      x.j0 = wins[k].j0;
      x.w = w;
      x.win = (int32_t)k;
      x.rows = (int32_t *)d_mx + at[k - k0];
      x.begl = nullptr;
    }
    HIPCHK(hipMemcpyAsync(d_w, wd.data(), wd.size() * sizeof(QdbWinDev), hipMemcpyHostToDevice, c->stream));
    int64_t cap = (int64_t)((double)res * c->qdb_cand_rate * 1.25) + 4096;
    int maxlen = 0;
    for (int64_t k = k0; k < k1; k++) maxlen = std::max(maxlen, wins[k].j0 - wins[k].i0 + 1);
    std::vector<QdbCand> hc;
    for (int attempt = 0; attempt < 2; attempt++) {
      QdbCand *d_c;
      unsigned long long *d_cnt;
      RCHK(scratch(c, S_CM_ROOT, (size_t)cap * sizeof(QdbCand), (void **)&d_c));
      RCHK(scratch(c, S_COUNTER, 8, (void **)&d_cnt));
      HIPCHK(hipMemsetAsync(d_cnt, 0, 8, c->stream));
      HIPCHK(ktime_begin(c));
      HIPCHK(launch_qdb_scan(q, tabs, inside, d_w, k1 - k0, cutoff, d_c, cap, d_cnt, maxlen, c->stream));
      HIPCHK(ktime_end(c, IXG_K_QDB_SCAN));
      unsigned long long cnt = 0;
      RCHK(d2h(c, &cnt, d_cnt, 8));
      c->qdb_cand_rate = std::max(c->qdb_cand_rate, (double)cnt / (double)std::max<int64_t>(res, 1));
      if ((int64_t)cnt > cap) {
        cap = (int64_t)cnt + 1024;
        continue;
      }
      hc.resize(cnt);
      if (cnt) RCHK(d2h(c, hc.data(), d_c, cnt * sizeof(QdbCand)));
      break;
    }
    std::vector<std::vector<HostCand>> per(k1 - k0);
    for (const QdbCand &x : hc) per[x.win - k0].push_back({x.j, x.d, x.sc});
    for (int64_t k = k0; k < k1; k++)
      qdb_report(*cm, wins[k].host_dsq, wins[k].i0, wins[k].j0, per[k - k0], cutoff, do_null3, (int32_t)k, out);
    k0 = k1;
  }
  return IXG_OK;
}

// ---- truncated QDB scanners (cm_dpsearch_trunc.rs:1800 generic_tr_scan) ----

struct TrqHostCand {
  int32_t j, d;
  float sc;
  int32_t root, mode;
};

// Greedy reporting per j (report_hits_greedily with cm_hit_AllowTruncation
// and null3, cm_dpsearch_trunc.rs:2322-2422), then overlap removal
// (remove_overlaps_greedy_tr, :2470-2495); appends to out.
void trq_report(const ixg_cm &cm, const uint8_t *dsq, int32_t i0, int32_t j0, std::vector<TrqHostCand> &cand,
                int pass_idx, int local, float cutoff, int do_null3, int32_t env, std::vector<ixg_trhit> &out) {
  std::sort(cand.begin(), cand.end(), [](const TrqHostCand &a, const TrqHostCand &b) {
    return a.j != b.j ? a.j < b.j : a.d < b.d;
  });
  std::vector<double> act;
  if (do_null3) {
    const int L = j0 - i0 + 1;
    act.assign((size_t)(L + 1) * 4, 0.0);
    for (int jp = 1; jp <= L; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      if (dsq[i0 + jp - 1] < 4) act[jp * 4 + dsq[i0 + jp - 1]] += 1.0;
    }
  }
  std::vector<ixg_trhit> tmp;
  size_t k = 0;
  while (k < cand.size()) {
    const int j = cand[k].j;
    float maxrep = IMPOSSIBLE;
    for (; k < cand.size() && cand[k].j == j; k++) {
      const int d = cand[k].d, i = j - d + 1;
      float hs = cand[k].sc, bias = 0.f;
      if (!(hs > maxrep && hs >= cutoff && hs > IMPOSSIBLE + 1.0f)) continue;
      bool rep = tr_allow(cm.h, pass_idx, local, i, j, i0, j0, cand[k].mode, cand[k].root);
      if (rep && do_null3) {
        float comp[4];
        const int ip = i - i0 + 1, jj = j - i0 + 1;
        for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
        float sum = 0.f;
        for (int a = 0; a < 4; a++) sum += comp[a];
        if (sum > 0.f)
          for (int a = 0; a < 4; a++) comp[a] /= sum;
        else
          for (int a = 0; a < 4; a++) comp[a] = 0.25f;
        const float corr = null3_score(cm.h, cm.ftbl_host, comp, d);
        hs -= corr;
        bias = corr;
        rep = hs > maxrep && hs >= cutoff;
      }
      if (rep) {
        tmp.push_back({i, j, hs, bias, cand[k].mode, cand[k].root, env});
        maxrep = hs;
      }
    }
  }
  std::stable_sort(tmp.begin(), tmp.end(), [](const ixg_trhit &a, const ixg_trhit &b) {
    if (a.score != b.score) return a.score > b.score;
    return a.i < b.i;
  });
  std::vector<char> removed(tmp.size(), 0);
  for (size_t a = 0; a < tmp.size(); a++) {
    if (removed[a]) continue;
    for (size_t b = a + 1; b < tmp.size(); b++)
      if (!removed[b] && !(tmp[b].j < tmp[a].i) && !(tmp[a].j < tmp[b].i)) removed[b] = 1;
  }
  for (size_t a = 0; a < tmp.size(); a++)
    if (!removed[a]) out.push_back(tmp[a]);
}

// Runs the truncated scan of every window (over the device sequence base +
// seq_off); appends each window's hits (env = window index) to out and its
// summary to summ.
int trq_run(ixg_ctx *c, const ixg_cm *cm, const uint8_t *d_base, const std::vector<QWin> &wins, const int32_t *dmax,
            int inside, int pass_idx, int local, float cutoff, int do_null3, std::vector<ixg_trhit> &out,
            std::vector<TrqSummary> &summ) {
  const CmHost &h = cm->h;
  const int M = h.M, W = h.W;
  const int nbegl = (int)std::count_if(h.qbegl.begin(), h.qbegl.end(), [](int b) { return b >= 0; });
  int row, fill;
  switch (pass_idx) {  // cm_trunc.rs:99-107, cm_dpsearch_trunc.rs:184-199
    case IXG_PASS_5P_ONLY_FORCE: row = 1, fill = 4; break;
    case IXG_PASS_3P_ONLY_FORCE: row = 2, fill = 2; break;
    case IXG_PASS_5P_AND_3P_FORCE:
    case IXG_PASS_5P_AND_3P_ANY: row = 0, fill = 2 | 4 | 8; break;
    default:
      g_err = "truncated QDB scan: pass_idx must be a truncated pass";
      return IXG_EINVAL;
  }
  if (inside) RCHK(check_ilog16(h));
  if (inside ? h.trp_ig.empty() : h.trp_g.empty()) {
    g_err = inside ? "truncated QDB Inside: ixg_cm_set_itrpenalties has not been called"
                   : "truncated QDB CYK: ixg_cm_set_trpenalties has not been called";
    return IXG_EINVAL;
  }
  std::vector<int32_t> dxcap(M);
  for (int v = 0; v < M; v++) {
    if (dmax && dmax[v] < 0) {
      g_err = "truncated QDB scan: negative band";
      return IXG_EINVAL;
    }
    dxcap[v] = std::min(dmax ? dmax[v] : W, W);
  }
  int32_t *d_dx;
  RCHK(scratch(c, S_CM_BANDS, (size_t)M * 4, (void **)&d_dx));
  HIPCHK(hipMemcpyAsync(d_dx, dxcap.data(), (size_t)M * 4, hipMemcpyHostToDevice, c->stream));
  const size_t pq = (size_t)((local ? 3 : 0) + row) * M;
  // the states with a truncated begin in this pass (cm_dpsearch_trunc.rs:2220-2289)
  std::vector<char> tbegin(M, 0);
  std::vector<int32_t> tpen(M, 0);
  for (int y = 1; y < M; y++) {
    if (inside) {
      const std::vector<int32_t> &pt = local ? h.trp_il : h.trp_ig;
      tpen[y] = pt[(size_t)row * M + y];
      tbegin[y] = tpen[y] != -INFTY_I;
    } else {
      const std::vector<float> &pt = local ? h.trp_l : h.trp_g;
      memcpy(&tpen[y], &pt[(size_t)row * M + y], 4);
      tbegin[y] = pt[(size_t)row * M + y] > IMPOSSIBLE + 1.0f;
    }
  }
  std::vector<int32_t> bands(3 * (size_t)M, 0);
  for (int v = 0; v < M; v++) {
    bands[v] = 1;
    bands[M + v] = dxcap[v];
  }
  QdbSched sc;
  qdb_schedule(h, bands, local, &tbegin, tpen.data(), dmax ? 0 : 1, sc);
  int32_t *d_sched;
  RCHK(scratch(c, S_CM_SCHED, sc.blob.size() * 4, (void **)&d_sched));
  HIPCHK(hipMemcpyAsync(d_sched, sc.blob.data(), sc.blob.size() * 4, hipMemcpyHostToDevice, c->stream));
  const QdbView qv = qview(cm);
  const CmDev cd = cm->dev();
  TrqDev q;
  q.sttype = cd.sttype;
  q.cfirst = cd.cfirst;
  q.cnum = cd.cnum;
  q.begl = qv.begl;
  q.dxcap = d_dx;
  q.ilog = qv.ilog;
  q.lvl = d_sched + sc.o_lvl;
  q.ring = (const int2 *)(d_sched + sc.o_ring);
  q.row_state = d_sched + sc.o_row;
  q.deck = (const int2 *)(d_sched + sc.o_deck);
  q.item = (const int4 *)(d_sched + sc.o_item);
  q.unit = (const QdbUnit *)(d_sched + sc.o_unit);
  q.wave_item = d_sched + sc.o_wave;
  q.M = M;
  q.W = W;
  q.nbegl = nbegl;
  q.fill = fill;
  q.nonbanded = dmax ? 0 : 1;
  q.nw = sc.nw;
  q.nrows = sc.nrows;
  q.deck_extra = sc.deck_extra;
  q.lroot = sc.lroot;
  q.tblist = (const int4 *)(d_sched + sc.o_tb);
  q.ntb = sc.ntb;
  TrqTabs<float> tf{cd.tsc, cd.oesc, cd.endsc, qv.fel, qv.fbase0, qv.fbbase0, cd.lmesc, cd.rmesc, cm->d_tr + pq};
  const int32_t *tq = cm->d_tq;
  TrqTabs<int32_t> ti{qv.itsc, qv.ioesc, qv.iendsc, qv.iel, qv.ibase0, qv.ibbase0,
                      tq, tq + (size_t)M * KP, tq + 2 * (size_t)M * KP + pq};
  const void *tabs = inside ? (const void *)&ti : (const void *)&tf;
  const int64_t n = (int64_t)wins.size();
  summ.assign(n, TrqSummary{IMPOSSIBLE, IXG_TRMODE_UNKNOWN, IMPOSSIBLE});
  const int64_t budget = arena_elems(c, S_CM_ALPHA);  // plane and deck elements per launch group
  int64_t k0 = 0;
  while (k0 < n) {
    int64_t k1 = k0, elems = 0, res = 0;
    std::vector<int64_t> at;
    while (k1 < n) {
      const int L = wins[k1].j0 - wins[k1].i0 + 1, w = std::min(W, L), W1 = w + 1;
      const int64_t need = (4 * (int64_t)sc.nrows + 3 * ((int64_t)nbegl * W1 + sc.deck_extra)) * W1;
      if (k1 > k0 && elems + need > budget) break;
      at.push_back(elems);
      elems += need;
      res += L;
      k1++;
    }
    void *d_mx;
    QdbWinDev *d_w;
    TrqSummary *d_s;
    RCHK(scratch(c, S_CM_ALPHA, (size_t)elems * 4, &d_mx));
    RCHK(scratch(c, S_CM_ENV, (size_t)(k1 - k0) * sizeof(QdbWinDev), (void **)&d_w));
    RCHK(scratch(c, S_SC1, (size_t)(k1 - k0) * sizeof(TrqSummary), (void **)&d_s));
    std::vector<QdbWinDev> wd(k1 - k0);
    for (int64_t k = k0; k < k1; k++) {
      QdbWinDev &x = wd[k - k0];
      const int L = wins[k].j0 - wins[k].i0 + 1, w = std::min(W, L), W1 = w + 1;
      x.dsq = d_base + wins[k].seq_off;
      x.i0 = wins[k].i0;
      x.j0 = wins[k].j0;
      x.w = w;
      x.win = (int32_t)(k - k0);
      x.rows = (int32_t *)d_mx + at[k - k0];
      x.begl = (int32_t *)d_mx + at[k - k0] + 4 * (int64_t)sc.nrows * W1;
    }
    HIPCHK(hipMemcpyAsync(d_w, wd.data(), wd.size() * sizeof(QdbWinDev), hipMemcpyHostToDevice, c->stream));
    int64_t cap = (int64_t)((double)res * c->trq_cand_rate * 1.25) + 4096;
    int maxlen = 0;
    for (int64_t k = k0; k < k1; k++) maxlen = std::max(maxlen, wins[k].j0 - wins[k].i0 + 1);
    std::vector<TrqCand> hc;
    for (int attempt = 0; attempt < 2; attempt++) {
      TrqCand *d_c;
      unsigned long long *d_cnt;
      RCHK(scratch(c, S_CM_ROOT, (size_t)cap * sizeof(TrqCand), (void **)&d_c));
      RCHK(scratch(c, S_COUNTER, 8, (void **)&d_cnt));
      HIPCHK(hipMemsetAsync(d_cnt, 0, 8, c->stream));
      HIPCHK(ktime_begin(c));
      HIPCHK(launch_trqdb_scan(q, tabs, inside, d_w, k1 - k0, cutoff, d_c, cap, d_cnt, d_s, maxlen, c->stream));
      HIPCHK(ktime_end(c, IXG_K_TRQDB_SCAN));
      unsigned long long cnt = 0;
      RCHK(d2h(c, &cnt, d_cnt, 8));
      c->trq_cand_rate = std::max(c->trq_cand_rate, (double)cnt / (double)std::max<int64_t>(res, 1));
      if ((int64_t)cnt > cap) {
        cap = (int64_t)cnt + 1024;
        continue;
      }
      hc.resize(cnt);
      if (cnt) RCHK(d2h(c, hc.data(), d_c, cnt * sizeof(TrqCand)));
      break;
    }
    RCHK(d2h(c, summ.data() + k0, d_s, (size_t)(k1 - k0) * sizeof(TrqSummary)));
    std::vector<std::vector<TrqHostCand>> per(k1 - k0);
    for (const TrqCand &x : hc) per[x.win].push_back({x.j, x.d, x.sc, x.root, x.mode});
    for (int64_t k = k0; k < k1; k++)
      trq_report(*cm, wins[k].host_dsq, wins[k].i0, wins[k].j0, per[k - k0], pass_idx, local, cutoff, do_null3,
                 (int32_t)k, out);
    k0 = k1;
  }
  return IXG_OK;
}

int check_window(const ixg_cm *cm, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0) {
  if (!cm || !cm->d_q || !dsq || i0 < 1 || j0 > L || j0 < i0) {
    g_err = "QDB scan: bad arguments (1 <= i0 <= j0 <= L, device CM)";
    return IXG_EINVAL;
  }
  for (int64_t i = i0; i <= j0; i++)
    if (dsq[i] >= KP) {
      g_err = "digitized residue code >= 18 in the window";
      return IXG_EINVAL;
    }
  return IXG_OK;
}

}  // namespace

const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L);
const uint8_t *ixg_target_host_dsq(const ixg_target *t, int strand);

extern "C" {

int ixg_cm_itables(const ixg_cm *cm, int32_t *itsc, int32_t *ioesc, int32_t *ibeginsc, int32_t *iendsc,
                   int32_t *iel_selfsc) {
  if (!cm) return IXG_EINVAL;
  const CmHost &h = cm->h;
  if (itsc) memcpy(itsc, h.itsc.data(), h.itsc.size() * 4);
  if (ioesc) memcpy(ioesc, h.ioesc.data(), h.ioesc.size() * 4);
  if (ibeginsc) memcpy(ibeginsc, h.ibeginsc.data(), h.ibeginsc.size() * 4);
  if (iendsc) memcpy(iendsc, h.iendsc.data(), h.iendsc.size() * 4);
  if (iel_selfsc) *iel_selfsc = h.iel_selfsc;
  return IXG_OK;
}

int ixg_qdb_scan(ixg_ctx *c, const ixg_cm *cm, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0,
                 const int32_t *dmin, const int32_t *dmax, int inside, float cutoff, int do_null3, ixg_hit *hits,
                 int64_t cap, int64_t *nhits) {
  if (!c || !nhits || (cap > 0 && !hits)) return IXG_EINVAL;
  RCHK(check_window(cm, dsq, L, i0, j0));
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)L + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)L + 1, hipMemcpyHostToDevice, c->stream));
  std::vector<ixg_hit> out;
  RCHK(qdb_run(c, cm, d, {QWin{i0, j0, 0, dsq}}, dmin, dmax, inside ? 1 : 0, cutoff, do_null3, out));
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_qdb_scan_batch(ixg_ctx *c, const ixg_cm *cm, const ixg_target *tgt, const ixg_qdbwin *win, int64_t n,
                       const int32_t *dmin, const int32_t *dmax, int inside, float cutoff, int do_null3,
                       ixg_hit *hits, int64_t cap, int64_t *nhits) {
  if (!c || !cm || !cm->d_q || !tgt || (n && !win) || !nhits || (cap > 0 && !hits)) {
    g_err = "ixg_qdb_scan_batch: bad arguments";
    return IXG_EINVAL;
  }
  HIPCHK(hipSetDevice(c->device));
  int64_t L0 = 0;
  const uint8_t *base = ixg_target_device_base(tgt, &L0);
  std::vector<QWin> w(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t off = 0, L = 0;
    const uint8_t *host = nullptr;
    if ((win[k].strand != 0 && win[k].strand != 1) || !ixg_target_view(tgt, win[k].seq, win[k].strand, &off, &L, &host) ||
        win[k].i0 < 1 || win[k].j0 > L || win[k].j0 < win[k].i0) {
      g_err = "ixg_qdb_scan_batch: window outside the target";
      return IXG_EINVAL;
    }
    w[k] = QWin{win[k].i0, win[k].j0, off, host};
  }
  std::vector<ixg_hit> out;
  RCHK(qdb_run(c, cm, base, w, dmin, dmax, inside ? 1 : 0, cutoff, do_null3, out));
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_cm_set_itrpenalties(ixg_ctx *c, ixg_cm *cm, const int32_t *ig, const int32_t *il) {
  if (!c || !cm || !cm->d_tq || !ig || !il) {
    g_err = "ixg_cm_set_itrpenalties: bad arguments (device CM required)";
    return IXG_EINVAL;
  }
  const int M = cm->h.M;
  cm->h.trp_ig.assign(ig, ig + 3 * (size_t)M);
  cm->h.trp_il.assign(il, il + 3 * (size_t)M);
  HIPCHK(hipSetDevice(c->device));
  int32_t *dst = cm->d_tq + 2 * (size_t)M * KP;
  HIPCHK(hipMemcpy(dst, ig, 3 * (size_t)M * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dst + 3 * (size_t)M, il, 3 * (size_t)M * 4, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(nullptr));  // the null-stream DMA has landed before any context stream reads it
  return IXG_OK;
}

int ixg_cm_imtables(const ixg_cm *cm, int32_t *ilmesc, int32_t *irmesc) {
  if (!cm) return IXG_EINVAL;
  if (ilmesc) memcpy(ilmesc, cm->h.ilmesc.data(), cm->h.ilmesc.size() * 4);
  if (irmesc) memcpy(irmesc, cm->h.irmesc.data(), cm->h.irmesc.size() * 4);
  return IXG_OK;
}

int ixg_trqdb_scan(ixg_ctx *c, const ixg_cm *cm, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0,
                   const int32_t *dmax, int inside, int pass_idx, int local, float cutoff, int do_null3, float *vsc,
                   int32_t *vmode, float *bsc_full, ixg_trhit *hits, int64_t cap, int64_t *nhits) {
  if (!c || !vsc || !vmode || !nhits || (cap > 0 && !hits)) return IXG_EINVAL;
  RCHK(check_window(cm, dsq, L, i0, j0));
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)L + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)L + 1, hipMemcpyHostToDevice, c->stream));
  std::vector<ixg_trhit> out;
  std::vector<TrqSummary> sm;
  RCHK(trq_run(c, cm, d, {QWin{i0, j0, 0, dsq}}, dmax, inside ? 1 : 0, pass_idx, local ? 1 : 0, cutoff, do_null3,
               out, sm));
  *vsc = sm[0].vsc;
  *vmode = sm[0].vmode;
  if (bsc_full) *bsc_full = sm[0].bsc_full;
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_trqdb_scan_batch(ixg_ctx *c, const ixg_cm *cm, const ixg_target *tgt, const ixg_qdbwin *win, int64_t n,
                         const int32_t *dmax, int inside, int pass_idx, int local, float cutoff, int do_null3,
                         float *vsc, int32_t *vmode, float *bsc_full, ixg_trhit *hits, int64_t cap,
                         int64_t *nhits) {
  if (!c || !cm || !cm->d_q || !tgt || (n && (!win || !vsc || !vmode)) || !nhits || (cap > 0 && !hits)) {
    g_err = "ixg_trqdb_scan_batch: bad arguments";
    return IXG_EINVAL;
  }
  HIPCHK(hipSetDevice(c->device));
  int64_t L0 = 0;
  const uint8_t *base = ixg_target_device_base(tgt, &L0);
  std::vector<QWin> w(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t off = 0, L = 0;
    const uint8_t *host = nullptr;
    if ((win[k].strand != 0 && win[k].strand != 1) || !ixg_target_view(tgt, win[k].seq, win[k].strand, &off, &L, &host) ||
        win[k].i0 < 1 || win[k].j0 > L || win[k].j0 < win[k].i0) {
      g_err = "ixg_trqdb_scan_batch: window outside the target";
      return IXG_EINVAL;
    }
    w[k] = QWin{win[k].i0, win[k].j0, off, host};
  }
  std::vector<ixg_trhit> out;
  std::vector<TrqSummary> sm;
  RCHK(trq_run(c, cm, base, w, dmax, inside ? 1 : 0, pass_idx, local ? 1 : 0, cutoff, do_null3, out, sm));
  for (int64_t k = 0; k < n; k++) {
    vsc[k] = sm[k].vsc;
    vmode[k] = sm[k].vmode;
    if (bsc_full) bsc_full[k] = sm[k].bsc_full;
  }
  *nhits = (int64_t)out.size();
  for (int64_t k = 0; k < std::min<int64_t>(cap, out.size()); k++) hits[k] = out[k];
  return (int64_t)out.size() > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_cyk_seq_filter(ixg_ctx *c, const ixg_cm *cm, const uint8_t *dsq, int64_t n, const int32_t *dmin1,
                       const int32_t *dmax1, float cutoff, int do_null3, int32_t *es, int32_t *ee, int64_t cap,
                       int64_t *nenv) {
  if (!c || !nenv || (cap > 0 && (!es || !ee)) || n < 1) return IXG_EINVAL;
  RCHK(check_window(cm, dsq, n, 1, (int32_t)n));
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_CM_SEQ, (size_t)n + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)n + 1, hipMemcpyHostToDevice, c->stream));
  std::vector<ixg_hit> hits;
  RCHK(qdb_run(c, cm, d, {QWin{1, (int32_t)n, 0, dsq}}, dmin1, dmax1, 0, cutoff, do_null3, hits));
  const int W = cm->h.W;
  std::vector<std::pair<int32_t, int32_t>> pad;  // (end, start) for the sort
  for (const ixg_hit &h : hits)
    pad.push_back({std::min<int32_t>((int32_t)n, h.i + (W - 1)), std::max<int32_t>(1, h.j - (W - 1))});
  std::sort(pad.begin(), pad.end());
  int64_t no = 0;
  for (size_t k = 0; k < pad.size(); k++) {
    const int32_t iw = pad[k].second;
    int32_t jw = pad[k].first;
    while (k + 1 < pad.size() && pad[k + 1].second <= jw) {
      k++;
      jw = std::max(jw, pad[k].first);
    }
    if (no < cap) {
      es[no] = iw;
      ee[no] = jw;
    }
    no++;
  }
  *nenv = no;
  return no > cap ? IXG_ERANGE : IXG_OK;
}

}  // extern "C"
