// F4/F5 envelope definition of the STD pass (cm_search.rs:2449-2600
// trunc_env_def with the glocal profile gm): per window the glocal Forward
// (F4) and composition-bias (F4b) gates, the glocal Backward, posterior
// domain decoding, region finding, the stochastic-trace clustering of
// multidomain regions (p7_generic.rs:1352-1760) and the Forward rescoring of
// each domain, then the envelope (F5, F5b) gates.
//
// Every Forward and Backward runs on the GPU (gfb_kernels.hip), batched over
// all windows of a call: the window Forwards, the Backwards of the F4/F4b
// survivors, the region Forwards of multidomain regions (with their matrix,
// read back for the traces) and the domain rescoring Forwards are four
// launches. The host keeps what is sequential and cheap: gates, decoding
// from the special states, region finding, the traces (a seeded RNG whose
// draw order defines the result) and the clustering.
#include <algorithm>
#include <cmath>
#include <cstring>
#include "ctx.h"

using namespace ixg;

namespace ixg {

namespace {

// Uploads the items, runs the glocal Forward (or, for every item with
// backward set, else for the items flagged bit 4, the Backward) and returns
// the scores, the special rows of every item ([(len+1)][E N J B C],
// concatenated) and, for Forward items with flag bit 1, their matrices.
int run_gfb(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, std::vector<GfbItem> &items, bool backward,
            std::vector<float> &sc, std::vector<float> &xall, std::vector<float> *mats) {
  bool mixed = false;
  for (auto &it : items) {
    if (backward) it.flags |= 16;
    mixed |= ((it.flags ^ items[0].flags) & 16) != 0;
  }
  const int M = p7_gm_host(p).M;
  int64_t xtot = 0, mtot = 0;
  for (auto &it : items) {
    it.xoff = xtot;
    xtot += 5 * ((int64_t)it.len + 1);
    it.moff = mtot;
    if (it.flags & 2) mtot += 3 * ((int64_t)it.len + 1) * (M + 1);
  }
  const int64_t n = (int64_t)items.size();
  sc.assign(n, 0.f);
  xall.assign(xtot, 0.f);
  if (mats) mats->assign(mtot, 0.f);
  if (n == 0) return IXG_OK;
  GfbItem *d_items;
  float *d_x, *d_m = nullptr, *d_sc;
  RCHK(scratch(c, S_GFB_ITEMS, items.size() * sizeof(GfbItem), (void **)&d_items));
  RCHK(scratch(c, S_GFB_X, (size_t)xtot * 4, (void **)&d_x));
  if (mtot) RCHK(scratch(c, S_GFB_M, (size_t)mtot * 4, (void **)&d_m));
  RCHK(scratch(c, S_GFB_SC, (size_t)n * 4, (void **)&d_sc));
  HIPCHK(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(GfbItem), hipMemcpyHostToDevice, c->stream));
  HIPCHK(ktime_begin(c));
  HIPCHK(launch_gfb(p7_gm_dev(p), d_seq, d_items, n, d_x, d_m, d_sc, c->stream));
  HIPCHK(ktime_end(c, mixed ? IXG_K_GFB : ((items[0].flags & 16) ? IXG_K_GBWD : IXG_K_GFWD)));
  HIPCHK(hipMemcpyAsync(sc.data(), d_sc, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(xall.data(), d_x, (size_t)xtot * 4, hipMemcpyDeviceToHost, c->stream));
  if (mats && mtot) HIPCHK(hipMemcpyAsync(mats->data(), d_m, (size_t)mtot * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return IXG_OK;
}

// prof: the profile index (0 glocal, 1 5', 2 3', 3 5'&3')
GfbItem make_item(int64_t off, int len, const GmHost &g, int flags, int prof = 0) {
  GfbItem it{};
  it.off = off;
  it.len = len;
  it.flags = flags | (g.local_end ? 1 : 0) | (prof << 2);
  g.xsc8(it.xsc);
  return it;
}

// ---- stochastic traces of a multidomain region (p7_generic.rs:1352-1760) ----

// Easel's fast generator (Knuth's LCG after a Jenkins mix of the seed).
struct FastRng {
  uint32_t x;
  explicit FastRng(uint32_t seed) {
    uint32_t a = seed, b = 87654321u, c = 12345678u;
    a -= b; a -= c; a ^= c >> 13;
    b -= c; b -= a; b ^= a << 8;
    c -= a; c -= b; c ^= b >> 13;
    a -= b; a -= c; a ^= c >> 12;
    b -= c; b -= a; b ^= a << 16;
    c -= a; c -= b; c ^= b >> 5;
    a -= b; a -= c; a ^= c >> 3;
    b -= c; b -= a; b ^= a << 10;
    c -= a; c -= b; c ^= b >> 15;
    x = c ? c : 42u;
  }
  double uniform() {
    x = x * 69069u + 1u;
    return (double)x / 4294967296.0;
  }
  // one draw against the (unnormalised) weights, walking the cumulative sum
  int choose(const float *p, int n) {
    const double roll = uniform();
    double norm = 0.0, acc = 0.0;
    for (int i = 0; i < n; i++) norm += (double)p[i];
    for (int i = 0; i < n; i++) {
      acc += (double)p[i];
      if (roll < acc / norm) return i;
    }
    return n - 1;
  }
};

// log-space weights -> probabilities (esl_vec_FLogNorm: log-sum-exp with a
// 50-nat cut, exp, Kahan-compensated normalisation)
void log_normalise(float *v, int n) {
  float mx = -INFINITY;
  for (int i = 0; i < n; i++) mx = std::fmax(mx, v[i]);
  float denom;
  if (mx == INFINITY) {
    denom = INFINITY;
  } else {
    float s = 0.f;
    for (int i = 0; i < n; i++)
      if (v[i] > mx - 50.0f) s += expf(v[i] - mx);
    denom = logf(s) + mx;
  }
  for (int i = 0; i < n; i++) v[i] += -denom;
  for (int i = 0; i < n; i++) v[i] = expf(v[i]);
  float sum = 0.f, comp = 0.f;
  for (int i = 0; i < n; i++) {
    const float y = v[i] - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  if (sum != 0.f)
    for (int i = 0; i < n; i++) v[i] /= sum;
  else
    for (int i = 0; i < n; i++) v[i] = 1.0f / (float)n;
}

struct Seg {
  int idx, i, j, k, m;
};

// One sampled trace of the region's Forward matrix (states walked from T back
// to S); appends each domain's (first/last matched residue, first/last model
// position) in forward order.
void sample_trace(FastRng &r, int lr, const GmHost &g, const float *dp, const float *xm, int t, std::vector<Seg> &out) {
  const int M = g.M;
  const size_t W = (size_t)(M + 1);
  auto Mx = [&](int i, int k) { return dp[(size_t)i * 3 * W + k]; };
  auto Ix = [&](int i, int k) { return dp[(size_t)i * 3 * W + W + k]; };
  auto Dx = [&](int i, int k) { return dp[(size_t)i * 3 * W + 2 * W + k]; };
  auto X = [&](int i, int s) { return xm[(size_t)i * 5 + s]; };
  auto T = [&](int s, int k) { return g.tsc[(size_t)k * 8 + s]; };
  enum { sT, sC, sE, sM, sD, sI, sN, sB, sJ, sS };
  enum { E = 0, N = 1, J = 2, B = 3, C = 4 };
  struct Step {
    int st, k, i;
  };
  std::vector<Step> tr;
  int k = 0, i = lr;
  tr.push_back({sT, k, i});
  tr.push_back({sC, k, i});
  int prv = sC;
  float w[4];
  while (prv != sS) {
    int cur;
    switch (tr.back().st) {
      case sC:
        w[0] = X(i - 1, C) + g.xsc[GmHost::XC][0];
        w[1] = X(i, E) + g.xsc[GmHost::XE][1];
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sC : sE;
        break;
      case sE:
        k = M;
        w[0] = Mx(i, M);
        w[1] = Dx(i, M);
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sM : sD;
        break;
      case sM: {
        w[0] = X(i - 1, B) + T(GmHost::BM, k - 1);
        w[1] = Mx(i - 1, k - 1) + T(GmHost::MM, k - 1);
        w[2] = Ix(i - 1, k - 1) + T(GmHost::IM, k - 1);
        w[3] = Dx(i - 1, k - 1) + T(GmHost::DM, k - 1);
        log_normalise(w, 4);
        static const int to[4] = {sB, sM, sI, sD};
        cur = to[r.choose(w, 4)];
        k--;
        i--;
        break;
      }
      case sD:
        w[0] = Mx(i, k - 1) + T(GmHost::MD, k - 1);
        w[1] = Dx(i, k - 1) + T(GmHost::DD, k - 1);
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sM : sD;
        k--;
        break;
      case sI:
        w[0] = Mx(i - 1, k) + T(GmHost::MI, k);
        w[1] = Ix(i - 1, k) + T(GmHost::II, k);
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sM : sI;
        i--;
        break;
      case sN:
        cur = i == 0 ? sS : sN;
        break;
      case sB:
        w[0] = X(i, N) + g.xsc[GmHost::XN][1];
        w[1] = X(i, J) + g.xsc[GmHost::XJ][1];
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sN : sJ;
        break;
      case sJ:
        w[0] = X(i - 1, J) + g.xsc[GmHost::XJ][0];
        w[1] = X(i, E) + g.xsc[GmHost::XE][0];
        log_normalise(w, 2);
        cur = r.choose(w, 2) == 0 ? sJ : sE;
        break;
      default:
        return;
    }
    tr.push_back({cur, k, i});
    if ((cur == sN || cur == sJ || cur == sC) && cur == prv) i--;
    prv = cur;
  }
  bool open = false;
  Seg s{};
  for (auto it = tr.rbegin(); it != tr.rend(); ++it) {
    if (it->st == sB) {
      if (open) out.push_back(s);
      open = true;
      s = Seg{t, 0, 0, 0, 0};
    } else if (it->st == sM && open) {
      if (s.i == 0) s.i = it->i;
      if (s.k == 0) s.k = it->k;
      s.j = it->i;
      s.m = it->k;
    }
  }
  if (open) out.push_back(s);
}

// Sampled domains that overlap by 80% in sequence and model and lie on a
// nearby diagonal are linked (p7_spensemble.c link rule as restated at
// p7_generic.rs:1594-1616).
bool linked(const Seg &a, const Seg &b) {
  int ov = std::min(a.j, b.j) - std::max(a.i, b.i) + 1;
  int ln = std::min(a.j - a.i + 1, b.j - b.i + 1);
  if ((float)ov / (float)ln < 0.8f) return false;
  ov = std::min(a.m, b.m) - std::max(a.k, b.k);
  ln = std::min(a.m - a.k + 1, b.m - b.k + 1);
  if ((float)ov / (float)ln < 0.8f) return false;
  return std::abs((a.i - a.k) - (b.i - b.k)) <= 4 || std::abs((a.j - a.m) - (b.j - b.m)) <= 4;
}

// Single-linkage clusters in Easel's visiting order (p7_generic.rs:1619-1648).
int single_linkage(const std::vector<Seg> &s, std::vector<int> &cl) {
  const int n = (int)s.size();
  std::vector<int> a(n), b;
  for (int v = 0; v < n; v++) a[v] = n - v - 1;
  int na = n, nc = 0;
  cl.assign(n, 0);
  while (na > 0) {
    b.push_back(a[--na]);
    while (!b.empty()) {
      const int v = b.back();
      b.pop_back();
      cl[v] = nc;
      for (int q = na - 1; q >= 0; q--)
        if (linked(s[v], s[a[q]])) {
          b.push_back(a[q]);
          a[q] = a[--na];
        }
    }
    nc++;
  }
  return nc;
}

// Consensus endpoints of the significant clusters (posterior >= 0.25), in
// order of start, with the 80%-overlap dominance rule of
// glocal_region_trace_ensemble (p7_generic.rs:1660-1752).
std::vector<std::pair<int, int>> cluster_domains(const std::vector<Seg> &s, int nsamples) {
  std::vector<std::pair<int, int>> res;
  if (s.empty()) return res;
  std::vector<int> cl;
  const int nc = single_linkage(s, cl);
  struct Sig {
    int i, j;
    float p;
  };
  std::vector<Sig> sig;
  for (int c = 0; c < nc; c++) {
    int ninc = 0, last = -1;
    int lo[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX}, hi[4] = {INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN};
    for (size_t h = 0; h < s.size(); h++) {
      if (cl[h] != c) continue;
      if (s[h].idx != last) ninc++;
      last = s[h].idx;
      const int v[4] = {s[h].i, s[h].j, s[h].k, s[h].m};
      for (int q = 0; q < 4; q++) lo[q] = std::min(lo[q], v[q]), hi[q] = std::max(hi[q], v[q]);
    }
    if ((float)ninc / (float)nsamples < 0.25f) continue;
    const int thr = (int)ceilf((float)ninc * 0.02f);
    int best[4];
    for (int q = 0; q < 4; q++) {  // i and k leftmost, j and m rightmost
      std::vector<int> cnt(hi[q] - lo[q] + 1, 0);
      for (size_t h = 0; h < s.size(); h++)
        if (cl[h] == c) {
          const int v[4] = {s[h].i, s[h].j, s[h].k, s[h].m};
          cnt[v[q] - lo[q]]++;
        }
      const int argmax = (int)(std::max_element(cnt.begin(), cnt.end()) - cnt.begin());
      if (q == 0 || q == 2) {
        int x = lo[q];
        while (x <= hi[q] && cnt[x - lo[q]] < thr) x++;
        best[q] = x > hi[q] ? lo[q] + argmax : x;
      } else {
        int x = hi[q];
        while (x >= lo[q] && cnt[x - lo[q]] < thr) x--;
        best[q] = x < lo[q] ? lo[q] + argmax : x;
      }
    }
    if (best[0] > best[1] || best[2] > best[3]) continue;
    sig.push_back({best[0], best[1], (float)ninc / (float)nsamples});
  }
  std::stable_sort(sig.begin(), sig.end(), [](const Sig &a, const Sig &b) { return a.i < b.i; });
  std::vector<char> dom(sig.size(), 0);
  for (size_t d = 0; d < sig.size(); d++)
    for (size_t d2 = d + 1; d2 < sig.size(); d2++) {
      const int ov = std::min(sig[d].j, sig[d2].j) - std::max(sig[d].i, sig[d2].i) + 1;
      if (ov == 0) break;
      const int ln = std::min(sig[d].j - sig[d].i + 1, sig[d2].j - sig[d2].i + 1);
      if ((float)ov / (float)ln >= 0.8f) {
        if (sig[d].p > sig[d2].p) dom[d2] = 1;
        else dom[d] = 1;
      }
    }
  for (size_t d = 0; d < sig.size(); d++)
    if (!dom[d]) res.push_back({sig[d].i, sig[d].j});
  return res;
}

}  // namespace

namespace {

// A pass's profile as trunc_env_def configures it for one window
// (cm_search.rs:2475-2520): the glocal gm with the standard length model in
// the STD pass; the truncated passes' Rgm / Lgm with their 5' / 3' length
// models and Tgm unchanged. A truncated pass's profile is one mutable copy
// per (chunk, pass): once a window has resolved a multidomain region, the
// trace ensemble has left it unihit with the standard length model at that
// window's length (p7_generic.rs:1733-1760), and later windows start from
// there (dirty_len, 0 while clean).
struct PassProfile {
  GmHost g;
  int prof;
};
PassProfile window_profile(const ixg_p7 *p, int pass, int wlen, int dirty_len) {
  PassProfile r;
  if (pass == IXG_PASS_STD_ANY) {
    r.g = p7_gm_host(p);
    r.g.set_length(wlen);
    r.prof = 0;
    return r;
  }
  const int k = pass == IXG_PASS_5P_ONLY_FORCE ? 0 : (pass == IXG_PASS_3P_ONLY_FORCE ? 1 : 2);
  r.g = p7_gm_trunc(p, k);
  r.prof = k + 1;
  if (dirty_len) {
    r.g.set_unihit();
    r.g.set_length(dirty_len);
  }
  if (pass == IXG_PASS_5P_ONLY_FORCE) r.g.set_length_5p(wlen);
  else if (pass == IXG_PASS_3P_ONLY_FORCE) r.g.set_length_3p(wlen);
  return r;
}

// One wave of the envelope definition: every window w of `todo` with the
// profile its group's dirty length gives. Per window: envelopes, accounting
// and, for truncated passes, whether it resolved a multidomain region (and
// at which length).
int env_def_wave(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                 const std::vector<EnvWin> &wins, const std::vector<size_t> &todo, const std::vector<int> &dirty,
                 const ixg_envopts &o, std::vector<std::vector<std::pair<int32_t, int32_t>>> &envs,
                 std::vector<ixg_envacct> &wacct, std::vector<int> &dirtied) {
  const double LN2v = 0.693147180559945309417232121458176568;
  const size_t nw = todo.size();
  auto W = [&](size_t t) -> const EnvWin & { return wins[todo[t]]; };
  auto wlen = [&](size_t t) { return (int)(W(t).end - W(t).start + 1); };
  auto woff = [&](size_t t) { return tasks[W(t).task].base + W(t).start - 1; };
  auto is_std = [&](size_t t) { return W(t).pass == IXG_PASS_STD_ANY; };
  // per-pass statistics and score correction (cm_search.rs:2475-2489)
  auto corr = [&](size_t t) -> float {
    if (W(t).pass == IXG_PASS_5P_ONLY_FORCE) return 0.0f;
    if (W(t).pass == IXG_PASS_3P_ONLY_FORCE) return logf(1.0f / (float)o.clen);
    return NAN;  // none
  };
  auto mu = [&](size_t t) { return is_std(t) ? o.gfmu : o.lftau; };
  auto lam = [&](size_t t) { return is_std(t) ? o.gflambda : o.lflambda; };
  std::vector<PassProfile> pp(nw);
  std::vector<GfbItem> items(nw);
  for (size_t t = 0; t < nw; t++) {
    pp[t] = window_profile(p, W(t).pass, wlen(t), dirty[todo[t]]);
    items[t] = make_item(woff(t), wlen(t), pp[t].g, 0, pp[t].prof);
  }
  // F4: the Forward of every window, and in the same launch its Backward
  // (needed only past the F4 / F4b gates; computing it for every window
  // takes it off the critical path of the stage)
  std::vector<GfbItem> both(items);
  for (size_t t = 0; t < nw; t++) {
    both.push_back(items[t]);
    both.back().flags |= 16;
  }
  std::vector<float> fbsc, fbx;
  RCHK(run_gfb(c, p, d_seq, both, false, fbsc, fbx, nullptr));
  std::vector<float> &fsc = fbsc, &fx = fbx;
  for (size_t t = 0; t < nw; t++) items[t].xoff = both[t].xoff;
  std::vector<double> nullsc(nw);
  std::vector<size_t> s4;
  for (size_t t = 0; t < nw; t++) {
    nullsc[t] = bg_null_one(wlen(t));
    const float cr = corr(t);
    const double s = std::isnan(cr) ? ((double)fsc[t] - nullsc[t]) / LN2v : ((double)(fsc[t] + cr) - nullsc[t]) / LN2v;
    if (exp_surv(s, mu(t), lam(t)) > o.f4) continue;
    wacct[todo[t]].n_past_gfwd++;
    wacct[todo[t]].pos_past_gfwd += (uint64_t)wlen(t);
    s4.push_back(t);
  }
  // F4b (and F5b): the window's composition-bias filter score
  std::vector<float> bias(nw, 0.f);
  if ((o.do_gfwdbias || o.do_edefbias) && !s4.empty()) {
    std::vector<WinRef> sub;
    for (size_t t : s4) sub.push_back({W(t).task, W(t).start, W(t).end});
    std::vector<float> sc;
    std::vector<uint8_t> ov;
    RCHK(run_scores(c, p, d_seq, tasks, sub, 1, sc, ov));
    for (size_t q = 0; q < s4.size(); q++) bias[s4[q]] = sc[q];
  }
  std::vector<size_t> s4b;
  for (size_t t : s4) {
    if (o.do_gfwdbias) {
      const float cr = corr(t);
      const double s =
          std::isnan(cr) ? ((double)fsc[t] - (double)bias[t]) / LN2v : ((double)(fsc[t] + cr) - nullsc[t]) / LN2v;
      if (exp_surv(s, mu(t), lam(t)) > o.f4b) continue;
      wacct[todo[t]].n_past_gfwdbias++;
      wacct[todo[t]].pos_past_gfwdbias += (uint64_t)wlen(t);
    }
    s4b.push_back(t);
  }
  if (s4b.empty()) return IXG_OK;
  // the survivors' Backward rows (same configuration), from the launch above
  std::vector<GfbItem> bitems;
  for (size_t t : s4b) bitems.push_back(both[nw + t]);
  const std::vector<float> &bx = fbx;
  // decoding and region finding (p7_generic.rs:1765-1836); each region is a
  // single domain or goes through the trace ensemble
  struct Region {
    size_t t;
    int i, j;
    bool multi;
  };
  std::vector<Region> regions;
  for (size_t q = 0; q < s4b.size(); q++) {
    const size_t t = s4b[q];
    const int L = wlen(t);
    std::vector<float> btot(L + 1), etot(L + 1), mocc(L + 1);
    domain_decoding(pp[t].g, L, fx.data() + items[t].xoff, bx.data() + bitems[q].xoff, btot.data(), etot.data(),
                    mocc.data());
    int i = -1;
    bool trig = false;
    for (int j = 1; j <= L; j++) {
      if (!trig) {
        if (mocc[j] - (btot[j] - btot[j - 1]) < o.rt2) i = j;
        else if (i == -1) i = j;
        if (mocc[j] >= o.rt1) trig = true;
      } else if (mocc[j] - (etot[j] - etot[j - 1]) < o.rt2) {
        float mx = -1.0f;  // expected number of domains the region holds
        for (int z = i; z <= j; z++) {
          const float e = etot[z] - etot[i - 1], b = btot[j] - btot[z - 1];
          mx = std::max(mx, e < b ? e : b);
        }
        regions.push_back({t, i, j, mx >= o.rt3});
        i = -1;
        trig = false;
      }
    }
  }
  // multidomain regions: the region Forward (multihit at the profile's
  // length, with its matrix), then the traces on the host
  auto multihit_cfg = [&](size_t t) {
    GmHost g = pp[t].g;
    g.set_multihit();
    g.set_length(pp[t].g.len);
    return g;
  };
  std::vector<GfbItem> ritems;
  std::vector<size_t> ridx;
  for (size_t r = 0; r < regions.size(); r++)
    if (regions[r].multi) {
      const size_t t = regions[r].t;
      ritems.push_back(make_item(woff(t) + regions[r].i - 1, regions[r].j - regions[r].i + 1, multihit_cfg(t), 2,
                                 pp[t].prof));
      ridx.push_back(r);
    }
  std::vector<std::vector<std::pair<int, int>>> rdoms(regions.size());
  if (!ritems.empty()) {
    std::vector<float> rsc, rx, rm;
    RCHK(run_gfb(c, p, d_seq, ritems, false, rsc, rx, &rm));
    for (size_t q = 0; q < ritems.size(); q++) {
      const Region &R = regions[ridx[q]];
      const GmHost g = multihit_cfg(R.t);
      FastRng rng(181u);  // re-seeded per region
      std::vector<Seg> segs;
      for (int t = 0; t < o.ns; t++) {
        const size_t n0 = segs.size();
        sample_trace(rng, ritems[q].len, g, rm.data() + ritems[q].moff, rx.data() + ritems[q].xoff, t, segs);
        for (size_t h = n0; h < segs.size(); h++) segs[h].i += R.i - 1, segs[h].j += R.i - 1;
      }
      rdoms[ridx[q]] = cluster_domains(segs, o.ns);
    }
  }
  for (size_t r = 0; r < regions.size(); r++)
    if (!regions[r].multi) rdoms[r].push_back({regions[r].i, regions[r].j});
  // rescoring: the Forward of each domain with the profile as it stands
  // then: unihit at the profile's length after the STD pass's reconfiguration
  // or after a trace ensemble; a truncated pass's own (unihit) profile before
  std::vector<GfbItem> ditems;
  struct Dom {
    size_t t;
    int i, j;
  };
  std::vector<Dom> doms;
  size_t cur = SIZE_MAX;
  bool after_multi = false;
  for (size_t r = 0; r < regions.size(); r++) {
    const size_t t = regions[r].t;
    if (t != cur) cur = t, after_multi = false;
    if (regions[r].multi) {
      after_multi = true;
      if (!is_std(t)) dirtied[todo[t]] = pp[t].g.len;
    }
    GmHost g = pp[t].g;
    if (is_std(t) || after_multi) {
      g.set_unihit();
      g.set_length(pp[t].g.len);
    }
    for (auto &d : rdoms[r]) {
      ditems.push_back(make_item(woff(t) + d.first - 1, d.second - d.first + 1, g, 0, pp[t].prof));
      doms.push_back({t, d.first, d.second});
    }
  }
  std::vector<float> dsc, dx;
  RCHK(run_gfb(c, p, d_seq, ditems, false, dsc, dx, nullptr));
  // F5 / F5b in domain order
  for (size_t q = 0; q < doms.size(); q++) {
    const size_t t = doms[q].t;
    const int L = wlen(t), env_len = doms[q].j - doms[q].i + 1;
    const float env_sc = dsc[q] + ((float)L - (float)env_len) * logf((float)L / ((float)L + 3.0f));
    const float cr = corr(t);
    double pv = ((double)env_sc - (nullsc[t] + 0.0)) / LN2v;
    if (!std::isnan(cr)) pv += (double)cr / LN2v;
    if (exp_surv(pv, mu(t), lam(t)) > o.f5) continue;
    ixg_envacct &a = wacct[todo[t]];
    a.n_past_edef++;
    a.pos_past_edef += (uint64_t)env_len;
    if (o.do_edefbias) {
      double sb = ((double)env_sc - (double)bias[t]) / LN2v;
      if (!std::isnan(cr)) sb += (double)cr / LN2v;
      if (exp_surv(sb, mu(t), lam(t)) > o.f5b) continue;
      a.n_past_edefbias++;
      a.pos_past_edefbias += (uint64_t)env_len;
    }
    envs[todo[t]].push_back(
        {(int32_t)(doms[q].i + W(t).start - 1), (int32_t)(doms[q].j + W(t).start - 1)});
  }
  return IXG_OK;
}

}  // namespace

// trunc_env_def (cm_search.rs:2449) over the windows of several (chunk,
// pass) groups, each group's windows consecutive and in the reference's
// order, windows in the coordinates of their task's sequence view. The
// enforce rules of the 5' / 3' passes (the window must start at the view's
// first / end at its last residue, cm_search.rs:2497-2502) are applied here.
// envs[w]: the window's surviving envelopes; wacct[w]: its accounting.
// Every window runs in one wave; a group whose truncated-pass profile a
// multidomain region changed runs its later windows again in the next wave
// with the changed profile.
int env_def_passes(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                   const std::vector<EnvWin> &wins, const ixg_envopts &o,
                   std::vector<std::vector<std::pair<int32_t, int32_t>>> &envs, std::vector<ixg_envacct> &wacct) {
  const size_t nw = wins.size();
  envs.assign(nw, {});
  wacct.assign(nw, ixg_envacct{});
  std::vector<size_t> todo;
  for (size_t w = 0; w < nw; w++) {
    const EnvWin &x = wins[w];
    const bool first = x.pass == IXG_PASS_5P_ONLY_FORCE || x.pass == IXG_PASS_5P_AND_3P_FORCE;
    const bool final = x.pass == IXG_PASS_3P_ONLY_FORCE || x.pass == IXG_PASS_5P_AND_3P_FORCE;
    if (first && x.start != 1) continue;
    if (final && x.end != tasks[x.task].L) continue;
    todo.push_back(w);
  }
  std::vector<int> dirty(nw, 0), dirtied(nw, 0);
  auto same_group = [&](size_t a, size_t b) { return wins[a].task == wins[b].task && wins[a].pass == wins[b].pass; };
  while (!todo.empty()) {
    for (size_t w : todo) envs[w].clear(), wacct[w] = ixg_envacct{}, dirtied[w] = 0;
    RCHK(env_def_wave(c, p, d_seq, tasks, wins, todo, dirty, o, envs, wacct, dirtied));
    // the windows after a group's first dirtying window go again
    std::vector<size_t> next;
    for (size_t q = 0; q < todo.size(); q++) {
      const size_t w = todo[q];
      if (!dirtied[w]) continue;
      size_t r = q + 1;
      for (; r < todo.size() && same_group(todo[r], w); r++) {
        dirty[todo[r]] = dirtied[w];
        next.push_back(todo[r]);
      }
      q = r - 1;
    }
    todo.swap(next);
  }
  return IXG_OK;
}

// The STD-pass envelope definition over windows of several tasks (windows
// in chunk coordinates of their task). envs[w]: the window's surviving
// envelopes (chunk coordinates), in the reference's order; acct per task.
int env_def_batch(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                  const std::vector<WinRef> &wins, const ixg_envopts &o,
                  std::vector<std::vector<std::pair<int32_t, int32_t>>> &envs, std::vector<ixg_envacct> &acct) {
  std::vector<EnvWin> ew(wins.size());
  for (size_t w = 0; w < wins.size(); w++) ew[w] = {wins[w].task, wins[w].start, wins[w].end, IXG_PASS_STD_ANY};
  std::vector<ixg_envacct> wa;
  RCHK(env_def_passes(c, p, d_seq, tasks, ew, o, envs, wa));
  acct.assign(tasks.size(), ixg_envacct{});
  for (size_t w = 0; w < wins.size(); w++) {
    const uint64_t *s = (const uint64_t *)&wa[w];
    uint64_t *d = (uint64_t *)&acct[wins[w].task];
    for (size_t q = 0; q < sizeof(ixg_envacct) / 8; q++) d[q] += s[q];
  }
  return IXG_OK;
}

}  // namespace ixg

extern "C" {

int ixg_gfb_scores(ixg_ctx *c, const ixg_p7 *p, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                   const int64_t *len, const int32_t *cfg_len, int64_t n, int mode, float *sc, float *xmx,
                   float *mat) {
  if (!c || !p || !seq || (n && (!off || !len || !sc)) || n < 0 || ((mode & IXG_GFB_MATRIX) && (mode & IXG_GFB_BACKWARD))) {
    g_err = "ixg_gfb_scores: bad arguments";
    return IXG_EINVAL;
  }
  if (!p7_has_device(p) || p7_gm_host(p).M < 2) {
    g_err = "ixg_gfb_scores: host-only profile or M < 2";
    return IXG_EINVAL;
  }
  for (int64_t i = 0; i < n; i++) {
    if (off[i] < 0 || len[i] < 1 || off[i] + len[i] >= seq_len || len[i] > INT32_MAX ||
        (cfg_len && cfg_len[i] < 1)) {
      g_err = "ixg_gfb_scores: item outside the sequence buffer";
      return IXG_EINVAL;
    }
    for (int64_t k = 1; k <= len[i]; k++)
      if (seq[off[i] + k] >= KP) {
        g_err = "digitized residue code >= 18 in an item";
        return IXG_EINVAL;
      }
  }
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_SEQ, (size_t)seq_len + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, seq, (size_t)seq_len, hipMemcpyHostToDevice, c->stream));
  std::vector<GfbItem> items(n);
  for (int64_t i = 0; i < n; i++) {
    GmHost g = p7_gm_host(p);
    if (mode & IXG_GFB_UNIHIT) g.set_unihit();
    g.set_length(cfg_len ? cfg_len[i] : (int)len[i]);
    items[i] = make_item(off[i], (int)len[i], g, (mode & IXG_GFB_MATRIX) ? 2 : 0);
  }
  std::vector<float> s, x, m;
  RCHK(run_gfb(c, p, d, items, (mode & IXG_GFB_BACKWARD) != 0, s, x, (mode & IXG_GFB_MATRIX) ? &m : nullptr));
  memcpy(sc, s.data(), (size_t)n * 4);
  if (xmx) memcpy(xmx, x.data(), x.size() * 4);
  if (mat && (mode & IXG_GFB_MATRIX)) memcpy(mat, m.data(), m.size() * 4);
  return IXG_OK;
}

int ixg_env_def_pass(ixg_ctx *c, const ixg_p7 *p, const uint8_t *dsq, int64_t L, const ixg_window *wins,
                     int64_t nwin, int pass_idx, const ixg_envopts *o, int32_t *es, int32_t *ee, int64_t cap,
                     int64_t *nenv, ixg_envacct *acct) {
  if (!c || !p || !dsq || (nwin && !wins) || nwin < 0 || !o || !nenv || (cap && (!es || !ee)) ||
      pass_idx < IXG_PASS_STD_ANY || pass_idx > IXG_PASS_5P_AND_3P_FORCE ||
      (pass_idx == IXG_PASS_3P_ONLY_FORCE && o->clen < 1)) {
    g_err = "ixg_env_def: bad arguments";
    return IXG_EINVAL;
  }
  if (!p7_has_device(p) || p7_gm_host(p).M < 2) {
    g_err = "ixg_env_def: host-only profile or M < 2";
    return IXG_EINVAL;
  }
  for (int64_t w = 0; w < nwin; w++)
    if (wins[w].start < 1 || wins[w].end > L || wins[w].end < wins[w].start) {
      g_err = "ixg_env_def: window outside the sequence";
      return IXG_EINVAL;
    }
  for (int64_t i = 1; i <= L; i++)
    if (dsq[i] >= KP) {
      g_err = "digitized residue code >= 18 in the sequence";
      return IXG_EINVAL;
    }
  HIPCHK(hipSetDevice(c->device));
  uint8_t *d;
  RCHK(scratch(c, S_SEQ, (size_t)L + 2, (void **)&d));
  HIPCHK(hipMemcpyAsync(d, dsq, (size_t)L + 2, hipMemcpyHostToDevice, c->stream));
  std::vector<HostTask> tasks{{0, L}};
  std::vector<EnvWin> ew(nwin);
  for (int64_t w = 0; w < nwin; w++) ew[w] = {0, wins[w].start, wins[w].end, pass_idx};
  std::vector<std::vector<std::pair<int32_t, int32_t>>> envs;
  std::vector<ixg_envacct> wa;
  RCHK(env_def_passes(c, p, d, tasks, ew, *o, envs, wa));
  int64_t n = 0;
  ixg_envacct tot{};
  for (int64_t w = 0; w < nwin; w++) {
    for (auto &e : envs[w]) {
      if (n < cap) es[n] = e.first, ee[n] = e.second;
      n++;
    }
    const uint64_t *s = (const uint64_t *)&wa[w];
    uint64_t *dd = (uint64_t *)&tot;
    for (size_t q = 0; q < sizeof(ixg_envacct) / 8; q++) dd[q] += s[q];
  }
  *nenv = n;
  if (acct) *acct = tot;
  return n > cap ? IXG_ERANGE : IXG_OK;
}

int ixg_env_def(ixg_ctx *c, const ixg_p7 *p, const uint8_t *dsq, int64_t L, const ixg_window *wins, int64_t nwin,
                const ixg_envopts *o, int32_t *es, int32_t *ee, int64_t cap, int64_t *nenv, ixg_envacct *acct) {
  return ixg_env_def_pass(c, p, dsq, L, wins, nwin, IXG_PASS_STD_ANY, o, es, ee, cap, nenv, acct);
}

}  // extern "C"
