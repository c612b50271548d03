// The STD pass of a cmsearch over one resident sequence, F1 through F7, as
// FaithfulSearcher::search runs it for every (chunk x strand) task
// (cm_search.rs:1300, :1774 window_tasks, :1813 run_window_task, :2214
// trunc_pipeline_one_strand with the STD pass, :2707 cm_std_stage, :1915
// finalize), with every task's stage batched into one device pass:
//
//   F1-F3b   ixg_f3_filter_tasks: one SSV sweep over all tasks, one
//            Viterbi/Forward/bias launch over all surviving windows
//   F4-F5    env_def_batch: glocal Forward/Backward launches over all windows,
//            domain definition on the host
//   F6       CP9 Forward/Backward/posterior bands of every envelope (one launch
//            per tau round), HMM->CM band conversion on host workers, tau
//            doubled for envelopes whose matrix exceeds the size limit, then
//            one HB CYK launch; P-value gate and CYK envelope redefinition
//   F7       the same band step at the final tau, one HB Inside launch with
//            greedy hit reporting and null3; per-envelope overlap removal
//
// then genome coordinates, E-values, global overlap removal and the final
// sort on the host. Envelope bands and scans work in the coordinates of the
// task's strand buffer on the device, so no sequence is copied per task.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include "ctx.h"

using namespace ixg;

namespace ixg {
int env_def_batch(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                  const std::vector<WinRef> &wins, const ixg_envopts &o,
                  std::vector<std::vector<std::pair<int32_t, int32_t>>> &envs, std::vector<ixg_envacct> &acct);
int cp9_posterior_bands(ixg_ctx *c, const ixg_cp9 *h, const std::vector<const uint8_t *> &dsq,
                        const std::vector<int32_t> &i0, const std::vector<int32_t> &j0, double tau, int mode,
                        std::vector<int32_t> &pn);
int cp9_nodes(const ixg_cp9 *h);
void hmm2ij(const ixg_hbmap &m, int32_t *pn, int i0, int j0, bool search, int32_t *imin, int32_t *imax,
            int32_t *jmin, int32_t *jmax);
int64_t ij2d(const ixg_hbmap &m, const int32_t *imin, const int32_t *imax, const int32_t *jmin, const int32_t *jmax,
             bool trunc, int32_t *hd_off, std::vector<int32_t> &hdmin, std::vector<int32_t> &hdmax);
float hb_mx_mb(const ixg_hbmap &m, const int32_t *jmin, const int32_t *jmax, const int32_t *hd_off,
               const int32_t *hdmin, const int32_t *hdmax, int L, int64_t *ncells);
int hbmap_M(const ixg_hbmap *m);
int hbmap_hmm_m(const ixg_hbmap *m);
}  // namespace ixg

const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L);

namespace {

// HB bands of one envelope in the scanners' layout
struct EnvBands {
  std::vector<int32_t> jmin, jmax, hd_off, hdmin, hdmax;
  ixg_bands view() const { return ixg_bands{jmin.data(), jmax.data(), hd_off.data(), hdmin.data(), hdmax.data()}; }
};

// cp9_iterate_seq2bands (cp9.rs:4075) over a batch of envelopes on the
// device target: every envelope starts at tau; those whose banded matrix is
// over the size limit go to the next round at twice the tau (capped at
// maxtau, after which the bands stand). skip[k] set when the final size is
// still over the limit (the caller then drops the envelope, as
// cm_std_stage does).
int iterate_bands(ixg_ctx *c, const ixg_cp9 *cp9, const ixg_hbmap *map, const uint8_t *dbase, int64_t L,
                  const std::vector<ixg_envelope> &env, double tau0, double maxtau, float limit,
                  std::vector<EnvBands> &out, std::vector<char> &skip) {
  const size_t n = env.size();
  const int M = hbmap_M(map), H = cp9_nodes(cp9);
  out.assign(n, {});
  skip.assign(n, 0);
  std::vector<size_t> todo(n);
  for (size_t k = 0; k < n; k++) todo[k] = k;
  std::vector<double> tau(n, tau0);
  std::vector<char> at_limit(n, 0);
  while (!todo.empty()) {
    // one CP9 launch per distinct tau among the pending envelopes
    std::vector<double> taus;
    for (size_t k : todo) taus.push_back(tau[k]);
    std::sort(taus.begin(), taus.end());
    taus.erase(std::unique(taus.begin(), taus.end()), taus.end());
    std::vector<size_t> next;
    for (double t : taus) {
      std::vector<size_t> grp;
      std::vector<const uint8_t *> dsq;
      std::vector<int32_t> i0, j0;
      for (size_t k : todo)
        if (tau[k] == t) {
          grp.push_back(k);
          dsq.push_back(dbase + (env[k].strand ? L + 2 : 0));
          i0.push_back(env[k].i0);
          j0.push_back(env[k].j0);
        }
      std::vector<int32_t> pn;
      RCHK(cp9_posterior_bands(c, cp9, dsq, i0, j0, t, IXG_CP9_FWD_SCAN | IXG_CP9_BCK_SCAN, pn));
      std::vector<float> mb(grp.size());
      par_for((int64_t)grp.size(), [&](int64_t g) {
        const size_t k = grp[g];
        EnvBands &b = out[k];
        std::vector<int32_t> imin(M), imax(M);
        b.jmin.assign(M, 0);
        b.jmax.assign(M, 0);
        b.hd_off.assign(M + 1, 0);
        hmm2ij(*map, pn.data() + g * 6 * (size_t)(H + 1), i0[g], j0[g], true, imin.data(), imax.data(),
               b.jmin.data(), b.jmax.data());
        ij2d(*map, imin.data(), imax.data(), b.jmin.data(), b.jmax.data(), false, b.hd_off.data(), b.hdmin, b.hdmax);
        mb[g] = hb_mx_mb(*map, b.jmin.data(), b.jmax.data(), b.hd_off.data(), b.hdmin.data(), b.hdmax.data(),
                         j0[g] - i0[g] + 1, nullptr);
      });
      for (size_t g = 0; g < grp.size(); g++) {
        const size_t k = grp[g];
        if (mb[g] < limit || at_limit[k]) {  // final: dropped when still over the limit
          skip[k] = mb[g] > limit;
          continue;
        }
        tau[k] *= 2.0;
        if (tau[k] >= maxtau) tau[k] = maxtau, at_limit[k] = 1;
        next.push_back(k);
      }
    }
    todo.swap(next);
  }
  return IXG_OK;
}

struct Hit {
  int32_t task;
  int64_t start, stop;
  float score, bias;
  double evalue;
  int32_t strand;
  bool removed = false;
};

}  // namespace

extern "C" int ixg_search_std(ixg_ctx *c, const ixg_p7 *p, const ixg_cm *cm, const ixg_cp9 *cp9,
                              const ixg_hbmap *map, const ixg_target *tgt, const ixg_searchopts *o,
                              ixg_searchhit *out, int64_t cap, int64_t *nout, ixg_searchacct *acct) {
  if (!c || !p || !cm || !cp9 || !map || !tgt || !o || !nout || (cap && !out)) {
    g_err = "ixg_search_std: bad arguments";
    return IXG_EINVAL;
  }
  if (cp9_nodes(cp9) != hbmap_hmm_m(map)) {
    g_err = "ixg_search_std: the CP9 is not the band map's";
    return IXG_EINVAL;
  }
  HIPCHK(hipSetDevice(c->device));
  using clk = std::chrono::steady_clock;
  const clk::time_point t_begin = clk::now();
  clk::time_point t_mark = t_begin;
  for (double &x : c->search_ms) x = 0;
  c->search_cells[0] = c->search_cells[1] = 0;
  auto stage = [&](int k) {  // every stage ends with its results on the host
    const clk::time_point t = clk::now();
    c->search_ms[k] += std::chrono::duration<double, std::milli>(t - t_mark).count();
    t_mark = t;
  };
  int64_t L = 0;
  const uint8_t *dbase = ixg_target_device_base(tgt, &L);
  ixg_searchacct A{};
  // (chunk x strand) tasks (cm_search.rs:1774)
  const int64_t nt = ixg_window_tasks(L, o->f3.maxw, o->do_top, o->do_bot, nullptr, 0);
  std::vector<ixg_task> tasks(nt);
  ixg_window_tasks(L, o->f3.maxw, o->do_top, o->do_bot, tasks.data(), nt);
  std::vector<HostTask> ht(nt);
  for (int64_t t = 0; t < nt; t++) {
    const ixg_task &k = tasks[t];
    ht[t].base = k.strand == 0 ? (k.start - 1) : (L + 2) + (L - k.end);
    ht[t].L = k.end - k.start + 1;
    A.nres += (uint64_t)(ht[t].L - k.ctx);  // STD pass: new residues (cm_search.rs:2306-2311)
  }
  // F1-F3b
  std::vector<ixg_acct> fa(nt);
  int64_t nw = 0;
  std::vector<ixg_window> wins(4096);
  int rc = ixg_f3_filter_tasks(c, p, tgt, tasks.data(), nt, &o->f3, wins.data(), (int64_t)wins.size(), &nw, fa.data());
  if (rc == IXG_ERANGE) {
    wins.resize(nw);
    rc = ixg_f3_filter_tasks(c, p, tgt, tasks.data(), nt, &o->f3, wins.data(), nw, &nw, fa.data());
  }
  RCHK(rc);
  wins.resize(nw);
  stage(IXG_SS_F1_F3B);
  for (const ixg_acct &a : fa) {
    const uint64_t *s = (const uint64_t *)&a;
    uint64_t *d = (uint64_t *)&A.f;
    for (size_t q = 0; q < sizeof(ixg_acct) / 8; q++) d[q] += s[q];
  }
  // F4-F5
  std::vector<WinRef> wl(nw);
  for (int64_t w = 0; w < nw; w++) wl[w] = {wins[w].task, wins[w].start, wins[w].end};
  std::vector<std::vector<std::pair<int32_t, int32_t>>> envs;
  std::vector<ixg_envacct> ea;
  RCHK(env_def_batch(c, p, dbase, ht, wl, o->env, envs, ea));
  stage(IXG_SS_F4_F5);
  for (const ixg_envacct &a : ea) {
    const uint64_t *s = (const uint64_t *)&a;
    uint64_t *d = (uint64_t *)&A.e;
    for (size_t q = 0; q < sizeof(ixg_envacct) / 8; q++) d[q] += s[q];
  }
  // envelopes in the coordinates of their task's strand buffer
  std::vector<ixg_envelope> env;
  std::vector<int32_t> etask;
  for (int64_t w = 0; w < nw; w++) {
    const ixg_task &k = tasks[wins[w].task];
    const int64_t off = k.strand == 0 ? k.start - 1 : L - k.end;
    for (auto &e : envs[w]) {
      ixg_envelope x{};
      x.i0 = (int32_t)(off + e.first);
      x.j0 = (int32_t)(off + e.second);
      x.strand = k.strand;
      env.push_back(x);
      etask.push_back(wins[w].task);
    }
  }
  // F6: CYK filter over HMM bands at the filter tau
  std::vector<ixg_envelope> surv;
  std::vector<int32_t> stask;
  if (!o->do_fcyk) {
    surv = env;
    stask = etask;
  } else if (!env.empty()) {
    std::vector<EnvBands> eb;
    std::vector<char> skip;
    RCHK(iterate_bands(c, cp9, map, dbase, L, env, o->fcyk_tau, o->maxtau, o->size_limit, eb, skip));
    stage(IXG_SS_F6_BANDS);
    std::vector<ixg_envelope> run;
    std::vector<size_t> ridx;
    for (size_t k = 0; k < env.size(); k++)
      if (!skip[k]) {
        ixg_envelope x = env[k];
        x.bands = eb[k].view();
        run.push_back(x);
        ridx.push_back(k);
      }
    std::vector<float> vsc(run.size());
    std::vector<int64_t> ei(run.size()), ej(run.size());
    int64_t nh = 0, cells = 0;
    if (!run.empty())  // every envelope may exceed the size limit
      RCHK(ixg_hb_scan_batch(c, cm, tgt, run.data(), (int64_t)run.size(), 0, o->cyk_env_cutoff, 0, 0.f, o->do_null3,
                             vsc.data(), ei.data(), ej.data(), nullptr, 0, &nh, &cells));
    c->search_cells[0] = cells;
    stage(IXG_SS_F6_CYK);
    for (size_t r = 0; r < run.size(); r++) {
      if (exp_surv((double)vsc[r], o->cyk_mu, o->cyk_lambda) > o->f6) continue;
      ixg_envelope x = env[ridx[r]];
      if (o->do_fcykenv && ei[r] != -1 && ej[r] != -1) x.i0 = (int32_t)ei[r], x.j0 = (int32_t)ej[r];
      A.n_past_cyk++;
      A.pos_past_cyk += (uint64_t)(x.j0 - x.i0 + 1);
      surv.push_back(x);
      stask.push_back(etask[ridx[r]]);
    }
  }
  // F7: Inside (or CYK) at the final tau, hits above pli_t
  std::vector<Hit> hits;
  if (!surv.empty()) {
    std::vector<EnvBands> eb;
    std::vector<char> skip;
    RCHK(iterate_bands(c, cp9, map, dbase, L, surv, o->final_tau, o->maxtau, o->size_limit, eb, skip));
    stage(IXG_SS_F7_BANDS);
    std::vector<ixg_envelope> run;
    std::vector<size_t> ridx;
    for (size_t k = 0; k < surv.size(); k++)
      if (!skip[k]) {
        ixg_envelope x = surv[k];
        x.bands = eb[k].view();
        run.push_back(x);
        ridx.push_back(k);
      }
    std::vector<float> vsc(run.size());
    std::vector<int64_t> ei(run.size()), ej(run.size());
    int64_t nh = 0, cells = 0;
    // hits buffer as large as the root cells, the most ReportHitsGreedily
    // can report, so the scan never has to run again for room
    int64_t rcells = 0;
    for (const ixg_envelope &x : run)
      for (int j = x.bands.jmin[0]; j <= x.bands.jmax[0]; j++) {
        const int r = x.bands.hd_off[0] + (j - x.bands.jmin[0]);
        rcells += std::max(0, x.bands.hdmax[r] - x.bands.hdmin[r] + 1);
      }
    std::vector<ixg_hit> hb(std::max<int64_t>(rcells, 1));
    if (run.empty()) hb.clear();
    else rc = ixg_hb_scan_batch(c, cm, tgt, run.data(), (int64_t)run.size(), o->do_cyk ? 0 : 1, 0.0f, 1, o->pli_t,
                           o->do_null3, vsc.data(), ei.data(), ej.data(), hb.data(), (int64_t)hb.size(), &nh, &cells);
    if (rc == IXG_ERANGE) {
      hb.resize(nh);
      rc = ixg_hb_scan_batch(c, cm, tgt, run.data(), (int64_t)run.size(), o->do_cyk ? 0 : 1, 0.0f, 1, o->pli_t,
                             o->do_null3, vsc.data(), ei.data(), ej.data(), hb.data(), nh, &nh, &cells);
    }
    RCHK(rc);
    hb.resize(nh);
    c->search_cells[1] = cells;
    stage(IXG_SS_F7_INSIDE);
    // per envelope: greedy overlap removal (cp9.rs:6544), score desc then start asc
    std::vector<std::vector<ixg_hit>> per(run.size());
    for (const ixg_hit &h : hb) per[h.env].push_back(h);
    for (size_t r = 0; r < run.size(); r++) {
      auto &v = per[r];
      std::stable_sort(v.begin(), v.end(), [](const ixg_hit &a, const ixg_hit &b) {
        if (a.score != b.score) return a.score > b.score;
        return a.i < b.i;
      });
      std::vector<char> gone(v.size(), 0);
      for (size_t a = 0; a < v.size(); a++) {
        if (gone[a]) continue;
        for (size_t b = a + 1; b < v.size(); b++)
          if (!gone[b] && !(v[b].j < v[a].i) && !(v[a].j < v[b].i)) gone[b] = 1;
      }
      const int32_t t = stask[ridx[r]];
      const ixg_task &k = tasks[t];
      const int64_t off = k.strand == 0 ? k.start - 1 : L - k.end;
      for (size_t a = 0; a < v.size(); a++) {
        if (gone[a]) continue;
        const int64_t ws = v[a].i - off, we = v[a].j - off;  // chunk-local
        Hit h;
        h.task = t;
        h.strand = k.strand;
        if (k.strand) h.start = k.end - ws + 1, h.stop = k.end - we + 1;
        else h.start = k.start + ws - 1, h.stop = k.start + we - 1;
        h.score = v[a].score;
        h.bias = v[a].bias;
        h.evalue = exp_surv((double)h.score, o->fin_mu, o->fin_lambda) * o->ez;
        hits.push_back(h);
      }
    }
  }
  // finalize (cm_search.rs:1915): global overlap removal per strand, then the
  // E-value sort and the reporting threshold
  {
    std::vector<size_t> ord(hits.size());
    for (size_t q = 0; q < ord.size(); q++) ord[q] = q;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
      const Hit &x = hits[a], &y = hits[b];
      if (x.strand != y.strand) return x.strand < y.strand;
      if (x.score != y.score) return x.score > y.score;
      if (x.start != y.start) return x.start < y.start;
      return x.stop < y.stop;
    });
    for (size_t a = 0; a < ord.size(); a++) {
      Hit &ha = hits[ord[a]];
      if (ha.removed) continue;
      for (size_t b = a + 1; b < ord.size(); b++) {
        Hit &hb = hits[ord[b]];
        if (hb.strand != ha.strand) break;
        if (hb.removed) continue;
        const bool ov = ha.strand == 0 ? (!(hb.stop < ha.start) && !(ha.stop < hb.start))
                                       : (!(hb.start < ha.stop) && !(ha.start < hb.stop));
        if (ov) hb.removed = true;
      }
    }
  }
  std::vector<Hit> fin;
  for (const Hit &h : hits)
    if (!h.removed) fin.push_back(h);
  std::sort(fin.begin(), fin.end(), [](const Hit &a, const Hit &b) {
    if (a.evalue != b.evalue) return a.evalue < b.evalue;
    if (a.score != b.score) return a.score > b.score;
    if (a.start != b.start) return a.start < b.start;
    return a.stop < b.stop;
  });
  int64_t n = 0;
  for (const Hit &h : fin) {
    if (!(h.evalue <= o->e_report)) continue;
    if (n < cap) {
      ixg_searchhit &x = out[n];
      x.start = h.start;
      x.stop = h.stop;
      x.score = h.score;
      x.bias = h.bias;
      x.evalue = h.evalue;
      x.strand = h.strand;
      x.pad = 0;
    }
    n++;
  }
  A.n_hits = (uint64_t)n;
  *nout = n;
  if (acct) *acct = A;
  stage(IXG_SS_FINALIZE);
  c->search_ms[IXG_SS_TOTAL] = std::chrono::duration<double, std::milli>(clk::now() - t_begin).count();
  return n > cap ? IXG_ERANGE : IXG_OK;
}

extern "C" int ixg_search_stats(ixg_ctx *c, double *ms, int64_t *cells) {
  if (!c || !ms || !cells) {
    g_err = "ixg_search_stats: bad arguments";
    return IXG_EINVAL;
  }
  for (int k = 0; k < IXG_SS_N; k++) ms[k] = c->search_ms[k];
  cells[0] = c->search_cells[0];
  cells[1] = c->search_cells[1];
  return IXG_OK;
}
