// A cmsearch over the resident sequences of a target, F1 through F7, as
// FaithfulSearcher::search runs it for every (chunk x strand) task
// (cm_search.rs:1300, :1774 window_tasks, :1813 run_window_task, :2214
// trunc_pipeline_one_strand, :2707 cm_std_stage, :2863 cm_trunc_stage, :1915
// finalize), with every task's stage batched into one device pass:
//
//   F1-F3b   ixg_f3_filter_tasks: one SSV sweep over all chunks and the
//            truncated passes' subsequences, one Viterbi/Forward/bias launch
//            over all surviving windows
//   F4-F5    env_def_passes: glocal Forward/Backward launches over all
//            windows of every pass (each item with its pass's profile),
//            domain definition on the host
//   F6       CP9 Forward/Backward/posterior bands of every envelope (one launch
//            per tau round), HMM->CM band conversion on host workers, tau
//            doubled for envelopes whose matrix exceeds the size limit, then
//            one HB CYK launch; P-value gate and CYK envelope redefinition
//   F7       the same band step at the final tau, one HB Inside launch with
//            greedy hit reporting and null3; per-envelope overlap removal
//
// then, per truncated pass, cp9_seq2bands_trunc, the TrCYK filter and
// FTrInside, each one launch over the pass's envelopes of every chunk; then
// genome coordinates, E-values, global overlap removal and the final sort
// on the host. Envelope bands and scans work in the coordinates of the
// task's strand buffer on the device, so no sequence is copied per task.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include "ctx.h"

using namespace ixg;

namespace ixg {
int env_def_passes(ixg_ctx *c, const ixg_p7 *p, const uint8_t *d_seq, const std::vector<HostTask> &tasks,
                   const std::vector<EnvWin> &wins, const ixg_envopts &o,
                   std::vector<std::vector<std::pair<int32_t, int32_t>>> &envs, std::vector<ixg_envacct> &wacct);
int cp9_posterior_bands(ixg_ctx *c, const ixg_cp9 *h, const std::vector<const uint8_t *> &dsq,
                        const std::vector<int32_t> &i0, const std::vector<int32_t> &j0, double tau, int mode,
                        std::vector<int32_t> &pn, std::vector<int32_t> *occ = nullptr);
int cp9_nodes(const ixg_cp9 *h);
void hmm2ij(const ixg_hbmap &m, int32_t *pn, int i0, int j0, bool search, int32_t *imin, int32_t *imax,
            int32_t *jmin, int32_t *jmax, TrBand *tb);
void tr_predict(const int32_t *pn, const int32_t *occ, int H, int i0, int j0, TrBand &tb);
void tr_candidates(const ixg_hbmap &m, int pass, TrBand &tb);
int64_t ij2d(const ixg_hbmap &m, const int32_t *imin, const int32_t *imax, const int32_t *jmin, const int32_t *jmax,
             bool trunc, int32_t *hd_off, std::vector<int32_t> &hdmin, std::vector<int32_t> &hdmax);
float hb_mx_mb(const ixg_hbmap &m, const int32_t *jmin, const int32_t *jmax, const int32_t *hd_off,
               const int32_t *hdmin, const int32_t *hdmax, int L, int64_t *ncells);
int hbmap_M(const ixg_hbmap *m);
int hbmap_hmm_m(const ixg_hbmap *m);
int f3_filter_views(ixg_ctx *c, const ixg_p7 *p, const ixg_target *tgt, const std::vector<HostTask> &tasks,
                    const ixg_f3opts &o, std::vector<ixg_window> &wins, std::vector<ixg_acct> &accts);
}  // namespace ixg

const uint8_t *ixg_target_device_base(const ixg_target *t, int64_t *L);

namespace {
// the device sequence an envelope's coordinates refer to: its sequence's
// strand buffer
const uint8_t *env_dsq(const ixg_target *tgt, const uint8_t *dbase, const ixg_envelope &e) {
  int64_t off = 0, L = 0;
  ixg_target_view(tgt, e.seq, e.strand, &off, &L, nullptr);
  return dbase + off;
}
}  // namespace

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
int iterate_bands(ixg_ctx *c, const ixg_cp9 *cp9, const ixg_hbmap *map, const ixg_target *tgt,
                  const uint8_t *dbase, const std::vector<ixg_envelope> &env, double tau0, double maxtau, float limit,
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
          dsq.push_back(env_dsq(tgt, dbase, env[k]));
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
               b.jmin.data(), b.jmax.data(), nullptr);
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

// cp9_seq2bands_trunc (cp9.rs:4130) over a batch of envelopes of one pass:
// the pass's CP9 Forward / Backward (each scanning unless the pass forces
// its first / final residue) with posterior bands and occupancy on the
// device, then on host workers the start / end prediction, the J/L/R/T
// candidacy, cp9_hmm2ijbands and ij2d_bands with do_trunc. No size limit
// applies (cm_trunc_stage calls it once per envelope and tau).
struct TrEnvBands {
  EnvBands b;
  std::vector<uint8_t> jv, lv, rv, tv;
  ixg_trvalid valid() const { return ixg_trvalid{jv.data(), lv.data(), rv.data(), tv.data()}; }
};
int trunc_bands(ixg_ctx *c, const ixg_cp9 *cp9, const ixg_hbmap *map, const ixg_target *tgt,
                const uint8_t *dbase, const std::vector<ixg_envelope> &env, int pass, double tau, std::vector<TrEnvBands> &out) {
  const size_t n = env.size();
  const int M = hbmap_M(map), H = cp9_nodes(cp9);
  out.assign(n, {});
  if (n == 0) return IXG_OK;
  const bool first = pass == IXG_PASS_5P_ONLY_FORCE || pass == IXG_PASS_5P_AND_3P_FORCE;
  const bool final = pass == IXG_PASS_3P_ONLY_FORCE || pass == IXG_PASS_5P_AND_3P_FORCE;
  const int mode = (first ? 0 : IXG_CP9_FWD_SCAN) | (final ? 0 : IXG_CP9_BCK_SCAN);
  std::vector<const uint8_t *> dsq(n);
  std::vector<int32_t> i0(n), j0(n), pn, occ;
  for (size_t k = 0; k < n; k++) {
    dsq[k] = env_dsq(tgt, dbase, env[k]);
    i0[k] = env[k].i0;
    j0[k] = env[k].j0;
  }
  RCHK(cp9_posterior_bands(c, cp9, dsq, i0, j0, tau, mode, pn, &occ));
  par_for((int64_t)n, [&](int64_t k) {
    TrEnvBands &t = out[k];
    int32_t *pk = pn.data() + k * 6 * (size_t)(H + 1);
    TrBand tb;
    tr_predict(pk, occ.data() + k * (size_t)(H + 1), H, i0[k], j0[k], tb);
    tr_candidates(*map, pass, tb);
    std::vector<int32_t> imin(M), imax(M);
    t.b.jmin.assign(M, 0);
    t.b.jmax.assign(M, 0);
    t.b.hd_off.assign(M + 1, 0);
    hmm2ij(*map, pk, i0[k], j0[k], true, imin.data(), imax.data(), t.b.jmin.data(), t.b.jmax.data(), &tb);
    ij2d(*map, imin.data(), imax.data(), t.b.jmin.data(), t.b.jmax.data(), true, t.b.hd_off.data(), t.b.hdmin,
         t.b.hdmax);
    t.jv = tb.jv, t.lv = tb.lv, t.rv = tb.rv, t.tv = tb.tv;
  });
  return IXG_OK;
}

struct Hit {
  int32_t task;
  int64_t seq;
  int64_t start, stop;
  float score, bias;
  double evalue;
  int32_t strand, pass;
  bool removed = false;
};

// Genome coordinates of a hit at chunk-local ws..we of task k (cm_search.rs:1877-1883).
void genome_coords(const ixg_task &k, int64_t ws, int64_t we, Hit &h) {
  h.strand = k.strand;
  if (k.strand) h.start = k.end - ws + 1, h.stop = k.end - we + 1;
  else h.start = k.start + ws - 1, h.stop = k.start + we - 1;
}

}  // namespace

extern "C" int ixg_search(ixg_ctx *c, const ixg_p7 *p, const ixg_cm *cm, const ixg_cp9 *cp9, const ixg_hbmap *map,
                          const ixg_cp9 *const *trcp9, const ixg_hbmap *const *trmap, const ixg_target *tgt,
                          const ixg_searchopts *o, ixg_searchhit *out, int64_t cap, int64_t *nout,
                          ixg_searchacct *acct) {
  if (!c || !p || !cm || !cp9 || !map || !tgt || !o || !nout || (cap && !out) || o->trunc_mode < 0 ||
      o->trunc_mode > 1 || (o->trunc_mode && (!trcp9 || !trmap))) {
    g_err = "ixg_search: bad arguments";
    return IXG_EINVAL;
  }
  if (cp9_nodes(cp9) != hbmap_hmm_m(map)) {
    g_err = "ixg_search: the CP9 is not the band map's";
    return IXG_EINVAL;
  }
  const bool do_trunc = o->trunc_mode == 1;
  if (do_trunc)
    for (int q = 0; q < 3; q++)
      if (!trcp9[q] || !trmap[q] || cp9_nodes(trcp9[q]) != hbmap_hmm_m(trmap[q]) || hbmap_M(trmap[q]) != hbmap_M(map)) {
        g_err = "ixg_search: a truncated pass's CP9 or band map is missing or does not match";
        return IXG_EINVAL;
      }
  HIPCHK(hipSetDevice(c->device));
  using clk = std::chrono::steady_clock;
  const clk::time_point t_begin = clk::now();
  clk::time_point t_mark = t_begin;
  for (double &x : c->search_ms) x = 0;
  for (int i = 0; i < 5; i++) c->ms[i] = 0, c->cells[i] = 0;
  c->search_cells[0] = c->search_cells[1] = 0;
  auto stage = [&](int k) {  // every stage ends with its results on the host
    const clk::time_point t = clk::now();
    c->search_ms[k] += std::chrono::duration<double, std::milli>(t - t_mark).count();
    t_mark = t;
  };
  int64_t L0 = 0;
  const uint8_t *dbase = ixg_target_device_base(tgt, &L0);
  const int64_t nseq = ixg_target_nseq(tgt);
  ixg_searchacct A{};
  const int64_t maxw = o->f3.maxw;
  // the flat (sequence x chunk x strand) task list (cm_search.rs:1774-1808),
  // chunk coordinates in their sequence; seq_len[s] its length
  std::vector<ixg_task> tasks;
  std::vector<int64_t> tseq, seq_len(nseq);
  for (int64_t s = 0; s < nseq; s++) {
    int64_t off = 0;
    ixg_target_view(tgt, s, 0, &off, &seq_len[s], nullptr);
    const int64_t k = ixg_window_tasks(seq_len[s], maxw, o->do_top, o->do_bot, nullptr, 0);
    const size_t at = tasks.size();
    tasks.resize(at + k);
    ixg_window_tasks(seq_len[s], maxw, o->do_top, o->do_bot, tasks.data() + at, k);
    tseq.resize(at + k, s);
  }
  const int64_t nt = (int64_t)tasks.size();
  // the truncated passes a task runs (cm_search.rs:2265-2280) and the
  // sequence each searches (:2380-2395): the chunk when it is at most maxW
  // long (or for 5'&3'), else its first (5') / last (3') maxW residues
  struct TrRun {
    int32_t task, pass, sub;  // sub: its task in the filter's task list
    int64_t start_offset;
  };
  std::vector<ixg_task> ftasks(tasks);  // chunk tasks, then the truncated passes' subsequences
  std::vector<int64_t> fseq(tseq);
  std::vector<TrRun> truns;
  for (int64_t t = 0; t < nt && do_trunc; t++) {
    const ixg_task &k = tasks[t];
    const int64_t wl = k.end - k.start + 1, L = seq_len[tseq[t]];
    const bool h5 = k.strand ? k.end == L : k.start == 1, h3 = k.strand ? k.start == 1 : k.end == L;
    const bool run[3] = {h5, h3, h5 && h3 && wl <= maxw};
    for (int q = 0; q < 3; q++) {
      if (!run[q]) continue;
      const int pass = IXG_PASS_5P_ONLY_FORCE + q;
      A.tr_nres[q] += (uint64_t)std::min<int64_t>(maxw, wl);
      if (wl <= maxw || pass == IXG_PASS_5P_AND_3P_FORCE) {
        truns.push_back({(int32_t)t, pass, (int32_t)t, 0});
        continue;
      }
      // a view of the chunk's first / last maxW residues on its strand
      ixg_task sk{};
      sk.strand = k.strand;
      sk.ctx = 0;
      const bool front = pass == IXG_PASS_5P_ONLY_FORCE;  // first maxW residues of the strand's chunk
      if ((k.strand == 0) == front) sk.start = k.start, sk.end = k.start + maxw - 1;
      else sk.start = k.end - maxw + 1, sk.end = k.end;
      truns.push_back({(int32_t)t, pass, (int32_t)ftasks.size(), front ? 0 : wl - maxw});
      ftasks.push_back(sk);
      fseq.push_back(tseq[t]);
    }
  }
  // view_off[t]: the offset of task t's residues in its sequence's strand
  // buffer (top strand: dsq[1] = seq[start]; bottom: dsq[k] =
  // comp(seq[end-k+1]) = rc[L-end+k]); ht[t].base adds the strand buffer's
  // offset in the target
  const int64_t nft = (int64_t)ftasks.size();
  std::vector<HostTask> ht(nft);
  std::vector<int64_t> view_off(nft);
  for (int64_t t = 0; t < nft; t++) {
    const ixg_task &k = ftasks[t];
    int64_t soff = 0, L = 0;
    ixg_target_view(tgt, fseq[t], k.strand, &soff, &L, nullptr);
    view_off[t] = k.strand == 0 ? k.start - 1 : L - k.end;
    ht[t].base = soff + view_off[t];
    ht[t].L = k.end - k.start + 1;
    if (t < nt) A.nres += (uint64_t)(ht[t].L - k.ctx);  // STD pass: new residues (cm_search.rs:2306-2311)
  }
  // F1-F3b over every chunk and truncated-pass subsequence. With few
  // subsequences, one call over all of them; with many (a database of short
  // sequences, where every sequence has both ends), the chunks first and
  // then only the subsequences of chunks with STD survivors, the ones whose
  // results are used (cm_search.rs:2333-2335): the same windows and
  // counters for less filter work.
  std::vector<ixg_acct> fa;
  std::vector<ixg_window> wins;
  int rc = IXG_OK;
  const int64_t nsub = nft - nt;
  if (nsub <= 256) {
    RCHK(f3_filter_views(c, p, tgt, ht, o->f3, wins, fa));
  } else {
    std::vector<HostTask> h1(ht.begin(), ht.begin() + nt);
    RCHK(f3_filter_views(c, p, tgt, h1, o->f3, wins, fa));
    fa.resize(nft);
    std::vector<char> has(nt, 0);
    for (const ixg_window &w : wins) has[w.task] = 1;
    std::vector<int32_t> pick;  // subsequence tasks to filter, ascending
    for (const TrRun &R : truns)
      if (R.sub >= nt && has[R.task]) pick.push_back(R.sub);
    std::sort(pick.begin(), pick.end());
    if (!pick.empty()) {
      std::vector<HostTask> h2(pick.size());
      for (size_t q = 0; q < pick.size(); q++) h2[q] = ht[pick[q]];
      std::vector<ixg_window> w2;
      std::vector<ixg_acct> a2;
      RCHK(f3_filter_views(c, p, tgt, h2, o->f3, w2, a2));
      for (size_t q = 0; q < pick.size(); q++) fa[pick[q]] = a2[q];
      for (ixg_window w : w2) {
        w.task = pick[w.task];
        wins.push_back(w);
      }
    }
  }
  stage(IXG_SS_F1_F3B);
  auto add_acct = [](const void *src, void *dst, size_t bytes) {
    const uint64_t *s = (const uint64_t *)src;
    uint64_t *d = (uint64_t *)dst;
    for (size_t q = 0; q < bytes / 8; q++) d[q] += s[q];
  };
  // windows come in task order: task t's are wins[wbeg[t] .. wbeg[t+1])
  std::vector<int64_t> nwin_of(nft, 0), wbeg(nft + 1, 0);
  for (const ixg_window &w : wins) nwin_of[w.task]++;
  for (int64_t t = 0; t < nft; t++) wbeg[t + 1] = wbeg[t] + nwin_of[t];
  for (int64_t t = 0; t < nt; t++) add_acct(&fa[t], &A.f, sizeof(ixg_acct));
  // envelope definition of the STD pass and of every truncated pass whose
  // chunk had STD survivors (cm_search.rs:2333-2335) and its own
  std::vector<EnvWin> ew;
  std::vector<int32_t> ew_run;  // -1: STD (task = the chunk), else the index into truns
  for (const ixg_window &w : wins)
    if (w.task < nt) ew.push_back({w.task, w.start, w.end, IXG_PASS_STD_ANY}), ew_run.push_back(-1);
  for (size_t r = 0; r < truns.size(); r++) {
    const TrRun &R = truns[r];
    if (nwin_of[R.task] == 0) continue;  // no STD survivors: the truncated passes are skipped
    add_acct(&fa[R.sub], &A.tr_f[R.pass - IXG_PASS_5P_ONLY_FORCE], sizeof(ixg_acct));
    for (int64_t q = wbeg[R.sub]; q < wbeg[R.sub + 1]; q++)
      ew.push_back({R.sub, wins[q].start, wins[q].end, R.pass}), ew_run.push_back((int32_t)r);
  }
  std::vector<std::vector<std::pair<int32_t, int32_t>>> envs;
  std::vector<ixg_envacct> ea;
  RCHK(env_def_passes(c, p, dbase, ht, ew, o->env, envs, ea));
  stage(IXG_SS_F4_F5);
  for (size_t w = 0; w < ew.size(); w++)
    add_acct(&ea[w], ew_run[w] < 0 ? &A.e : &A.tr_e[truns[ew_run[w]].pass - IXG_PASS_5P_ONLY_FORCE],
             sizeof(ixg_envacct));
  // envelopes in the coordinates of their sequence's strand buffer
  std::vector<ixg_envelope> env;
  std::vector<int32_t> etask;
  std::vector<std::vector<ixg_envelope>> tenv(truns.size());
  for (size_t w = 0; w < ew.size(); w++) {
    const int64_t off = view_off[ew[w].task];
    for (auto &e : envs[w]) {
      ixg_envelope x{};
      x.i0 = (int32_t)(off + e.first);
      x.j0 = (int32_t)(off + e.second);
      x.strand = ftasks[ew[w].task].strand;
      x.seq = (int32_t)fseq[ew[w].task];
      if (ew_run[w] < 0) {
        env.push_back(x);
        etask.push_back(ew[w].task);
      } else {
        tenv[ew_run[w]].push_back(x);
      }
    }
  }
  // the truncated passes' CM stages (cm_search.rs:2863 cm_trunc_stage), one
  // pass per call (the pass selects the CP9 and the scanners' planes) on
  // context cc: TrCYK over bands at the filter tau with its gate and
  // envelope redefinition (not at an enforced end), then FTrInside (or
  // TrCYK) at the final tau; hits in the chunk's coordinates shifted by the
  // pass's offset. Writes only its own TrOut.
  struct TrOut {
    std::vector<Hit> hits;
    uint64_t n_past_cyk = 0, pos_past_cyk = 0, nhits = 0;
    int rc = IXG_OK;
    std::string err;
    double ms = 0;  // the pass's wall time on its thread
  };
  auto trunc_pass = [&](ixg_ctx *cc, int q, TrOut &out) -> int {
    HIPCHK(hipSetDevice(cc->device));
    const int pass = IXG_PASS_5P_ONLY_FORCE + q;
    int rc = IXG_OK;
    const bool ef = pass != IXG_PASS_3P_ONLY_FORCE, el = pass != IXG_PASS_5P_ONLY_FORCE;  // enforced first / final
    std::vector<ixg_envelope> pe;
    std::vector<int32_t> prun;
    for (size_t r = 0; r < truns.size(); r++)
      if (truns[r].pass == pass)
        for (const ixg_envelope &x : tenv[r]) pe.push_back(x), prun.push_back((int32_t)r);
    if (pe.empty()) return IXG_OK;
    std::vector<ixg_envelope> ps;
    std::vector<int32_t> psrun;
    if (!o->do_fcyk) {
      ps = pe, psrun = prun;
    } else {
      std::vector<TrEnvBands> tb;
      RCHK(trunc_bands(cc, trcp9[q], trmap[q], tgt, dbase, pe, pass, o->fcyk_tau, tb));
      std::vector<ixg_trenvelope> te(pe.size());
      for (size_t k = 0; k < pe.size(); k++) {
        te[k].i0 = pe[k].i0, te[k].j0 = pe[k].j0, te[k].strand = pe[k].strand, te[k].seq = pe[k].seq;
        te[k].bands = tb[k].b.view();
        te[k].valid = tb[k].valid();
      }
      std::vector<float> vsc(te.size());
      std::vector<int32_t> vmode(te.size());
      std::vector<int64_t> ei(te.size()), ej(te.size());
      int64_t nh = 0;
      // the filter's hits are not used: no room for them (IXG_ERANGE)
      rc = ixg_trhb_scan_batch(cc, cm, tgt, te.data(), (int64_t)te.size(), 0, pass, 1, o->cyk_env_cutoff, 0.0f,
                               o->do_null3, vsc.data(), vmode.data(), ei.data(), ej.data(), nullptr, 0, &nh, nullptr);
      if (rc != IXG_ERANGE) RCHK(rc);
      for (size_t k = 0; k < te.size(); k++) {
        if (exp_surv((double)vsc[k], o->cyk_mu, o->cyk_lambda) > o->f6) continue;
        ixg_envelope x = pe[k];
        if (o->do_fcykenv && ei[k] != -1 && ej[k] != -1) {
          if (!ef) x.i0 = (int32_t)ei[k];
          if (!el) x.j0 = (int32_t)ej[k];
        }
        out.n_past_cyk++;
        out.pos_past_cyk += (uint64_t)(x.j0 - x.i0 + 1);
        ps.push_back(x);
        psrun.push_back(prun[k]);
      }
    }
    if (ps.empty()) return IXG_OK;
    std::vector<TrEnvBands> tb;
    RCHK(trunc_bands(cc, trcp9[q], trmap[q], tgt, dbase, ps, pass, o->final_tau, tb));
    std::vector<ixg_trenvelope> te(ps.size());
    for (size_t k = 0; k < ps.size(); k++) {
      te[k].i0 = ps[k].i0, te[k].j0 = ps[k].j0, te[k].strand = ps[k].strand, te[k].seq = ps[k].seq;
      te[k].bands = tb[k].b.view();
      te[k].valid = tb[k].valid();
    }
    std::vector<float> vsc(te.size());
    std::vector<int32_t> vmode(te.size());
    std::vector<int64_t> ei(te.size()), ej(te.size());
    int64_t nh = 0;
    std::vector<ixg_trhit> th(1024);
    rc = ixg_trhb_scan_batch(cc, cm, tgt, te.data(), (int64_t)te.size(), o->do_cyk ? 0 : 1, pass, 1, 0.0f, o->pli_t,
                             o->do_null3, vsc.data(), vmode.data(), ei.data(), ej.data(), th.data(),
                             (int64_t)th.size(), &nh, nullptr);
    if (rc == IXG_ERANGE) {
      th.resize(nh);
      rc = ixg_trhb_scan_batch(cc, cm, tgt, te.data(), (int64_t)te.size(), o->do_cyk ? 0 : 1, pass, 1, 0.0f,
                               o->pli_t, o->do_null3, vsc.data(), vmode.data(), ei.data(), ej.data(), th.data(), nh,
                               &nh, nullptr);
    }
    RCHK(rc);
    th.resize(nh);
    for (const ixg_trhit &x : th) {
      const TrRun &R = truns[psrun[x.env]];
      const int64_t off = view_off[R.task];  // the chunk's offset in its strand buffer
      Hit h;
      h.task = R.task;
      h.seq = tseq[R.task];
      h.pass = pass;
      genome_coords(tasks[R.task], x.i - off, x.j - off, h);
      h.score = x.score;
      h.bias = x.bias;
      h.evalue = exp_surv((double)h.score, o->fin_mu, o->fin_lambda) * o->ez;
      out.hits.push_back(h);
      out.nhits++;
    }
    return IXG_OK;
  };
  // each truncated pass with envelopes runs on a thread of its own with an
  // auxiliary context (its own stream and scratch), overlapping the STD
  // pass's F6 / F7 launches on this context's stream; joined before finalize
  TrOut tro[3];
  std::vector<std::thread> trth;
  std::vector<int> trq;
  for (int q = 0; q < 3 && do_trunc; q++) {
    bool any = false;
    for (size_t r = 0; r < truns.size() && !any; r++) any = truns[r].pass == IXG_PASS_5P_ONLY_FORCE + q && !tenv[r].empty();
    if (!any) continue;
    if (!c->aux[q]) RCHK(ixg_init(c->device, c->budget, &c->aux[q]));
    trq.push_back(q);
  }
  struct Joiner {  // every return path waits for the threads before the state they read goes away
    std::vector<std::thread> &t;
    ~Joiner() {
      for (auto &x : t)
        if (x.joinable()) x.join();
    }
  } joiner{trth};
  for (int q : trq)
    trth.emplace_back([&, q] {
      const clk::time_point t0 = clk::now();
      tro[q].rc = trunc_pass(c->aux[q], q, tro[q]);
      if (tro[q].rc != IXG_OK) tro[q].err = g_err;
      tro[q].ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    });
  auto join_trunc = [&]() -> int {
    for (auto &t : trth) t.join();
    trth.clear();
    int first = IXG_OK;
    for (int q : trq) {
      ixg_ctx *x = c->aux[q];
      HIPCHK(ktime_flush(x));
      for (int k = 0; k < IXG_NKERNELS; k++) c->kms[k] += x->kms[k], c->kn[k] += x->kn[k], x->kms[k] = 0, x->kn[k] = 0;
      if (tro[q].rc != IXG_OK && first == IXG_OK) first = tro[q].rc, g_err = tro[q].err;
    }
    return first;
  };
  std::vector<Hit> hits;
  // F6: CYK filter over HMM bands at the filter tau
  std::vector<ixg_envelope> surv;
  std::vector<int32_t> stask;
  if (!o->do_fcyk) {
    surv = env;
    stask = etask;
  } else if (!env.empty()) {
    std::vector<EnvBands> eb;
    std::vector<char> skip;
    RCHK(iterate_bands(c, cp9, map, tgt, dbase, env, o->fcyk_tau, o->maxtau, o->size_limit, eb, skip));
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
  if (!surv.empty()) {
    std::vector<EnvBands> eb;
    std::vector<char> skip;
    RCHK(iterate_bands(c, cp9, map, tgt, dbase, surv, o->final_tau, o->maxtau, o->size_limit, eb, skip));
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
      const int64_t off = view_off[t];
      for (size_t a = 0; a < v.size(); a++) {
        if (gone[a]) continue;
        Hit h;
        h.task = t;
        h.seq = tseq[t];
        h.pass = IXG_PASS_STD_ANY;
        genome_coords(tasks[t], v[a].i - off, v[a].j - off, h);
        h.score = v[a].score;
        h.bias = v[a].bias;
        h.evalue = exp_surv((double)h.score, o->fin_mu, o->fin_lambda) * o->ez;
        hits.push_back(h);
      }
    }
  }
  {
    const int trc = join_trunc();
    for (int q : trq) c->search_ms[IXG_SS_TRUNC] = std::max(c->search_ms[IXG_SS_TRUNC], tro[q].ms);
    stage(IXG_SS_FINALIZE);
    RCHK(trc);
  }
  for (int q : trq) {
    hits.insert(hits.end(), tro[q].hits.begin(), tro[q].hits.end());
    A.tr_n_past_cyk[q] += tro[q].n_past_cyk;
    A.tr_pos_past_cyk[q] += tro[q].pos_past_cyk;
    A.tr_hits[q] += tro[q].nhits;
  }
  // finalize (cm_search.rs:1915): global overlap removal per sequence and
  // strand (:3260), then the E-value sort and the reporting threshold
  {
    std::vector<size_t> ord(hits.size());
    for (size_t q = 0; q < ord.size(); q++) ord[q] = q;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
      const Hit &x = hits[a], &y = hits[b];
      if (x.seq != y.seq) return x.seq < y.seq;
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
        if (hb.seq != ha.seq || hb.strand != ha.strand) break;
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
    if (a.seq != b.seq) return a.seq < b.seq;
    if (a.start != b.start) return a.start < b.start;
    if (a.pass != b.pass) return a.pass > b.pass;
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
      x.pass = h.pass;
      x.seq = h.seq;
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

extern "C" int ixg_search_std(ixg_ctx *c, const ixg_p7 *p, const ixg_cm *cm, const ixg_cp9 *cp9,
                              const ixg_hbmap *map, const ixg_target *tgt, const ixg_searchopts *o,
                              ixg_searchhit *out, int64_t cap, int64_t *nout, ixg_searchacct *acct) {
  if (!o) {
    g_err = "ixg_search_std: bad arguments";
    return IXG_EINVAL;
  }
  ixg_searchopts so = *o;
  so.trunc_mode = 0;
  return ixg_search(c, p, cm, cp9, map, nullptr, nullptr, tgt, &so, out, cap, nout, acct);
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
