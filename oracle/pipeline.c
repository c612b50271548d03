/*
 * pipeline.c -- CPU oracle (test infrastructure only) for the STD pass of a
 * search over one sequence, F1 through F7: FaithfulSearcher::search
 * (cm_search.rs:1300) with window_tasks (:1774), run_window_task (:1813),
 * trunc_pipeline_one_strand's STD pass (:2214), trunc_env_def (:2449),
 * cm_std_stage (:2707) with cp9_iterate_seq2bands (cp9.rs:4075) and
 * remove_overlaps_greedy (cp9.rs:6544), and finalize (cm_search.rs:1915).
 * Each task is searched on its own copy of the chunk (reverse-complemented
 * for the bottom strand), as the reference does. Not linked into the product.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "ixo_pipe.h"

#define CM_MAX_RESIDUE_COUNT 100000

typedef struct {
  int64_t start, stop;
  float score, bias;
  double evalue;
  int strand, removed, pass;
  int64_t seq; /* index of the target sequence */
} phit;

/* cp9.rs:4053 cp9_seq2bands + :4075 the tau iteration; fills the bands of
 * dsq[i0..j0]; returns the final size in Mb. The arrays are allocated here. */
static double iterate_bands(const ixo_pipe *P, const uint8_t *dsq, int i0, int j0, double tau, double maxtau,
                            float limit, int32_t **jmin, int32_t **jmax, int32_t **hd_off, int32_t **hdmin,
                            int32_t **hdmax) {
  const int H = P->cp9->M, M = P->hbcm->M, L = j0 - i0 + 1;
  const size_t cells = (size_t)(L + 1) * (H + 1);
  int32_t *mx[8];
  for (int q = 0; q < 8; q++) mx[q] = (int32_t *)malloc(sizeof(int32_t) * cells);
  int32_t *sca = (int32_t *)malloc(sizeof(int32_t) * (size_t)(L + 1));
  int32_t *pn = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)(H + 1));
  int32_t *imin = (int32_t *)malloc(sizeof(int32_t) * M), *imax = (int32_t *)malloc(sizeof(int32_t) * M);
  *jmin = (int32_t *)malloc(sizeof(int32_t) * M);
  *jmax = (int32_t *)malloc(sizeof(int32_t) * M);
  *hd_off = (int32_t *)malloc(sizeof(int32_t) * (M + 1));
  *hdmin = *hdmax = NULL;
  int at_limit = 0;
  double mb = 0.0;
  for (;;) {
    int64_t bp;
    ixo_cp9_forward(P->cp9, dsq, i0, j0, 1, 0, mx[0], mx[1], mx[2], mx[3], sca, &bp);
    ixo_cp9_backward(P->cp9, dsq, i0, j0, 1, 0, mx[4], mx[5], mx[6], mx[7], sca, &bp);
    ixo_cp9_fb2hmmbands(P->cp9, dsq, mx[0], mx[1], mx[2], mx[4], mx[5], mx[6], i0, j0, 1.0 - tau, 1, pn);
    ixo_hmm2ij_bands(P->hbcm, P->hbcp9, pn, i0, j0, 1, imin, imax, *jmin, *jmax);
    const int64_t rows = ixo_hb_ij2d(P->hbcm, imin, imax, *jmin, *jmax, 0, *hd_off, NULL, NULL, 0);
    free(*hdmin);
    free(*hdmax);
    *hdmin = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows + 1));
    *hdmax = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows + 1));
    ixo_hb_ij2d(P->hbcm, imin, imax, *jmin, *jmax, 0, *hd_off, *hdmin, *hdmax, rows);
    mb = ixo_hb_mx_mb(P->hbcm, *jmin, *jmax, *hd_off, *hdmin, *hdmax, L, NULL);
    if (mb < limit || at_limit) break;
    tau *= 2.0;
    if (tau >= maxtau) {
      tau = maxtau;
      at_limit = 1;
    }
  }
  for (int q = 0; q < 8; q++) free(mx[q]);
  free(sca);
  free(pn);
  free(imin);
  free(imax);
  return mb;
}

static void free_bands(int32_t *a, int32_t *b, int32_t *c, int32_t *d, int32_t *e) {
  free(a);
  free(b);
  free(c);
  free(d);
  free(e);
}

/* cp9.rs:6544 remove_overlaps_greedy: score descending, then start ascending;
 * the reference's sort is stable, so the reporting order breaks ties */
typedef struct { ixo_hit h; int64_t at; } ohit;
static int hit_cmp(const void *x, const void *y) {
  const ohit *a = (const ohit *)x, *b = (const ohit *)y;
  if (a->h.score != b->h.score) return a->h.score > b->h.score ? -1 : 1;
  if (a->h.i != b->h.i) return (a->h.i > b->h.i) - (a->h.i < b->h.i);
  return (a->at > b->at) - (a->at < b->at);
}

static void push(phit **v, int64_t *n, int64_t *cap, phit h) {
  if (*n == *cap) {
    *cap = *cap ? 2 * *cap : 64;
    *v = (phit *)realloc(*v, sizeof(phit) * (size_t)*cap);
  }
  (*v)[(*n)++] = h;
}

/* cp9.rs:4130 cp9_seq2bands_trunc with the pass's CP9 (q) for
 * dsq[i0..j0]: Forward / Backward scanning unless the pass forces the first
 * / final residue, posterior bands and occupancy, start / end prediction,
 * candidacy, hmm2ij with do_trunc, ij2d with do_trunc. Arrays allocated here. */
static void trunc_bands(const ixo_pipe *P, int q, int pass, const uint8_t *dsq, int i0, int j0, double tau,
                        int32_t **jmin, int32_t **jmax, int32_t **hd_off, int32_t **hdmin, int32_t **hdmax,
                        uint8_t *jv, uint8_t *lv, uint8_t *rv, uint8_t *tv) {
  const ixo_cp9 *h = P->trcp9[q];
  const int H = h->M, M = P->hbcm->M, L = j0 - i0 + 1;
  const int fscan = !(pass == 2 || pass == 4), bscan = !(pass == 3 || pass == 4);
  const size_t cells = (size_t)(L + 1) * (H + 1);
  int32_t *mx[8];
  for (int k = 0; k < 8; k++) mx[k] = (int32_t *)malloc(sizeof(int32_t) * cells);
  int32_t *sca = (int32_t *)malloc(sizeof(int32_t) * (size_t)(L + 1));
  int32_t *pn = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)(H + 1));
  int32_t *occ = (int32_t *)malloc(sizeof(int32_t) * (size_t)(H + 1));
  int64_t bp;
  ixo_cp9_forward(h, dsq, i0, j0, fscan, 0, mx[0], mx[1], mx[2], mx[3], sca, &bp);
  ixo_cp9_backward(h, dsq, i0, j0, bscan, 0, mx[4], mx[5], mx[6], mx[7], sca, &bp);
  ixo_cp9_fb2hmmbands(h, dsq, mx[0], mx[1], mx[2], mx[4], mx[5], mx[6], i0, j0, 1.0 - tau, fscan, pn);
  ixo_cp9_occupancy(h, dsq, mx[0], mx[2], mx[4], mx[6], i0, j0, fscan, occ);
  ixo_trband tb;
  tb.jvalid = jv, tb.lvalid = lv, tb.rvalid = rv, tb.tvalid = tv;
  ixo_tr_predict_start_end(pn, occ, H, i0, j0, 0.01f, 0.98f, &tb);
  ixo_tr_candidates(P->hbcm, pass, &tb);
  int32_t *imin = (int32_t *)malloc(sizeof(int32_t) * M), *imax = (int32_t *)malloc(sizeof(int32_t) * M);
  *jmin = (int32_t *)malloc(sizeof(int32_t) * M);
  *jmax = (int32_t *)malloc(sizeof(int32_t) * M);
  *hd_off = (int32_t *)malloc(sizeof(int32_t) * (M + 1));
  ixo_hmm2ij_bands_tr(P->hbcm, P->trhbcp9[q], pn, i0, j0, 1, &tb, imin, imax, *jmin, *jmax);
  const int64_t rows = ixo_hb_ij2d(P->hbcm, imin, imax, *jmin, *jmax, 1, *hd_off, NULL, NULL, 0);
  *hdmin = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows + 1));
  *hdmax = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rows + 1));
  ixo_hb_ij2d(P->hbcm, imin, imax, *jmin, *jmax, 1, *hd_off, *hdmin, *hdmax, rows);
  for (int k = 0; k < 8; k++) free(mx[k]);
  free(sca);
  free(pn);
  free(occ);
  free(imin);
  free(imax);
}

/* cm_search.rs:2380-2420 + :2863 cm_trunc_stage: the truncated pass `pass`
 * (q = pass - 2) of a chunk dsq[1..wl] whose STD pass had survivors: its
 * subsequence (the chunk, or its first / last maxW residues), the filter
 * (the STD windows when the subsequence is the chunk), trunc_env_def, the
 * TrCYK filter and FTrInside; hits shifted back into chunk coordinates. */
static void trunc_pass(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *dsq, int64_t wl, int64_t wgs,
                       int64_t wge, int strand, int pass, const ixo_window *std_wins, int64_t std_nw,
                       const ixo_acct *std_acct, ixo_pipeacct *A, phit **hits, int64_t *nh, int64_t *hcap) {
  const int q = pass - 2;
  const int64_t maxw = o->f3.maxw;
  const uint8_t *sub = dsq;
  int64_t sublen = wl, off = 0;
  uint8_t *own = NULL;
  if (!(wl <= maxw || pass == 4)) {
    sublen = maxw;
    off = pass == 2 ? 0 : wl - maxw;
    own = (uint8_t *)malloc((size_t)maxw + 2);
    own[0] = own[maxw + 1] = 255;
    memcpy(own + 1, dsq + 1 + off, (size_t)maxw);
    sub = own;
  }
  const int64_t wcap = 1 << 16;
  ixo_window *wins = (ixo_window *)malloc(sizeof(ixo_window) * wcap);
  ixo_acct fa;
  int64_t nw;
  if (sub == dsq) {
    nw = std_nw < wcap ? std_nw : wcap;
    memcpy(wins, std_wins, sizeof(ixo_window) * (size_t)nw);
    fa = *std_acct;
  } else {
    memset(&fa, 0, sizeof(fa));
    nw = ixo_f3_filter_sequence(P->msv, P->vit, P->fwd, P->bias, P->p7, sub, sublen, &o->f3, wins, wcap, &fa);
    if (nw > wcap) nw = wcap;
  }
  {
    const uint64_t *s = (const uint64_t *)&fa;
    uint64_t *d = (uint64_t *)&A->tr_f[q];
    for (size_t k = 0; k < sizeof(ixo_acct) / 8; k++) d[k] += s[k];
  }
  if (nw > 0) {
    const int64_t ecap = 4096;
    int32_t *es = (int32_t *)malloc(sizeof(int32_t) * ecap), *ee = (int32_t *)malloc(sizeof(int32_t) * ecap);
    ixo_envacct ea;
    int64_t ne = ixo_env_def_pass(P->gmt[q], pass, P->bias, sub, sublen, wins, nw, &o->env, es, ee, ecap, &ea);
    if (ne > ecap) ne = ecap;
    {
      const uint64_t *s = (const uint64_t *)&ea;
      uint64_t *d = (uint64_t *)&A->tr_e[q];
      for (size_t k = 0; k < sizeof(ixo_envacct) / 8; k++) d[k] += s[k];
    }
    const int M = P->cm->M;
    const float *pty = P->trp_l + (size_t)(pass == 4 ? 0 : (pass == 2 ? 1 : 2)) * M;
    uint8_t *jv = (uint8_t *)malloc((size_t)M), *lv = (uint8_t *)malloc((size_t)M), *rv = (uint8_t *)malloc((size_t)M),
            *tv = (uint8_t *)malloc((size_t)M);
    ixo_trvalid val = {jv, lv, rv, tv};
    const int ef = pass != 3, el = pass != 2;
    for (int64_t k = 0; k < ne; k++) {
      int e0 = es[k], e1 = ee[k];
      int32_t *jmin, *jmax, *hd_off, *hdmin, *hdmax;
      if (o->do_fcyk) {
        trunc_bands(P, q, pass, sub, e0, e1, o->fcyk_tau, &jmin, &jmax, &hd_off, &hdmin, &hdmax, jv, lv, rv, tv);
        ixo_bands b = {jmin, jmax, hd_off, hdmin, hdmax};
        float vsc;
        int vmode;
        int64_t ei, ej;
        ixo_trhb_scan(P->cm, P->sc, P->lmesc, P->rmesc, pty, &b, &val, sub, e0, e1, pass, 1, 0, 0.0f, o->do_null3,
                      o->cyk_env_cutoff, &vsc, &vmode, &ei, &ej, NULL, 0, NULL);
        free_bands(jmin, jmax, hd_off, hdmin, hdmax);
        if (ixo_exp_surv((double)vsc, o->cyk_mu, o->cyk_lambda) > o->f6) continue;
        if (o->do_fcykenv && ei != -1 && ej != -1) {
          if (!ef) e0 = (int)ei;
          if (!el) e1 = (int)ej;
        }
        A->tr_n_past_cyk[q]++;
        A->tr_pos_past_cyk[q] += (uint64_t)(e1 - e0 + 1);
      }
      es[k] = e0, ee[k] = e1;
      /* F7 for this envelope (the reference collects the survivors first;
       * every survivor's final round is independent of the others) */
      trunc_bands(P, q, pass, sub, e0, e1, o->final_tau, &jmin, &jmax, &hd_off, &hdmin, &hdmax, jv, lv, rv, tv);
      ixo_bands b = {jmin, jmax, hd_off, hdmin, hdmax};
      const int64_t hc = 4096;
      ixo_trhit *th = (ixo_trhit *)malloc(sizeof(ixo_trhit) * hc);
      float vsc;
      int vmode;
      int64_t ei, ej;
      int64_t nt = ixo_trhb_scan(P->cm, P->sc, P->lmesc, P->rmesc, pty, &b, &val, sub, e0, e1, pass, 1,
                                 o->do_cyk ? 0 : 1, o->pli_t, o->do_null3, 0.0f, &vsc, &vmode, &ei, &ej, th, hc, NULL);
      if (nt > hc) nt = hc;
      for (int64_t a = 0; a < nt; a++) {
        phit h;
        memset(&h, 0, sizeof(h));
        const int64_t hi = th[a].i + off, hj = th[a].j + off;
        if (strand) h.start = wge - hi + 1, h.stop = wge - hj + 1;
        else h.start = wgs + hi - 1, h.stop = wgs + hj - 1;
        h.score = th[a].score;
        h.bias = th[a].bias;
        h.evalue = ixo_exp_surv((double)h.score, o->fin_mu, o->fin_lambda) * o->ez;
        h.strand = strand;
        h.pass = pass;
        push(hits, nh, hcap, h);
        A->tr_hits[q]++;
      }
      free(th);
      free_bands(jmin, jmax, hd_off, hdmin, hdmax);
    }
    free(jv);
    free(lv);
    free(rv);
    free(tv);
    free(es);
    free(ee);
  }
  free(wins);
  free(own);
}

/* the STD pass of one (chunk x strand) task on its chunk dsq[1..wl], then
 * its truncated passes (h5 / h3: the chunk holds the sequence's 5' / 3' end
 * on its strand) */
static void one_task(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *dsq, int64_t wl, int64_t wgs,
                     int64_t wge, int strand, int h5, int h3, ixo_pipeacct *A, phit **hits, int64_t *nh,
                     int64_t *hcap) {
  const int64_t maxw = o->f3.maxw;
  const int run[3] = {o->trunc_mode && h5, o->trunc_mode && h3, o->trunc_mode && h5 && h3 && wl <= maxw};
  for (int q = 0; q < 3; q++)  /* charged before any early out (cm_search.rs:2283-2300) */
    if (run[q]) A->tr_nres[q] += (uint64_t)(maxw < wl ? maxw : wl);
  const int64_t wcap = 1 << 16;
  ixo_window *wins = (ixo_window *)malloc(sizeof(ixo_window) * wcap);
  ixo_acct fa;
  memset(&fa, 0, sizeof(fa));
  int64_t nw = ixo_f3_filter_sequence(P->msv, P->vit, P->fwd, P->bias, P->p7, dsq, wl, &o->f3, wins, wcap, &fa);
  if (nw > wcap) nw = wcap;
  {
    const uint64_t *s = (const uint64_t *)&fa;
    uint64_t *d = (uint64_t *)&A->f;
    for (size_t q = 0; q < sizeof(ixo_acct) / 8; q++) d[q] += s[q];
  }
  if (nw == 0) {
    free(wins);
    return;
  }
  const int64_t ecap = 4096;
  int32_t *es = (int32_t *)malloc(sizeof(int32_t) * ecap), *ee = (int32_t *)malloc(sizeof(int32_t) * ecap);
  ixo_envacct ea;
  int64_t ne = ixo_env_def_std(P->gm, P->bias, dsq, wl, wins, nw, &o->env, es, ee, ecap, &ea);
  if (ne > ecap) ne = ecap;
  {
    const uint64_t *s = (const uint64_t *)&ea;
    uint64_t *d = (uint64_t *)&A->e;
    for (size_t q = 0; q < sizeof(ixo_envacct) / 8; q++) d[q] += s[q];
  }
  /* F6 */
  int32_t *ss = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne + 1)), *se = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne + 1));
  int64_t ns = 0;
  for (int64_t k = 0; k < ne; k++) {
    int e0 = es[k], e1 = ee[k];
    if (!o->do_fcyk) {
      ss[ns] = e0, se[ns] = e1, ns++;
      continue;
    }
    int32_t *jmin, *jmax, *hd_off, *hdmin, *hdmax;
    const double mb = iterate_bands(P, dsq, e0, e1, o->fcyk_tau, o->maxtau, o->size_limit, &jmin, &jmax, &hd_off,
                                    &hdmin, &hdmax);
    if (mb > o->size_limit) {
      free_bands(jmin, jmax, hd_off, hdmin, hdmax);
      continue;
    }
    ixo_bands b = {jmin, jmax, hd_off, hdmin, hdmax};
    float vsc;
    int64_t ei, ej;
    ixo_hb_scan(P->cm, P->sc, &b, dsq, e0, e1, 0, o->cyk_env_cutoff, &vsc, &ei, &ej, NULL, NULL);
    free_bands(jmin, jmax, hd_off, hdmin, hdmax);
    if (ixo_exp_surv((double)vsc, o->cyk_mu, o->cyk_lambda) > o->f6) continue;
    if (o->do_fcykenv && ei != -1 && ej != -1) e0 = (int)ei, e1 = (int)ej;
    A->n_past_cyk++;
    A->pos_past_cyk += (uint64_t)(e1 - e0 + 1);
    ss[ns] = e0, se[ns] = e1, ns++;
  }
  /* F7 */
  for (int64_t k = 0; k < ns; k++) {
    const int e0 = ss[k], e1 = se[k];
    int32_t *jmin, *jmax, *hd_off, *hdmin, *hdmax;
    const double mb = iterate_bands(P, dsq, e0, e1, o->final_tau, o->maxtau, o->size_limit, &jmin, &jmax, &hd_off,
                                    &hdmin, &hdmax);
    if (mb > o->size_limit) {
      free_bands(jmin, jmax, hd_off, hdmin, hdmax);
      continue;
    }
    ixo_bands b = {jmin, jmax, hd_off, hdmin, hdmax};
    const int64_t nc = ixo_hb_ncells(&b, P->cm->M);
    float *root = (float *)malloc(sizeof(float) * (size_t)(nc + 1));
    float vsc;
    int64_t ei, ej;
    ixo_hb_scan(P->cm, P->sc, &b, dsq, e0, e1, o->do_cyk ? 0 : 1, 0.0f, &vsc, &ei, &ej, root, NULL);
    const int64_t rcap = 4096;
    ixo_hit *raw = (ixo_hit *)malloc(sizeof(ixo_hit) * rcap);
    int64_t nr = ixo_hb_report_hits(P->cm, &b, dsq, e0, e1, root, o->pli_t, o->do_null3, raw, rcap);
    if (nr > rcap) nr = rcap;
    ohit *srt = (ohit *)malloc(sizeof(ohit) * (size_t)(nr + 1));
    for (int64_t a = 0; a < nr; a++) srt[a].h = raw[a], srt[a].at = a;
    qsort(srt, (size_t)nr, sizeof(ohit), hit_cmp);
    for (int64_t a = 0; a < nr; a++) raw[a] = srt[a].h;
    free(srt);
    char *gone = (char *)calloc((size_t)(nr + 1), 1);
    for (int64_t a = 0; a < nr; a++) {
      if (gone[a]) continue;
      for (int64_t c = a + 1; c < nr; c++)
        if (!gone[c] && !(raw[c].j < raw[a].i) && !(raw[a].j < raw[c].i)) gone[c] = 1;
    }
    for (int64_t a = 0; a < nr; a++) {
      if (gone[a]) continue;
      phit h;
      memset(&h, 0, sizeof(h));
      if (strand) h.start = wge - raw[a].i + 1, h.stop = wge - raw[a].j + 1;
      else h.start = wgs + raw[a].i - 1, h.stop = wgs + raw[a].j - 1;
      h.score = raw[a].score;
      h.bias = raw[a].bias;
      h.evalue = ixo_exp_surv((double)h.score, o->fin_mu, o->fin_lambda) * o->ez;
      h.strand = strand;
      h.pass = 1;
      push(hits, nh, hcap, h);
    }
    free(gone);
    free(raw);
    free(root);
    free_bands(jmin, jmax, hd_off, hdmin, hdmax);
  }
  free(ss);
  free(se);
  free(es);
  free(ee);
  for (int q = 0; q < 3; q++)
    if (run[q]) trunc_pass(P, o, dsq, wl, wgs, wge, strand, 2 + q, wins, nw, &fa, A, hits, nh, hcap);
  free(wins);
}

static const phit *g_h;
static int ov_cmp(const void *x, const void *y) {
  const phit *a = &g_h[*(const int64_t *)x], *b = &g_h[*(const int64_t *)y];
  if (a->seq != b->seq) return a->seq < b->seq ? -1 : 1;
  if (a->strand != b->strand) return a->strand - b->strand;
  if (a->score != b->score) return a->score > b->score ? -1 : 1;
  if (a->start != b->start) return a->start < b->start ? -1 : 1;
  return (a->stop > b->stop) - (a->stop < b->stop);
}
static int fin_cmp(const void *x, const void *y) {
  const phit *a = (const phit *)x, *b = (const phit *)y;
  if (a->evalue != b->evalue) return a->evalue < b->evalue ? -1 : 1;
  if (a->score != b->score) return a->score > b->score ? -1 : 1;
  if (a->seq != b->seq) return a->seq < b->seq ? -1 : 1;
  if (a->start != b->start) return a->start < b->start ? -1 : 1;
  if (a->pass != b->pass) return a->pass > b->pass ? -1 : 1;
  return (a->stop > b->stop) - (a->stop < b->stop);
}

typedef struct {
  int64_t wgs, wge, wl, ctx, L, seq;
  int strand;
} otask;

/* cm_search.rs:1774-1808 window_tasks for sequence `seq` of length L */
static int64_t make_tasks(const ixo_pipeopts *o, int64_t L, int64_t seq, otask *out) {
  const int64_t maxw = o->f3.maxw;
  int64_t n = 0, new_start = 1;
  int first = 1;
  while (new_start <= L) {
    const int64_t ctx = first ? 0 : (maxw < new_start - 1 ? maxw : new_start - 1);
    const int64_t wgs = new_start - ctx;
    const int64_t cnt = CM_MAX_RESIDUE_COUNT < L - new_start + 1 ? CM_MAX_RESIDUE_COUNT : L - new_start + 1;
    const int64_t wl = ctx + cnt, wge = wgs + wl - 1;
    for (int strand = 0; strand < 2; strand++) {
      if ((strand == 0 && !o->do_top) || (strand == 1 && !o->do_bot)) continue;
      if (out) out[n] = (otask){wgs, wge, wl, ctx, L, seq, strand};
      n++;
    }
    new_start = wge + 1;
    first = 0;
  }
  return n;
}

static void run_task(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *const *seqs, const otask *t,
                     uint8_t *dsq, ixo_pipeacct *A, phit **hits, int64_t *nh, int64_t *hcap) {
  const uint8_t *seq = seqs[t->seq];
  const int64_t nh0 = *nh;
  dsq[0] = dsq[t->wl + 1] = 255;
  for (int64_t k = 1; k <= t->wl; k++) {
    const uint8_t x = t->strand ? seq[t->wge - (k - 1)] : seq[t->wgs + k - 1];
    dsq[k] = t->strand ? (x < 4 ? (uint8_t)(3 - x) : x) : x;
  }
  A->nres += (uint64_t)(t->wl - t->ctx);
  const int h5 = t->strand ? t->wge == t->L : t->wgs == 1, h3 = t->strand ? t->wgs == 1 : t->wge == t->L;
  one_task(P, o, dsq, t->wl, t->wgs, t->wge, t->strand, h5, h3, A, hits, nh, hcap);
  for (int64_t k = nh0; k < *nh; k++) (*hits)[k].seq = t->seq;
}

/* finalize (cm_search.rs:1915): global overlap removal, E-value sort,
 * reporting threshold; returns the number reported */
static int64_t finalize(const ixo_pipeopts *o, phit *hits, int64_t nh, ixo_pipehit *out, int64_t cap) {
  int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nh + 1));
  for (int64_t q = 0; q < nh; q++) ord[q] = q;
  g_h = hits;
  qsort(ord, (size_t)nh, sizeof(int64_t), ov_cmp);
  for (int64_t a = 0; a < nh; a++) {
    phit *ha = &hits[ord[a]];
    if (ha->removed) continue;
    for (int64_t b = a + 1; b < nh; b++) {
      phit *hb = &hits[ord[b]];
      if (hb->seq != ha->seq || hb->strand != ha->strand) break;
      if (hb->removed) continue;
      const int ov = ha->strand == 0 ? (!(hb->stop < ha->start) && !(ha->stop < hb->start))
                                     : (!(hb->start < ha->stop) && !(ha->start < hb->stop));
      if (ov) hb->removed = 1;
    }
  }
  free(ord);
  int64_t nk = 0;
  for (int64_t q = 0; q < nh; q++)
    if (!hits[q].removed) hits[nk++] = hits[q];
  qsort(hits, (size_t)nk, sizeof(phit), fin_cmp);
  int64_t n = 0;
  for (int64_t q = 0; q < nk; q++) {
    if (!(hits[q].evalue <= o->e_report)) continue;
    if (n < cap) {
      out[n].start = hits[q].start;
      out[n].stop = hits[q].stop;
      out[n].score = hits[q].score;
      out[n].bias = hits[q].bias;
      out[n].evalue = hits[q].evalue;
      out[n].strand = hits[q].strand;
      out[n].pass = hits[q].pass;
      out[n].seq = hits[q].seq;
    }
    n++;
  }
  return n;
}

int64_t ixo_search_std(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *seq, int64_t L, ixo_pipehit *out,
                       int64_t cap, ixo_pipeacct *A) {
  return ixo_search_multi_mt(P, o, &seq, &L, 1, out, cap, A, 1);
}

int64_t ixo_search_std_mt(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *seq, int64_t L,
                          ixo_pipehit *out, int64_t cap, ixo_pipeacct *A, int nthreads) {
  return ixo_search_multi_mt(P, o, &seq, &L, 1, out, cap, A, nthreads);
}

/* the tasks run on nthreads threads (the reference's par_iter over tasks,
 * cm_search.rs:1312); per-thread hit lists and counters are merged before
 * finalize, so the result does not depend on the thread count */
typedef struct {
  const ixo_pipe *P;
  const ixo_pipeopts *o;
  const uint8_t *const *seqs;
  const otask *tasks;
  int64_t nt;
  int64_t *next;
  pthread_mutex_t *mu;
  ixo_pipeacct A;
  phit *hits;
  int64_t nh, hcap;
} worker_t;

static void *worker(void *arg) {
  worker_t *w = (worker_t *)arg;
  uint8_t *dsq = (uint8_t *)malloc((size_t)(CM_MAX_RESIDUE_COUNT + w->o->f3.maxw + 2));
  for (;;) {
    pthread_mutex_lock(w->mu);
    const int64_t t = (*w->next)++;
    pthread_mutex_unlock(w->mu);
    if (t >= w->nt) break;
    run_task(w->P, w->o, w->seqs, &w->tasks[t], dsq, &w->A, &w->hits, &w->nh, &w->hcap);
  }
  free(dsq);
  return NULL;
}

int64_t ixo_search_multi_mt(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *const *seqs, const int64_t *L,
                            int64_t nseq, ixo_pipehit *out, int64_t cap, ixo_pipeacct *A, int nthreads) {
  memset(A, 0, sizeof(*A));
  (void)ixo_flogsum(0.f, 0.f); /* the lazily built tables, before any thread reads them */
  (void)ixo_ilogsum(0, 0);
  (void)ixo_p7_flogsum(0.f, 0.f);
  /* the flat task list of every sequence in order (cm_search.rs:1787) */
  int64_t nt = 0;
  for (int64_t s = 0; s < nseq; s++) nt += make_tasks(o, L[s], s, NULL);
  otask *tasks = (otask *)malloc(sizeof(otask) * (size_t)(nt + 1));
  for (int64_t s = 0, k = 0; s < nseq; s++) k += make_tasks(o, L[s], s, tasks + k);
  if (nthreads < 1) nthreads = 1;
  worker_t *w = (worker_t *)calloc((size_t)nthreads, sizeof(worker_t));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  int64_t next = 0;
  for (int q = 0; q < nthreads; q++) {
    w[q].P = P, w[q].o = o, w[q].seqs = seqs, w[q].tasks = tasks, w[q].nt = nt, w[q].next = &next, w[q].mu = &mu;
    if (q) pthread_create(&th[q], NULL, worker, &w[q]);
  }
  worker(&w[0]);
  for (int q = 1; q < nthreads; q++) pthread_join(th[q], NULL);
  phit *hits = NULL;
  int64_t nh = 0, hcap = 0;
  for (int q = 0; q < nthreads; q++) {
    const uint64_t *s = (const uint64_t *)&w[q].A;
    uint64_t *d = (uint64_t *)A;
    for (size_t k = 0; k < sizeof(ixo_pipeacct) / 8; k++) d[k] += s[k];
    for (int64_t k = 0; k < w[q].nh; k++) push(&hits, &nh, &hcap, w[q].hits[k]);
    free(w[q].hits);
  }
  const int64_t n = finalize(o, hits, nh, out, cap);
  A->n_hits = (uint64_t)n;
  free(hits);
  free(w);
  free(th);
  free(tasks);
  return n;
}
