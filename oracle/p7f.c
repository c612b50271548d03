/*
 * p7f.c -- CPU oracle (test infrastructure only): restatement of the
 * infernox p7 filter path (F1 SSV, F1b MSV, F2 Viterbi, F3 Forward, F3b
 * composition bias) and of the filter cascade that strings them together.
 *
 * Not part of the product: only tests, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, as the checker.
 *
 * Each function names the reference file:line it follows. Float code is
 * compiled with -ffp-contract=off; double->float casts follow the reference's
 * `as f32` placement exactly.
 */
#include "ixo.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KP IXO_KP
#define KCANON 4
#define NEGINF_W ((int16_t)-32768)
static const double LN2 = 0.693147180559945309417232121458176568;
static const double WINDOW_BETA = 1e-7;          /* cm_pipeline.rs:119 */
static const int64_t CM_MAX_RESIDUE_COUNT = 100000; /* cm_pipeline.rs:858 */

/* cm_pipeline.rs:124-141 degen_set: canonical residues each code expands to. */
static int degen_set(int code, int out[4]) {
  static const int sets[16][5] = {
    {1, 0}, {1, 1}, {1, 2}, {1, 3}, {0},
    {2, 0, 2}, {2, 1, 3}, {2, 0, 1}, {2, 2, 3}, {2, 1, 2}, {2, 0, 3},
    {3, 0, 1, 3}, {3, 1, 2, 3}, {3, 0, 1, 2}, {3, 0, 2, 3}, {4, 0, 1, 2, 3}};
  if (code < 0 || code > 15) return 0;
  int n = sets[code][0];
  for (int i = 0; i < n; i++) out[i] = sets[code][1 + i];
  return n;
}

/* Match scores MSC[k][x] = ln(mat/0.25) as f32, degenerate = mean over set
 * (cm_pipeline.rs:216-239, p7_vitfilter.rs:134-150, cm_pipeline.rs:974-996). */
static void match_scores(const ixo_p7 *p7, int k, float sc[KP]) {
  for (int x = 0; x < KP; x++) sc[x] = -INFINITY;
  for (int x = 0; x < KCANON; x++) sc[x] = (float)log((double)p7->mat[k * 4 + x] / 0.25);
  for (int code = KCANON; code < KP; code++) {
    int set[4];
    int n = degen_set(code, set);
    if (n == 0) continue;
    float s = 0.0f;
    for (int i = 0; i < n; i++) s += sc[set[i]];
    sc[code] = s / (float)n;
  }
}

/* cm_pipeline.rs:157-160 unbiased_byteify */
static uint8_t unbiased_byteify(float scale_b, float sc) {
  float v = -roundf(scale_b * sc);
  if (v >= 255.0f) return 255;
  if (v <= 0.0f) return 0;
  return (uint8_t)v;
}

/* cm_pipeline.rs:172-179 biased_byteify */
static uint8_t biased_byteify(float scale_b, uint8_t bias_b, float sc) {
  float v = -roundf(scale_b * sc);
  if (v > (float)(255 - (int)bias_b)) return 255;
  int r = (int)v + (int)bias_b;
  if (r < 0) r = 0;
  if (r > 255) r = 255;
  return (uint8_t)r;
}

static uint8_t sat_add_u8(uint8_t a, uint8_t b) { int r = a + b; return r > 255 ? 255 : (uint8_t)r; }
static uint8_t sat_sub_u8(uint8_t a, uint8_t b) { int r = a - b; return r < 0 ? 0 : (uint8_t)r; }

/* cm_pipeline.rs:204-344 build_msv_filter */
int ixo_build_msv(const ixo_p7 *p7, int max_length, ixo_msv *f) {
  int m = p7->M;
  memset(f, 0, sizeof(*f));
  f->m = m;
  f->scale_b = (float)(3.0 / LN2);
  f->base_b = 190;
  f->max_length = max_length;
  f->mmu = p7->lmmu;
  f->mlambda = p7->lmlambda;
  float *msc = (float *)malloc(sizeof(float) * (size_t)(m + 1) * KP);
  for (int x = 0; x < KP; x++) msc[x] = -INFINITY;
  float maxsc = 0.0f;
  for (int k = 1; k <= m; k++) {
    match_scores(p7, k, msc + (size_t)k * KP);
    for (int x = 0; x < KCANON; x++)
      if (msc[k * KP + x] > maxsc) maxsc = msc[k * KP + x];
  }
  f->bias_b = unbiased_byteify(f->scale_b, -maxsc);
  f->rbv = (uint8_t *)malloc((size_t)(m + 1) * KP);
  memset(f->rbv, 255, (size_t)(m + 1) * KP);
  for (int k = 1; k <= m; k++)
    for (int x = 0; x < KP; x++)
      f->rbv[k * KP + x] = biased_byteify(f->scale_b, f->bias_b, msc[k * KP + x]);
  free(msc);
  f->tbm_b = unbiased_byteify(f->scale_b, (float)log(2.0 / ((double)m * ((double)m + 1.0))));
  f->tec_b = unbiased_byteify(f->scale_b, logf(0.5f));
  f->tjb_b = unbiased_byteify(f->scale_b, (float)log(3.0 / ((double)max_length + 3.0)));

  /* prefix/suffix lengths, cm_pipeline.rs:281-324 */
  f->prefix = (float *)calloc((size_t)m + 1, sizeof(float));
  f->suffix = (float *)calloc((size_t)m + 1, sizeof(float));
  float sum = 0.0f;
  for (int k = 1; k < m; k++) {
    double t_mi = (double)p7->trans[k * 7 + 1];
    double t_ii = (double)p7->trans[k * 7 + 4];
    float pk;
    if (t_mi == 0.0) pk = 1.0f;
    else {
      double lii = log(t_ii);
      if (lii == 0.0 || !isfinite(lii)) pk = 1.0f;
      else pk = (float)(1.0 + trunc(log(WINDOW_BETA / t_mi) / lii));
    }
    f->prefix[k] = pk;
    sum += pk;
  }
  f->prefix[0] = 0.0f;
  f->prefix[m] = 0.0f;
  if (sum > 0.0f)
    for (int k = 1; k < m; k++) f->prefix[k] /= sum;
  if (m >= 1) {
    f->suffix[m] = f->prefix[m - 1];
    for (int k = m - 1; k >= 1; k--) f->suffix[k] = f->suffix[k + 1] + f->prefix[k - 1];
    for (int k = 2; k < m; k++) f->prefix[k] += f->prefix[k - 1];
  }
  return 0;
}

void ixo_free_msv(ixo_msv *f) {
  free(f->rbv); free(f->prefix); free(f->suffix);
  memset(f, 0, sizeof(*f));
}

/* easel/gumbel.rs:168-179 */
double ixo_gumbel_invsurv(double p, double mu, double lambda) {
  double log_part;
  if (p < 5e-9) log_part = (pow(p, p) - 1.0) / p;
  else log_part = log(-log(1.0 - p));
  return mu - (log_part / lambda);
}

/* cm_pipeline.rs:1852-1861 */
double ixo_gumbel_surv(double x, double mu, double lambda) {
  double y = lambda * (x - mu);
  double ey = -exp(-y);
  if (fabs(ey) < 5e-9) return -ey;
  return 1.0 - exp(ey);
}

/* cp9.rs:6531-6537 */
double ixo_exp_surv(double x, double mu, double lambda) {
  if (x < mu) return 1.0;
  return exp(-lambda * (x - mu));
}

/* cm_pipeline.rs:1122-1125 */
double ixo_bg_null_one(int64_t L) {
  double p1 = (double)L / ((double)L + 1.0);
  return (double)L * log(p1) + log(1.0 - p1);
}

/* cm_pipeline.rs:522-533 sc_thresh */
int ixo_ssv_thresh(const ixo_msv *f, double f1) {
  double p1 = (double)f->max_length / ((double)f->max_length + 1.0);
  double nullsc = (double)f->max_length * log(p1) + log(1.0 - p1);
  double invp = ixo_gumbel_invsurv(f1, f->mmu, f->mlambda);
  return (int)ceil((nullsc + invp * LN2 + 3.0) * (double)f->scale_b + (double)f->base_b +
                   (double)f->tec_b + (double)f->tjb_b);
}

static uint8_t ssv_cost(const ixo_msv *f, int64_t k, int code) {
  return code < KP ? f->rbv[k * KP + code] : 255;
}

/* cm_pipeline.rs:521-717 ssv_diagonals (scalar form; the reference's AVX2 and
 * SSE row paths are element-wise identical to it, cm_pipeline.rs:472-480). */
int64_t ixo_ssv_diagonals(const ixo_msv *f, const uint8_t *dsq, int64_t L, double f1,
                          ixo_rawwin *out, int64_t cap) {
  int sc_thresh = ixo_ssv_thresh(f, f1);
  uint8_t xb = sat_sub_u8(f->base_b, sat_add_u8(f->tjb_b, f->tbm_b));
  int m = f->m;
  uint8_t *dp = (uint8_t *)calloc((size_t)m + 1, 1);
  uint8_t *nw = (uint8_t *)calloc((size_t)m + 1, 1);
  int64_t nwin = 0;
  int64_t i = 1;
  while (i <= L) {
    int x = dsq[i];
    uint8_t xe = 0;
    for (int k = 1; k <= m; k++) {
      uint8_t c = ssv_cost(f, k, x);
      uint8_t prev = dp[k - 1] > xb ? dp[k - 1] : xb;
      uint8_t sv = sat_sub_u8(sat_add_u8(prev, f->bias_b), c);
      nw[k] = sv;
      if (sv > xe) xe = sv;
    }
    uint8_t *t = dp; dp = nw; nw = t;

    int end_k = 0;
    if ((int)xe >= sc_thresh) {
      uint8_t xe2 = 0;
      for (int k = 1; k <= m; k++)
        if (dp[k] > xe2) { xe2 = dp[k]; end_k = k; }
    }
    if ((int)xe >= sc_thresh && end_k > 0) {
      int baseline = (int)f->base_b - (int)f->tjb_b - (int)f->tbm_b;
      int64_t start = end_k, tstart = i;
      int rem = xe;
      while (rem > baseline && start >= 1 && tstart >= 1) {
        rem -= (int)f->bias_b - (int)ssv_cost(f, start, dsq[tstart]);
        start--; tstart--;
      }
      start++; tstart++;
      int sc2 = xe, max_sc = xe, since = 0;
      int64_t max_end = i;
      int64_t k = end_k + 1, n = i + 1;
      while (k < m && n <= L) {
        sc2 += (int)f->bias_b - (int)ssv_cost(f, k, dsq[n]);
        if (sc2 >= max_sc) { max_sc = sc2; max_end = n; since = 0; }
        else { since++; if (since == 5) break; }
        k++; n++;
      }
      int64_t end_model = end_k + (max_end - i);
      float ret_sc = ((float)(max_sc - (int)f->tjb_b) - (float)f->base_b) / f->scale_b - 3.0f;
      if (nwin < cap) {
        out[nwin].n = tstart;
        out[nwin].length = end_model - start + 1;
        out[nwin].k = end_model;
        out[nwin].score = ret_sc;
      }
      nwin++;
      memset(dp, 0, (size_t)m + 1);
      i = max_end + 1;
      continue;
    }
    i++;
  }
  free(dp); free(nw);
  return nwin;
}

/* The SSV row maxima of cm_pipeline.rs:558-650 with the trigger disabled
 * (no reset): xe[i] for i in 1..L. Used to check the recurrence against the
 * closed-form segment-sum characterization. */
void ixo_ssv_rowmax(const ixo_msv *f, const uint8_t *dsq, int64_t L, uint8_t *xe_out) {
  uint8_t xb = sat_sub_u8(f->base_b, sat_add_u8(f->tjb_b, f->tbm_b));
  int m = f->m;
  uint8_t *dp = (uint8_t *)calloc((size_t)m + 1, 1);
  uint8_t *nw = (uint8_t *)calloc((size_t)m + 1, 1);
  for (int64_t i = 1; i <= L; i++) {
    uint8_t xe = 0;
    for (int k = 1; k <= m; k++) {
      uint8_t prev = dp[k - 1] > xb ? dp[k - 1] : xb;
      uint8_t sv = sat_sub_u8(sat_add_u8(prev, f->bias_b), ssv_cost(f, k, dsq[i]));
      nw[k] = sv;
      if (sv > xe) xe = sv;
    }
    uint8_t *t = dp; dp = nw; nw = t;
    xe_out[i] = xe;
  }
  free(dp); free(nw);
}

/* cm_pipeline.rs:722-793 extend_and_merge; raws is modified in place. */
int64_t ixo_extend_and_merge(const ixo_msv *f, ixo_rawwin *raws, int64_t n, int64_t target_len,
                             float pct_overlap, ixo_window *out) {
  if (n == 0) return 0;
  double maxl = (double)f->max_length;
  for (int64_t w = 0; w < n; w++) {
    int64_t ks = raws[w].k - raws[w].length + 1;
    if (ks < 0) ks = 0;
    if (ks > f->m) ks = f->m;
    int64_t ke = raws[w].k;
    if (ke < 0) ke = 0;
    if (ke > f->m) ke = f->m;
    double pref = (double)f->prefix[ks];
    double suff = (double)f->suffix[ke];
    double ws = (double)raws[w].n - maxl * (0.1 + pref);
    if (ws < 1.0) ws = 1.0;
    double we = (double)raws[w].n + (double)raws[w].length + maxl * (0.1 + suff);
    if (we > (double)target_len) we = (double)target_len;
    int64_t s = (int64_t)ws, e = (int64_t)we;
    raws[w].n = s;
    raws[w].length = e - s + 1;
  }
  int64_t nk = 0;
  ixo_rawwin *kept = (ixo_rawwin *)malloc(sizeof(ixo_rawwin) * (size_t)n);
  kept[nk++] = raws[0];
  for (int64_t i = 1; i < n; i++) {
    ixo_rawwin cur = raws[i];
    ixo_rawwin *prev = &kept[nk - 1];
    int64_t ws = prev->n > cur.n ? prev->n : cur.n;
    int64_t pe = prev->n + prev->length - 1, ce = cur.n + cur.length - 1;
    int64_t we = pe < ce ? pe : ce;
    int64_t wl = we - ws + 1;
    int64_t minlen = prev->length < cur.length ? prev->length : cur.length;
    if ((float)wl / (float)minlen > pct_overlap) {
      int64_t ms = prev->n < cur.n ? prev->n : cur.n;
      int64_t me = pe > ce ? pe : ce;
      prev->n = ms;
      prev->length = me - ms + 1;
    } else {
      kept[nk++] = cur;
    }
  }
  for (int64_t i = 0; i < nk; i++) {
    out[i].start = kept[i].n;
    out[i].end = kept[i].n + kept[i].length - 1;
    out[i].score = kept[i].score;
  }
  free(kept);
  return nk;
}

/* cm_pipeline.rs:817-837 split_windows */
int64_t ixo_split_windows(const ixo_window *in, int64_t n, int64_t cmw, ixo_window *out, int64_t cap) {
  int64_t no = 0;
  for (int64_t i = 0; i < n; i++) {
    ixo_window w = in[i];
    int64_t wlen = w.end - w.start + 1;
    if (wlen > 2 * cmw) {
      int64_t nws = w.start;
      int64_t nwe = w.start + 2 * cmw - 1 < w.end ? w.start + 2 * cmw - 1 : w.end;
      if (no < cap) { out[no].start = nws; out[no].end = nwe; out[no].score = w.score; }
      no++;
      while (nwe < w.end) {
        nws = nws + cmw < w.end ? nws + cmw : w.end;
        nwe = nwe + cmw < w.end ? nwe + cmw : w.end;
        if (no < cap) { out[no].start = nws; out[no].end = nwe; out[no].score = w.score; }
        no++;
      }
    } else {
      if (no < cap) out[no] = w;
      no++;
    }
  }
  return no;
}

/* cm_pipeline.rs:842-855 f1_msv_windows (pct_overlap 0.0) */
int64_t ixo_f1_msv_windows(const ixo_msv *f, const uint8_t *dsq, int64_t L, double f1,
                           int64_t cmw, ixo_window *out, int64_t cap) {
  int64_t nr = ixo_ssv_diagonals(f, dsq, L, f1, NULL, 0);
  if (nr == 0) return 0;
  ixo_rawwin *raws = (ixo_rawwin *)malloc(sizeof(ixo_rawwin) * (size_t)nr);
  ixo_ssv_diagonals(f, dsq, L, f1, raws, nr);
  ixo_window *merged = (ixo_window *)malloc(sizeof(ixo_window) * (size_t)nr);
  int64_t nm = ixo_extend_and_merge(f, raws, nr, L, 0.0f, merged);
  int64_t ns = ixo_split_windows(merged, nm, cmw, out, cap);
  free(raws); free(merged);
  return ns;
}

/* cm_pipeline.rs:396-452 msv_score / msv_score_reconfig / msv_score_tjb.
 * wlen_reconfig > 0 reconfigures tjb_b for that length. Returns 1 on the
 * eslERANGE overflow (score unset), 0 otherwise. */
int ixo_msv_score(const ixo_msv *f, const uint8_t *dsq, int64_t L, int wlen_reconfig, float *out) {
  uint8_t tjb_b = f->tjb_b;
  if (wlen_reconfig > 0)
    tjb_b = unbiased_byteify(f->scale_b, (float)log(3.0 / ((double)wlen_reconfig + 3.0)));
  int m = f->m;
  uint8_t tjbm = (uint8_t)(tjb_b + f->tbm_b);   /* wrapping_add */
  uint8_t *dp = (uint8_t *)calloc((size_t)m + 1, 1);
  uint8_t xj = 0;
  uint8_t xb = sat_sub_u8(f->base_b, tjbm);
  for (int64_t i = 1; i <= L; i++) {
    int x = dsq[i];
    uint8_t xe = 0, mpv = 0;
    for (int k = 1; k <= m; k++) {
      uint8_t sv = mpv > xb ? mpv : xb;
      sv = sat_add_u8(sv, f->bias_b);
      sv = sat_sub_u8(sv, f->rbv[k * KP + x]);
      if (sv > xe) xe = sv;
      mpv = dp[k];
      dp[k] = sv;
    }
    if (sat_add_u8(xe, f->bias_b) == 255) { free(dp); return 1; }
    xe = sat_sub_u8(xe, f->tec_b);
    if (xe > xj) xj = xe;
    xb = sat_sub_u8(f->base_b > xj ? f->base_b : xj, tjbm);
  }
  free(dp);
  *out = ((float)((int)xj - (int)tjb_b) - (float)f->base_b) / f->scale_b - 3.0f;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* F2 Viterbi filter                                                         */
/* ------------------------------------------------------------------------ */

/* p7_vitfilter.rs:57-69 wordify */
static int16_t wordify(float scale_w, float sc) {
  if (sc == -INFINITY) return NEGINF_W;
  float v = roundf(scale_w * sc);
  if (v >= 32767.0f) return 32767;
  if (v <= -32768.0f) return NEGINF_W;
  return (int16_t)v;
}
static int16_t adds16(int16_t a, int16_t b) {
  int r = (int)a + (int)b;
  if (r > 32767) r = 32767;
  if (r < -32768) r = -32768;
  return (int16_t)r;
}
static int16_t min16(int16_t a, int16_t b) { return a < b ? a : b; }
static int16_t max16(int16_t a, int16_t b) { return a > b ? a : b; }

/* Occupancy and local-entry Z (cm_pipeline.rs:1009-1026, p7_vitfilter.rs:177-190). */
static void occupancy(const ixo_p7 *p7, float *occ, float *z) {
  int m = p7->M;
  occ[0] = 0.0f;
  if (m >= 1) occ[1] = p7->trans[0 * 7 + 1] + p7->trans[0 * 7 + 0];
  for (int k = 2; k <= m; k++) {
    const float *t = p7->trans + (size_t)(k - 1) * 7;
    occ[k] = occ[k - 1] * (t[0] + t[1]) + (1.0f - occ[k - 1]) * t[5];
  }
  float zz = 0.0f;
  for (int k = 1; k <= m; k++) zz += occ[k] * (float)(m - k + 1);
  *z = zz;
}

/* p7_vitfilter.rs:121-278 build_vit_filter */
int ixo_build_vit(const ixo_p7 *p7, ixo_vit *f) {
  int m = p7->M;
  memset(f, 0, sizeof(*f));
  f->m = m;
  f->scale_w = (float)(500.0 / LN2);
  f->base_w = 12000;
  f->rwv = (int16_t *)malloc(sizeof(int16_t) * (size_t)(m + 1) * KP);
  for (size_t i = 0; i < (size_t)(m + 1) * KP; i++) f->rwv[i] = NEGINF_W;
  for (int k = 1; k <= m; k++) {
    float sc[KP];
    match_scores(p7, k, sc);
    for (int x = 0; x < KP; x++) f->rwv[k * KP + x] = wordify(f->scale_w, sc[x]);
  }
  int16_t **arrs[8] = {&f->tmm, &f->tim, &f->tdm, &f->vbm, &f->tmi, &f->tii, &f->tmd, &f->tdd};
  for (int a = 0; a < 8; a++) {
    *arrs[a] = (int16_t *)malloc(sizeof(int16_t) * (size_t)(m + 1));
    for (int k = 0; k <= m; k++) (*arrs[a])[k] = NEGINF_W;
  }
  for (int j = 1; j < m; j++) {
    const float *t = p7->trans + (size_t)j * 7;
    f->tmm[j] = min16(wordify(f->scale_w, (float)log((double)t[0])), 0);
    f->tim[j] = min16(wordify(f->scale_w, (float)log((double)t[3])), 0);
    f->tdm[j] = min16(wordify(f->scale_w, (float)log((double)t[5])), 0);
    f->tmi[j] = min16(wordify(f->scale_w, (float)log((double)t[1])), 0);
    f->tii[j] = min16(wordify(f->scale_w, (float)log((double)t[4])), -1);
    f->tmd[j] = min16(wordify(f->scale_w, (float)log((double)t[2])), 0);
    f->tdd[j] = min16(wordify(f->scale_w, (float)log((double)t[6])), 0);
  }
  float *occ = (float *)malloc(sizeof(float) * (size_t)(m + 1));
  float z;
  occupancy(p7, occ, &z);
  for (int k = 1; k <= m; k++) {
    float logv = (float)log((double)(occ[k] / z));
    f->vbm[k] = min16(wordify(f->scale_w, logv), 0);
  }
  free(occ);
  f->xw_e_move = wordify(f->scale_w, -(float)LN2);
  f->xw_e_loop = wordify(f->scale_w, -(float)LN2);

  int q = (m - 1) / 8 + 1;
  if (q < 2) q = 2;
  f->q = q;
  f->rw = (int16_t *)malloc(sizeof(int16_t) * (size_t)KP * q * 8);
  for (int x = 0; x < KP; x++)
    for (int qi = 0; qi < q; qi++)
      for (int zz = 0; zz < 8; zz++) {
        int k = qi + 1 + zz * q;
        f->rw[(x * q + qi) * 8 + zz] = k <= m ? f->rwv[k * KP + x] : NEGINF_W;
      }
  f->tw = (int16_t *)malloc(sizeof(int16_t) * (size_t)(8 * q) * 8);
  for (int i = 0; i < 8 * q * 8; i++) f->tw[i] = NEGINF_W;
  for (int qi = 0; qi < q; qi++)
    for (int zz = 0; zz < 8; zz++) {
      int k = qi + 1 + zz * q;
      int vb = qi * 7;
      int16_t bm = NEGINF_W, mm = NEGINF_W, im = NEGINF_W, dm = NEGINF_W;
      if (k <= m) { bm = f->vbm[k]; mm = f->tmm[k - 1]; im = f->tim[k - 1]; dm = f->tdm[k - 1]; }
      f->tw[(vb + 0) * 8 + zz] = bm;
      f->tw[(vb + 1) * 8 + zz] = mm;
      f->tw[(vb + 2) * 8 + zz] = im;
      f->tw[(vb + 3) * 8 + zz] = dm;
      int16_t md = NEGINF_W, mi = NEGINF_W, ii = NEGINF_W;
      if (k < m) { md = f->tmd[k]; mi = f->tmi[k]; ii = f->tii[k]; }
      f->tw[(vb + 4) * 8 + zz] = md;
      f->tw[(vb + 5) * 8 + zz] = mi;
      f->tw[(vb + 6) * 8 + zz] = ii;
      f->tw[(7 * q + qi) * 8 + zz] = k < m ? f->tdd[k] : NEGINF_W;
    }
  int16_t ddb = NEGINF_W;
  for (int k = 2; k < m - 1; k++) {
    int dd = (int)f->tdd[k] + (int)f->tdm[k + 1] - (int)f->vbm[k + 2];
    int mx = (int)ddb > dd ? (int)ddb : dd;
    ddb = (int16_t)mx;   /* `as i16` wraps */
  }
  f->ddbound_w = ddb;
  return 0;
}

void ixo_free_vit(ixo_vit *f) {
  free(f->rwv); free(f->tmm); free(f->tim); free(f->tdm); free(f->vbm);
  free(f->tmi); free(f->tii); free(f->tmd); free(f->tdd); free(f->tw); free(f->rw);
  memset(f, 0, sizeof(*f));
}

/* p7_vitfilter.rs:305-386 vit_score_scalar. Returns 1 on overflow. */
int ixo_vit_score_scalar(const ixo_vit *f, const uint8_t *dsq, int64_t L, float *out) {
  int m = f->m;
  float pmove = 3.0f / ((float)L + 3.0f);
  int16_t xw_move = wordify(f->scale_w, logf(pmove));
  size_t n = (size_t)m + 1;
  int16_t *buf = (int16_t *)malloc(sizeof(int16_t) * n * 6);
  for (size_t i = 0; i < n * 6; i++) buf[i] = NEGINF_W;
  int16_t *mp = buf, *ip = buf + n, *dp = buf + 2 * n, *mc = buf + 3 * n, *ic = buf + 4 * n, *dc = buf + 5 * n;
  int16_t xn = f->base_w, xb = adds16(xn, xw_move), xj = NEGINF_W, xc = NEGINF_W;
  for (int64_t i = 1; i <= L; i++) {
    int x = dsq[i];
    mc[0] = ic[0] = dc[0] = NEGINF_W;
    for (int k = 1; k <= m; k++) {
      int16_t sc = adds16(xb, f->vbm[k]);
      sc = max16(sc, adds16(mp[k - 1], f->tmm[k - 1]));
      sc = max16(sc, adds16(ip[k - 1], f->tim[k - 1]));
      sc = max16(sc, adds16(dp[k - 1], f->tdm[k - 1]));
      mc[k] = adds16(sc, f->rwv[k * KP + x]);
    }
    for (int k = 1; k <= m; k++) ic[k] = max16(adds16(mp[k], f->tmi[k]), adds16(ip[k], f->tii[k]));
    int16_t xe = NEGINF_W;
    for (int k = 1; k <= m; k++) if (mc[k] > xe) xe = mc[k];
    for (int k = 1; k <= m; k++) dc[k] = max16(adds16(mc[k - 1], f->tmd[k - 1]), adds16(dc[k - 1], f->tdd[k - 1]));
    if (xe >= 32767) { free(buf); return 1; }
    xn = adds16(xn, 0);
    xc = max16(xc, adds16(xe, f->xw_e_move));
    xj = max16(xj, adds16(xe, f->xw_e_loop));
    xb = max16(adds16(xj, xw_move), adds16(xn, xw_move));
    int16_t *t;
    t = mp; mp = mc; mc = t;
    t = ip; ip = ic; ic = t;
    t = dp; dp = dc; dc = t;
  }
  free(buf);
  if (xc > NEGINF_W) {
    float ret = (float)xc + (float)xw_move - (float)f->base_w;
    ret /= f->scale_w;
    *out = ret - 3.0f;
  } else {
    *out = -INFINITY;
  }
  return 0;
}

/* 8-lane i16 vector helpers for the striped restatement. */
typedef struct { int16_t v[8]; } v8;
static v8 v8_set(int16_t a) { v8 r; for (int i = 0; i < 8; i++) r.v[i] = a; return r; }
static v8 v8_adds(v8 a, v8 b) { v8 r; for (int i = 0; i < 8; i++) r.v[i] = adds16(a.v[i], b.v[i]); return r; }
static v8 v8_max(v8 a, v8 b) { v8 r; for (int i = 0; i < 8; i++) r.v[i] = max16(a.v[i], b.v[i]); return r; }
static v8 v8_shr_neginf(v8 a) { v8 r; r.v[0] = NEGINF_W; for (int i = 1; i < 8; i++) r.v[i] = a.v[i - 1]; return r; }
static v8 v8_ld(const int16_t *base, int v) { v8 r; memcpy(r.v, base + v * 8, 16); return r; }
static int16_t v8_hmax(v8 a) { int16_t m = a.v[0]; for (int i = 1; i < 8; i++) m = max16(m, a.v[i]); return m; }
static int v8_anygt(v8 a, v8 b) { for (int i = 0; i < 8; i++) if (a.v[i] > b.v[i]) return 1; return 0; }

/* p7_vitfilter.rs:410-525 vit_score_sse (Farrar-striped, lazy-F). */
int ixo_vit_score_striped(const ixo_vit *f, const uint8_t *dsq, int64_t L, float *out) {
  int q = f->q;
  float pmove = 3.0f / ((float)L + 3.0f);
  int16_t xw_move = wordify(f->scale_w, logf(pmove));
  v8 neg = v8_set(NEGINF_W);
  v8 *mmx = (v8 *)malloc(sizeof(v8) * (size_t)q * 3);
  v8 *imx = mmx + q, *dmx = mmx + 2 * q;
  for (int i = 0; i < 3 * q; i++) mmx[i] = neg;
  int16_t xn = f->base_w, xb = adds16(xn, xw_move), xj = NEGINF_W, xc = NEGINF_W;
  int dd_base = 7 * q;
  for (int64_t i = 1; i <= L; i++) {
    int x = dsq[i];
    int rbase = x * q;
    v8 dcv = neg, xev = neg, dmaxv = neg, xbv = v8_set(xb);
    v8 mpv = v8_shr_neginf(mmx[q - 1]);
    v8 dpv = v8_shr_neginf(dmx[q - 1]);
    v8 ipv = v8_shr_neginf(imx[q - 1]);
    for (int qi = 0; qi < q; qi++) {
      int vb = qi * 7;
      v8 sv = v8_adds(xbv, v8_ld(f->tw, vb));
      sv = v8_max(sv, v8_adds(mpv, v8_ld(f->tw, vb + 1)));
      sv = v8_max(sv, v8_adds(ipv, v8_ld(f->tw, vb + 2)));
      sv = v8_max(sv, v8_adds(dpv, v8_ld(f->tw, vb + 3)));
      sv = v8_adds(sv, v8_ld(f->rw, rbase + qi));
      xev = v8_max(xev, sv);
      mpv = mmx[qi]; dpv = dmx[qi]; ipv = imx[qi];
      mmx[qi] = sv;
      dmx[qi] = dcv;
      dcv = v8_adds(sv, v8_ld(f->tw, vb + 4));
      dmaxv = v8_max(dcv, dmaxv);
      v8 svi = v8_adds(mpv, v8_ld(f->tw, vb + 5));
      imx[qi] = v8_max(svi, v8_adds(ipv, v8_ld(f->tw, vb + 6)));
    }
    int16_t xe = v8_hmax(xev);
    if (xe >= 32767) { free(mmx); return 1; }
    xn = adds16(xn, 0);
    xc = max16(xc, adds16(xe, f->xw_e_move));
    xj = max16(xj, adds16(xe, f->xw_e_loop));
    xb = max16(adds16(xj, xw_move), adds16(xn, xw_move));
    int16_t dmax = v8_hmax(dmaxv);
    if ((int)dmax + (int)f->ddbound_w > (int)xb) {
      dcv = v8_shr_neginf(dcv);
      for (int qi = 0; qi < q; qi++) {
        v8 d = v8_max(dcv, dmx[qi]);
        dmx[qi] = d;
        dcv = v8_adds(d, v8_ld(f->tw, dd_base + qi));
      }
      for (;;) {
        dcv = v8_shr_neginf(dcv);
        int qi = 0;
        while (qi < q) {
          if (!v8_anygt(dcv, dmx[qi])) break;
          v8 d = v8_max(dcv, dmx[qi]);
          dmx[qi] = d;
          dcv = v8_adds(d, v8_ld(f->tw, dd_base + qi));
          qi++;
        }
        if (qi != q) break;
      }
    } else {
      dmx[0] = v8_shr_neginf(dcv);
    }
  }
  free(mmx);
  if (xc > NEGINF_W) {
    float ret = (float)xc + (float)xw_move - (float)f->base_w;
    ret /= f->scale_w;
    *out = ret - 3.0f;
  } else {
    *out = -INFINITY;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* F3 Forward filter                                                         */
/* ------------------------------------------------------------------------ */

/* cm_pipeline.rs:969-1113 build_forward_filter */
int ixo_build_fwd(const ixo_p7 *p7, ixo_fwd *f) {
  int m = p7->M;
  memset(f, 0, sizeof(*f));
  f->m = m;
  f->ftau = p7->lftau;
  f->flambda = p7->lflambda;
  size_t n = (size_t)m + 1;
  f->rfv = (float *)calloc(n * KP, sizeof(float));
  for (int k = 1; k <= m; k++) {
    float sc[KP];
    match_scores(p7, k, sc);
    for (int x = 0; x < KP; x++) f->rfv[k * KP + x] = expf(sc[x]);
  }
  float **arrs[8] = {&f->tbm, &f->amm, &f->aim, &f->adm, &f->tmi, &f->tii, &f->tmd, &f->tdd};
  for (int a = 0; a < 8; a++) *arrs[a] = (float *)calloc(n, sizeof(float));
  float *occ = (float *)malloc(sizeof(float) * n);
  float z;
  occupancy(p7, occ, &z);
  for (int k = 1; k <= m; k++) f->tbm[k] = (float)exp(log((double)(occ[k] / z)));
  free(occ);
  for (int k = 1; k <= m; k++) {
    const float *t = p7->trans + (size_t)(k - 1) * 7;
    f->amm[k] = t[0];
    f->aim[k] = t[3];
    f->adm[k] = t[5];
  }
  for (int k = 1; k < m; k++) {
    const float *t = p7->trans + (size_t)k * 7;
    f->tmi[k] = t[1];
    f->tii[k] = t[4];
    f->tmd[k] = t[2];
    f->tdd[k] = t[6];
  }
  int nq = m >= 1 ? ((m - 1) / 4 + 1) : 2;
  if (nq < 2) nq = 2;
  f->nq = nq;
  f->tmain = (float *)calloc((size_t)nq * 28, sizeof(float));
  f->tddv = (float *)calloc((size_t)nq * 4, sizeof(float));
  f->rfvv = (float *)calloc((size_t)KP * nq * 4, sizeof(float));
  for (int qi = 0; qi < nq; qi++)
    for (int zz = 0; zz < 4; zz++) {
      int p = qi + 1 + zz * nq;
      int ok = p <= m;
      f->tmain[qi * 28 + 0 * 4 + zz] = ok ? f->tbm[p] : 0.0f;
      f->tmain[qi * 28 + 1 * 4 + zz] = ok ? f->amm[p] : 0.0f;
      f->tmain[qi * 28 + 2 * 4 + zz] = ok ? f->aim[p] : 0.0f;
      f->tmain[qi * 28 + 3 * 4 + zz] = ok ? f->adm[p] : 0.0f;
      f->tmain[qi * 28 + 4 * 4 + zz] = ok ? f->tmd[p] : 0.0f;
      f->tmain[qi * 28 + 5 * 4 + zz] = ok ? f->tmi[p] : 0.0f;
      f->tmain[qi * 28 + 6 * 4 + zz] = ok ? f->tii[p] : 0.0f;
      f->tddv[qi * 4 + zz] = ok ? f->tdd[p] : 0.0f;
      for (int x = 0; x < KP; x++) f->rfvv[x * (nq * 4) + qi * 4 + zz] = ok ? f->rfv[p * KP + x] : 0.0f;
    }
  return 0;
}

void ixo_free_fwd(ixo_fwd *f) {
  free(f->rfv); free(f->tbm); free(f->amm); free(f->aim); free(f->adm);
  free(f->tmi); free(f->tii); free(f->tmd); free(f->tdd);
  free(f->tmain); free(f->tddv); free(f->rfvv);
  memset(f, 0, sizeof(*f));
}

typedef struct { float v[4]; } v4;
static v4 v4_set(float a) { v4 r = {{a, a, a, a}}; return r; }
static v4 v4_add(v4 a, v4 b) { v4 r; for (int i = 0; i < 4; i++) r.v[i] = a.v[i] + b.v[i]; return r; }
static v4 v4_mul(v4 a, v4 b) { v4 r; for (int i = 0; i < 4; i++) r.v[i] = a.v[i] * b.v[i]; return r; }
static v4 v4_ld(const float *p) { v4 r; memcpy(r.v, p, 16); return r; }
static void v4_st(float *p, v4 a) { memcpy(p, a.v, 16); }
static v4 v4_rsz(v4 a) { v4 r = {{0.0f, a.v[0], a.v[1], a.v[2]}}; return r; }
static int v4_anygt(v4 a, v4 b) { for (int i = 0; i < 4; i++) if (a.v[i] > b.v[i]) return 1; return 0; }

/* cm_pipeline.rs:1268-1418 forward_filter_score_striped (the default F3). */
float ixo_fwd_score_striped(const ixo_fwd *f, const uint8_t *dsq, int64_t L) {
  int m = f->m, q = f->nq;
  float nj = 1.0f;
  float pmove = (2.0f + nj) / ((float)L + 2.0f + nj);
  float ploop = 1.0f - pmove;
  float xf_e_move = 0.5f, xf_e_loop = 0.5f;
  float *buf = (float *)calloc((size_t)q * 12, sizeof(float));
  float *mm = buf, *dm = buf + 4 * q, *im = buf + 8 * q;
  const float *tp = f->tmain, *tdd = f->tddv;
  v4 zero = v4_set(0.0f);
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  for (int64_t i = 1; i <= L; i++) {
    int x = dsq[i];
    const float *rp = f->rfvv + (size_t)x * (q * 4);
    v4 dcv = zero, xev = zero, xbv = v4_set(xb);
    v4 mpv = v4_rsz(v4_ld(mm + (q - 1) * 4));
    v4 dpv = v4_rsz(v4_ld(dm + (q - 1) * 4));
    v4 ipv = v4_rsz(v4_ld(im + (q - 1) * 4));
    for (int qi = 0; qi < q; qi++) {
      int tb = qi * 28;
      v4 sv = v4_mul(xbv, v4_ld(tp + tb));
      sv = v4_add(sv, v4_mul(mpv, v4_ld(tp + tb + 4)));
      sv = v4_add(sv, v4_mul(ipv, v4_ld(tp + tb + 8)));
      sv = v4_add(sv, v4_mul(dpv, v4_ld(tp + tb + 12)));
      sv = v4_mul(sv, v4_ld(rp + qi * 4));
      xev = v4_add(xev, sv);
      mpv = v4_ld(mm + qi * 4);
      dpv = v4_ld(dm + qi * 4);
      ipv = v4_ld(im + qi * 4);
      v4_st(mm + qi * 4, sv);
      v4_st(dm + qi * 4, dcv);
      dcv = v4_mul(sv, v4_ld(tp + tb + 16));
      v4 svi = v4_mul(mpv, v4_ld(tp + tb + 20));
      v4_st(im + qi * 4, v4_add(svi, v4_mul(ipv, v4_ld(tp + tb + 24))));
    }
    dcv = v4_rsz(dcv);
    v4_st(dm, zero);
    for (int qi = 0; qi < q; qi++) {
      v4 d = v4_add(dcv, v4_ld(dm + qi * 4));
      v4_st(dm + qi * 4, d);
      dcv = v4_mul(d, v4_ld(tdd + qi * 4));
    }
    if (m < 100) {
      for (int j = 1; j < 4; j++) {
        dcv = v4_rsz(dcv);
        for (int qi = 0; qi < q; qi++) {
          v4 d = v4_add(dcv, v4_ld(dm + qi * 4));
          v4_st(dm + qi * 4, d);
          dcv = v4_mul(dcv, v4_ld(tdd + qi * 4));
        }
      }
    } else {
      for (int j = 1; j < 4; j++) {
        int cv = 0;
        dcv = v4_rsz(dcv);
        for (int qi = 0; qi < q; qi++) {
          v4 old = v4_ld(dm + qi * 4);
          v4 d = v4_add(dcv, old);
          if (v4_anygt(d, old)) cv = 1;
          v4_st(dm + qi * 4, d);
          dcv = v4_mul(dcv, v4_ld(tdd + qi * 4));
        }
        if (!cv) break;
      }
    }
    for (int qi = 0; qi < q; qi++) xev = v4_add(xev, v4_ld(dm + qi * 4));
    /* _MM_SHUFFLE(0,3,2,1) then _MM_SHUFFLE(1,0,3,2): lane0 = (x0+x1)+(x2+x3) */
    float s01 = xev.v[0] + xev.v[1];
    float s23 = xev.v[2] + xev.v[3];
    float xe = s01 + s23;
    xn = xn * ploop;
    xc = (xc * ploop) + (xe * xf_e_move);
    xj = (xj * ploop) + (xe * xf_e_loop);
    xb = (xj * pmove) + (xn * pmove);
    if (xe > 1.0e4f) {
      float inv = 1.0f / xe;
      xn *= inv; xc *= inv; xj *= inv; xb *= inv;
      for (int k = 0; k < 4 * q; k++) { mm[k] *= inv; dm[k] *= inv; im[k] *= inv; }
      totscale += (float)log((double)xe);
    }
  }
  free(buf);
  double sc = (double)totscale + log((double)xc * (double)pmove);
  return (float)sc;
}

/* p7_fwdback.rs:194-288 XFactors and :301-523 p7_forward: the full-matrix
 * odds-space Forward (rows 0..L kept, natural order k = 0..m), M and I per
 * cell from the previous row, the DD paths in the striped lane layout with
 * the M >= 100 lazy early exit, the 4-lane xE reduction, specials and sparse
 * rescaling; p7_fwdback.rs:529-532 forward_score. unihit selects the unihit
 * length model (:236-252). xmx rows hold E, N, J, B, C, SCALE. */
float ixo_p7_forward(const ixo_fwd *f, const uint8_t *dsq, int64_t L, int unihit, float *mmx, float *imx,
                     float *dmx, float *xmx, float *totscale_out) {
  const int m = f->m, q = f->nq;
  const float nj = unihit ? 0.0f : 1.0f;
  const float pmove = (2.0f + nj) / ((float)L + 2.0f + nj), ploop = 1.0f - pmove;
  const float e_move = unihit ? 1.0f : 0.5f, e_loop = unihit ? 0.0f : 0.5f;
  const size_t W = (size_t)m + 1;
  for (size_t k = 0; k < W; k++) mmx[k] = imx[k] = dmx[k] = 0.0f;
  xmx[0] = 0.0f; xmx[1] = 1.0f; xmx[2] = 0.0f; xmx[3] = pmove; xmx[4] = 0.0f; xmx[5] = 1.0f;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  float *dmp = (float *)malloc(sizeof(float) * (size_t)q * 4);
  for (int64_t i = 1; i <= L; i++) {
    const int x = dsq[i];
    const float *mp = mmx + (size_t)(i - 1) * W, *ip = imx + (size_t)(i - 1) * W, *dp = dmx + (size_t)(i - 1) * W;
    float *mc = mmx + (size_t)i * W, *ic = imx + (size_t)i * W, *dc = dmx + (size_t)i * W;
    mc[0] = ic[0] = dc[0] = 0.0f;
    for (int k = 1; k <= m; k++) {
      const float sv = xb * f->tbm[k] + mp[k - 1] * f->amm[k] + ip[k - 1] * f->aim[k] + dp[k - 1] * f->adm[k];
      mc[k] = sv * f->rfv[k * KP + x];
    }
    for (int k = 1; k <= m; k++) ic[k] = mp[k] * f->tmi[k] + ip[k] * f->tii[k];
    /* striped DD: lane z of quad qi is position k = z*q + qi + 1 */
#define AT(arr, k) (((k) >= 1 && (k) <= m) ? (arr)[k] : 0.0f)
    float dcv[4] = {0, 0, 0, 0};
    for (int qi = 0; qi < q; qi++) {
      for (int z = 0; z < 4; z++) dmp[qi * 4 + z] = dcv[z];
      for (int z = 0; z < 4; z++) {
        const int k = z * q + qi + 1;
        dcv[z] = AT(mc, k) * AT(f->tmd, k);
      }
    }
    { float r[4] = {0.0f, dcv[0], dcv[1], dcv[2]}; memcpy(dcv, r, sizeof r); }
    for (int z = 0; z < 4; z++) dmp[z] = 0.0f;
    for (int qi = 0; qi < q; qi++)
      for (int z = 0; z < 4; z++) {
        const int k = z * q + qi + 1;
        const float d = dcv[z] + dmp[qi * 4 + z];
        dmp[qi * 4 + z] = d;
        dcv[z] = d * AT(f->tdd, k);
      }
    for (int j = 1; j < 4; j++) {
      int changed = 0;
      { float r[4] = {0.0f, dcv[0], dcv[1], dcv[2]}; memcpy(dcv, r, sizeof r); }
      for (int qi = 0; qi < q; qi++)
        for (int z = 0; z < 4; z++) {
          const int k = z * q + qi + 1;
          const float old = dmp[qi * 4 + z], d = dcv[z] + old;
          if (d > old) changed = 1;
          dmp[qi * 4 + z] = d;
          dcv[z] = dcv[z] * AT(f->tdd, k);
        }
      if (m >= 100 && !changed) break;
    }
#undef AT
    for (int qi = 0; qi < q; qi++)
      for (int z = 0; z < 4; z++) {
        const int k = z * q + qi + 1;
        if (k <= m) dc[k] = dmp[qi * 4 + z];
      }
    float lane[4] = {0, 0, 0, 0};
    for (int r = 0; r < 4; r++)
      for (int qi = 0; qi < q; qi++) {
        const int k = r * q + qi + 1;
        if (k <= m) lane[r] += mc[k];
      }
    for (int r = 0; r < 4; r++)
      for (int qi = 0; qi < q; qi++) {
        const int k = r * q + qi + 1;
        if (k <= m) lane[r] += dc[k];
      }
    float xe = (lane[0] + lane[1]) + (lane[2] + lane[3]);
    xn *= ploop;
    xc = (xc * ploop) + (xe * e_move);
    xj = (xj * ploop) + (xe * e_loop);
    xb = (xj * pmove) + (xn * pmove);
    float scale = 1.0f;
    if (xe > 1.0e4f) {
      const float inv = 1.0f / xe;
      xn *= inv; xc *= inv; xj *= inv; xb *= inv;
      for (int k = 0; k <= m; k++) { mc[k] *= inv; dc[k] *= inv; ic[k] *= inv; }
      scale = xe;
      totscale += (float)log((double)xe);
      xe = 1.0f;
    }
    float *xr = xmx + (size_t)i * 6;
    xr[0] = xe; xr[1] = xn; xr[2] = xj; xr[3] = xb; xr[4] = xc; xr[5] = scale;
  }
  free(dmp);
  if (totscale_out) *totscale_out = totscale;
  return (float)((double)totscale + log((double)xc * (double)pmove));
}

/* cm_pipeline.rs:1420-1657 forward_filter_score, scalar order (the --strict
 * path: M/I per element, serial DD sweep, left-to-right xE sums). */
float ixo_fwd_score_scalar(const ixo_fwd *f, const uint8_t *dsq, int64_t L) {
  int m = f->m;
  size_t n = (size_t)m + 1;
  float pmove = 3.0f / ((float)L + 3.0f);
  float ploop = 1.0f - pmove;
  float *buf = (float *)calloc(n * 6, sizeof(float));
  float *mp = buf, *ip = buf + n, *dp = buf + 2 * n, *mc = buf + 3 * n, *ic = buf + 4 * n, *dc = buf + 5 * n;
  float xn = 1.0f, xj = 0.0f, xb = pmove, xc = 0.0f, totscale = 0.0f;
  for (int64_t i = 1; i <= L; i++) {
    int x = dsq[i];
    mc[0] = ic[0] = dc[0] = 0.0f;
    for (int k = 1; k <= m; k++) {
      float sv = xb * f->tbm[k] + mp[k - 1] * f->amm[k] + ip[k - 1] * f->aim[k] + dp[k - 1] * f->adm[k];
      mc[k] = sv * f->rfv[k * KP + x];
      ic[k] = mp[k] * f->tmi[k] + ip[k] * f->tii[k];
    }
    dc[1] = 0.0f;
    for (int k = 2; k <= m; k++) dc[k] = mc[k - 1] * f->tmd[k - 1] + dc[k - 1] * f->tdd[k - 1];
    float xe = 0.0f;
    for (int k = 1; k <= m; k++) xe += mc[k];
    for (int k = 1; k <= m; k++) xe += dc[k];
    xn = xn * ploop;
    xc = (xc * ploop) + (xe * 0.5f);
    xj = (xj * ploop) + (xe * 0.5f);
    xb = (xj * pmove) + (xn * pmove);
    if (xe > 1.0e4f) {
      float inv = 1.0f / xe;
      xn *= inv; xc *= inv; xj *= inv; xb *= inv;
      for (int k = 0; k <= m; k++) { mc[k] *= inv; dc[k] *= inv; ic[k] *= inv; }
      totscale += (float)log((double)xe);
    }
    float *t;
    t = mp; mp = mc; mc = t;
    t = ip; ip = ic; ic = t;
    t = dp; dp = dc; dc = t;
  }
  free(buf);
  double sc = (double)totscale + log((double)xc * (double)pmove);
  return (float)sc;
}

/* ------------------------------------------------------------------------ */
/* F3b composition-bias filter                                               */
/* ------------------------------------------------------------------------ */

/* cm_pipeline.rs:1683-1716 build_bias_filter */
void ixo_build_bias(const float compo[4], int m, ixo_bias *b) {
  float bg = 0.25f;
  float l1 = (float)m / 8.0f;
  for (int code = 0; code < KP; code++) {
    b->eo[code][0] = 1.0f;
    int set[4];
    int n = degen_set(code, set);
    if (n == 0) b->eo[code][1] = 1.0f;
    else {
      float num = 0.0f, den = 0.0f;
      for (int i = 0; i < n; i++) { num += compo[set[i]]; den += bg; }
      b->eo[code][1] = den > 0.0f ? num / den : 0.0f;
    }
  }
  b->t1[0] = 1.0f / (l1 + 1.0f);
  b->t1[1] = l1 / (l1 + 1.0f);
  b->t1[2] = 1.0f;
  b->pi[0] = 0.999f;
  b->pi[1] = 0.001f;
}

/* cm_pipeline.rs:1722-1774 bias_filter_score */
float ixo_bias_score(const ixo_bias *b, const uint8_t *dsq, int64_t L) {
  float p1 = (float)L / ((float)L + 1.0f);
  float t[2][3] = {{p1, 1.0f - p1, 1.0f}, {b->t1[0], b->t1[1], b->t1[2]}};
  if (L == 0) return 0.0f;
  float logsc = 0.0f;
  int x1 = dsq[1];
  float dp[2] = {b->eo[x1][0] * b->pi[0], b->eo[x1][1] * b->pi[1]};
  float mx = dp[0] > dp[1] ? dp[0] : dp[1];
  /* Rust f32::max: returns the larger; for equal values either. */
  dp[0] /= mx;
  dp[1] /= mx;
  logsc += (float)log((double)mx);
  for (int64_t i = 2; i <= L; i++) {
    int x = dsq[i];
    float nd[2];
    for (int k = 0; k < 2; k++) {
      float s = 0.0f;
      for (int mm = 0; mm < 2; mm++) s += dp[mm] * t[mm][k];
      nd[k] = s * b->eo[x][k];
    }
    mx = nd[0] > nd[1] ? nd[0] : nd[1];
    dp[0] = nd[0] / mx;
    dp[1] = nd[1] / mx;
    logsc += (float)log((double)mx);
  }
  float end = dp[0] * t[0][2] + dp[1] * t[1][2];
  logsc += (float)log((double)end);
  return logsc + (float)L * logf(p1) + logf(1.0f - p1);
}

/* ------------------------------------------------------------------------ */
/* Filter cascade                                                            */
/* ------------------------------------------------------------------------ */

/* cm_pipeline.rs:1863-2164 f3_filter_sequence. Windows are chunk-local and
 * 1-based. The deterministic 2*maxW tiling is used when do_msv is off. */
int64_t ixo_f3_filter_sequence(const ixo_msv *mf, const ixo_vit *vf, const ixo_fwd *ff,
                               const ixo_bias *bf, const ixo_p7 *p7,
                               const uint8_t *dsq, int64_t L, const ixo_f3opts *o,
                               ixo_window *out, int64_t cap, ixo_acct *acct) {
  int do_msv = o->opts & 1, do_fwd = (o->opts >> 1) & 1, do_fwdbias = (o->opts >> 2) & 1;
  int do_vit = (o->opts >> 3) & 1, do_vitbias = (o->opts >> 4) & 1, do_msvbias = (o->opts >> 5) & 1;
  int64_t nwin;
  ixo_window *wins;
  if (do_msv) {
    nwin = ixo_f1_msv_windows(mf, dsq, L, o->f1, o->cmw, NULL, 0);
    wins = (ixo_window *)malloc(sizeof(ixo_window) * (size_t)(nwin > 0 ? nwin : 1));
    ixo_f1_msv_windows(mf, dsq, L, o->f1, o->cmw, wins, nwin);
  } else {
    int64_t mw = o->maxw;
    nwin = 1;
    if (L > 2 * mw) {
      nwin += (L - 2 * mw) / (mw + 1);
      if ((L - 2 * mw) % (mw + 1) > 0) nwin++;
    }
    wins = (ixo_window *)malloc(sizeof(ixo_window) * (size_t)nwin);
    for (int64_t i = 0; i < nwin; i++) {
      int64_t ws = 1 + i * (mw + 1);
      int64_t we = ws + 2 * mw - 1 < L ? ws + 2 * mw - 1 : L;
      wins[i].start = ws; wins[i].end = we; wins[i].score = 0.0f;
    }
  }
  size_t nb = (size_t)(nwin > 0 ? nwin : 1);
  char *s_f1b = (char *)malloc(nb), *s_f2 = (char *)malloc(nb), *s_f2b = (char *)malloc(nb);
  char *s_f3 = (char *)calloc(nb, 1), *s_f3b = (char *)calloc(nb, 1);
  float *wb = (float *)malloc(sizeof(float) * nb);
  for (int64_t i = 0; i < nwin; i++) { s_f1b[i] = s_f2[i] = s_f2b[i] = 1; wb[i] = -999.0f; }
  uint8_t *sub = (uint8_t *)malloc((size_t)L + 2);
  for (int64_t i = 0; i < nwin; i++) {
    int64_t ws = wins[i].start, we = wins[i].end, wlen = we - ws + 1;
    sub[0] = 255;
    memcpy(sub + 1, dsq + ws, (size_t)wlen);
    sub[wlen + 1] = 255;
    double nullsc = ixo_bg_null_one(wlen);
    if (do_msvbias) {
      float mfsc;
      if (ixo_msv_score(mf, sub, wlen, (int)wlen, &mfsc) == 0) {
        float filtersc = ixo_bias_score(bf, sub, wlen);
        double w = ((double)mfsc - (double)filtersc) / LN2;
        if (ixo_gumbel_surv(w, p7->lmmu, p7->lmlambda) > o->f1b) { s_f1b[i] = 0; continue; }
      }
    }
    if (do_vit) {
      float vfsc;
      if (ixo_vit_score_scalar(vf, sub, wlen, &vfsc) == 0) {
        double w = ((double)vfsc - nullsc) / LN2;
        if (ixo_gumbel_surv(w, p7->lvmu, p7->lvlambda) > o->f2) { s_f2[i] = 0; s_f2b[i] = 0; continue; }
        if (do_vitbias) {
          float filtersc = ixo_bias_score(bf, sub, wlen);
          double wbb = ((double)vfsc - (double)filtersc) / LN2;
          if (ixo_gumbel_surv(wbb, p7->lvmu, p7->lvlambda) > o->f2b) { s_f2b[i] = 0; continue; }
        }
      }
    }
    float fwdsc = 0.0f;
    if (do_fwd) {
      fwdsc = ixo_fwd_score_striped(ff, sub, wlen);
      double w = ((double)fwdsc - nullsc) / LN2;
      double p = ixo_exp_surv(w, ff->ftau, ff->flambda);
      wb[i] = (float)w;
      if (p > o->f3) continue;
    }
    s_f3[i] = 1;
    if (do_fwd && do_fwdbias) {
      float filtersc = ixo_bias_score(bf, sub, wlen);
      double w = ((double)fwdsc - (double)filtersc) / LN2;
      double p = ixo_exp_surv(w, ff->ftau, ff->flambda);
      wb[i] = (float)w;
      if (p > o->f3b) continue;
    }
    s_f3b[i] = 1;
  }
  free(sub);
  /* survivors + merge (cm_pipeline.rs:2069-2114) */
  int64_t ns = 0;
  ixo_window *sv = (ixo_window *)malloc(sizeof(ixo_window) * nb);
  for (int64_t i = 0; i < nwin; i++)
    if (s_f3b[i]) { sv[ns].start = wins[i].start; sv[ns].end = wins[i].end; sv[ns].score = wb[i]; ns++; }
  int64_t no = 0;
  int64_t i = 0;
  while (i < ns) {
    int64_t i2 = i + 1;
    while (i2 < ns && sv[i].end + 1 >= sv[i2].start) {
      sv[i].end = sv[i2].end;
      if (sv[i2].score > sv[i].score) sv[i].score = sv[i2].score;
      i2++;
    }
    if (no < cap) out[no] = sv[i];
    no++;
    i = i2;
  }
  if (acct) {
    memset(acct, 0, sizeof(*acct));
    acct->n_past_msv = (uint64_t)nwin;
    for (int64_t k = 0; k < nwin; k++) {
      uint64_t wl = (uint64_t)(wins[k].end - wins[k].start + 1);
      acct->pos_past_msv += wl;
      if (s_f1b[k]) { acct->n_past_msvbias++; acct->pos_past_msvbias += wl; }
      if (s_f2[k]) { acct->n_past_vit++; acct->pos_past_vit += wl; }
      if (s_f2b[k]) { acct->n_past_vitbias++; acct->pos_past_vitbias += wl; }
      if (s_f3[k]) { acct->n_past_fwd++; acct->pos_past_fwd += wl; }
      if (s_f3b[k]) { acct->n_past_fwdbias++; acct->pos_past_fwdbias += wl; }
      if (k > 0) {
        int64_t ov = wins[k - 1].end - wins[k].start + 1;
        if (ov > 0) {
          acct->pos_past_msv -= (uint64_t)ov;
          if (s_f1b[k] && s_f1b[k - 1]) acct->pos_past_msvbias -= (uint64_t)ov;
          if (s_f2[k] && s_f2[k - 1]) acct->pos_past_vit -= (uint64_t)ov;
          if (s_f2b[k] && s_f2b[k - 1]) acct->pos_past_vitbias -= (uint64_t)ov;
          if (s_f3[k] && s_f3[k - 1]) acct->pos_past_fwd -= (uint64_t)ov;
          if (s_f3b[k] && s_f3b[k - 1]) acct->pos_past_fwdbias -= (uint64_t)ov;
        }
      }
    }
  }
  free(wins); free(s_f1b); free(s_f2); free(s_f2b); free(s_f3); free(s_f3b); free(wb); free(sv);
  (void)CM_MAX_RESIDUE_COUNT;
  return no;
}

/* ------------------------------------------------------------------------ */
/* Log-sum tables                                                            */
/* ------------------------------------------------------------------------ */

static int ilogsum_tbl[23000];
static float flogsum_tbl[23000];
static int tables_ready = 0;
static void init_tables(void) {
  if (tables_ready) return;
  /* cp9.rs:28-41: rint(1000*log2(1+2^(-i/1000))) */
  for (int i = 0; i < 23000; i++)
    ilogsum_tbl[i] = (int)rint(1000.0 * log2(1.0 + pow(2.0, -(double)i / 1000.0)));
  /* cp9.rs:5391-5401 with sreLOG2/sreEXP2 (cp9.rs:961-973) */
  for (int i = 0; i < 23000; i++) {
    double x = 1.0 + exp((-(double)i / 1000.0) * 0.69314718);
    flogsum_tbl[i] = (float)(x > 0.0 ? log(x) * 1.44269504 : -1e36);
  }
  tables_ready = 1;
}
int ixo_ilogsum_table(int i) { init_tables(); return ilogsum_tbl[i]; }
float ixo_flogsum_table(int i) { init_tables(); return flogsum_tbl[i]; }

/* cp9.rs:62-74 ilogsum_with */
int ixo_ilogsum(int s1, int s2) {
  init_tables();
  const int INFTY = 987654321;
  int mx = s1 > s2 ? s1 : s2;
  if (mx < -INFTY) mx = -INFTY;
  int mn = s1 < s2 ? s1 : s2;
  int64_t diff = (int64_t)mx - (int64_t)mn;
  if (mn <= -INFTY || diff >= 23000) return mx;
  return mx + ilogsum_tbl[diff];
}

/* cp9.rs:5405-5413 flogsum */
float ixo_flogsum(float s1, float s2) {
  init_tables();
  float mx = s1 > s2 ? s1 : s2;
  float mn = s1 < s2 ? s1 : s2;
  if (mn == -INFINITY || (mx - mn) >= 23.0f) return mx;
  return mx + flogsum_tbl[(size_t)((mx - mn) * 1000.0f)];
}
