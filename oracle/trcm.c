/*
 * trcm.c -- CPU oracle (test infrastructure only): restatement of infernox's
 * truncated HMM-banded scanners, TrCYK and FTrInside, with the truncated
 * greedy hit reporting, and of the model/band inputs they read (marginal
 * emission scores, the consensus emit map, marginal-mode validity).
 *
 * The scanners follow the reference's loop order state by state
 * (cm_dpsearch_trunc.rs:264-1029 and :1055-1794) so that the FLogsum
 * accumulation order of FTrInside is the reference's.
 */
#include "ixo_cm.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KP 18
#define K4 4
#define MAXC IXO_MAXCONNECT
static const float IMPOSSIBLE = -1e36f; /* cm_dpsearch_trunc.rs:33 */

float ixo_flogsum(float s1, float s2);
float ixo_null3_correction(const float *null, const float *comp, int len, float omega);

static double sre_log2(double x) { return x > 0.0 ? log(x) * 1.44269504 : -1e36; } /* cp9.rs:961 */
static int not_impossible(float x) { return x > IMPOSSIBLE + 1.0f; }               /* cm_dpsearch_trunc.rs:36 */

static const unsigned char DEGEN[KP][4] = { /* cm.rs:24-43 */
  {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {0, 0, 0, 0}, {1, 0, 1, 0}, {0, 1, 0, 1},
  {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 1, 1, 0}, {1, 0, 0, 1}, {1, 1, 0, 1}, {0, 1, 1, 1}, {1, 1, 1, 0},
  {1, 0, 1, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}};
static int is_residue(int x) { return x < K4 || (x > K4 && x < KP - 2); }

/* cp9.rs:4375-4388 f_expect_score: null-weighted mean over a degenerate code */
static float expect_score(int x, const float *sc, const float *p) {
  if (!is_residue(x)) return 0.0f;
  float r = 0.0f, den = 0.0f;
  for (int i = 0; i < K4; i++)
    if (DEGEN[x][i]) {
      r += sc[i] * p[i];
      den += p[i];
    }
  return r / den;
}
static void expect_vec(float *sc, const float *p) { /* cp9.rs:4416-4422 */
  for (int x = K4 + 1; x <= KP - 3; x++) sc[x] = expect_score(x, sc, p);
}

/* cp9.rs:4441-4527 fill_marginal_emissions (float part). */
void ixo_cm_marginals(const ixo_cm *cm, float *lmesc, float *rmesc) {
  for (int v = 0; v < cm->M; v++) {
    float *lm = lmesc + (size_t)v * KP, *rm = rmesc + (size_t)v * KP;
    for (int x = 0; x < KP; x++) lm[x] = rm[x] = 0.0f;
    lm[K4] = rm[K4] = lm[KP - 2] = rm[KP - 2] = lm[KP - 1] = rm[KP - 1] = IMPOSSIBLE;
    const int st = cm->sttype[v];
    const float *e = cm->e + (size_t)v * 16;
    if (st == IXO_MP_ST) {
      float l4[4] = {0, 0, 0, 0}, r4[4] = {0, 0, 0, 0};
      for (int x = 0; x < K4; x++)
        for (int y = 0; y < K4; y++) {
          l4[x] += e[x * 4 + y];
          r4[y] += e[x * 4 + y];
        }
      for (int x = 0; x < K4; x++) {
        lm[x] = (float)sre_log2((double)(l4[x] / cm->null[x]));
        rm[x] = (float)sre_log2((double)(r4[x] / cm->null[x]));
      }
      expect_vec(lm, cm->null);
      expect_vec(rm, cm->null);
    } else if (st == IXO_ML_ST || st == IXO_IL_ST || st == IXO_MR_ST || st == IXO_IR_ST) {
      float *emit = (st == IXO_ML_ST || st == IXO_IL_ST) ? lm : rm;
      for (int x = 0; x < K4; x++) emit[x] = (float)sre_log2((double)(e[x] / cm->null[x]));
      expect_vec(emit, cm->null);
    }
  }
}

/* cm_nohmm.rs:27-33 prob2score: floor(0.5 + 1000 log2(p / null)), -INFTY for p == 0 */
#define TR_INFTY 987654321
static int32_t tr_prob2score(float p, float null) {
  if (p == 0.0f) return -TR_INFTY;
  return (int32_t)floor(0.5 + 1000.0 * log2((double)p / (double)null));
}
/* cp9.rs:4391-4408 i_expect_score (esl_abc_IExpectScore): the null-weighted
 * mean of the canonical integer scores over a degenerate code, rounded half
 * away from zero; 0 for non-residue codes. */
static int32_t i_expect_score(int x, const int32_t *sc, const float *p) {
  if (!is_residue(x)) return 0;
  float r = 0.0f, den = 0.0f;
  for (int i = 0; i < K4; i++)
    if (DEGEN[x][i]) {
      r += (float)sc[i] * p[i];
      den += p[i];
    }
  r /= den;
  return r < 0.0f ? (int32_t)(r - 0.5f) : (int32_t)(r + 0.5f);
}
static void i_expect_vec(int32_t *sc, const float *p) { /* cp9.rs:4423-4429 */
  for (int x = K4 + 1; x <= KP - 3; x++) sc[x] = i_expect_score(x, sc, p);
}

/* cp9.rs:4441-4527 fill_marginal_emissions (integer part ilmesc/irmesc). */
void ixo_cm_imarginals(const ixo_cm *cm, int32_t *ilmesc, int32_t *irmesc) {
  for (int v = 0; v < cm->M; v++) {
    int32_t *lm = ilmesc + (size_t)v * KP, *rm = irmesc + (size_t)v * KP;
    for (int x = 0; x < KP; x++) lm[x] = rm[x] = 0;
    lm[K4] = rm[K4] = lm[KP - 2] = rm[KP - 2] = lm[KP - 1] = rm[KP - 1] = -TR_INFTY;
    const int st = cm->sttype[v];
    const float *e = cm->e + (size_t)v * 16;
    if (st == IXO_MP_ST) {
      float l4[4] = {0, 0, 0, 0}, r4[4] = {0, 0, 0, 0};
      for (int x = 0; x < K4; x++)
        for (int y = 0; y < K4; y++) {
          l4[x] += e[x * 4 + y];
          r4[y] += e[x * 4 + y];
        }
      for (int x = 0; x < K4; x++) {
        lm[x] = tr_prob2score(l4[x], cm->null[x]);
        rm[x] = tr_prob2score(r4[x], cm->null[x]);
      }
      i_expect_vec(lm, cm->null);
      i_expect_vec(rm, cm->null);
    } else if (st == IXO_ML_ST || st == IXO_IL_ST || st == IXO_MR_ST || st == IXO_IR_ST) {
      int32_t *emit = (st == IXO_ML_ST || st == IXO_IL_ST) ? lm : rm;
      for (int x = 0; x < K4; x++) emit[x] = tr_prob2score(e[x], cm->null[x]);
      i_expect_vec(emit, cm->null);
    }
  }
}

static int node_of_state(const ixo_cm *cm, int v) {
  int nd = 0;
  while (nd + 1 < cm->nodes && cm->nodemap[nd + 1] <= v) nd++;
  return nd;
}

/* cp9.rs:144-204 create_emit_map (lpos/rpos): a left-to-right walk of the
 * guide tree counting consensus columns; a node's lpos is the column count
 * after its left emission, its rpos one past the count when its subtree ends. */
void ixo_emit_map(const ixo_cm *cm, int32_t *lpos, int32_t *rpos) {
  const int n = cm->nodes;
  int *stk = (int *)malloc(sizeof(int) * (size_t)(2 * n + 4));
  int sp = 0, cpos = 0;
  for (int nd = 0; nd < n; nd++) lpos[nd] = rpos[nd] = -1;
  stk[sp++] = 0; /* (node << 1) | on_right */
  while (sp) {
    const int it = stk[--sp], nd = it >> 1, nt = cm->ndtype[nd];
    if (it & 1) {
      rpos[nd] = cpos + 1;
      if (nt == IXO_MATP_ND || nt == IXO_MATR_ND) cpos++;
      continue;
    }
    if (nt == IXO_MATP_ND || nt == IXO_MATL_ND) cpos++;
    lpos[nd] = cpos;
    stk[sp++] = (nd << 1) | 1;
    if (nt == IXO_BIF_ND) {
      const int v = cm->nodemap[nd];
      stk[sp++] = node_of_state(cm, cm->cnum[v]) << 1;
      stk[sp++] = node_of_state(cm, cm->cfirst[v]) << 1;
    } else if (nt != IXO_END_ND) {
      stk[sp++] = (nd + 1) << 1;
    }
  }
  free(stk);
}

static int fill_l(int pass) { return pass == IXO_PASS_5P_AND_3P_FORCE || pass == IXO_PASS_5P_AND_3P_ANY || pass == IXO_PASS_3P_ONLY_FORCE; }
static int fill_r(int pass) { return pass == IXO_PASS_5P_AND_3P_FORCE || pass == IXO_PASS_5P_AND_3P_ANY || pass == IXO_PASS_5P_ONLY_FORCE; }
static int fill_t(int pass) { return pass == IXO_PASS_5P_AND_3P_FORCE || pass == IXO_PASS_5P_AND_3P_ANY; }

static void state_span(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int v, int *lp, int *rp) {
  const int nd = node_of_state(cm, v), nt = cm->ndtype[nd];
  *lp = (nt == IXO_MATP_ND || nt == IXO_MATL_ND) ? lpos[nd] : lpos[nd] + 1;
  *rp = (nt == IXO_MATP_ND || nt == IXO_MATR_ND) ? rpos[nd] : rpos[nd] - 1;
}

/* cp9.rs:1925-1990 (the per-state rules and the root propagation). */
void ixo_tr_validity(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int pass_idx, int sp1, int sp2,
                     int ep1, int ep2, uint8_t *jv, uint8_t *lv, uint8_t *rv, uint8_t *tv) {
  const int al = fill_l(pass_idx), ar = fill_r(pass_idx), at = fill_t(pass_idx);
  for (int v = 0; v < cm->M; v++) {
    int lp, rp;
    state_span(cm, lpos, rpos, v, &lp, &rp);
    jv[v] = lp >= sp1 && rp <= ep1;
    lv[v] = al && lp >= sp1 && lp <= ep1 && rp > ep2;
    rv[v] = ar && rp <= ep1 && rp >= sp1 && lp < sp2;
    tv[v] = cm->sttype[v] == IXO_B_ST ? (at && lp < sp2 && rp > ep2) : 0;
  }
  for (int v = 0; v < cm->M; v++) {
    const int st = cm->sttype[v];
    const int pj = st == IXO_B_ST || st == IXO_MP_ST || st == IXO_ML_ST || st == IXO_IL_ST || st == IXO_MR_ST ||
                   st == IXO_IR_ST;
    const int pl = st == IXO_B_ST || st == IXO_MP_ST || st == IXO_ML_ST || st == IXO_IL_ST;
    const int pr = st == IXO_B_ST || st == IXO_MP_ST || st == IXO_MR_ST || st == IXO_IR_ST;
    if (pj && jv[v]) jv[0] = 1;
    if (pl && lv[v]) lv[0] = 1;
    if (pr && rv[v]) rv[0] = 1;
    if (st == IXO_B_ST && tv[v]) tv[0] = 1;
  }
}

/* cm_trunc.rs:1-38 state deltas */
static int sdelta(int st) { return st == IXO_MP_ST ? 2 : ((st >= IXO_ML_ST && st <= IXO_IR_ST) ? 1 : 0); }
static int sldelta(int st) { return (st == IXO_MP_ST || st == IXO_ML_ST || st == IXO_IL_ST) ? 1 : 0; }
static int srdelta(int st) { return (st == IXO_MP_ST || st == IXO_MR_ST || st == IXO_IR_ST) ? 1 : 0; }

/* cm_dpsearch_trunc.rs:2376-2422 allow_truncation (cm_hit_AllowTruncation) */
static int allow_truncation(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int pass_idx, int local,
                            int start, int stop, int i0, int j0, int mode, int b) {
  if (pass_idx == IXO_PASS_STD_ANY || pass_idx == IXO_PASS_5P_AND_3P_ANY) return 1;
  if (start == i0 && stop == j0) return 1;
  int lp, rp;
  state_span(cm, lpos, rpos, b, &lp, &rp);
  if (local) {
    if (mode == IXO_TRMODE_J) return 1;
    if (mode == IXO_TRMODE_L) return stop == j0;
    if (mode == IXO_TRMODE_R) return start == i0;
    return 0;
  }
  if (mode == IXO_TRMODE_J) return lp == 1 && rp == cm->clen;
  if (mode == IXO_TRMODE_L) return stop == j0 && lp == 1;
  if (mode == IXO_TRMODE_R) return start == i0 && rp == cm->clen;
  return 0;
}

int ixo_tr_allow(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int pass_idx, int local, int start,
                 int stop, int i0, int j0, int mode, int b) {
  return allow_truncation(cm, lpos, rpos, pass_idx, local, start, stop, i0, j0, mode, b);
}

typedef struct {
  int64_t *start, *roff, ncells;
} lay_t;
static void lay_make(const ixo_bands *b, int M, lay_t *L) {
  L->start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
  L->roff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(b->hd_off[M] + 1));
  int64_t nc = 0, r = 0;
  for (int v = 0; v < M; v++) {
    L->start[v] = r;
    const int nj = b->jmax[v] >= b->jmin[v] ? b->jmax[v] - b->jmin[v] + 1 : 0;
    for (int jp = 0; jp < nj; jp++) {
      L->roff[r++] = nc;
      const int w = b->hdmax[b->hd_off[v] + jp] - b->hdmin[b->hd_off[v] + jp] + 1;
      if (w > 0) nc += w;
    }
  }
  L->start[M] = r;
  L->ncells = nc;
}

#define HMN(v, jp) b->hdmin[b->hd_off[v] + (jp)]
#define HMX(v, jp) b->hdmax[b->hd_off[v] + (jp)]
#define RB(v, jp) Lo.roff[Lo.start[v] + (jp)]
#define MAX2(a, b) ((a) > (b) ? (a) : (b))
#define MIN2(a, b) ((a) < (b) ? (a) : (b))
/* the child combine: `if cand > cur { cur = cand }` (TrCYK) or FLogsum (FTrInside) */
#define COMB(dst, cand)                                   \
  do {                                                    \
    const float c_ = (cand);                              \
    if (inside) (dst) = ixo_flogsum((dst), c_);           \
    else if (c_ > (dst)) (dst) = c_;                      \
  } while (0)



static int cmp_hits(const void *pa, const void *pb) {
  const ixo_trhit *a = (const ixo_trhit *)pa, *b = (const ixo_trhit *)pb;
  if (a->score > b->score) return -1;
  if (a->score < b->score) return 1;
  if (a->i != b->i) return a->i < b->i ? -1 : 1;
  return 0;
}

/* remove_overlaps_greedy_tr (cm_dpsearch_trunc.rs:2470-2495): score
 * descending, then start ascending (a stable order: insertion sort keeps equal
 * keys in place); then every hit overlapping a kept one is dropped. tmp is
 * reordered; returns the number kept (beyond cap: counted only). */
int64_t ixo_tr_remove_overlaps(ixo_trhit *tmp, int64_t ntmp, ixo_trhit *hits, int64_t cap) {
  for (int64_t a = 1; a < ntmp; a++) {
    ixo_trhit h = tmp[a];
    int64_t k = a;
    while (k > 0 && cmp_hits(&tmp[k - 1], &h) > 0) {
      tmp[k] = tmp[k - 1];
      k--;
    }
    tmp[k] = h;
  }
  char *rem = (char *)calloc((size_t)(ntmp + 1), 1);
  int64_t nk = 0;
  for (int64_t a = 0; a < ntmp; a++) {
    if (rem[a]) continue;
    for (int64_t k = a + 1; k < ntmp; k++)
      if (!rem[k] && !(tmp[k].j < tmp[a].i) && !(tmp[a].j < tmp[k].i)) rem[k] = 1;
  }
  for (int64_t a = 0; a < ntmp; a++)
    if (!rem[a]) {
      if (nk < cap) hits[nk] = tmp[a];
      nk++;
    }
  free(rem);
  return nk;
}

int64_t ixo_trhb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const float *lmesc, const float *rmesc, const float *pty,
                      const ixo_bands *b, const ixo_trvalid *val, const uint8_t *dsq, int i0, int j0, int pass_idx,
                      int local, int inside, float cutoff, int do_null3, float env_cutoff, float *vsc, int *vmode,
                      int64_t *envi, int64_t *envj, ixo_trhit *hits, int64_t cap, float *planes_out) {
  const int M = cm->M, ww = j0 - i0 + 1;
  const int FL = fill_l(pass_idx), FR = fill_r(pass_idx), FT = fill_t(pass_idx);
  lay_t Lo;
  lay_make(b, M, &Lo);
  const int64_t nc = Lo.ncells;
  float *ja = (float *)malloc(sizeof(float) * (size_t)(4 * nc + 4));
  float *la = ja + nc, *ra = ja + 2 * nc, *ta = ja + 3 * nc;
  for (int64_t c = 0; c < 4 * nc; c++) ja[c] = IMPOSSIBLE;
  float *el = (float *)malloc(sizeof(float) * (size_t)(ww + 1));
  for (int d = 0; d <= ww; d++) el[d] = sc->el_selfsc * (float)d;
  const float ninf = IMPOSSIBLE;
#define JV(x) (val->jvalid[x] != 0)
#define LV(x) (val->lvalid[x] && FL)
#define RV(x) (val->rvalid[x] && FR)
#define TV(x) (val->tvalid[x] && FT)

  for (int v = M - 1; v >= 1; v--) {
    const int st = cm->sttype[v], sd = sdelta(st), sdl = sldelta(st), sdr = srdelta(st);
    const int cf = cm->cfirst[v], cn = cm->cnum[v];
    const float *tsc = sc->tsc + (size_t)v * MAXC;
    const float *esc = sc->oesc + (size_t)v * KP * KP;
    const float *lme = lmesc + (size_t)v * KP, *rme = rmesc + (size_t)v * KP;
    const int dj = JV(v), dl = LV(v), dr = RV(v), dt = TV(v);
    const int jmn = b->jmin[v], jmx = b->jmax[v];
    /* local-end re-initialisation per plane (:357-409) */
    if (not_impossible(sc->endsc[v])) {
      for (int j = jmn; j <= jmx; j++) {
        const int jp = j - jmn;
        const int64_t base = RB(v, jp);
        for (int d = HMN(v, jp); d <= HMX(v, jp); d++) {
          const int64_t c = base + (d - HMN(v, jp));
          if (dj && d >= sd) ja[c] = el[d - sd] + sc->endsc[v];
          if (dl && d >= sdl) la[c] = el[d - sdl] + sc->endsc[v];
          if (dr && d >= sdr) ra[c] = el[d - sdr] + sc->endsc[v];
        }
      }
    }
    if (st == IXO_E_ST) { /* :411-419 */
      for (int j = jmn; j <= jmx; j++) {
        const int64_t base = RB(v, j - jmn);
        if (dj) ja[base] = 0.0f;
        if (dl) la[base] = 0.0f;
        if (dr) ra[base] = 0.0f;
      }
    } else if (st == IXO_ML_ST || st == IXO_IL_ST) { /* :420-495 */
      for (int j = jmn; j <= jmx; j++) {
        const int jp = j - jmn, jsdr = j - sdr;
        const int64_t bv = RB(v, jp);
        for (int d = HMN(v, jp); d <= HMX(v, jp); d++) {
          const int i = j - d + 1;
          const int64_t c = bv + (d - HMN(v, jp));
          if (dj || dl) {
            for (int yo = 0; yo < cn; yo++) {
              const int y = cf + yo;
              if (jsdr < b->jmin[y] || jsdr > b->jmax[y]) continue;
              const int djy = JV(y), dly = LV(y);
              if (!(djy || dly)) continue;
              const int jpy = j - b->jmin[y] - sdr;
              if (d - sd >= HMN(y, jpy) && d - sd <= HMX(y, jpy)) {
                const int64_t cy = RB(y, jpy) + (d - sd - HMN(y, jpy));
                if (dj && djy) COMB(ja[c], ja[cy] + tsc[yo]);
                if (dl && dly) COMB(la[c], la[cy] + tsc[yo]);
              }
            }
            if (dj) {
              ja[c] += esc[dsq[i]];
              if (ja[c] < ninf) ja[c] = ninf;
            }
            if (dl) {
              const float e = esc[dsq[i]];
              la[c] = d >= 2 ? la[c] + e : e;
              if (la[c] < ninf) la[c] = ninf;
            }
          }
          if (dr) {
            float rsc = ra[c];
            for (int yo = 0; yo < cn; yo++) {
              const int y = cf + yo;
              if (jsdr < b->jmin[y] || jsdr > b->jmax[y]) continue;
              const int dry = RV(y), djy = JV(y);
              if (!((djy || dry) && y != v)) continue;
              const int jpy = j - b->jmin[y] - sdr;
              if (d >= HMN(y, jpy) && d <= HMX(y, jpy)) {
                const int64_t cy = RB(y, jpy) + (d - HMN(y, jpy));
                if (djy) COMB(rsc, ja[cy] + tsc[yo]);
                if (dry) COMB(rsc, ra[cy] + tsc[yo]);
              }
            }
            ra[c] = rsc;
          }
        }
      }
    } else if (st == IXO_MR_ST || st == IXO_IR_ST) { /* :496-592 */
      if (dj || dr) {
        for (int j = jmn; j <= jmx; j++) {
          const int jp = j - jmn, jsdr = j - sdr;
          const int64_t bv = RB(v, jp);
          for (int d = HMN(v, jp); d <= HMX(v, jp); d++) {
            const int64_t c = bv + (d - HMN(v, jp));
            for (int yo = 0; yo < cn; yo++) {
              const int y = cf + yo;
              if (jsdr < b->jmin[y] || jsdr > b->jmax[y]) continue;
              const int djy = JV(y), dry = RV(y);
              if (!(djy || dry)) continue;
              const int jpy = j - b->jmin[y] - sdr;
              if (d - sd >= HMN(y, jpy) && d - sd <= HMX(y, jpy)) {
                const int64_t cy = RB(y, jpy) + (d - sd - HMN(y, jpy));
                if (dj && djy) COMB(ja[c], ja[cy] + tsc[yo]);
                if (dr && dry) COMB(ra[c], ra[cy] + tsc[yo]);
              }
            }
            if (dj) {
              ja[c] += esc[dsq[j]];
              if (ja[c] < ninf) ja[c] = ninf;
            }
            if (dr) {
              const float e = esc[dsq[j]];
              ra[c] = d >= 2 ? ra[c] + e : e;
              if (ra[c] < ninf) ra[c] = ninf;
            }
          }
        }
      }
      if (dl) {
        for (int j = jmn; j <= jmx; j++) {
          const int jp = j - jmn;
          const int64_t bv = RB(v, jp);
          for (int d = HMN(v, jp); d <= HMX(v, jp); d++) {
            const int64_t c = bv + (d - HMN(v, jp));
            float lsc = la[c];
            for (int yo = 0; yo < cn; yo++) {
              const int y = cf + yo;
              if (y == v || j < b->jmin[y] || j > b->jmax[y]) continue;
              const int dly = LV(y), djy = JV(y);
              if (!(dly || djy)) continue;
              const int jpy = j - b->jmin[y];
              if (d >= HMN(y, jpy) && d <= HMX(y, jpy)) {
                const int64_t cy = RB(y, jpy) + (d - HMN(y, jpy));
                if (djy) COMB(lsc, ja[cy] + tsc[yo]);
                if (dly) COMB(lsc, la[cy] + tsc[yo]);
              }
            }
            la[c] = lsc;
          }
        }
      }
    } else if (st == IXO_MP_ST) { /* :593-708 */
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo;
        const int djy = JV(y), dly = LV(y), dry = RV(y);
        const float t = tsc[yo];
        if ((dj && djy) || (dr && (djy || dry))) {
          const int jn = MAX2(jmn, b->jmin[y] + sdr), jx = MIN2(jmx, b->jmax[y] + sdr);
          for (int j = jn; j <= jx; j++) {
            const int jpv = j - jmn, jpy = j - b->jmin[y] - sdr;
            if (dj && djy) {
              const int dn = MAX2(HMN(v, jpv), HMN(y, jpy) + sd), dx = MIN2(HMX(v, jpv), HMX(y, jpy) + sd);
              for (int d = dn; d <= dx; d++) COMB(ja[RB(v, jpv) + (d - HMN(v, jpv))], ja[RB(y, jpy) + (d - sd - HMN(y, jpy))] + t);
            }
            if (dr && (dry || djy)) {
              const int dn = MAX2(HMN(v, jpv), HMN(y, jpy) + sdr), dx = MIN2(HMX(v, jpv), HMX(y, jpy) + sdr);
              for (int d = dn; d <= dx; d++) {
                const int64_t c = RB(v, jpv) + (d - HMN(v, jpv)), cy = RB(y, jpy) + (d - sdr - HMN(y, jpy));
                if (djy) COMB(ra[c], ja[cy] + t);
                if (dry) COMB(ra[c], ra[cy] + t);
              }
            }
          }
        }
        if (dl && (dly || djy)) {
          const int jn = MAX2(jmn, b->jmin[y]), jx = MIN2(jmx, b->jmax[y]);
          for (int j = jn; j <= jx; j++) {
            const int jpv = j - jmn, jpy = j - b->jmin[y];
            const int dn = MAX2(HMN(v, jpv), HMN(y, jpy) + sdr), dx = MIN2(HMX(v, jpv), HMX(y, jpy) + sdr);
            for (int d = dn; d <= dx; d++) {
              const int64_t c = RB(v, jpv) + (d - HMN(v, jpv)), cy = RB(y, jpy) + (d - sdr - HMN(y, jpy));
              if (djy) COMB(la[c], ja[cy] + t);
              if (dly) COMB(la[c], la[cy] + t);
            }
          }
        }
      }
      for (int j = jmn; j <= jmx; j++) { /* emissions (:680-708) and the clamp (:709-722) */
        const int jp = j - jmn;
        const int64_t bv = RB(v, jp);
        for (int d = HMN(v, jp); d <= HMX(v, jp); d++) {
          const int i = j - d + 1;
          const int64_t c = bv + (d - HMN(v, jp));
          if (d >= 2) {
            if (dj) ja[c] += esc[dsq[i] * KP + dsq[j]];
            if (dl) la[c] += lme[dsq[i]];
            if (dr) ra[c] += rme[dsq[j]];
          } else {
            if (dj) ja[c] = ninf;
            if (dl) la[c] = lme[dsq[i]];
            if (dr) ra[c] = rme[dsq[j]];
          }
          if (dj && ja[c] < ninf) ja[c] = ninf;
          if (dl && la[c] < ninf) la[c] = ninf;
          if (dr && ra[c] < ninf) ra[c] = ninf;
        }
      }
    } else if (st != IXO_B_ST) { /* D, S (:723-772) */
      for (int yo = 0; yo < cn; yo++) {
        const int y = cf + yo;
        const int djy = JV(y), dly = LV(y), dry = RV(y);
        const float t = tsc[yo];
        if (!((dj && djy) || (dl && dly) || (dr && dry))) continue;
        const int jn = MAX2(jmn, b->jmin[y] + sdr), jx = MIN2(jmx, b->jmax[y] + sdr);
        for (int j = jn; j <= jx; j++) {
          const int jpv = j - jmn, jpy = j - b->jmin[y] - sdr;
          const int dn = MAX2(HMN(v, jpv), HMN(y, jpy) + sd), dx = MIN2(HMX(v, jpv), HMX(y, jpy) + sd);
          for (int d = dn; d <= dx; d++) {
            const int64_t c = RB(v, jpv) + (d - HMN(v, jpv)), cy = RB(y, jpy) + (d - sd - HMN(y, jpy));
            if (dj && djy) COMB(ja[c], ja[cy] + t);
            if (dl && dly) COMB(la[c], la[cy] + t);
            if (dr && dry) COMB(ra[c], ra[cy] + t);
            if (d == dn && dn == 0) {
              if (dl) la[c] = ninf;
              if (dr) ra[c] = ninf;
            }
          }
        }
      }
    } else { /* B (:773-879) */
      const int y = cf, z = cn;
      const int djy = JV(y), dly = LV(y), dry = RV(y), djz = JV(z), dlz = LV(z), drz = RV(z);
      const int jn = MAX2(jmn, b->jmin[z]), jx = MIN2(jmx, b->jmax[z]);
      for (int j = jn; j <= jx; j++) {
        const int jpv = j - jmn, jpy = j - b->jmin[y], jpz = j - b->jmin[z];
        const int hz0 = HMN(z, jpz);
        int kn = MAX2(j - b->jmax[y], hz0);
        kn = MAX2(kn, 0);
        const int kx = MIN2(jpy, HMX(z, jpz));
        for (int d = HMN(v, jpv); d <= HMX(v, jpv); d++) {
          const int64_t c = RB(v, jpv) + (d - HMN(v, jpv));
          for (int k = kn; k <= kx; k++) {
            const int jpymk = jpy - k;
            if (k >= d - HMX(y, jpymk) && k <= d - HMN(y, jpymk)) {
              const int64_t cy = RB(y, jpymk) + (d - HMN(y, jpymk) - k), cz = RB(z, jpz) + (k - hz0);
              if (dj && djy && djz) COMB(ja[c], ja[cy] + ja[cz]);
              if (dl && djy && dlz) COMB(la[c], ja[cy] + la[cz]);
              if (dr && dry && djz) COMB(ra[c], ra[cy] + ja[cz]);
              if (k != 0 && k != d && dt && dry && dlz) COMB(ta[c], ra[cy] + la[cz]);
            }
          }
        }
      }
      if (dl && (djy || dly)) { /* L from the left child alone (:847-862) */
        const int jn2 = MAX2(jmn, b->jmin[y]), jx2 = MIN2(jmx, b->jmax[y]);
        for (int j = jn2; j <= jx2; j++) {
          const int jpv = j - jmn, jpy = j - b->jmin[y];
          const int dn = MAX2(HMN(v, jpv), HMN(y, jpy)), dx = MIN2(HMX(v, jpv), HMX(y, jpy));
          for (int d = dn; d <= dx; d++) {
            const int64_t c = RB(v, jpv) + (d - HMN(v, jpv)), cy = RB(y, jpy) + (d - HMN(y, jpy));
            if (djy) COMB(la[c], ja[cy]);
            if (dly) COMB(la[c], la[cy]);
          }
        }
      }
      if (dr && (djz || drz)) { /* R from the right child alone (:863-878) */
        const int jn2 = MAX2(jmn, b->jmin[z]), jx2 = MIN2(jmx, b->jmax[z]);
        for (int j = jn2; j <= jx2; j++) {
          const int jpv = j - jmn, jpz = j - b->jmin[z];
          const int dn = MAX2(HMN(v, jpv), HMN(z, jpz)), dx = MIN2(HMX(v, jpv), HMX(z, jpz));
          for (int d = dn; d <= dx; d++) {
            const int64_t c = RB(v, jpv) + (d - HMN(v, jpv)), cz = RB(z, jpz) + (d - HMN(z, jpz));
            if (djz) COMB(ra[c], ja[cz]);
            if (drz) COMB(ra[c], ra[cz]);
          }
        }
      }
    }
  }

  /* ROOT: truncated begins by MAX with per-d best (:882-987), greedy reporting */
  const int dj0 = JV(0), dl0 = LV(0), dr0 = RV(0), dt0 = TV(0);
  int32_t *lpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  int32_t *rpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  ixo_emit_map(cm, lpos, rpos);
  double *act = (double *)calloc((size_t)(ww + 1) * 4, sizeof(double));
  if (do_null3)
    for (int jp = 1; jp <= ww; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      const int r = dsq[i0 + jp - 1];
      if (r < 4) act[jp * 4 + r] += 1.0;
    }
  float *bestsc = (float *)malloc(sizeof(float) * (size_t)(ww + 1));
  int *bestr = (int *)malloc(sizeof(int) * (size_t)(ww + 1));
  int *bestmode = (int *)malloc(sizeof(int) * (size_t)(ww + 1));
  ixo_trhit *tmp = NULL;
  int64_t ntmp = 0, tcap = 0;
  for (int j = b->jmin[0]; j <= b->jmax[0]; j++) {
    const int jpv = j - b->jmin[0];
    const int64_t bv = RB(0, jpv);
    for (int d = 0; d <= ww; d++) {
      bestr[d] = 0;
      bestsc[d] = IMPOSSIBLE;
      bestmode[d] = IXO_TRMODE_UNKNOWN;
    }
    for (int y = 1; y < M; y++) {
      const float tp = pty[y];
      if (!not_impossible(tp) || j < b->jmin[y] || j > b->jmax[y]) continue;
      const int jpy = j - b->jmin[y];
      const int dn = MAX2(HMN(0, jpv), HMN(y, jpy)), dx = MIN2(HMX(0, jpv), HMX(y, jpy));
      const int64_t cy0 = RB(y, jpy);
      struct { int on; float *pl; int mode; } pass[4] = {
        {dj0 && JV(y), ja, IXO_TRMODE_J}, {dl0 && LV(y), la, IXO_TRMODE_L},
        {dr0 && RV(y), ra, IXO_TRMODE_R}, {dt0 && TV(y) && cm->sttype[y] == IXO_B_ST, ta, IXO_TRMODE_T}};
      for (int p = 0; p < 4; p++) {
        if (!pass[p].on) continue;
        for (int d = dn; d <= dx; d++) {
          const float s = pass[p].pl[cy0 + (d - HMN(y, jpy))] + tp;
          float *cell = &pass[p].pl[bv + (d - HMN(0, jpv))];
          if (s > *cell) {
            *cell = s;
            if (s > bestsc[d]) {
              bestsc[d] = s;
              bestmode[d] = pass[p].mode;
              bestr[d] = y;
            }
          }
        }
      }
    }
    /* report_hits_greedily (:2322-2372) */
    const int dmin = HMN(0, jpv), dmax = HMX(0, jpv);
    if (dmin > dmax) continue;
    float maxrep = IMPOSSIBLE;
    for (int d = MAX2(dmin, 1); d <= dmax; d++) {
      const int i = j - d + 1;
      const int mode = bestmode[d];
      float hs = bestsc[d], bias = 0.0f;
      if (!(hs > maxrep && hs >= cutoff && not_impossible(hs))) continue;
      int rep = allow_truncation(cm, lpos, rpos, pass_idx, local, i, j, i0, j0, mode, bestr[d]);
      if (rep && do_null3) {
        float comp[4];
        const int ip = i - i0 + 1, jj = j - i0 + 1;
        for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
        float s = 0.0f;
        for (int a = 0; a < 4; a++) s += comp[a];
        if (s > 0.0f)
          for (int a = 0; a < 4; a++) comp[a] /= s;
        else
          for (int a = 0; a < 4; a++) comp[a] = 1.0f / 4.0f;
        const float corr = ixo_null3_correction(cm->null, comp, d, (float)cm->n3_omega);
        hs -= corr;
        bias = corr;
        rep = hs > maxrep && hs >= cutoff;
      }
      if (rep) {
        if (ntmp == tcap) {
          tcap = tcap ? 2 * tcap : 64;
          tmp = (ixo_trhit *)realloc(tmp, sizeof(ixo_trhit) * (size_t)tcap);
        }
        tmp[ntmp].i = i;
        tmp[ntmp].j = j;
        tmp[ntmp].score = hs;
        tmp[ntmp].bias = bias;
        tmp[ntmp].mode = mode;
        tmp[ntmp].root = bestr[d];
        ntmp++;
        maxrep = hs;
      }
    }
  }
  /* vsc_root over the four planes and the envelope (:989-1022) */
  float vr = IMPOSSIBLE;
  int vm = IXO_TRMODE_UNKNOWN;
  int64_t ei = (int64_t)j0 + 1, ej = (int64_t)i0 - 1;
  for (int j = b->jmin[0]; j <= b->jmax[0]; j++) {
    const int jpv = j - b->jmin[0];
    for (int d = HMN(0, jpv); d <= HMX(0, jpv); d++) {
      const int64_t c = RB(0, jpv) + (d - HMN(0, jpv));
      if (dj0 && ja[c] > vr) { vr = ja[c]; vm = IXO_TRMODE_J; }
      if (dl0 && la[c] > vr) { vr = la[c]; vm = IXO_TRMODE_L; }
      if (dr0 && ra[c] > vr) { vr = ra[c]; vm = IXO_TRMODE_R; }
      if (dt0 && ta[c] > vr) { vr = ta[c]; vm = IXO_TRMODE_T; }
    }
    for (int d = HMN(0, jpv); d <= HMX(0, jpv); d++) {
      const int64_t c = RB(0, jpv) + (d - HMN(0, jpv));
      if ((dj0 && ja[c] >= env_cutoff) || (dl0 && la[c] >= env_cutoff) || (dr0 && ra[c] >= env_cutoff) ||
          (dt0 && ta[c] >= env_cutoff)) {
        const int i = j - d + 1;
        if (i < ei) ei = i;
        if (j > ej) ej = j;
      }
    }
  }
  *vsc = vr;
  *vmode = vm;
  *envi = ei == (int64_t)j0 + 1 ? -1 : ei;
  *envj = ej == (int64_t)i0 - 1 ? -1 : ej;
  const int64_t nk = ixo_tr_remove_overlaps(tmp, ntmp, hits, cap);
  if (planes_out) memcpy(planes_out, ja, sizeof(float) * (size_t)(4 * nc));
  free(tmp);
  free(bestsc);
  free(bestr);
  free(bestmode);
  free(act);
  free(lpos);
  free(rpos);
  free(el);
  free(ja);
  free(Lo.start);
  free(Lo.roff);
  return nk;
}
