/*
 * cm.c -- CPU oracle (test infrastructure only): restatement of the
 * infernox covariance-model score configuration and of the HMM-banded CYK
 * and Inside scanners (F6/F7) with their greedy hit reporting.
 *
 * Each function names the reference file:line it follows. The scanners follow
 * the reference's scalar loops, which the reference documents as
 * bit-identical to its AVX2 paths (cp9.rs:4565-4646, 5421-5458).
 */
#include "ixo_cm.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KP 18
#define K4 4
#define MAXC IXO_MAXCONNECT
static const float IMPOSSIBLE = -1e36f; /* cp9.rs:4189 */

/* cp9.rs:961-967 sreLOG2 */
static double sre_log2(double x) { return x > 0.0 ? log(x) * 1.44269504 : -1e36; }

/* cm.rs:24-43 RNA_DEGEN (sym order "ACGU-RYMKSWHBVDN*~") */
static const unsigned char DEGEN[KP][4] = {
  {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {0, 0, 0, 0}, {1, 0, 1, 0}, {0, 1, 0, 1},
  {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 1, 1, 0}, {1, 0, 0, 1}, {1, 1, 0, 1}, {0, 1, 1, 1}, {1, 1, 1, 0},
  {1, 0, 1, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}};
static int ndegen(int x) { return DEGEN[x][0] + DEGEN[x][1] + DEGEN[x][2] + DEGEN[x][3]; }
static int is_residue(int x) { return x < K4 || (x > K4 && x < KP - 2); }       /* cm.rs:68-71 */
static float fcount_frac(int x, int l) { return DEGEN[x][l] ? 1.0f / (float)ndegen(x) : 0.0f; } /* cm.rs:77-83 */
static float favg_score(int x, const float *sc) {                                /* cm.rs:88-99 */
  if (!is_residue(x)) return 0.0f;
  float r = 0.0f;
  for (int i = 0; i < K4; i++)
    if (DEGEN[x][i]) r += sc[i];
  return r / (float)ndegen(x);
}

/* cp9.rs:451-461 esl_vec_fsum (Kahan) */
static float vec_fsum(const float *v, int n) {
  float sum = 0.0f, c = 0.0f;
  for (int i = 0; i < n; i++) {
    float y = v[i] - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

static int is_exit_node(const ixo_cm *cm, int nd) { /* cp9.rs:4231-4236 */
  int t = cm->ndtype[nd];
  return (t == IXO_MATP_ND || t == IXO_MATL_ND || t == IXO_MATR_ND || t == IXO_BEGL_ND || t == IXO_BEGR_ND) &&
         cm->ndtype[nd + 1] != IXO_END_ND;
}

static void int_scores(const ixo_cm *cm, const float *t, const float *begin, const float *end, float el,
                       ixo_cmisc *is);

/* cp9.rs:4194-4256 cm_config_local (local != 0) + cp9.rs:4263-4386
 * cm_logoddsify + the EL bound of cp9.rs:4538-4545 cm_configure_scores; with
 * local == 0 the global configuration of cm_nohmm.rs:264
 * cm_configure_scores_global (no local begins or ends). With is != NULL also
 * the integer scores of cm_nohmm.rs:279 build_integer_scores. */
int ixo_cm_configure_full(const ixo_cm *cm, int local, ixo_cmsc *s, ixo_cmisc *is) {
  const int M = cm->M;
  memset(s, 0, sizeof(*s));
  s->M = M;
  s->tsc = (float *)malloc(sizeof(float) * (size_t)M * MAXC);
  s->oesc = (float *)malloc(sizeof(float) * (size_t)M * KP * KP);
  s->beginsc = (float *)malloc(sizeof(float) * (size_t)M);
  s->endsc = (float *)malloc(sizeof(float) * (size_t)M);
  float *t = (float *)malloc(sizeof(float) * (size_t)M * MAXC);
  float *begin = (float *)calloc((size_t)M, sizeof(float));
  float *end = (float *)calloc((size_t)M, sizeof(float));
  memcpy(t, cm->t, sizeof(float) * (size_t)M * MAXC);
  if (local) {
    /* local begins */
    int nstarts = 0;
    for (int nd = 2; nd < cm->nodes; nd++) {
      int ty = cm->ndtype[nd];
      if (ty == IXO_MATP_ND || ty == IXO_MATL_ND || ty == IXO_MATR_ND || ty == IXO_BIF_ND) nstarts++;
    }
    begin[cm->nodemap[1]] = 1.0f - cm->pbegin;
    float pb = cm->pbegin / (float)nstarts;
    for (int nd = 2; nd < cm->nodes; nd++) {
      int ty = cm->ndtype[nd];
      if (ty == IXO_MATP_ND || ty == IXO_MATL_ND || ty == IXO_MATR_ND || ty == IXO_BIF_ND) begin[cm->nodemap[nd]] = pb;
    }
    for (int y = 0; y < cm->cnum[0]; y++) t[y] = 0.0f;
    /* local ends */
    int nexits = 0;
    for (int nd = 1; nd < cm->nodes; nd++)
      if (is_exit_node(cm, nd)) nexits++;
    for (int nd = 1; nd < cm->nodes; nd++)
      if (is_exit_node(cm, nd)) {
        int v = cm->nodemap[nd];
        end[v] = cm->pend / (float)nexits;
        int cn = cm->cnum[v];
        float denom = vec_fsum(t + (size_t)v * MAXC, cn);
        denom += end[v];
        float inv = 1.0f / denom;
        for (int x = 0; x < cn; x++) t[(size_t)v * MAXC + x] *= inv;
      }
  }
  s->local_begin = local ? 1 : 0;
  /* el_selfsc adjustment (cp9.rs:4541-4543) */
  s->el_selfsc = cm->el_selfsc;
  if (s->el_selfsc * (float)cm->W < IMPOSSIBLE) s->el_selfsc = IMPOSSIBLE / ((float)cm->W + 1.0f);
  /* log-odds */
  for (int v = 0; v < M; v++) {
    int st = cm->sttype[v];
    float *ts = s->tsc + (size_t)v * MAXC;
    for (int x = 0; x < MAXC; x++) ts[x] = 0.0f;
    if (st != IXO_B_ST && st != IXO_E_ST)
      for (int x = 0; x < cm->cnum[v]; x++) ts[x] = (float)sre_log2((double)t[(size_t)v * MAXC + x]);
    float *oe = s->oesc + (size_t)v * KP * KP;
    for (int i = 0; i < KP * KP; i++) oe[i] = 0.0f;
    const float *e = cm->e + (size_t)v * 16;
    if (st == IXO_MP_ST) {
      float esc[16];
      for (int x = 0; x < K4; x++)
        for (int y = 0; y < K4; y++)
          esc[x * 4 + y] = (float)sre_log2((double)(e[x * 4 + y] / (cm->null[x] * cm->null[y])));
      for (int i = 0; i < KP * KP; i++) oe[i] = IMPOSSIBLE;
      for (int a = 0; a < K4; a++)
        for (int b = 0; b < K4; b++) oe[a * KP + b] = esc[a * 4 + b];
      for (int a = K4 + 1; a < KP - 1; a++)
        for (int b = 0; b < K4; b++) {
          float sc = 0.0f;
          for (int l = 0; l < K4; l++) sc += esc[l * 4 + b] * fcount_frac(a, l);
          oe[a * KP + b] = sc;
        }
      for (int a = 0; a < K4; a++)
        for (int b = K4 + 1; b < KP - 1; b++) {
          float sc = 0.0f;
          for (int r = 0; r < K4; r++) sc += esc[a * 4 + r] * fcount_frac(b, r);
          oe[a * KP + b] = sc;
        }
      for (int a = K4 + 1; a < KP - 1; a++)
        for (int b = K4 + 1; b < KP - 1; b++) {
          float sc = 0.0f;
          for (int l = 0; l < K4; l++)
            for (int r = 0; r < K4; r++) sc += esc[l * 4 + r] * fcount_frac(a, l) * fcount_frac(b, r);
          oe[a * KP + b] = sc;
        }
    } else if (st == IXO_ML_ST || st == IXO_MR_ST || st == IXO_IL_ST || st == IXO_IR_ST) {
      float esc[4];
      for (int x = 0; x < K4; x++) esc[x] = (float)sre_log2((double)(e[x] / cm->null[x]));
      for (int a = 0; a < KP; a++) oe[a] = IMPOSSIBLE;
      for (int a = 0; a < K4; a++) oe[a] = esc[a];
      for (int a = K4 + 1; a < KP - 1; a++) oe[a] = favg_score(a, esc);
    }
    s->beginsc[v] = (float)sre_log2((double)begin[v]);
    s->endsc[v] = (float)sre_log2((double)end[v]);
  }
  if (is) int_scores(cm, t, begin, end, s->el_selfsc, is);
  free(t);
  free(begin);
  free(end);
  return 0;
}

int ixo_cm_configure(const ixo_cm *cm, ixo_cmsc *s) { return ixo_cm_configure_full(cm, 1, s, NULL); }

#define INFTY 987654321 /* cp9.rs:21 */
/* cm_nohmm.rs:27-33 prob2score: floor(0.5 + 1000 log2(p / null)), -INFTY for p == 0 */
static int32_t prob2score(float p, float null) {
  if (p == 0.0f) return -INFTY;
  return (int32_t)floor(0.5 + 1000.0 * log2((double)p / (double)null));
}
/* cm.rs:104-120 abc_iavg_score: mean over a degenerate code, rounded half away */
static int32_t iavg_score(int x, const int32_t *sc) {
  if (!is_residue(x)) return 0;
  float r = 0.0f;
  for (int i = 0; i < K4; i++)
    if (DEGEN[x][i]) r += (float)sc[i];
  r /= (float)ndegen(x);
  return r < 0.0f ? (int32_t)(r - 0.5f) : (int32_t)(r + 0.5f);
}

/* cm_nohmm.rs:279-368 build_integer_scores */
static void int_scores(const ixo_cm *cm, const float *t, const float *begin, const float *end, float el,
                       ixo_cmisc *is) {
  const int M = cm->M;
  is->itsc = (int32_t *)calloc((size_t)M * MAXC, sizeof(int32_t));
  is->ioesc = (int32_t *)malloc(sizeof(int32_t) * (size_t)M * KP * KP);
  is->ibeginsc = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
  is->iendsc = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
  for (int v = 0; v < M; v++) {
    const int st = cm->sttype[v];
    int32_t *oe = is->ioesc + (size_t)v * KP * KP;
    for (int i = 0; i < KP * KP; i++) oe[i] = 0;
    if (st != IXO_B_ST && st != IXO_E_ST)
      for (int x = 0; x < cm->cnum[v]; x++) is->itsc[(size_t)v * MAXC + x] = prob2score(t[(size_t)v * MAXC + x], 1.0f);
    const float *e = cm->e + (size_t)v * 16;
    if (st == IXO_MP_ST) {
      int32_t ie[16];
      for (int a = 0; a < K4; a++)
        for (int b = 0; b < K4; b++) ie[a * 4 + b] = prob2score(e[a * 4 + b], cm->null[a] * cm->null[b]);
      for (int i = 0; i < KP * KP; i++) oe[i] = -INFTY;
      for (int a = 0; a < K4; a++)
        for (int b = 0; b < K4; b++) oe[a * KP + b] = ie[a * 4 + b];
      for (int a = K4 + 1; a < KP - 1; a++)
        for (int b = 0; b < K4; b++) {
          float sc = 0.0f;
          for (int l = 0; l < K4; l++) sc += (float)ie[l * 4 + b] * fcount_frac(a, l);
          oe[a * KP + b] = (int32_t)sc;
        }
      for (int a = 0; a < K4; a++)
        for (int b = K4 + 1; b < KP - 1; b++) {
          float sc = 0.0f;
          for (int r = 0; r < K4; r++) sc += (float)ie[a * 4 + r] * fcount_frac(b, r);
          oe[a * KP + b] = (int32_t)sc;
        }
      for (int a = K4 + 1; a < KP - 1; a++)
        for (int b = K4 + 1; b < KP - 1; b++) {
          float sc = 0.0f;
          for (int l = 0; l < K4; l++)
            for (int r = 0; r < K4; r++) sc += (float)ie[l * 4 + r] * fcount_frac(a, l) * fcount_frac(b, r);
          oe[a * KP + b] = (int32_t)sc;
        }
    } else if (st == IXO_ML_ST || st == IXO_MR_ST || st == IXO_IL_ST || st == IXO_IR_ST) {
      int32_t ie[4];
      for (int a = 0; a < K4; a++) ie[a] = prob2score(e[a], cm->null[a]);
      for (int a = 0; a < KP; a++) oe[a] = -INFTY;
      for (int a = 0; a < K4; a++) oe[a] = ie[a];
      for (int a = K4 + 1; a < KP - 1; a++) oe[a] = iavg_score(a, ie);
    }
    is->ibeginsc[v] = prob2score(begin[v], 1.0f);
    is->iendsc[v] = prob2score(end[v], 1.0f);
  }
  is->iel_selfsc = prob2score(powf(2.0f, el), 1.0f);
}

void ixo_cm_free_iscores(ixo_cmisc *is) {
  free(is->itsc); free(is->ioesc); free(is->ibeginsc); free(is->iendsc);
  memset(is, 0, sizeof(*is));
}

void ixo_cm_free_scores(ixo_cmsc *s) {
  free(s->tsc); free(s->oesc); free(s->beginsc); free(s->endsc);
  memset(s, 0, sizeof(*s));
}

/* cp9.rs:2844-2856 state_delta and the right delta (0/1). */
static int state_delta(int st) {
  switch (st) {
    case IXO_MP_ST: return 2;
    case IXO_ML_ST: case IXO_MR_ST: case IXO_IL_ST: case IXO_IR_ST: return 1;
    default: return 0;
  }
}
static int state_left_delta(int st) { return (st == IXO_MP_ST || st == IXO_ML_ST || st == IXO_IL_ST) ? 1 : 0; }
static int state_right_delta(int st) { return (st == IXO_MP_ST || st == IXO_MR_ST || st == IXO_IR_ST) ? 1 : 0; }

/* cp9.rs:3815-3850 ij2d_bands: d bands from i and j bands. */
void ixo_ij2d_bands(const int8_t *sttype, int M, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                    const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax) {
  int32_t off = 0;
  for (int v = 0; v < M; v++) {
    hd_off[v] = off;
    int jbw = jmax[v] - jmin[v];
    int n = jbw >= 0 ? jbw + 1 : 0;
    int st = sttype[v];
    for (int jp = 0; jp < n; jp++) {
      if (st == IXO_E_ST) {
        hdmin[off + jp] = 0;
        hdmax[off + jp] = 0;
      } else {
        int sd = state_delta(st);
        int sl = state_left_delta(st), sr = state_right_delta(st);
        int dn = do_trunc ? (sl > sr ? sl : sr) : sd;
        int j = jp + jmin[v];
        int hdn = j - imax[v] + 1, hdx = j - imin[v] + 1;
        if (hdx < dn) {
          hdmin[off + jp] = -1;
          hdmax[off + jp] = -2;
        } else {
          hdmin[off + jp] = hdn > dn ? hdn : dn;
          hdmax[off + jp] = hdx;
        }
      }
    }
    off += n;
  }
  hd_off[M] = off;
}

/* cp9.rs:5405-5413 flogsum, with the table of cp9.rs:5391-5401 */
float ixo_flogsum(float s1, float s2);

static inline float cmax(float a, float b) { return b > a ? b : a; } /* `if cand > best { best = cand }` */

/* Common banded layout (cp9.rs:4741-4763). */
typedef struct {
  int64_t *roff_start; /* per state: index into roff */
  int64_t *roff;       /* per (v, jp): cell offset */
  int64_t ncells;
} layout_t;
static void make_layout(const ixo_bands *b, int M, layout_t *L) {
  L->roff_start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
  L->roff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(b->hd_off[M] + 1));
  int64_t nc = 0, r = 0;
  for (int v = 0; v < M; v++) {
    L->roff_start[v] = r;
    int nj = b->jmax[v] >= b->jmin[v] ? b->jmax[v] - b->jmin[v] + 1 : 0;
    for (int jp = 0; jp < nj; jp++) {
      L->roff[r++] = nc;
      int w = b->hdmax[b->hd_off[v] + jp] - b->hdmin[b->hd_off[v] + jp] + 1;
      if (w > 0) nc += w;
    }
  }
  L->roff_start[M] = r;
  L->ncells = nc;
}
static void free_layout(layout_t *L) { free(L->roff_start); free(L->roff); }

int64_t ixo_hb_ncells(const ixo_bands *b, int M) {
  layout_t L;
  make_layout(b, M, &L);
  int64_t n = L.ncells;
  free_layout(&L);
  return n;
}

#define HDMIN(v, jp) b->hdmin[b->hd_off[v] + (jp)]
#define HDMAX(v, jp) b->hdmax[b->hd_off[v] + (jp)]
#define ROFF(v, jp) Lo.roff[Lo.roff_start[v] + (jp)]

/* cp9.rs:4717-5254 (CYK, inside = 0) and cp9.rs:5465-5885 (Inside,
 * inside = 1): fills the banded matrix for states M-1..1 and state 0's
 * children. The root's local begins and the reporting are done by
 * finish_root(). */
static void hb_fill(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_bands *b, const uint8_t *dsq, int i0, int j0,
                    int inside, float *alpha, layout_t Lo) {
  const int M = cm->M;
  const int W = j0 - i0 + 1;
  float *el = (float *)malloc(sizeof(float) * (size_t)(W + 1));
  for (int d = 0; d <= W; d++) el[d] = sc->el_selfsc * (float)d;
  for (int64_t c = 0; c < Lo.ncells; c++) alpha[c] = IMPOSSIBLE;
  const float ninf = IMPOSSIBLE;
  for (int v = M - 1; v >= 0; v--) {
    const int st = cm->sttype[v], sd = state_delta(st), sdr = state_right_delta(st);
    const int cfirst = cm->cfirst[v], cnum = cm->cnum[v];
    const float *tsc = sc->tsc + (size_t)v * MAXC;
    const float *oesc = sc->oesc + (size_t)v * KP * KP;
    const int jmn = b->jmin[v], jmx = b->jmax[v];
    if (sc->endsc[v] > -9.999e35f) {
      for (int j = jmn; j <= jmx; j++) {
        int jp = j - jmn, dpv = 0;
        for (int d = HDMIN(v, jp); d <= HDMAX(v, jp); d++, dpv++) {
          int dp = d - sd > 0 ? d - sd : 0;
          alpha[ROFF(v, jp) + dpv] = el[dp] + sc->endsc[v];
        }
      }
    }
    if (st == IXO_E_ST) {
      for (int j = jmn; j <= jmx; j++) alpha[ROFF(v, j - jmn)] = 0.0f;
    } else if (st == IXO_IL_ST || st == IXO_IR_ST) {
      for (int j = jmn; j <= jmx; j++) {
        const int jp = j - jmn, jsdr = j - sdr;
        const int64_t base = ROFF(v, jp);
        for (int d = HDMIN(v, jp); d <= HDMAX(v, jp); d++) {
          const int dpv = d - HDMIN(v, jp);
          float cur = alpha[base + dpv];
          for (int yo = 0; yo < cnum; yo++) {
            const int y = cfirst + yo;
            if (jsdr < b->jmin[y] || jsdr > b->jmax[y]) continue;
            const int jpy = j - b->jmin[y] - sdr;
            if (d - sd >= HDMIN(y, jpy) && d - sd <= HDMAX(y, jpy)) {
              const float cand = alpha[ROFF(y, jpy) + (d - sd - HDMIN(y, jpy))] + tsc[yo];
              cur = inside ? ixo_flogsum(cur, cand) : cmax(cur, cand);
            }
          }
          const int i = j - d + 1;
          cur += oesc[st == IXO_IL_ST ? dsq[i] : dsq[j]];
          if (cur < ninf) cur = ninf;
          alpha[base + dpv] = cur;
        }
      }
    } else if (st != IXO_B_ST) {
      for (int yo = 0; yo < cnum; yo++) {
        const int y = cfirst + yo;
        const float t = tsc[yo];
        int jn = jmn > b->jmin[y] + sdr ? jmn : b->jmin[y] + sdr;
        int jx = jmx < b->jmax[y] + sdr ? jmx : b->jmax[y] + sdr;
        for (int j = jn; j <= jx; j++) {
          const int jpv = j - jmn, jpy = j - b->jmin[y] - sdr;
          int dn = HDMIN(v, jpv) > HDMIN(y, jpy) + sd ? HDMIN(v, jpv) : HDMIN(y, jpy) + sd;
          int dx = HDMAX(v, jpv) < HDMAX(y, jpy) + sd ? HDMAX(v, jpv) : HDMAX(y, jpy) + sd;
          const int64_t iv0 = ROFF(v, jpv) + (dn - HDMIN(v, jpv));
          const int64_t iy0 = ROFF(y, jpy) + (dn - HDMIN(y, jpy) - sd);
          for (int k = 0; k <= dx - dn; k++) {
            const float cand = alpha[iy0 + k] + t;
            alpha[iv0 + k] = inside ? ixo_flogsum(alpha[iv0 + k], cand) : cmax(alpha[iv0 + k], cand);
          }
        }
      }
      for (int j = jmn; j <= jmx; j++) {
        const int jp = j - jmn;
        const int w = HDMAX(v, jp) - HDMIN(v, jp);
        const int64_t base = ROFF(v, jp);
        if (st == IXO_ML_ST || st == IXO_MP_ST) {
          int i = j - HDMIN(v, jp) + 1;
          for (int dpv = 0; dpv <= w; dpv++, i--)
            alpha[base + dpv] += oesc[st == IXO_ML_ST ? dsq[i] : dsq[i] * KP + dsq[j]];
        } else if (st == IXO_MR_ST) {
          const float em = oesc[dsq[j]];
          for (int dpv = 0; dpv <= w; dpv++) alpha[base + dpv] += em;
        }
        for (int dpv = 0; dpv <= w; dpv++)
          if (alpha[base + dpv] < ninf) alpha[base + dpv] = ninf;
      }
    } else { /* B */
      const int y = cfirst, z = cnum;
      int jn = jmn > b->jmin[z] ? jmn : b->jmin[z];
      int jx = jmx < b->jmax[z] ? jmx : b->jmax[z];
      for (int j = jn; j <= jx; j++) {
        const int jpv = j - jmn, jpy = j - b->jmin[y], jpz = j - b->jmin[z];
        const int hdmnz = HDMIN(z, jpz);
        int kn = j - b->jmax[y] > hdmnz ? j - b->jmax[y] : hdmnz;
        if (kn < 0) kn = 0;
        const int kx = jpy < HDMAX(z, jpz) ? jpy : HDMAX(z, jpz);
        const int64_t basev = ROFF(v, jpv), basez = ROFF(z, jpz);
        for (int d = HDMIN(v, jpv); d <= HDMAX(v, jpv); d++) {
          const int dpv = d - HDMIN(v, jpv);
          float cur = alpha[basev + dpv];
          for (int k = kn; k <= kx; k++) {
            const int jpymk = jpy - k;
            if (k >= d - HDMAX(y, jpymk) && k <= d - HDMIN(y, jpymk)) {
              const float cand = alpha[ROFF(y, jpymk) + (d - HDMIN(y, jpymk) - k)] + alpha[basez + (k - hdmnz)];
              cur = inside ? ixo_flogsum(cur, cand) : cmax(cur, cand);
            }
          }
          alpha[basev + dpv] = cur;
        }
      }
    }
  }
  free(el);
}

/* cp9.rs:5264-5304 / 5908-5937: local begins into the root deck (MAX), then
 * copies the root rows out: root[(jp, dp)] laid out like the matrix rows. */
static void finish_root(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_bands *b, float *alpha, layout_t Lo) {
  const int M = cm->M;
  const int v = 0;
  if (!sc->local_begin) return;
  for (int j = b->jmin[v]; j <= b->jmax[v]; j++) {
    const int jpv = j - b->jmin[v];
    for (int y = 1; y < M; y++) {
      if (!(sc->beginsc[y] > -9.999e35f) || j < b->jmin[y] || j > b->jmax[y]) continue;
      const int jpy = j - b->jmin[y];
      int dn = HDMIN(v, jpv) > HDMIN(y, jpy) ? HDMIN(v, jpv) : HDMIN(y, jpy);
      int dx = HDMAX(v, jpv) < HDMAX(y, jpy) ? HDMAX(v, jpv) : HDMAX(y, jpy);
      for (int d = dn; d <= dx; d++) {
        const float s = alpha[ROFF(y, jpy) + (d - HDMIN(y, jpy))] + sc->beginsc[y];
        float *av = &alpha[ROFF(v, jpv) + (d - HDMIN(v, jpv))];
        if (s > *av) *av = s;
      }
    }
  }
}

int64_t ixo_hb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_bands *b, const uint8_t *dsq, int i0, int j0,
                    int inside, float env_cutoff, float *vsc, int64_t *envi, int64_t *envj, float *root_out,
                    float *alpha_out) {
  layout_t Lo;
  make_layout(b, cm->M, &Lo);
  float *alpha = (float *)malloc(sizeof(float) * (size_t)(Lo.ncells > 0 ? Lo.ncells : 1));
  hb_fill(cm, sc, b, dsq, i0, j0, inside, alpha, Lo);
  finish_root(cm, sc, b, alpha, Lo);
  /* vsc_root and the envelope (cp9.rs:5305-5325) */
  float vr = IMPOSSIBLE;
  int64_t ei = (int64_t)j0 + 1, ej = (int64_t)i0 - 1;
  const int v = 0;
  for (int j = b->jmin[v]; j <= b->jmax[v]; j++) {
    const int jp = j - b->jmin[v];
    for (int d = HDMIN(v, jp); d <= HDMAX(v, jp); d++) {
      const float s = alpha[ROFF(v, jp) + (d - HDMIN(v, jp))];
      if (s > vr) vr = s;
      if (s >= env_cutoff) {
        const int i = j - d + 1;
        if (i < ei) ei = i;
        if (j > ej) ej = j;
      }
    }
  }
  *vsc = vr;
  *envi = ei == (int64_t)j0 + 1 ? -1 : ei;
  *envj = ej == (int64_t)i0 - 1 ? -1 : ej;
  int64_t nroot = 0;
  for (int j = b->jmin[0]; j <= b->jmax[0]; j++) {
    const int w = HDMAX(0, j - b->jmin[0]) - HDMIN(0, j - b->jmin[0]) + 1;
    if (w > 0) nroot += w;
  }
  if (root_out && nroot) memcpy(root_out, alpha + Lo.roff[0], sizeof(float) * (size_t)nroot);
  if (alpha_out) memcpy(alpha_out, alpha, sizeof(float) * (size_t)Lo.ncells);
  free(alpha);
  free_layout(&Lo);
  return nroot;
}

/* cp9.rs:6516-6529 score_correction_null3 */
float ixo_null3_correction(const float *null, const float *comp, int len, float omega) {
  float score = 0.0f;
  for (int a = 0; a < K4; a++) {
    double term = sre_log2((double)(comp[a] / null[a])) * (double)comp[a] * (double)len;
    score = (float)((double)score + term);
  }
  score = (float)((double)score + sre_log2((double)omega));
  return ixo_flogsum(0.0f, score);
}

/* cp9.rs:5887-5900 act prefix counts + cp9.rs:5950-5997 ReportHitsGreedily,
 * from the root deck rows (as returned by ixo_hb_scan). */
int64_t ixo_hb_report_hits(const ixo_cm *cm, const ixo_bands *b, const uint8_t *dsq, int i0, int j0,
                           const float *root, float cutoff, int do_null3, ixo_hit *out, int64_t cap) {
  const int W = j0 - i0 + 1;
  double *act = (double *)calloc((size_t)(W + 1) * 4, sizeof(double));
  if (do_null3)
    for (int jp = 1; jp <= W; jp++) {
      for (int a = 0; a < 4; a++) act[jp * 4 + a] = act[(jp - 1) * 4 + a];
      int r = dsq[i0 + jp - 1];
      if (r < 4) act[jp * 4 + r] += 1.0;
    }
  float *bestsc = (float *)malloc(sizeof(float) * (size_t)(W + 1));
  int64_t nh = 0, off = 0;
  for (int j = b->jmin[0]; j <= b->jmax[0]; j++) {
    const int jp = j - b->jmin[0];
    for (int d = 0; d <= W; d++) bestsc[d] = IMPOSSIBLE;
    const int dmin = HDMIN(0, jp), dmax = HDMAX(0, jp);
    for (int d = dmin; d <= dmax; d++) bestsc[d] = root[off + (d - dmin)];
    if (dmax >= dmin) off += dmax - dmin + 1;
    if (dmin <= dmax) {
      float maxrep = IMPOSSIBLE;
      for (int d = dmin > 1 ? dmin : 1; d <= dmax; d++) {
        const int i = j - d + 1;
        float hs = bestsc[d];
        if (hs > maxrep && hs >= cutoff && hs > -9.999e35f) {
          int rep = 1;
          float bias = 0.0f;
          if (do_null3) {
            float comp[4];
            const int ip = i - i0 + 1, jj = j - i0 + 1;
            for (int a = 0; a < 4; a++) comp[a] = (float)(act[jj * 4 + a] - act[(ip - 1) * 4 + a]);
            float s = vec_fsum(comp, 4);
            if (s != 0.0f)
              for (int a = 0; a < 4; a++) comp[a] /= s;
            else
              for (int a = 0; a < 4; a++) comp[a] = 1.0f / 4.0f;
            float corr = ixo_null3_correction(cm->null, comp, d, (float)cm->n3_omega);
            hs -= corr;
            bias = corr;
            rep = hs > maxrep && hs >= cutoff;
          }
          if (rep) {
            if (nh < cap) {
              out[nh].i = i;
              out[nh].j = j;
              out[nh].score = hs;
              out[nh].bias = bias;
            }
            nh++;
            maxrep = hs;
          }
        }
      }
    }
  }
  free(act);
  free(bestsc);
  return nh;
}
