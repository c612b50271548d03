/*
 * qdb.c -- CPU oracle (test infrastructure only): restatement of infernox's
 * query-dependent-banded (or non-banded) scanners, cm_nohmm.rs:406
 * generic_scan<f32> (CYK, cyk_seq_filter :855) and generic_scan<i32>
 * (integer Inside with ILogsum, final_stage_inside_opt :919), with their
 * greedy reporting (:769) and overlap removal (cp9.rs:6544), and the
 * envelope padding/merge of cyk_seq_filter (:855-894).
 */
#include "ixo_cm.h"
#include "ixo.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KP 18
#define MAXC IXO_MAXCONNECT
#define INFTY 987654321 /* cp9.rs:21 */

float ixo_null3_correction(const float *null, const float *comp, int len, float omega);

static int not_impossible(float x) { return x > -1e36f + 1.0f; } /* cm_nohmm.rs:21 */
static int sdelta_q(int st) { return st == IXO_MP_ST ? 2 : ((st >= IXO_ML_ST && st <= IXO_IR_ST) ? 1 : 0); }
static int srdelta_q(int st) { return (st == IXO_MP_ST || st == IXO_MR_ST || st == IXO_IR_ST) ? 1 : 0; }

static int node_of(const ixo_cm *cm, int v) {
  int nd = 0;
  while (nd + 1 < cm->nodes && cm->nodemap[nd + 1] <= v) nd++;
  return nd;
}
/* BEGL_S: the S state of a BEGL node (constants.rs stid BEGL_S) */
static int is_begl(const ixo_cm *cm, int v) {
  return cm->sttype[v] == IXO_S_ST && cm->ndtype[node_of(cm, v)] == IXO_BEGL_ND;
}

/* cp9.rs:6544-6574 remove_overlaps_greedy: score descending, then start
 * ascending (stable), then drop every hit overlapping a kept one. */
static int64_t remove_overlaps(ixo_hit *h, int64_t n, ixo_hit *out, int64_t cap) {
  for (int64_t a = 1; a < n; a++) {
    ixo_hit x = h[a];
    int64_t k = a;
    while (k > 0 && (h[k - 1].score < x.score || (h[k - 1].score == x.score && h[k - 1].i > x.i))) {
      h[k] = h[k - 1];
      k--;
    }
    h[k] = x;
  }
  char *rem = (char *)calloc((size_t)n + 1, 1);
  int64_t nk = 0;
  for (int64_t a = 0; a < n; a++) {
    if (rem[a]) continue;
    for (int64_t k = a + 1; k < n; k++)
      if (!rem[k] && !(h[k].j < h[a].i) && !(h[a].j < h[k].i)) rem[k] = 1;
  }
  for (int64_t a = 0; a < n; a++)
    if (!rem[a]) {
      if (nk < cap) out[nk] = h[a];
      nk++;
    }
  free(rem);
  return nk;
}

/* float semiring: IMPOSSIBLE / ESL_MAX (cm_nohmm.rs:98-138) */
#define S float
#define S_ZERO (-1e36f)
#define S_COMB(a, b) ((a) >= (b) ? (a) : (b))
#define S_VMAX(a, b) ((a) >= (b) ? (a) : (b))
#define S_ADD(a, b) ((a) + (b))
#define S_BITS(a) (a)
#define S_POSSIBLE(a) not_impossible(a)
#define QDB_SCAN qdb_scan_f
#define QDB_REDUCE qdb_reduce_f
#include "qdb_impl.h"
#undef S
#undef S_ZERO
#undef S_COMB
#undef S_VMAX
#undef S_ADD
#undef S_BITS
#undef S_POSSIBLE
#undef QDB_SCAN
#undef QDB_REDUCE

/* integer semiring: -INFTY / ILogsum, wrapping adds (cm_nohmm.rs:140-185) */
static int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t imax(int32_t a, int32_t b) { return a >= b ? a : b; }
static float scorify(int32_t sc) { return sc == -INFTY ? -1e36f : (float)((double)sc / 1000.0); } /* :37 */
#define S int32_t
#define S_ZERO (-INFTY)
#define S_COMB(a, b) ixo_ilogsum((a), (b))
#define S_VMAX(a, b) imax((a), (b))
#define S_ADD(a, b) iadd((a), (b))
#define S_BITS(a) scorify(a)
#define S_POSSIBLE(a) ((a) != -INFTY)
#define QDB_SCAN qdb_scan_i
#define QDB_REDUCE qdb_reduce_i
#include "qdb_impl.h"

int64_t ixo_qdb_scan(const ixo_cm *cm, const ixo_cmsc *fs, const ixo_cmisc *is, const int32_t *dmin,
                     const int32_t *dmax, const uint8_t *dsq, int i0, int j0, int inside, float cutoff, int do_null3,
                     ixo_hit *out, int64_t cap, float *root_out) {
  if (!inside)
    return qdb_scan_f(cm, fs->tsc, fs->oesc, fs->endsc, fs->beginsc, fs->el_selfsc, fs->local_begin, dmin, dmax,
                      dsq, i0, j0, cutoff, do_null3, out, cap, root_out);
  return qdb_scan_i(cm, is->itsc, is->ioesc, is->iendsc, is->ibeginsc, is->iel_selfsc, fs->local_begin, dmin, dmax,
                    dsq, i0, j0, cutoff, do_null3, out, cap, root_out);
}

/* cm_nohmm.rs:855-894 cyk_seq_filter: CYK hits padded to +-(W-1), sorted by
 * end then start, merged when the next start <= the current end. */
int64_t ixo_cyk_seq_filter(const ixo_cm *cm, const ixo_cmsc *fs, const int32_t *dmin1, const int32_t *dmax1,
                           const uint8_t *dsq, int n, float cutoff, int do_null3, int32_t *es, int32_t *ee,
                           int64_t cap) {
  int64_t nh = qdb_scan_f(cm, fs->tsc, fs->oesc, fs->endsc, fs->beginsc, fs->el_selfsc, fs->local_begin, dmin1, dmax1,
                          dsq, 1, n, cutoff, do_null3, NULL, 0, NULL);
  ixo_hit *h = (ixo_hit *)malloc(sizeof(ixo_hit) * (size_t)(nh + 1));
  qdb_scan_f(cm, fs->tsc, fs->oesc, fs->endsc, fs->beginsc, fs->el_selfsc, fs->local_begin, dmin1, dmax1, dsq, 1, n,
             cutoff, do_null3, h, nh, NULL);
  int32_t *pi = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1)), *pj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh + 1));
  for (int64_t k = 0; k < nh; k++) {
    pi[k] = h[k].j - (cm->W - 1) > 1 ? h[k].j - (cm->W - 1) : 1;
    pj[k] = h[k].i + (cm->W - 1) < n ? h[k].i + (cm->W - 1) : n;
  }
  for (int64_t a = 1; a < nh; a++) { /* insertion sort by (end, start) */
    int32_t xi = pi[a], xj = pj[a];
    int64_t k = a;
    while (k > 0 && (pj[k - 1] > xj || (pj[k - 1] == xj && pi[k - 1] > xi))) {
      pi[k] = pi[k - 1];
      pj[k] = pj[k - 1];
      k--;
    }
    pi[k] = xi;
    pj[k] = xj;
  }
  int64_t no = 0;
  for (int64_t k = 0; k < nh; k++) {
    int32_t iw = pi[k], jw = pj[k];
    while (k + 1 < nh && pi[k + 1] <= jw) {
      k++;
      if (pj[k] > jw) jw = pj[k];
    }
    if (no < cap) {
      es[no] = iw;
      ee[no] = jw;
    }
    no++;
  }
  free(h);
  free(pi);
  free(pj);
  return no;
}
