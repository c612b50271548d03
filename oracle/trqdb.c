/*
 * trqdb.c -- CPU oracle (test infrastructure only): restatement of infernox's
 * QDB / non-banded truncated scanners, cm_dpsearch_trunc.rs:1800
 * generic_tr_scan<f32> (ref_tr_cyk_scan, :207) and generic_tr_scan<i32>
 * (ref_itr_inside_scan, :226, ILogsum), with the truncated greedy reporting
 * (:2322-2372), cm_hit_AllowTruncation (:2376) and overlap removal (:2470).
 */
#include "ixo_cm.h"
#include "ixo.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KP 18
#define MAXC IXO_MAXCONNECT
#define TQ_INFTY 987654321 /* cp9.rs:21 */
#define TQ_IMPOSSIBLE (-1e36f)

float ixo_null3_correction(const float *null, const float *comp, int len, float omega);

/* cm_dpsearch_trunc.rs:184-199 fill_from_pass_idx */
static int tq_fill_l(int p) { return p == IXO_PASS_5P_AND_3P_FORCE || p == IXO_PASS_5P_AND_3P_ANY || p == IXO_PASS_3P_ONLY_FORCE; }
static int tq_fill_r(int p) { return p == IXO_PASS_5P_AND_3P_FORCE || p == IXO_PASS_5P_AND_3P_ANY || p == IXO_PASS_5P_ONLY_FORCE; }
static int tq_fill_t(int p) { return p == IXO_PASS_5P_AND_3P_FORCE || p == IXO_PASS_5P_AND_3P_ANY; }
/* cm_dpsearch_trunc.rs:66-86 state_delta, state_right_delta */
static int tq_sd(int st) { return st == IXO_MP_ST ? 2 : ((st >= IXO_ML_ST && st <= IXO_IR_ST) ? 1 : 0); }
static int tq_sdr(int st) { return (st == IXO_MP_ST || st == IXO_MR_ST || st == IXO_IR_ST) ? 1 : 0; }
static int tq_node(const ixo_cm *cm, int v) {
  int nd = 0;
  while (nd + 1 < cm->nodes && cm->nodemap[nd + 1] <= v) nd++;
  return nd;
}
static int tq_is_begl(const ixo_cm *cm, int v) {
  return cm->sttype[v] == IXO_S_ST && cm->ndtype[tq_node(cm, v)] == IXO_BEGL_ND;
}
/* cm_dpsearch_trunc.rs:88 state_is_detached: the next state is an END node's
 * E state (the only kind of E state). */
static int tq_detached(const ixo_cm *cm, int v) { return v + 1 < cm->M && cm->sttype[v + 1] == IXO_E_ST; }

/* float semiring: truncated CYK (cm_dpsearch_trunc.rs:106-115) */
#define S float
#define S_ZERO TQ_IMPOSSIBLE
#define S_COMB(a, b) ((a) >= (b) ? (a) : (b))
#define S_MAX2(a, b) ((a) >= (b) ? (a) : (b))
#define S_ADD(a, b) ((a) + (b))
#define S_BITS(a) (a)
#define S_POSSIBLE(a) ((a) > TQ_IMPOSSIBLE + 1.0f)
#define TRQDB_SCAN trqdb_scan_f
#include "trqdb_impl.h"
#undef S
#undef S_ZERO
#undef S_COMB
#undef S_MAX2
#undef S_ADD
#undef S_BITS
#undef S_POSSIBLE
#undef TRQDB_SCAN

/* integer semiring: truncated Inside with ILogsum and wrapping adds (:116-126) */
static int32_t tq_iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t tq_imax(int32_t a, int32_t b) { return a >= b ? a : b; }
static float tq_scorify(int32_t sc) { return sc == -TQ_INFTY ? TQ_IMPOSSIBLE : (float)((double)sc / 1000.0); }
#define S int32_t
#define S_ZERO (-TQ_INFTY)
#define S_COMB(a, b) ixo_ilogsum((a), (b))
#define S_MAX2(a, b) tq_imax((a), (b))
#define S_ADD(a, b) tq_iadd((a), (b))
#define S_BITS(a) tq_scorify(a)
#define S_POSSIBLE(a) ((a) != -TQ_INFTY)
#define TRQDB_SCAN trqdb_scan_i
#include "trqdb_impl.h"

int64_t ixo_trqdb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_cmisc *is, const float *lmesc,
                       const float *rmesc, const int32_t *ilmesc, const int32_t *irmesc, const float *pty,
                       const int32_t *ipty, const int32_t *dmax, const uint8_t *dsq, int i0, int j0, int inside,
                       int pass_idx, int local, float cutoff, int do_null3, float *vsc, int *vmode, float *bsc_full,
                       ixo_trhit *hits, int64_t cap, float *best_out) {
  if (!inside)
    return trqdb_scan_f(cm, sc->tsc, sc->oesc, lmesc, rmesc, sc->endsc, sc->el_selfsc, pty, dmax, dsq, i0, j0,
                        pass_idx, local, cutoff, do_null3, vsc, vmode, bsc_full, hits, cap, best_out);
  return trqdb_scan_i(cm, is->itsc, is->ioesc, ilmesc, irmesc, is->iendsc, is->iel_selfsc, ipty, dmax, dsq, i0, j0,
                      pass_idx, local, cutoff, do_null3, vsc, vmode, bsc_full, hits, cap, best_out);
}
