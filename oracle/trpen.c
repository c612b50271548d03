/*
 * trpen.c -- CPU oracle (test infrastructure only): the CM's expected state
 * occupancy (cm_expected_state_occupancy, cp9.rs:752-837) and the
 * truncated-begin penalties built from it (TrPenalties::new,
 * cm_trunc.rs:192-384, with the local begin probabilities of
 * cm_modelconfig.c:397 computed inline as there). Not linked into the
 * product.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include "ixo_cm.h"

#define INFTY 987654321
#define IMPOSSIBLE (-1e36f)

static int *state_nodes(const ixo_cm *cm) {
  int *nd = (int *)malloc(sizeof(int) * (size_t)cm->M);
  for (int n = 0; n < cm->nodes; n++) {
    const int end = n + 1 < cm->nodes ? cm->nodemap[n + 1] : cm->M;
    for (int v = cm->nodemap[n]; v < end; v++) nd[v] = n;
  }
  return nd;
}

/* cp9.rs:752 on the un-localized probabilities: psi[v] = expected number of
 * visits to v. A state's parents are the states listing it as a child, in
 * increasing order (cp9.rs:802-808: y = pnum-1 .. final_y), an insert's own
 * self transition excluded and added as the geometric factor afterwards;
 * the transition a parent x takes to v is t[x][v - cfirst[x]]. */
void ixo_cm_psi(const ixo_cm *cm, double *psi) {
  const int M = cm->M;
  for (int v = 0; v < M; v++) {
    const int st = cm->sttype[v];
    if (st == IXO_S_ST) {
      psi[v] = 1.0;
      continue;
    }
    const int ins = st == IXO_IL_ST || st == IXO_IR_ST;
    double p = 0.0;
    for (int x = 0; x < v + ins; x++) {
      const int sx = cm->sttype[x];
      if (sx == IXO_B_ST || sx == IXO_E_ST) continue;
      const int off = v - cm->cfirst[x];
      if (off < 0 || off >= cm->cnum[x]) continue;
      if (ins && x == v) continue;
      p += psi[x] * (double)cm->t[(size_t)x * 6 + off];
    }
    psi[v] = p;
    if (ins) {
      const double ts = (double)cm->t[(size_t)v * 6];
      psi[v] += psi[v] * (ts / (1.0 - ts));
    }
  }
}

/* cm_trunc.rs:111 inserts_given_node */
static void inserts_given_node(const ixo_cm *cm, int nd, int *i1, int *i2) {
  *i1 = *i2 = -1;
  if (nd < 0) return;
  const int v = cm->nodemap[nd];
  switch (cm->ndtype[nd]) {
    case IXO_MATP_ND: *i1 = v + 4, *i2 = v + 5; break;
    case IXO_MATL_ND:
    case IXO_MATR_ND: *i1 = v + 2; break;
    case IXO_BEGR_ND: *i1 = v + 1; break;
    case IXO_ROOT_ND: *i1 = v + 1, *i2 = v + 2; break;
    default: break;
  }
}

/* Prob2Score(p, 1) (cm_trunc.rs:175); a float-to-int cast saturates in the
 * reference */
static int32_t prob2score1(float p) {
  if (p == 0.0f) return -INFTY;
  const double l = p > 0.0f ? log((double)p) * 1.44269504 : -1e36;
  const double r = floor(0.5 + 1000.0 * l);
  return r < -2147483648.0 ? INT32_MIN : (r > 2147483647.0 ? INT32_MAX : (int32_t)r);
}
static float sre_log2f(float p) { return p > 0.0f ? (float)(log((double)p) * 1.44269504) : -1e36f; }

/* cm_trunc.rs:192-384 with psi of the global CM (ixo_cm_psi). g, l: [3][M]
 * log2 penalties (index 0 = 5'&3', 1 = 5' only, 2 = 3' only), ig, il their
 * Prob2Score integers. */
void ixo_trpenalties(const ixo_cm *cm, float *g, float *l, int32_t *ig, int32_t *il) {
  const int M = cm->M, clen = cm->clen;
  double *psi = (double *)malloc(sizeof(double) * (size_t)M);
  ixo_cm_psi(cm, psi);
  int *ndidx = state_nodes(cm);
  int32_t *lpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes),
          *rpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  ixo_emit_map(cm, lpos, rpos);
  for (int q = 0; q < 3 * M; q++) g[q] = l[q] = IMPOSSIBLE, ig[q] = il[q] = -INFTY;
  const float g53 = (float)(2.0 / (double)(clen * (clen + 1))), g5or3 = (float)(1.0 / (double)clen);
  /* the local begin probabilities (cm_modelconfig.c:397) */
  float *begin = (float *)calloc((size_t)M, sizeof(float));
  int nstarts = 0;
  for (int nd = 2; nd < cm->nodes; nd++) {
    const int t = cm->ndtype[nd];
    if (t == IXO_MATP_ND || t == IXO_MATL_ND || t == IXO_MATR_ND || t == IXO_BIF_ND) nstarts++;
  }
  if (cm->nodes > 1) begin[cm->nodemap[1]] = (float)(1.0 - (double)cm->pbegin);
  const float pb = cm->pbegin / (float)nstarts;
  for (int nd = 2; nd < cm->nodes; nd++) {
    const int t = cm->ndtype[nd];
    if (t == IXO_MATP_ND || t == IXO_MATL_ND || t == IXO_MATR_ND || t == IXO_BIF_ND) begin[cm->nodemap[nd]] = pb;
  }
  float prv5 = 0.f, prv3 = 0.f, prv53 = 0.f;
  for (int nd = 0; nd < cm->nodes; nd++) {
    const int t = cm->ndtype[nd];
    const int lp = (t == IXO_MATP_ND || t == IXO_MATL_ND) ? lpos[nd] : lpos[nd] + 1;
    const int rp = (t == IXO_MATP_ND || t == IXO_MATR_ND) ? rpos[nd] : rpos[nd] - 1;
    if (t == IXO_END_ND) {
      prv5 = prv3 = prv53 = 0.f;
    } else if (t == IXO_BEGL_ND || t == IXO_BEGR_ND) {
      const int s = cm->nodemap[nd];
      int b = -1; /* the bifurcation above (plast of the S state) */
      for (int x = 0; x < s; x++)
        if (cm->sttype[x] == IXO_B_ST && (cm->cfirst[x] == s || cm->cnum[x] == s)) b = x;
      prv5 = t == IXO_BEGL_ND ? 0.f : l[1 * M + b];
      prv3 = t == IXO_BEGR_ND ? 0.f : l[2 * M + b];
      prv53 = l[0 * M + b];
    } else if (t == IXO_MATP_ND || t == IXO_MATL_ND || t == IXO_MATR_ND || t == IXO_BIF_ND) {
      const int m = cm->nodemap[nd];
      int i1, i2;
      inserts_given_node(cm, nd - 1, &i1, &i2);
      const float m_psi = (float)psi[m];
      const float i1_psi = i1 >= 0 ? (float)psi[i1] : 0.f, i2_psi = i2 >= 0 ? (float)psi[i2] : 0.f;
      const float summed = m_psi + i1_psi + i2_psi;
#define SET(arr, idx, prob)                                            \
  do {                                                                 \
    (arr)[(idx) * M + m] = (m_psi / summed) * (prob);                  \
    if (i1 >= 0) (arr)[(idx) * M + i1] = (i1_psi / summed) * (prob);   \
    if (i2 >= 0) (arr)[(idx) * M + i2] = (i2_psi / summed) * (prob);   \
  } while (0)
      SET(g, 0, g53);
      if (rp == clen) SET(g, 1, g5or3);
      if (lp == 1) SET(g, 2, g5or3);
      const int sub = rp - lp + 1;
      const int nf53 = (sub * (sub + 1)) / 2;
      const float cur5 = begin[m] / (float)sub + prv5;
      const float cur3 = begin[m] / (float)sub + prv3;
      const float cur53 = begin[m] / (float)nf53 + prv53;
      SET(l, 0, cur53);
      SET(l, 1, cur5);
      SET(l, 2, cur3);
#undef SET
      prv5 = t == IXO_MATL_ND ? cur5 : 0.f;
      prv3 = t == IXO_MATR_ND ? cur3 : 0.f;
      prv53 = cur53;
    }
  }
  for (int v = 0; v < M; v++) {
    const int st = cm->sttype[v], nt = cm->ndtype[ndidx[v]];
    const int detached = v + 1 < M && cm->sttype[v + 1] == IXO_E_ST;
    const int q = (st == IXO_MP_ST) || (st == IXO_ML_ST && nt == IXO_MATL_ND) ||
                  (st == IXO_MR_ST && nt == IXO_MATR_ND) || st == IXO_B_ST ||
                  ((st == IXO_IL_ST || st == IXO_IR_ST) && !detached);
    if (!q) continue;
    if (st == IXO_IL_ST && nt == IXO_MATP_ND && cm->ndtype[ndidx[v] + 1] == IXO_END_ND) continue;
    for (int k = 0; k < 3; k++) {
      if (g[k * M + v] > -0.5e36f) {
        ig[k * M + v] = prob2score1(g[k * M + v]);
        g[k * M + v] = sre_log2f(g[k * M + v]);
      }
      il[k * M + v] = prob2score1(l[k * M + v]);
      l[k * M + v] = sre_log2f(l[k * M + v]);
    }
  }
  free(begin);
  free(lpos);
  free(rpos);
  free(ndidx);
  free(psi);
}
