/*
 * hmmband.c -- CPU oracle (test infrastructure only) for the HMM->CM band
 * conversion of the HMM-banded scanners: the emit map (cp9.rs:144
 * create_emit_map), hmm_bands_enforce_valid_parse (cp9.rs:2965) with
 * hmm_bands_fix_unreachable / hmm_bands_fill_gap (:2893, :2940),
 * cp9_hmm2ijbands (cp9.rs:3380, the non-truncated path), ij2d_bands
 * (:3815) and cm_hb_mx_size_needed (:3854). Not linked into the product.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ixo_band.h"

enum { D_ST = 0, MP_ST = 1, ML_ST = 2, MR_ST = 3, IL_ST = 4, IR_ST = 5, S_ST = 6, E_ST = 7, B_ST = 8 };
enum { BIF_ND = 0, MATP_ND = 1, MATL_ND = 2, MATR_ND = 3, BEGL_ND = 4, BEGR_ND = 5, ROOT_ND = 6, END_ND = 7 };
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* node of each state */
static int *node_index(const ixo_hbcm *cm) {
  int *ndidx = (int *)malloc(sizeof(int) * (size_t)cm->M);
  for (int nd = 0; nd < cm->nodes; nd++) {
    const int end = nd + 1 < cm->nodes ? cm->nodemap[nd + 1] : cm->M;
    for (int v = cm->nodemap[nd]; v < end; v++) ndidx[v] = nd;
  }
  return ndidx;
}

/* cp9.rs:144-207 create_emit_map (lpos / rpos only) */
void ixo_hb_emit_map(const ixo_hbcm *cm, int32_t *lpos, int32_t *rpos) {
  int *ndidx = node_index(cm);
  for (int n = 0; n < cm->nodes; n++) lpos[n] = rpos[n] = -1;
  int *pda = (int *)malloc(sizeof(int) * (size_t)(4 * cm->nodes + 8));
  int sp = 0, cpos = 0;
  pda[sp++] = 0; /* on_right */
  pda[sp++] = 0; /* nd */
  while (sp > 0) {
    const int nd = pda[--sp], on_right = pda[--sp];
    const int nt = cm->ndtype[nd];
    if (on_right) {
      rpos[nd] = cpos + 1;
      if (nt == MATP_ND || nt == MATR_ND) cpos++;
    } else {
      if (nt == MATP_ND || nt == MATL_ND) cpos++;
      lpos[nd] = cpos;
      pda[sp++] = 1;
      pda[sp++] = nd;
      if (nt == BIF_ND) {
        const int v = cm->nodemap[nd];
        pda[sp++] = 0;
        pda[sp++] = ndidx[cm->cnum[v]];
        pda[sp++] = 0;
        pda[sp++] = ndidx[cm->cfirst[v]];
      } else if (nt != END_ND) {
        pda[sp++] = 0;
        pda[sp++] = nd + 1;
      }
    }
  }
  free(pda);
  free(ndidx);
}

/* cp9.rs:2893-2938 (only with local begins/ends off) */
static void fix_unreachable(const ixo_hbcp9 *h, int32_t *pn, int k, int r_prv_min) {
  int32_t *mnm = pn, *mxm = pn + (h->hmm_m + 1), *mni = pn + 2 * (h->hmm_m + 1), *mxi = pn + 3 * (h->hmm_m + 1),
          *mnd = pn + 4 * (h->hmm_m + 1), *mxd = pn + 5 * (h->hmm_m + 1);
  int nxt_m = -1, nxt_d = -1;
  if (mnm[k] != -1 && mxm[k] != -1 && mxm[k] - 1 > r_prv_min) nxt_m = MAX(mnm[k] - 1, r_prv_min);
  if (mnd[k] != -1 && mxd[k] != -1 && mxd[k] > r_prv_min) nxt_d = MAX(mnd[k], r_prv_min);
  if (nxt_m != -1 || nxt_d != -1) {
    const int nxt_n = nxt_m == -1 ? nxt_d : nxt_d == -1 ? nxt_m : MIN(nxt_m, nxt_d);
    mni[k - 1] = mni[k - 1] != -1 ? MIN(mni[k - 1], r_prv_min + 1) : r_prv_min + 1;
    mxi[k - 1] = mxi[k - 1] != -1 ? MAX(mxi[k - 1], nxt_n) : nxt_n;
  } else {
    int kp = k;
    while (kp <= h->hmm_m && mxm[kp] < r_prv_min + 1 && mxd[kp] < r_prv_min) {
      mnd[kp] = r_prv_min;
      mxd[kp] = r_prv_min;
      kp++;
    }
  }
}

/* cp9.rs:2940-2963 */
static void fill_gap(const ixo_hbcp9 *h, int32_t *pn, int k, int min1, int min2, int prv_mn, int prv_dn) {
  int32_t *mni = pn + 2 * (h->hmm_m + 1), *mxi = pn + 3 * (h->hmm_m + 1);
  const int right_min = min1 <= min2 ? min2 : min1;
  int in = INT_MAX;
  if (prv_mn != INT_MAX) in = MIN(in, prv_mn + 1);
  if (prv_dn != INT_MAX) in = MIN(in, prv_dn);
  const int ix = right_min - 1;
  mni[k] = mni[k] != -1 ? MIN(mni[k], in) : in;
  mxi[k] = mxi[k] != -1 ? MAX(mxi[k], ix) : ix;
}

/* the reachable-residue arrays of cp9.rs:2878-2889 */
typedef struct {
  int *mn, *mx, *in, *ix, *dn, *dx, *nn_i, *nx_i, *nn_j, *nx_j;
} rarr;

/* a transition into a state with posterior band [pmin, pmax]: returns 1 and
 * the clipped (n, x) when the shifted source band [n0, x0] meets it */
static int clip(int n0, int x0, int pmin, int pmax, int *n, int *x) {
  if (MIN(x0, pmax) - MAX(n0, pmin) < 0) return 0;
  *n = MIN(MAX(n0, pmin), pmax);
  *x = MIN(x0, pmax);
  return 1;
}

/* cp9.rs:2965-3376 hmm_bands_enforce_valid_parse */
static void enforce_valid_parse(const ixo_hbcp9 *h, int32_t *pn, int i0, int j0, int doing_search, rarr *r) {
  const int M = h->hmm_m, sz = M + 1;
  int32_t *mnm = pn, *mxm = pn + sz, *mni = pn + 2 * sz, *mxi = pn + 3 * sz, *mnd = pn + 4 * sz, *mxd = pn + 5 * sz;
  const int lbe = h->local_begin && h->local_end;
  int *buf = (int *)malloc(sizeof(int) * (size_t)sz * 12);
  int *rr[12];
  for (int q = 0; q < 12; q++) {
    rr[q] = buf + (size_t)q * sz;
    for (int k = 0; k < sz; k++) rr[q][k] = q % 2 == 0 ? INT_MAX : INT_MIN;
  }
  int *mn = rr[0], *mx = rr[1], *in = rr[2], *ix = rr[3], *dn = rr[4], *dx = rr[5], *nn_i = rr[6], *nx_i = rr[7],
      *nn_j = rr[8], *nx_j = rr[9], *nn_h = rr[10], *nx_h = rr[11];
  int begn = INT_MAX, begx = INT_MIN, endn = INT_MAX, endx = INT_MIN;
  if (mnm[0] != -1) {
    mn[0] = mnm[0];
    mx[0] = mxm[0];
  }
  const int NINF = -987654321;
#define PRV(a) (k > 0 ? (a)[k - 1] : INT_MAX)
  int k = 0;
  while (k <= M) {
    int filled = 0, n, x;
    /* into I_k */
    if (mn[k] <= mx[k] && mni[k] != -1 && clip(mn[k] + 1, mx[k] + 1, mni[k], mxi[k], &n, &x)) {
      in[k] = MIN(in[k], n);
      ix[k] = MAX(ix[k], x);
    }
    if (dn[k] <= dx[k] && mni[k] != -1 && clip(dn[k] + 1, dx[k] + 1, mni[k], mxi[k], &n, &x)) {
      in[k] = MIN(in[k], n);
      ix[k] = MAX(ix[k], x);
    }
    if (in[k] <= ix[k] && mni[k] != -1) {
      if (in[k] <= mxi[k]) ix[k] = mxi[k];
      else {
        in[k] = INT_MAX;
        ix[k] = INT_MIN;
      }
    }
    if (k < M) {
      const int k1 = k + 1;
      /* into M_k+1 from M_k, D_k, I_k (the I_k gap test has no reach guard and
       * its INT_MIN - INT_MAX wraps, cp9.rs:3091-3104) */
      if (mn[k] <= mx[k] && mnm[k1] != -1 && clip(mn[k] + 1, mx[k] + 1, mnm[k1], mxm[k1], &n, &x)) {
        if (mn[k1] != INT_MAX && !lbe && MIN(x, mx[k1]) - MAX(n, mn[k1]) < -1) {
          fill_gap(h, pn, k, n, mn[k1], PRV(mn), PRV(dn));
          filled = 1;
        }
        mn[k1] = MIN(mn[k1], n);
        mx[k1] = MAX(mx[k1], x);
      }
      if (dn[k] <= dx[k] && mnm[k1] != -1 && clip(dn[k] + 1, dx[k] + 1, mnm[k1], mxm[k1], &n, &x)) {
        if (mn[k1] != INT_MAX && !lbe && MIN(x, mx[k1]) - MAX(n, mn[k1]) < -1) {
          fill_gap(h, pn, k, n, mn[k1], PRV(mn), PRV(dn));
          filled = 1;
        }
        mn[k1] = MIN(mn[k1], n);
        mx[k1] = MAX(mx[k1], x);
      }
      if (in[k] <= ix[k] && mnm[k1] != -1 && clip(in[k] + 1, ix[k] + 1, mnm[k1], mxm[k1], &n, &x)) {
        const int lo = MIN(x, mx[k1]), hi = MAX(n, mn[k1]);
        if (!lbe && (int)((unsigned)lo - (unsigned)hi) < -1) {
          fill_gap(h, pn, k, n, mn[k1], PRV(mn), PRV(dn));
          filled = 1;
        }
        mn[k1] = MIN(mn[k1], n);
        mx[k1] = MAX(mx[k1], x);
      }
      if (h->el && mnm[k1] != -1)
        for (int c = 0; c < h->el_from_ct[k1]; c++) {
          const int kpp = h->el_from_idx[h->el_from_off[k1] + c];
          if (mn[kpp] <= mx[kpp] && clip(mn[kpp], j0, mnm[k1], mxm[k1], &n, &x)) {
            if (mn[k1] != INT_MAX && !lbe && MIN(x, mx[k1]) - MAX(n, mn[k1]) < -1) {
              fill_gap(h, pn, k, n, mn[k1], PRV(mn), PRV(dn));
              filled = 1;
            }
            mn[k1] = MIN(mn[k1], n);
            mx[k1] = MAX(mx[k1], x);
          }
        }
      if (lbe && mnm[k1] != -1) {
        if (doing_search) {
          mn[k1] = MIN(mn[k1], mnm[k1]);
          mx[k1] = MAX(mx[k1], mxm[k1]);
        } else if (mnm[k1] == mn[0] + 1) {
          mn[k1] = MIN(mn[k1], mn[0] + 1);
          mx[k1] = MAX(mx[k1], mn[0] + 1);
        }
      }
    }
    /* into END */
    if ((k == M || lbe) && mn[k] <= mx[k] && h->esc[k] != NINF) {
      endn = MIN(endn, mn[k]);
      endx = MAX(endx, mx[k]);
    }
    if (k == M) {
      if (dn[k] <= dx[k] && h->tdm[k] != NINF) {
        endn = MIN(endn, dn[k]);
        endx = MAX(endx, dx[k]);
      }
      if (in[k] <= ix[k] && h->tim[k] != NINF) {
        endn = MIN(endn, in[k]);
        endx = MAX(endx, in[k]);
      }
      /* the EL-to-END term tests a CM flag bit the CP9 never carries
       * (cp9.rs:3185), so it never contributes */
    }
    /* into D_k+1 */
    if (k < M) {
      const int k1 = k + 1;
      const int src_n[3] = {mn[k], in[k], dn[k]}, src_x[3] = {mx[k], ix[k], dx[k]};
      for (int s = 0; s < 3; s++)
        if (src_n[s] <= src_x[s] && mnd[k1] != -1 && clip(src_n[s], src_x[s], mnd[k1], mxd[k1], &n, &x)) {
          if (dn[k1] != INT_MAX && !lbe && MIN(x, dx[k1]) - MAX(n, dn[k1]) < -1) {
            fill_gap(h, pn, k, n, dn[k1], PRV(mn), PRV(dn));
            filled = 1;
          }
          dn[k1] = MIN(dn[k1], n);
          dx[k1] = MAX(dx[k1], x);
        }
    }
    /* reachable-by-node bands */
    if (mn[k] <= mx[k]) {
      nn_h[k] = MIN(nn_h[k], mn[k]);
      nx_h[k] = MAX(nx_h[k], mx[k]);
      if (k != M) {
        nn_i[k + 1] = MIN(nn_i[k + 1], mn[k] + 1);
        nx_i[k + 1] = MAX(nx_i[k + 1], mx[k] + 1);
      }
      if (k != 0) {
        nn_j[k - 1] = MIN(nn_j[k - 1], mn[k] - 1);
        nx_j[k - 1] = MAX(nx_j[k - 1], mx[k] - 1);
      }
      if ((lbe && k > 0) || k == M) {
        if (doing_search) {
          nn_j[k] = MIN(nn_j[k], mn[k]);
          nx_j[k] = MAX(nx_j[k], mx[k]);
        } else if (mx[k] == j0) {
          nn_j[k] = MIN(nn_j[k], j0);
          nx_j[k] = MAX(nx_j[k], j0);
        }
      }
      if ((lbe && k > 0) || k == 1) {
        if (doing_search) {
          nn_i[k] = MIN(nn_i[k], mn[k]);
          nx_i[k] = MAX(nx_i[k], mx[k]);
          begn = MIN(begn, mn[k]);
          begx = MAX(begx, mx[k]);
        } else if (mn[k] == i0) {
          nn_i[k] = MIN(nn_i[k], i0);
          nx_i[k] = MAX(nx_i[k], i0);
        }
      }
    }
    if (in[k] <= ix[k]) {
      nn_h[k] = MIN(nn_h[k], in[k]);
      nx_h[k] = MAX(nx_h[k], ix[k]);
      if (k != M) {
        nn_i[k + 1] = MIN(nn_i[k + 1], in[k] + 1);
        nx_i[k + 1] = MAX(nx_i[k + 1], ix[k] + 1);
      }
      nn_j[k] = MIN(nn_j[k], in[k] - 1);
      nx_j[k] = MAX(nx_j[k], ix[k] - 1);
      if (k == 0) {
        begn = MIN(begn, in[k]);
        begx = MAX(begx, ix[k]);
      }
    }
    if (dn[k] <= dx[k]) {
      nn_h[k] = MIN(nn_h[k], dn[k]);
      nx_h[k] = MAX(nx_h[k], dx[k]);
      if (k != M) {
        nn_i[k + 1] = MIN(nn_i[k + 1], dn[k] + 1);
        nx_i[k + 1] = MAX(nx_i[k + 1], dx[k] + 1);
      }
      if (k != 0) {
        nn_j[k - 1] = MIN(nn_j[k - 1], dn[k]);
        nx_j[k - 1] = MAX(nx_j[k - 1], dx[k]);
      }
      if (k == 1) {
        begn = MIN(begn, dn[k] + 1);
        begx = MAX(begx, dx[k] + 1);
      }
    }
    if (!lbe && mn[k] > mx[k] && dn[k] > dx[k]) {
      fix_unreachable(h, pn, k, PRV(nn_h));
      k -= 2;
    } else if (filled) {
      k -= 1;
    }
    k++;
  }
  if (!doing_search) {
    begn = begx = i0;
    endn = endx = j0;
  }
  nn_i[1] = MIN(nn_i[1], begn);
  nx_i[1] = MAX(nx_i[1], begx);
  nn_j[M] = MIN(nn_j[M], endn);
  nx_j[M] = MAX(nx_j[M], endx);
  for (int q = 0; q <= M; q++) {
    if (mn[q] == INT_MAX) mn[q] = -1;
    if (mx[q] == INT_MIN) mx[q] = -2;
    if (in[q] == INT_MAX) in[q] = -1;
    if (ix[q] == INT_MIN) ix[q] = -2;
    if (dn[q] == INT_MAX) dn[q] = -1;
    if (dx[q] == INT_MIN) dx[q] = -2;
    if (lbe) {
      if (nn_i[q] == INT_MAX) nn_i[q] = -1;
      if (nx_i[q] == INT_MIN) nx_i[q] = -2;
      if (nn_j[q] == INT_MAX) nn_j[q] = -1;
      if (nx_j[q] == INT_MIN) nx_j[q] = -2;
    }
  }
  r->mn = mn, r->mx = mx, r->in = in, r->ix = ix, r->dn = dn, r->dx = dx;
  r->nn_i = nn_i, r->nx_i = nx_i, r->nn_j = nn_j, r->nx_j = nx_j;
  /* buf is owned by the caller through r->mn */
}

static int state_left_delta(int st) { return st == MP_ST || st == ML_ST || st == IL_ST; }
static int state_right_delta(int st) { return st == MP_ST || st == MR_ST || st == IR_ST; }
static int state_delta(int st) { return st == MP_ST ? 2 : (st == ML_ST || st == MR_ST || st == IL_ST || st == IR_ST); }

/* cp9.rs:3380-3794 cp9_hmm2ijbands; tb == NULL: do_trunc = FALSE */
void ixo_hmm2ij_bands_tr(const ixo_hbcm *cm, const ixo_hbcp9 *h, int32_t *pn, int i0, int j0, int doing_search,
                         ixo_trband *tb, int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax) {
  const int do_trunc = tb != NULL;
  const int M = cm->M, hmm_m = h->hmm_m;
  const int hmm_localized = h->local_begin || h->local_end || h->el;
  const int cm_localized = cm->local_begin && cm->local_end;
  int *ndidx = node_index(cm);
  int32_t *nd2l = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes),
          *nd2r = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  ixo_hb_emit_map(cm, nd2l, nd2r);
  for (int n = 0; n < cm->nodes; n++) {
    const int nt = cm->ndtype[n];
    if (!(nt == MATP_ND || nt == MATL_ND)) nd2l[n] = -1;
    if (!(nt == MATP_ND || nt == MATR_ND)) nd2r[n] = -1;
  }
  for (int v = 0; v < M; v++) imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
  rarr r;
  enforce_valid_parse(h, pn, i0, j0, doing_search, &r);
  const int *r_mn = r.mn, *r_mx = r.mx, *r_in = r.in, *r_ix = r.ix, *r_dn = r.dn, *r_dx = r.dx, *r_nn_i = r.nn_i,
            *r_nx_i = r.nx_i, *r_nn_j = r.nn_j, *r_nx_j = r.nx_j;
  int *pda = (int *)malloc(sizeof(int) * (size_t)(4 * cm->nodes + 8)), *lpda = (int *)malloc(sizeof(int) * (size_t)(cm->nodes + 8));
  int sp = 0, lsp = 0, lpos = 0, rpos = 0;
  pda[sp++] = 0;
  pda[sp++] = 0;
  while (sp > 0) {
    const int nd = pda[--sp], on_right = pda[--sp];
    const int ndt = cm->ndtype[nd];
    if (on_right) {
      if (ndt == BIF_ND) {
        const int v = cm->nodemap[nd], w = cm->cfirst[v], y = cm->cnum[v];
        imin[v] = imin[w] != -1 ? imin[w] : imin[y];
        imax[v] = imax[w] != -2 ? imax[w] : imax[y];
        jmin[v] = jmin[y] != -1 ? jmin[y] : jmin[w];
        jmax[v] = jmax[y] != -2 ? jmax[y] : jmax[w];
        if (!do_trunc &&
            (imin[w] == -1 || jmin[w] == -1 || imin[y] == -1 || jmin[y] == -1 || imax[w] == -2 || jmax[y] == -2)) {
          imin[v] = imin[w] = imin[y] = -1;
          jmin[v] = jmin[w] = jmin[y] = -1;
          imax[v] = imax[w] = imax[y] = -2;
          jmax[v] = jmax[w] = jmax[y] = -2;
          imin[y + 1] = jmin[y + 1] = -1;
          imax[y + 1] = jmax[y + 1] = -2;
        }
      } else if (ndt == MATP_ND) {
        lpos = nd2l[nd];
        rpos = nd2r[nd];
        int v = cm->nodemap[nd];
        jmin[v] = r_mn[rpos], jmax[v] = r_mx[rpos];
        v++;
        jmin[v] = r_dn[rpos], jmax[v] = r_dx[rpos];
        v++;
        jmin[v] = r_mn[rpos], jmax[v] = r_mx[rpos];
        v++;
        jmin[v] = r_dn[rpos], jmax[v] = r_dx[rpos];
        v++;
        jmin[v] = r_nn_j[rpos - 1], jmax[v] = r_nx_j[rpos - 1];
        v++;
        jmin[v] = r_in[rpos - 1], jmax[v] = r_ix[rpos - 1];
        imin[v] = r_nn_i[lpos + 1], imax[v] = r_nx_i[lpos + 1];
      } else if (ndt == MATL_ND) {
        lpos = nd2l[nd];
        int v = cm->nodemap[nd];
        for (int q = 0; q < 3; q++, v++) jmin[v] = r_nn_j[rpos], jmax[v] = r_nx_j[rpos];
      } else if (ndt == MATR_ND) {
        rpos = nd2r[nd];
        int v = cm->nodemap[nd];
        jmin[v] = r_mn[rpos], jmax[v] = r_mx[rpos];
        imin[v] = r_nn_i[lpos], imax[v] = r_nx_i[lpos];
        v++;
        jmin[v] = r_dn[rpos], jmax[v] = r_dx[rpos];
        imin[v] = r_nn_i[lpos], imax[v] = r_nx_i[lpos];
        v++;
        jmin[v] = r_in[rpos - 1], jmax[v] = r_ix[rpos - 1];
        imin[v] = r_nn_i[lpos], imax[v] = r_nx_i[lpos];
      } else if (ndt == BEGL_ND || ndt == BEGR_ND) {
        const int v = cm->nodemap[nd];
        imin[v] = jmin[v] = INT_MAX;
        imax[v] = jmax[v] = INT_MIN;
        for (int y = cm->cfirst[v]; y < cm->cfirst[v] + cm->cnum[v]; y++) {
          if (imin[y] != -1) imin[v] = MIN(imin[v], imin[y]), imax[v] = MAX(imax[v], imax[y]);
          if (jmin[y] != -1) jmin[v] = MIN(jmin[v], jmin[y]), jmax[v] = MAX(jmax[v], jmax[y]);
        }
        if (imin[v] == INT_MAX) imin[v] = -1, imax[v] = -2;
        if (jmin[v] == INT_MAX) jmin[v] = -1, jmax[v] = -2;
        if (ndt == BEGR_ND) {
          const int vil = v + 1;
          imin[vil] = r_in[lpos - 1];
          imax[vil] = r_ix[lpos - 1];
          if (imin[v] != -1 && imin[vil] != -1) {
            imin[v] = MIN(imin[v], imin[vil]);
            jmin[vil] = jmin[v] == -1 ? -1 : MAX(jmin[v], i0);
            jmax[vil] = jmax[v];
          } else {
            imin[vil] = jmin[vil] = -1;
            imax[vil] = jmax[vil] = -2;
          }
          lpos = lpda[--lsp];
        } else {
          lpda[lsp++] = lpos;
          lpos = rpos + 1;
        }
      } else if (ndt == END_ND) {
        const int v = cm->nodemap[nd];
        imin[v] = r_nn_i[lpos];
        imax[v] = r_nx_i[lpos] == -2 ? r_nx_i[lpos] : MIN(r_nx_i[lpos] + 1, j0 + 1);
        if (r_in[lpos] != -1) {
          imin[v] = MIN(imin[v], MAX(r_in[lpos] - 1, i0));
          imax[v] = MAX(imax[v], MAX(r_ix[lpos] - 1, i0));
        }
        rpos = lpos;
        if (imin[v] != -1) jmin[v] = imin[v] - 1, jmax[v] = imax[v] - 1;
      } else if (ndt == ROOT_ND) {
        int v = cm->nodemap[nd];
        imin[v] = r_nn_i[1], imax[v] = r_nx_i[1];
        jmin[v] = r_nn_j[hmm_m], jmax[v] = r_nx_j[hmm_m];
        v++;
        imin[v] = r_in[0], imax[v] = r_ix[0];
        jmin[v] = r_nn_j[hmm_m] == -1 ? -1 : MAX(r_nn_j[hmm_m], i0);
        jmax[v] = r_nx_j[hmm_m];
        if (r_in[hmm_m] != -1) {
          jmin[v] = MIN(jmin[v], r_in[hmm_m]);
          jmax[v] = MIN(jmax[v], r_ix[hmm_m]);
        }
        v++;
        if (r_in[hmm_m] != -1) {
          imin[v] = r_nn_i[1], imax[v] = r_nx_i[1];
          if (imin[v - 1] != -1) {
            imin[v] = MIN(imin[v], imin[v - 1] + 1);
            imax[v] = MAX(imax[v], imax[v - 1] + 1);
          }
          jmin[v] = r_in[hmm_m], jmax[v] = r_ix[hmm_m];
        }
      }
    } else {
      if (ndt == MATP_ND) {
        lpos = nd2l[nd];
        int v = cm->nodemap[nd];
        imin[v] = r_mn[lpos], imax[v] = r_mx[lpos];
        v++;
        imin[v] = r_mn[lpos], imax[v] = r_mx[lpos];
        v++;
        imin[v] = r_dn[lpos] == -1 ? -1 : r_dn[lpos] + 1, imax[v] = r_dx[lpos] == -2 ? -2 : r_dx[lpos] + 1;
        v++;
        imin[v] = r_dn[lpos] == -1 ? -1 : r_dn[lpos] + 1, imax[v] = r_dx[lpos] == -2 ? -2 : r_dx[lpos] + 1;
        v++;
        imin[v] = r_in[lpos], imax[v] = r_ix[lpos];
      } else if (ndt == MATL_ND) {
        lpos = nd2l[nd];
        int v = cm->nodemap[nd];
        imin[v] = r_mn[lpos], imax[v] = r_mx[lpos];
        v++;
        imin[v] = r_dn[lpos] == -1 ? -1 : r_dn[lpos] + 1, imax[v] = r_dx[lpos] == -2 ? -2 : r_dx[lpos] + 1;
        v++;
        imin[v] = r_in[lpos], imax[v] = r_ix[lpos];
      }
      pda[sp++] = 1;
      pda[sp++] = nd;
      if (ndt == BIF_ND) {
        const int v = cm->nodemap[nd];
        pda[sp++] = 0;
        pda[sp++] = ndidx[cm->cnum[v]];
        pda[sp++] = 0;
        pda[sp++] = ndidx[cm->cfirst[v]];
      } else if (ndt != END_ND) {
        pda[sp++] = 0;
        pda[sp++] = nd + 1;
      }
    }
  }
  if (do_trunc) /* cp9.rs:3646-3661: L / R / T marginal alignments widen the bands */
    for (int v = 0; v < M; v++) {
      if (tb->lvalid[v] || tb->tvalid[v]) {
        jmin[v] = jmin[v] == -1 ? tb->lmarg_jmin : MIN(jmin[v], tb->lmarg_jmin);
        jmax[v] = jmax[v] == -2 ? tb->lmarg_jmax : MAX(jmax[v], tb->lmarg_jmax);
      }
      if (tb->rvalid[v] || tb->tvalid[v]) {
        imin[v] = imin[v] == -1 ? tb->rmarg_imin : MIN(imin[v], tb->rmarg_imin);
        imax[v] = imax[v] == -2 ? tb->rmarg_imax : MAX(imax[v], tb->rmarg_imax);
      }
    }
  if (!doing_search) {
    imin[0] = i0;
    if (imin[1] != -1) imin[1] = i0;
    jmax[0] = j0;
    if (jmin[1] != -1) jmax[1] = j0;
    if (jmin[2] != -1) jmax[2] = j0;
  }
  for (int v = 0; v < M; v++) {
    if (imin[v] == -1 || jmin[v] == -1) imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
    if (v + 1 < M && cm->sttype[v + 1] == E_ST) /* detached (cp9.rs:446) */
      imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
    if (do_trunc) continue;
    if (cm->sttype[v] == MP_ST) {
      if (jmax[v] == i0) imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
      else if (jmin[v] == i0) jmin[v] += 1;
    }
    if (state_left_delta(cm->sttype[v]) && imin[v] != -1) {
      if (jmax[v] == i0 - 1) imin[v] = jmin[v] = -1, imax[v] = jmax[v] = -2;
      else if (jmin[v] == i0 - 1) jmin[v] = i0;
    }
  }
  if (hmm_localized && cm_localized) {
    if (do_trunc) tb->jvalid[0] = 1;
    if (imin[0] == -1) imin[0] = imax[0] = i0, jmin[0] = jmax[0] = j0;
    int nd1 = 1;
    if (i0 == j0)
      while (nd1 < cm->nodes && cm->ndtype[nd1] == MATP_ND) nd1++;
    if (cm->ndtype[nd1] == BIF_ND) {
      const int v = cm->nodemap[nd1], w = cm->cfirst[v], y = cm->cnum[v];
      if (do_trunc) tb->jvalid[v] = tb->jvalid[w] = tb->jvalid[y] = 1;
      if (imin[v] != -1 && imin[w] != -1 && imin[y] != -1) {
        if (!doing_search) {
          imin[v] = MIN(imin[v], i0), imax[v] = MAX(imax[v], i0);
          jmin[v] = MIN(jmin[v], j0), jmax[v] = MAX(jmax[v], j0);
          imin[w] = imin[v], imax[w] = imax[v];
          jmax[w] = MAX(jmax[w], MIN(j0, imax[w]));
          jmin[y] = jmin[v], jmax[y] = jmax[v];
          imin[y] = MIN(imin[y], MAX(i0, jmin[y]));
          imin[y] = MIN(imin[y], MAX(i0, jmax[w] + 1));
          imax[y] = MAX(imin[y], imax[y]);
        } else {
          imin[y] = MIN(imin[y], MAX(i0, jmax[w] + 1));
          imax[y] = MAX(imin[y], imax[y]);
        }
      } else if (!doing_search) {
        imin[v] = imax[v] = i0, jmin[v] = jmax[v] = j0;
        imin[w] = imax[w] = i0, jmin[w] = jmax[w] = j0 - 1;
        imin[y] = imax[y] = j0, jmin[y] = jmax[y] = j0;
      } else {
        imin[v] = imax[v] = imin[0], jmin[v] = jmax[v] = jmax[0];
        imin[w] = imax[w] = imin[0], jmin[w] = jmax[w] = jmax[0] - 1;
        imin[y] = imax[y] = jmax[0], jmin[y] = jmax[y] = jmax[0];
      }
    } else {
      const int v = cm->nodemap[nd1];
      if (do_trunc) tb->jvalid[v] = 1;
      if (!doing_search) {
        if (imin[v] == -1) imin[v] = imax[v] = i0, jmin[v] = jmax[v] = j0;
        else {
          imin[v] = MIN(imin[v], i0), imax[v] = MAX(imax[v], i0);
          jmin[v] = MIN(jmin[v], j0), jmax[v] = MAX(jmax[v], j0);
        }
      } else if (imin[v] == -1) {
        imin[v] = imin[0], imax[v] = imax[0], jmin[v] = jmin[0], jmax[v] = jmax[0];
      } else {
        imin[0] = MIN(imin[0], imin[v]), imax[0] = MAX(imax[0], imax[v]);
        jmin[0] = MIN(jmin[0], jmin[v]), jmax[0] = MAX(jmax[0], jmax[v]);
      }
    }
  }
  free((void *)r.mn);
  free(pda);
  free(lpda);
  free(nd2l);
  free(nd2r);
  free(ndidx);
  (void)state_right_delta;
}

void ixo_hmm2ij_bands(const ixo_hbcm *cm, const ixo_hbcp9 *h, int32_t *pn, int i0, int j0, int doing_search,
                      int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax) {
  ixo_hmm2ij_bands_tr(cm, h, pn, i0, j0, doing_search, NULL, imin, imax, jmin, jmax);
}

/* Score2Prob(sc, 1.0) (cp9.rs:1741) */
static float score2prob(int32_t sc) {
  return sc == -987654321 ? 0.0f : (float)exp((double)((float)sc / 1000.0f) * 0.69314718);
}

/* cp9.rs:1755-1916 cp9_predict_start_and_end_positions */
void ixo_tr_predict_start_end(const int32_t *pn, const int32_t *occ, int hmm_m, int i0, int j0, float thresh1,
                              float thresh2, ixo_trband *tb) {
  const int H1 = hmm_m + 1;
  const int32_t *pmin_m = pn, *pmax_m = pn + H1, *pmin_i = pn + 2 * H1, *pmax_i = pn + 3 * H1, *pmin_d = pn + 4 * H1,
                *pmax_d = pn + 5 * H1;
  int sp1 = -1, sp2 = -1, ep1, ep2, k = 1;
  while (k <= hmm_m && (sp1 == -1 || sp2 == -1)) {
    if (!(pmin_m[k] == -1 && pmin_i[k] == -1 && pmin_d[k] == -1)) {
      const float p = score2prob(occ[k]);
      if (sp1 == -1 && p > thresh1) sp1 = k;
      if (sp2 == -1 && p > thresh2) sp2 = k;
    }
    k++;
  }
  if (k == hmm_m + 1) {
    if (sp1 == -1) sp1 = hmm_m + 1;
    if (sp2 == -1) sp2 = hmm_m + 1;
  }
  if (sp1 == hmm_m + 1 && sp2 == hmm_m + 1) {
    ep1 = ep2 = 0;
  } else {
    ep1 = ep2 = -1;
    k = hmm_m;
    while (k >= 1 && (ep1 == -1 || ep2 == -1)) {
      if (!(pmin_m[k] == -1 && pmin_i[k] == -1 && pmin_d[k] == -1)) {
        const float p = score2prob(occ[k]);
        if (ep1 == -1 && p > thresh1) ep1 = k;
        if (ep2 == -1 && p > thresh2) ep2 = k;
      }
      k--;
    }
    if (k == 0) {
      if (ep1 == -1) ep1 = 0;
      if (ep2 == -1) ep2 = 0;
    }
  }
  tb->sp1 = sp1, tb->sp2 = sp2, tb->ep1 = ep1, tb->ep2 = ep2;
  const int mm1 = hmm_m + 1;
  /* the R marginal i band from the start nodes, the L marginal j band from
   * the end nodes (a delete's j is one less: cp9.rs:1866-1912) */
  if (sp1 == mm1) {
    tb->rmarg_imin = i0;
    tb->rmarg_imax = j0;
  } else {
    int lo = INT_MAX, hi = INT_MIN;
    const int s[2] = {sp1, sp2};
    for (int q = 0; q < 2; q++) {
      if (s[q] == mm1) continue;
      if (pmin_m[s[q]] >= 0) lo = MIN(lo, pmin_m[s[q]]);
      if (pmin_i[s[q]] >= 0) lo = MIN(lo, pmin_i[s[q]]);
      if (pmin_d[s[q]] >= 0) lo = MIN(lo, pmin_d[s[q]]);
      if (pmax_m[s[q]] >= 0) hi = MAX(hi, pmax_m[s[q]]);
      if (pmax_i[s[q]] >= 0) hi = MAX(hi, pmax_i[s[q]]);
      if (pmax_d[s[q]] >= 0) hi = MAX(hi, pmax_d[s[q]]);
    }
    if (lo == INT_MAX || sp2 == mm1) lo = i0;
    if (hi == INT_MIN || sp2 == mm1) hi = j0 + 1;
    tb->rmarg_imin = MIN(j0 + 1, MAX(i0, lo));
    tb->rmarg_imax = MIN(j0 + 1, MAX(i0, hi));
  }
  if (ep1 == 0) {
    tb->lmarg_jmin = i0 - 1;
    tb->lmarg_jmax = j0;
  } else {
    int lo = INT_MAX, hi = INT_MIN;
    const int e[2] = {ep1, ep2};
    for (int q = 0; q < 2; q++) {
      if (e[q] == 0) continue;
      if (pmin_m[e[q]] >= 0) lo = MIN(lo, pmin_m[e[q]]);
      if (pmin_i[e[q]] >= 0) lo = MIN(lo, pmin_i[e[q]]);
      if (pmin_d[e[q]] >= 0) lo = MIN(lo, pmin_d[e[q]] - 1);
      if (pmax_m[e[q]] >= 0) hi = MAX(hi, pmax_m[e[q]]);
      if (pmax_i[e[q]] >= 0) hi = MAX(hi, pmax_i[e[q]]);
      if (pmax_d[e[q]] >= 0) hi = MAX(hi, pmax_d[e[q]] - 1);
    }
    if (lo == INT_MAX || ep2 == 0) lo = i0 - 1;
    if (hi == INT_MIN || ep2 == 0) hi = j0;
    tb->lmarg_jmin = MIN(j0, MAX(i0 - 1, lo));
    tb->lmarg_jmax = MIN(j0, MAX(i0 - 1, hi));
  }
}

/* cp9.rs:1918-1990 cp9_marginal_candidates_from_start_end_positions */
void ixo_tr_candidates(const ixo_hbcm *cm, int pass_idx, ixo_trband *tb) {
  const int al = pass_idx == 3 || pass_idx == 4 || pass_idx == 5; /* fill L: 3P, 5P&3P */
  const int ar = pass_idx == 2 || pass_idx == 4 || pass_idx == 5; /* fill R: 5P, 5P&3P */
  const int at = pass_idx == 4 || pass_idx == 5;
  int *ndidx = node_index(cm);
  int32_t *lpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes),
          *rpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  ixo_hb_emit_map(cm, lpos, rpos);
  for (int v = 0; v < cm->M; v++) {
    const int nd = ndidx[v], nt = cm->ndtype[nd];
    const int lp = (nt == MATP_ND || nt == MATL_ND) ? lpos[nd] : lpos[nd] + 1;
    const int rp = (nt == MATP_ND || nt == MATR_ND) ? rpos[nd] : rpos[nd] - 1;
    tb->jvalid[v] = lp >= tb->sp1 && rp <= tb->ep1;
    tb->lvalid[v] = al && lp >= tb->sp1 && lp <= tb->ep1 && rp > tb->ep2;
    tb->rvalid[v] = ar && rp <= tb->ep1 && rp >= tb->sp1 && lp < tb->sp2;
    tb->tvalid[v] = cm->sttype[v] == B_ST ? (at && lp < tb->sp2 && rp > tb->ep2) : 0;
  }
  for (int v = 0; v < cm->M; v++) {
    const int st = cm->sttype[v];
    if (st == B_ST || st == MP_ST || st == ML_ST || st == IL_ST || st == MR_ST || st == IR_ST)
      if (tb->jvalid[v]) tb->jvalid[0] = 1;
    if (st == B_ST || st == MP_ST || st == ML_ST || st == IL_ST)
      if (tb->lvalid[v]) tb->lvalid[0] = 1;
    if (st == B_ST || st == MP_ST || st == MR_ST || st == IR_ST)
      if (tb->rvalid[v]) tb->rvalid[0] = 1;
    if (st == B_ST && tb->tvalid[v]) tb->tvalid[0] = 1;
    if (tb->jvalid[0] && tb->lvalid[0] && tb->rvalid[0] && tb->tvalid[0]) break;
  }
  free(lpos);
  free(rpos);
  free(ndidx);
}

/* cp9.rs:3815-3850 ij2d_bands; hd_off[v] = first row of state v (M+1
 * entries); returns the number of rows (sum of the j-band widths) */
int64_t ixo_hb_ij2d(const ixo_hbcm *cm, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                       const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax,
                       int64_t cap) {
  int64_t row = 0;
  for (int v = 0; v < cm->M; v++) {
    hd_off[v] = (int32_t)row;
    const int st = cm->sttype[v];
    const int n = jmax[v] - jmin[v] >= 0 ? jmax[v] - jmin[v] + 1 : 0;
    const int dn = do_trunc ? MAX(state_left_delta(st), state_right_delta(st)) : state_delta(st);
    for (int jp = 0; jp < n; jp++, row++) {
      if (row >= cap) continue;
      if (st == E_ST) {
        hdmin[row] = hdmax[row] = 0;
        continue;
      }
      const int j = jp + jmin[v];
      const int hdn = j - imax[v] + 1, hdx = j - imin[v] + 1;
      if (hdx < dn) hdmin[row] = -1, hdmax[row] = -2;
      else hdmin[row] = MAX(hdn, dn), hdmax[row] = hdx;
    }
  }
  hd_off[cm->M] = (int32_t)row;
  return row;
}

/* cp9.rs:3854-3879 cm_hb_mx_size_needed: (cells, Mb) with the EL deck when
 * the CM has local ends */
double ixo_hb_mx_mb(const ixo_hbcm *cm, const int32_t *jmin, const int32_t *jmax, const int32_t *hd_off,
                    const int32_t *hdmin, const int32_t *hdmax, int L, int64_t *ncells_out) {
  int64_t ncells = 0;
  float mb = (float)(64 + (cm->M + 1) * 8 + (cm->M + 1) * 4);
  for (int v = 0; v < cm->M; v++) {
    const int jbw = jmax[v] - jmin[v];
    mb += (float)(8 * (jbw + 1));
    for (int jp = 0; jp <= jbw; jp++) ncells += hdmax[hd_off[v] + jp] - hdmin[hd_off[v] + jp] + 1;
  }
  if (cm->local_end) ncells += (int64_t)((double)(L + 2) * (double)(L + 1) * 0.5);
  mb += 4.0f * (float)ncells;
  mb *= 0.000001f;
  if (ncells_out) *ncells_out = ncells;
  return mb;
}
