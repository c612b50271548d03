/*
 * qdb_impl.h -- body of the QDB scanner oracle, included by qdb.c once per
 * semiring with S (value type), S_ZERO, S_COMB, S_VMAX, S_ADD, S_BITS,
 * S_POSSIBLE and QDB_SCAN (the function name) defined. Follows
 * cm_nohmm.rs:406-767 generic_scan loop by loop: rolling rows alpha[2][M][W+1],
 * BEGL_S decks alpha_begl[W+1][nbegl][W+1], d = 0 base cases set once,
 * children combined in C's unrolled order (reduce_c_order, cm_nohmm.rs:187).
 */

static S QDB_REDUCE(int cnum, const S *terms, S init) { /* cm_nohmm.rs:187-244 */
  S s;
  switch (cnum) {
    case 1: return S_COMB(init, terms[0]);
    case 2: s = S_COMB(terms[1], init); return S_COMB(s, terms[0]);
    case 3: s = S_COMB(terms[2], terms[1]); s = S_COMB(s, init); return S_COMB(s, terms[0]);
    case 4:
      s = S_COMB(terms[3], terms[2]); s = S_COMB(s, init); s = S_COMB(s, terms[1]); return S_COMB(s, terms[0]);
    case 5:
      s = S_COMB(terms[4], terms[3]); s = S_COMB(s, init); s = S_COMB(s, terms[1]); s = S_COMB(s, terms[2]);
      return S_COMB(s, terms[0]);
    case 6:
      s = S_COMB(terms[5], init); s = S_COMB(s, terms[4]); s = S_COMB(s, terms[3]); s = S_COMB(s, terms[2]);
      s = S_COMB(s, terms[1]); return S_COMB(s, terms[0]);
    default:
      s = init;
      for (int k = 0; k < cnum; k++) s = S_COMB(s, terms[k]);
      return s;
  }
}

static int64_t QDB_SCAN(const ixo_cm *cm, const S *tsc, const S *oesc, const S *endsc, const S *beginsc, S el_self,
                        int local, const int32_t *dmin_in, const int32_t *dmax_in, const uint8_t *dsq, int i0,
                        int j0, float cutoff, int do_null3, ixo_hit *out, int64_t cap, float *root_out) {
  const int M = cm->M, smx_w = cm->W, L = j0 - i0 + 1;
  const int w = smx_w < L ? smx_w : L, W1 = w + 1;
  int32_t *dmin = (int32_t *)malloc(sizeof(int32_t) * (size_t)M), *dmax = (int32_t *)malloc(sizeof(int32_t) * (size_t)M);
  for (int v = 0; v < M; v++) {
    dmin[v] = dmin_in ? dmin_in[v] : 0;
    dmax[v] = dmax_in ? dmax_in[v] : smx_w;
  }
  int *dn_v = (int *)malloc(sizeof(int) * (size_t)M), *dxcap_v = (int *)malloc(sizeof(int) * (size_t)M);
  for (int v = 0; v < M; v++) {
    const int base = cm->sttype[v] == IXO_MP_ST ? 2 : 1;
    int dn = dmin[v] > base ? dmin[v] : base;
    if (dn > smx_w) dn = smx_w;
    dn_v[v] = dn;
    dxcap_v[v] = dmax[v] < smx_w ? dmax[v] : smx_w;
  }
  int *begl_idx = (int *)malloc(sizeof(int) * (size_t)M);
  int nbegl = 0;
  for (int v = 0; v < M; v++) begl_idx[v] = is_begl(cm, v) ? nbegl++ : -1;
  S *el_sca = (S *)malloc(sizeof(S) * (size_t)W1);
  el_sca[0] = (S)0;
  for (int d = 1; d <= w; d++) el_sca[d] = S_ADD(el_sca[d - 1], el_self);
  S *init_sc = (S *)malloc(sizeof(S) * (size_t)M * W1);
  for (int v = 0; v < M; v++)
    for (int d = 0; d <= w; d++)
      init_sc[(size_t)v * W1 + d] = not_impossible(S_BITS(endsc[v])) ? S_ADD(el_sca[d], endsc[v]) : S_ZERO;
  S *alpha = (S *)malloc(sizeof(S) * 2 * (size_t)M * W1);
  S *begl = (S *)malloc(sizeof(S) * (size_t)W1 * (nbegl ? nbegl : 1) * W1);
  for (size_t k = 0; k < 2 * (size_t)M * W1; k++) alpha[k] = S_ZERO;
  for (size_t k = 0; k < (size_t)W1 * (nbegl ? nbegl : 1) * W1; k++) begl[k] = S_ZERO;
#define A(r, v, d) alpha[((size_t)(r) * M + (v)) * W1 + (d)]
#define BG(jr, b, d) begl[((size_t)(jr) * nbegl + (b)) * W1 + (d)]
  /* d = 0 base cases (cm_nohmm.rs:489-532) */
  for (int v = M - 1; v >= 0; v--) {
    const int st = cm->sttype[v];
    if (begl_idx[v] < 0) {
      if (st == IXO_E_ST) {
        A(0, v, 0) = (S)0;
        A(1, v, 0) = (S)0;
      } else if (st == IXO_S_ST || st == IXO_D_ST) {
        S a0 = endsc[v];
        for (int yo = 0; yo < cm->cnum[v]; yo++) a0 = S_VMAX(a0, S_ADD(A(0, cm->cfirst[v] + yo, 0), tsc[v * MAXC + yo]));
        a0 = S_VMAX(a0, S_ZERO);
        A(0, v, 0) = a0;
        A(1, v, 0) = a0;
      } else if (st == IXO_B_ST) {
        const S a0 = S_ADD(BG(0, begl_idx[cm->cfirst[v]], 0), A(0, cm->cnum[v], 0));
        A(0, v, 0) = a0;
        A(1, v, 0) = a0;
      } else {
        A(1, v, 0) = A(0, v, 0);
      }
    } else {
      S a0 = endsc[v];
      for (int yo = 0; yo < cm->cnum[v]; yo++) a0 = S_VMAX(a0, S_ADD(A(0, cm->cfirst[v] + yo, 0), tsc[v * MAXC + yo]));
      a0 = S_VMAX(a0, S_ZERO);
      for (int j = 0; j <= w; j++) BG(j, begl_idx[v], 0) = a0;
    }
  }
  double *act = (double *)calloc((size_t)W1 * 4, sizeof(double));
  S *sc_v = (S *)malloc(sizeof(S) * (size_t)W1);
  float *bestsc = (float *)malloc(sizeof(float) * (size_t)W1);
  int *bestr = (int *)malloc(sizeof(int) * (size_t)W1);
  int *jp_wa = (int *)malloc(sizeof(int) * (size_t)W1);
  for (int d = 0; d <= w; d++) bestr[d] = -1;
  ixo_hit *tmp = NULL;
  int64_t ntmp = 0, tcap = 0;
  for (int j = i0; j <= j0; j++) {
    const int jpg = j - i0 + 1, cur = j & 1, prv = (j - 1) & 1;
    const int jrow = jpg < w ? jpg : w;
    for (int d = 0; d <= w; d++) jp_wa[d] = (((j - d) % W1) + W1) % W1;
    if (do_null3) {
      const int src = (jpg - 1) % W1, dst = jpg % W1;
      for (int a = 0; a < 4; a++) act[dst * 4 + a] = act[src * 4 + a];
      if (dsq[j] < 4) act[dst * 4 + dsq[j]] += 1.0;
    }
    for (int v = M - 1; v >= 1; v--) {
      const int st = cm->sttype[v];
      if (st == IXO_E_ST) continue;
      const int sd = sdelta_q(st), sdr = srdelta_q(st);
      const int isbegl = begl_idx[v] >= 0;
      const int jp_v = isbegl ? ((j % W1) + W1) % W1 : cur;
      const int jp_y = sdr > 0 ? prv : cur;
      const int dn = dn_v[v];
      const int dx = jrow < dxcap_v[v] ? jrow : dxcap_v[v];
      const S *esc = oesc + (size_t)v * KP * KP;
      const S esc_j = (st == IXO_IR_ST || st == IXO_MR_ST) ? esc[dsq[j]] : S_ZERO;
      const S *ts = tsc + (size_t)v * MAXC;
      const int cn = cm->cnum[v], y0 = cm->cfirst[v];
      if (st == IXO_B_ST) {
        const int wl = cm->cfirst[v], y = cm->cnum[v], bi = begl_idx[wl];
        for (int d = dn; d <= dx; d++) {
          const int dn_y = dmin[y] < smx_w ? dmin[y] : smx_w, dx_y = dmax[y] < smx_w ? dmax[y] : smx_w;
          const int dn_w = dmin[wl] < smx_w ? dmin[wl] : smx_w, dx_w = dmax[wl] < smx_w ? dmax[wl] : smx_w;
          int kmin = dn_y > d - dx_w ? dn_y : d - dx_w;
          if (kmin < 0) kmin = 0;
          const int kmax = dx_y < d - dn_w ? dx_y : d - dn_w;
          S sc = init_sc[(size_t)v * W1 + (d - sd)];
          for (int k = kmin; k <= kmax; k++) sc = S_COMB(sc, S_ADD(BG(jp_wa[k], bi, d - k), A(jp_y, y, k)));
          A(jp_v, v, d) = sc;
        }
      } else if (isbegl) {
        for (int d = dn; d <= dx; d++) {
          S sc = init_sc[(size_t)v * W1 + (d - sd)];
          for (int yo = 0; yo < cn; yo++) sc = S_COMB(sc, S_ADD(A(jp_y, y0 + yo, d - sd), ts[yo]));
          BG(jp_v, begl_idx[v], d) = sc;
        }
      } else if (st == IXO_IL_ST || st == IXO_IR_ST) {
        int i = j - dn + 1;
        for (int d = dn; d <= dx; d++) {
          S terms[6];
          for (int yo = 0; yo < cn; yo++) terms[yo] = S_ADD(A(jp_y, y0 + yo, d - sd), ts[yo]);
          S sc = QDB_REDUCE(cn, terms, init_sc[(size_t)v * W1 + (d - sd)]);
          if (st == IXO_IL_ST) sc = S_ADD(sc, esc[dsq[i--]]);
          else sc = S_ADD(sc, esc_j);
          A(jp_v, v, d) = sc;
        }
      } else {
        for (int d = dn; d <= dx; d++) {
          S terms[6];
          for (int yo = 0; yo < cn; yo++) terms[yo] = S_ADD(A(jp_y, y0 + yo, d - sd), ts[yo]);
          sc_v[d] = QDB_REDUCE(cn, terms, init_sc[(size_t)v * W1 + (d - sd)]);
        }
        int i = j - dn + 1;
        for (int d = dn; d <= dx; d++, i--) {
          if (st == IXO_ML_ST) A(jp_v, v, d) = S_ADD(sc_v[d], esc[dsq[i]]);
          else if (st == IXO_MR_ST) A(jp_v, v, d) = S_ADD(sc_v[d], esc_j);
          else if (st == IXO_MP_ST) A(jp_v, v, d) = S_ADD(sc_v[d], esc[dsq[i] * KP + dsq[j]]);
          else A(jp_v, v, d) = sc_v[d];
        }
      }
    }
    /* root: its own transitions, then local begins by MAX (cm_nohmm.rs:686-746) */
    const int dn0 = dn_v[0], dx0 = jrow < dxcap_v[0] ? jrow : dxcap_v[0];
    const int r0 = cm->cfirst[0];
    for (int d = dn0; d <= dx0; d++) {
      bestr[d] = 0;
      S a = S_COMB(S_ZERO, S_ADD(A(cur, r0, d), tsc[0]));
      for (int yo = 1; yo < cm->cnum[0]; yo++) a = S_COMB(a, S_ADD(A(cur, r0 + yo, d), tsc[yo]));
      A(cur, 0, d) = a;
    }
    if (local)
      for (int y = 1; y < M; y++) {
        if (!S_POSSIBLE(beginsc[y])) continue;
        int dn = dn0 > dn_v[y] ? dn0 : dn_v[y];
        int dx = dx0 < (jrow < dxcap_v[y] ? jrow : dxcap_v[y]) ? dx0 : (jrow < dxcap_v[y] ? jrow : dxcap_v[y]);
        for (int d = dn; d <= dx; d++) {
          const S cand = S_ADD(begl_idx[y] >= 0 ? BG(jp_wa[0], begl_idx[y], d) : A(cur, y, d), beginsc[y]);
          if (A(cur, 0, d) < cand) {
            A(cur, 0, d) = cand;
            bestr[d] = y;
          }
        }
      }
    for (int d = dn0; d <= dx0; d++) bestsc[d] = S_BITS(A(cur, 0, d));
    if (root_out) /* root scores of row j: [(j - i0) * (W + 1) + d], d <= W */
      for (int d = 0; d <= smx_w; d++)
        root_out[(size_t)(j - i0) * (smx_w + 1) + d] = (d >= dn0 && d <= dx0) ? bestsc[d] : -1e36f;
    /* report_hits_greedily (cm_nohmm.rs:769-813) */
    if (dn0 <= dx0) {
      float maxrep = -1e36f;
      for (int d = dn0 > 1 ? dn0 : 1; d <= dx0; d++) {
        const int i = j - d + 1;
        float hs = bestsc[d], bias = 0.0f;
        if (!(hs > maxrep && hs >= cutoff && not_impossible(hs))) continue;
        int rep = 1;
        if (do_null3) {
          float comp[4];
          const int ip = i - i0 + 1, jj = j - i0 + 1;
          for (int a = 0; a < 4; a++) comp[a] = (float)(act[(jj % W1) * 4 + a] - act[((ip - 1) % W1) * 4 + a]);
          float s = 0.0f;
          for (int a = 0; a < 4; a++) s += comp[a];
          if (s > 0.0f)
            for (int a = 0; a < 4; a++) comp[a] /= s;
          else
            for (int a = 0; a < 4; a++) comp[a] = 0.25f;
          const float corr = ixo_null3_correction(cm->null, comp, d, (float)cm->n3_omega);
          hs -= corr;
          bias = corr;
          rep = hs > maxrep && hs >= cutoff;
        }
        if (rep) {
          if (ntmp == tcap) {
            tcap = tcap ? 2 * tcap : 64;
            tmp = (ixo_hit *)realloc(tmp, sizeof(ixo_hit) * (size_t)tcap);
          }
          tmp[ntmp].i = i;
          tmp[ntmp].j = j;
          tmp[ntmp].score = hs;
          tmp[ntmp].bias = bias;
          ntmp++;
          maxrep = hs;
        }
      }
    }
  }
#undef A
#undef BG
  const int64_t nk = remove_overlaps(tmp, ntmp, out, cap);
  free(tmp); free(dmin); free(dmax); free(dn_v); free(dxcap_v); free(begl_idx); free(el_sca); free(init_sc);
  free(alpha); free(begl); free(act); free(sc_v); free(bestsc); free(bestr); free(jp_wa);
  return nk;
}
