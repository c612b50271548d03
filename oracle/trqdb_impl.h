/*
 * trqdb_impl.h -- body of the QDB / non-banded truncated scanner oracle
 * (test infrastructure only), included by trqdb.c once per semiring with
 * S (value type), S_ZERO, S_COMB (max / ILogsum), S_MAX2 (always max),
 * S_ADD, S_BITS (reported bits), S_POSSIBLE and TRQDB_SCAN (the function
 * name) defined. Follows cm_dpsearch_trunc.rs:1800-2316 generic_tr_scan loop
 * by loop: four marginal planes J/L/R/T in rolling rows [2][M][W+1], BEGL_S
 * decks for J/L/R [W+1][nbegl][W+1], d = 0 base cases set once, dn = 1 for
 * every state and dx = min(jrow, dmax, W), children folded in increasing
 * order, and the root rows never reset between j's (each truncated begin
 * must beat what the same-parity row already holds at d).
 */

static int64_t TRQDB_SCAN(const ixo_cm *cm, const S *tsc, const S *oesc, const S *lmesc, const S *rmesc,
                          const S *endsc, S el_self, const S *pty, const int32_t *dmax_in, const uint8_t *dsq,
                          int i0, int j0, int pass_idx, int local, float cutoff, int do_null3, float *vsc,
                          int *vmode, float *bsc_full_out, ixo_trhit *hits, int64_t cap, float *best_out) {
  const int M = cm->M, smx_w = cm->W, L = j0 - i0 + 1;
  const int w = smx_w < L ? smx_w : L, W1 = w + 1;
  const int fl = tq_fill_l(pass_idx), fr = tq_fill_r(pass_idx), ft = tq_fill_t(pass_idx);
  const int nonbanded = dmax_in == NULL;
  int *dmax = (int *)malloc(sizeof(int) * (size_t)M), *dxcap = (int *)malloc(sizeof(int) * (size_t)M);
  for (int v = 0; v < M; v++) {
    dmax[v] = nonbanded ? smx_w : dmax_in[v];
    dxcap[v] = dmax[v] < smx_w ? dmax[v] : smx_w;
  }
  const int dn_all = 1;
  int *begl_idx = (int *)malloc(sizeof(int) * (size_t)M);
  int nbegl = 0;
  for (int v = 0; v < M; v++) begl_idx[v] = tq_is_begl(cm, v) ? nbegl++ : -1;
  const int NB = nbegl ? nbegl : 1;
  /* init_sc[v][d] = el*d + endsc[v] where endsc is possible (:1851-1865) */
  S *el_sca = (S *)malloc(sizeof(S) * (size_t)W1);
  el_sca[0] = (S)0;
  for (int d = 1; d <= w; d++) el_sca[d] = S_ADD(el_sca[d - 1], el_self);
  S *init_sc = (S *)malloc(sizeof(S) * (size_t)M * W1);
  for (int v = 0; v < M; v++)
    for (int d = 0; d <= w; d++) init_sc[(size_t)v * W1 + d] = S_POSSIBLE(endsc[v]) ? S_ADD(el_sca[d], endsc[v]) : S_ZERO;
  const size_t plane = 2 * (size_t)M * W1, deck = (size_t)W1 * NB * W1;
  S *ja = (S *)malloc(sizeof(S) * plane * 4), *la = ja + plane, *ra = la + plane, *ta = ra + plane;
  S *jb = (S *)malloc(sizeof(S) * deck * 3), *lb = jb + deck, *rb = lb + deck;
  for (size_t k = 0; k < plane * 4; k++) ja[k] = S_ZERO;
  for (size_t k = 0; k < deck * 3; k++) jb[k] = S_ZERO;
#define PL(P, r, v, d) P[((size_t)(r) * M + (v)) * W1 + (d)]
#define DK(P, jr, b, d) P[((size_t)(jr) * NB + (b)) * W1 + (d)]
#define INIT(v, d) init_sc[(size_t)(v) * W1 + (d)]
  /* d = 0 base cases, set once (:1873-1922) */
  for (int v = M - 1; v >= 0; v--) {
    const int st = cm->sttype[v];
    if (begl_idx[v] < 0) {
      if (st == IXO_E_ST) {
        PL(ja, 0, v, 0) = PL(ja, 1, v, 0) = (S)0;
        PL(la, 0, v, 0) = PL(la, 1, v, 0) = (S)0;
        PL(ra, 0, v, 0) = PL(ra, 1, v, 0) = (S)0;
      } else if (st == IXO_S_ST || st == IXO_D_ST) {
        S a0 = endsc[v];
        for (int yo = 0; yo < cm->cnum[v]; yo++)
          a0 = S_MAX2(a0, S_ADD(PL(ja, 0, cm->cfirst[v] + yo, 0), tsc[(size_t)v * MAXC + yo]));
        a0 = S_MAX2(a0, S_ZERO);
        PL(ja, 0, v, 0) = PL(ja, 1, v, 0) = a0;
      } else if (st == IXO_B_ST) {
        const S a0 = S_ADD(DK(jb, 0, begl_idx[cm->cfirst[v]], 0), PL(ja, 0, cm->cnum[v], 0));
        PL(ja, 0, v, 0) = PL(ja, 1, v, 0) = a0;
      }
    } else {
      S a0 = endsc[v];
      for (int yo = 0; yo < cm->cnum[v]; yo++)
        a0 = S_MAX2(a0, S_ADD(PL(ja, 0, cm->cfirst[v] + yo, 0), tsc[(size_t)v * MAXC + yo]));
      a0 = S_MAX2(a0, S_ZERO);
      for (int j = 0; j <= w; j++) DK(jb, j, begl_idx[v], 0) = a0;
    }
  }
  int32_t *lpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  int32_t *rpos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cm->nodes);
  ixo_emit_map(cm, lpos, rpos);
  double *act = (double *)calloc((size_t)W1 * 4, sizeof(double));
  int *jp_wa = (int *)malloc(sizeof(int) * (size_t)W1);
  float *bestsc = (float *)malloc(sizeof(float) * (size_t)W1);
  int *bestr = (int *)malloc(sizeof(int) * (size_t)W1), *bestmode = (int *)malloc(sizeof(int) * (size_t)W1);
  float vsc_root = TQ_IMPOSSIBLE, bsc_full = TQ_IMPOSSIBLE;
  int vmode_root = IXO_TRMODE_UNKNOWN;
  ixo_trhit *tmp = NULL;
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
      const int sd = tq_sd(st), sdr = tq_sdr(st);
      const int isbegl = begl_idx[v] >= 0;
      const int jp_v = isbegl ? ((j % W1) + W1) % W1 : cur;
      const int jp_y = sdr > 0 ? prv : cur, jq_y = sdr > 0 ? cur : prv;
      const int dn = dn_all, dx = jrow < dxcap[v] ? jrow : dxcap[v];
      const S *esc = oesc + (size_t)v * KP * KP;
      const S *ts = tsc + (size_t)v * MAXC;
      const S *lm = lmesc + (size_t)v * KP, *rm = rmesc + (size_t)v * KP;
      const S esc_j = (st == IXO_IR_ST || st == IXO_MR_ST) ? esc[dsq[j]] : S_ZERO;
      const S rmesc_j = (st == IXO_IR_ST || st == IXO_MR_ST || st == IXO_MP_ST) ? rm[dsq[j]] : S_ZERO;
      const int cn = cm->cnum[v], y0 = cm->cfirst[v];
      if (st == IXO_B_ST) { /* :1978-2045 */
        const int wl = cm->cfirst[v], y = cm->cnum[v], bi = begl_idx[wl];
        const int dx_y = dmax[y] < smx_w ? dmax[y] : smx_w, dx_w = dmax[wl] < smx_w ? dmax[wl] : smx_w;
        for (int d = dn; d <= dx; d++) {
          int kmin = 0, kmax = d;
          if (!nonbanded) {
            kmin = d - dx_w > 0 ? d - dx_w : 0;
            kmax = dx_y < d ? dx_y : d;
          }
          S jsc = INIT(v, d - sd);
          S lsc = fl ? INIT(v, d - sd) : S_ZERO, rsc = fr ? INIT(v, d - sd) : S_ZERO;
          S tsv = ft ? INIT(v, d - sd) : S_ZERO;
          for (int k = kmin; k <= kmax; k++) {
            const S lft = DK(jb, jp_wa[k], bi, d - k);
            jsc = S_COMB(jsc, S_ADD(lft, PL(ja, jp_y, y, k)));
            if (fl) lsc = S_COMB(lsc, S_ADD(lft, PL(la, jp_y, y, k)));
            if (fr) rsc = S_COMB(rsc, S_ADD(DK(rb, jp_wa[k], bi, d - k), PL(ja, jp_y, y, k)));
          }
          if (ft) {
            const int kn = kmin > 1 ? kmin : 1, kx = d - 1 < kmax ? d - 1 : kmax;
            for (int k = kn; k <= kx; k++) tsv = S_COMB(tsv, S_ADD(DK(rb, jp_wa[k], bi, d - k), PL(la, jp_y, y, k)));
          }
          PL(ja, jp_v, v, d) = jsc;
          if (ft) PL(ta, jp_v, v, d) = tsv;
          if (fl)
            PL(la, jp_v, v, d) =
                kmin == 0 ? S_COMB(lsc, S_MAX2(DK(jb, jp_wa[0], bi, d), DK(lb, jp_wa[0], bi, d))) : lsc;
          if (fr)
            PL(ra, jp_v, v, d) = kmax == d ? S_COMB(rsc, S_MAX2(PL(ja, jp_y, y, d), PL(ra, jp_y, y, d))) : rsc;
        }
      } else if (isbegl) { /* :2046-2067 */
        const int bi = begl_idx[v];
        for (int d = dn; d <= dx; d++) {
          S jsc = INIT(v, d - sd), lsc = fl ? INIT(v, d - sd) : S_ZERO, rsc = fr ? INIT(v, d - sd) : S_ZERO;
          for (int yo = 0; yo < cn; yo++) {
            jsc = S_COMB(jsc, S_ADD(PL(ja, jp_y, y0 + yo, d - sd), ts[yo]));
            if (fl) lsc = S_COMB(lsc, S_ADD(PL(la, jp_y, y0 + yo, d - sd), ts[yo]));
            if (fr) rsc = S_COMB(rsc, S_ADD(PL(ra, jp_y, y0 + yo, d - sd), ts[yo]));
          }
          DK(jb, jp_v, bi, d) = jsc;
          if (fl) DK(lb, jp_v, bi, d) = lsc;
          if (fr) DK(rb, jp_v, bi, d) = rsc;
        }
      } else if (st == IXO_ML_ST || st == IXO_IL_ST) { /* EMITLEFT :2068-2106 */
        if (!tq_detached(cm, v)) {
          const int ryo0 = st == IXO_IL_ST ? 1 : 0;
          int i = j - dn + 1;
          for (int d = dn; d <= dx; d++, i--) {
            S jsc = INIT(v, d - sd), lsc = fl ? INIT(v, d - sd) : S_ZERO;
            for (int yo = 0; yo < cn; yo++) {
              jsc = S_COMB(jsc, S_ADD(PL(ja, jp_y, y0 + yo, d - sd), ts[yo]));
              if (fl) lsc = S_COMB(lsc, S_ADD(PL(la, jp_y, y0 + yo, d - sd), ts[yo]));
            }
            const S e = esc[dsq[i]];
            PL(ja, jp_v, v, d) = S_ADD(jsc, e);
            if (fl) PL(la, jp_v, v, d) = d >= 2 ? S_ADD(lsc, e) : e;
            if (fr) {
              S rsc = INIT(v, d);
              for (int yo = ryo0; yo < cn; yo++)
                rsc = S_COMB(rsc, S_MAX2(S_ADD(PL(ja, jp_y, y0 + yo, d), ts[yo]), S_ADD(PL(ra, jp_y, y0 + yo, d), ts[yo])));
              PL(ra, jp_v, v, d) = rsc;
            }
          }
        }
      } else if (st == IXO_MR_ST || st == IXO_IR_ST) { /* EMITRIGHT :2107-2141 */
        if (!tq_detached(cm, v)) {
          const int lyo0 = st == IXO_IR_ST ? 1 : 0;
          for (int d = dn; d <= dx; d++) {
            S jsc = INIT(v, d - sd), rsc = fr ? INIT(v, d - sd) : S_ZERO;
            for (int yo = 0; yo < cn; yo++) {
              jsc = S_COMB(jsc, S_ADD(PL(ja, jp_y, y0 + yo, d - sd), ts[yo]));
              if (fr) rsc = S_COMB(rsc, S_ADD(PL(ra, jp_y, y0 + yo, d - sd), ts[yo]));
            }
            PL(ja, jp_v, v, d) = S_ADD(jsc, esc_j);
            if (fr) PL(ra, jp_v, v, d) = d >= 2 ? S_ADD(rsc, esc_j) : esc_j;
            if (fl) {
              S lsc = INIT(v, d);
              for (int yo = lyo0; yo < cn; yo++)
                lsc = S_COMB(lsc, S_MAX2(S_ADD(PL(ja, jq_y, y0 + yo, d), ts[yo]), S_ADD(PL(la, jq_y, y0 + yo, d), ts[yo])));
              PL(la, jp_v, v, d) = lsc;
            }
          }
        }
      } else if (st == IXO_MP_ST) { /* EMITPAIR :2142-2190 */
        int i = j - dn + 1;
        for (int d = dn; d <= dx; d++, i--) {
          S jsc = d >= 2 ? INIT(v, d - sd) : S_ZERO;
          S lsc = fl ? INIT(v, d - 1) : S_ZERO, rsc = fr ? INIT(v, d - 1) : S_ZERO;
          for (int yo = 0; yo < cn; yo++) {
            if (d >= 2) jsc = S_COMB(jsc, S_ADD(PL(ja, jp_y, y0 + yo, d - 2), ts[yo]));
            if (fl)
              lsc = S_COMB(lsc, S_MAX2(S_ADD(PL(ja, jq_y, y0 + yo, d - 1), ts[yo]), S_ADD(PL(la, jq_y, y0 + yo, d - 1), ts[yo])));
            if (fr)
              rsc = S_COMB(rsc, S_MAX2(S_ADD(PL(ja, jp_y, y0 + yo, d - 1), ts[yo]), S_ADD(PL(ra, jp_y, y0 + yo, d - 1), ts[yo])));
          }
          PL(ja, jp_v, v, d) = d >= 2 ? S_ADD(jsc, esc[dsq[i] * KP + dsq[j]]) : S_ZERO;
          if (fl) PL(la, jp_v, v, d) = d >= 2 ? S_ADD(lsc, lm[dsq[i]]) : lm[dsq[i]];
          if (fr) PL(ra, jp_v, v, d) = d >= 2 ? S_ADD(rsc, rmesc_j) : rmesc_j;
        }
      } else { /* EMITNONE (D, S) :2191-2217 */
        for (int d = dn; d <= dx; d++) {
          S jsc = INIT(v, d - sd), lsc = fl ? INIT(v, d - sd) : S_ZERO, rsc = fr ? INIT(v, d - sd) : S_ZERO;
          for (int yo = 0; yo < cn; yo++) {
            jsc = S_COMB(jsc, S_ADD(PL(ja, jp_y, y0 + yo, d - sd), ts[yo]));
            if (fl) lsc = S_COMB(lsc, S_ADD(PL(la, jp_y, y0 + yo, d - sd), ts[yo]));
            if (fr) rsc = S_COMB(rsc, S_ADD(PL(ra, jp_y, y0 + yo, d - sd), ts[yo]));
          }
          PL(ja, jp_v, v, d) = jsc;
          if (fl) PL(la, jp_v, v, d) = lsc;
          if (fr) PL(ra, jp_v, v, d) = rsc;
        }
      }
    }
    /* ROOT: truncated begins into the persistent root rows (:2220-2289) */
    const int dn0 = dn_all, dx0 = jrow < dxcap[0] ? jrow : dxcap[0];
    for (int d = 0; d <= w; d++) {
      bestr[d] = 0;
      bestsc[d] = TQ_IMPOSSIBLE;
      bestmode[d] = IXO_TRMODE_UNKNOWN;
    }
    for (int y = 1; y < M; y++) {
      const S tp = pty[y];
      if (!S_POSSIBLE(tp)) continue;
      const int dxy = jrow < dxcap[y] ? jrow : dxcap[y];
      const int dx = dx0 < dxy ? dx0 : dxy;
      struct { int on; S *pl; int mode; } pass[4] = {{1, ja, IXO_TRMODE_J},
                                                     {fl, la, IXO_TRMODE_L},
                                                     {fr, ra, IXO_TRMODE_R},
                                                     {ft && cm->sttype[y] == IXO_B_ST, ta, IXO_TRMODE_T}};
      for (int p = 0; p < 4; p++) {
        if (!pass[p].on) continue;
        for (int d = dn0; d <= dx; d++) {
          const S sc = S_ADD(PL(pass[p].pl, cur, y, d), tp);
          if (sc > PL(pass[p].pl, cur, 0, d)) {
            PL(pass[p].pl, cur, 0, d) = sc;
            const float fsc = S_BITS(sc);
            if (fsc > bestsc[d]) {
              bestsc[d] = fsc;
              bestmode[d] = pass[p].mode;
              bestr[d] = y;
            }
          }
        }
      }
    }
    for (int d = dn0; d <= dx0; d++)
      if (bestsc[d] > vsc_root) {
        vsc_root = bestsc[d];
        vmode_root = bestmode[d];
      }
    if (j == j0 && j <= w && bestsc[j] > bsc_full) bsc_full = bestsc[j];
    if (best_out)
      for (int d = 0; d <= smx_w; d++)
        best_out[(size_t)(j - i0) * (smx_w + 1) + d] = d <= w ? bestsc[d] : TQ_IMPOSSIBLE;
    /* report_hits_greedily (:2322-2372) */
    if (dn0 <= dx0) {
      float maxrep = TQ_IMPOSSIBLE;
      for (int d = dn0 > 1 ? dn0 : 1; d <= dx0; d++) {
        const int i = j - d + 1, mode = bestmode[d];
        float hs = bestsc[d], bias = 0.0f;
        if (!(hs > maxrep && hs >= cutoff && hs > TQ_IMPOSSIBLE)) continue;
        int rep = ixo_tr_allow(cm, lpos, rpos, pass_idx, local, i, j, i0, j0, mode, bestr[d]);
        if (rep && do_null3) {
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
  }
#undef PL
#undef DK
#undef INIT
  *vsc = vsc_root;
  *vmode = vmode_root;
  if (bsc_full_out) *bsc_full_out = bsc_full;
  const int64_t nk = ixo_tr_remove_overlaps(tmp, ntmp, hits, cap);
  free(tmp); free(dmax); free(dxcap); free(begl_idx); free(el_sca); free(init_sc); free(ja); free(jb);
  free(lpos); free(rpos); free(act); free(jp_wa); free(bestsc); free(bestr); free(bestmode);
  return nk;
}
