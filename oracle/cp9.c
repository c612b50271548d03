/*
 * cp9.c -- CPU oracle (test infrastructure only): the CP9 HMM banding DPs of
 * infernox restated in C: integer Forward and Backward over a window
 * (cp9.rs:1275 cp9_forward, :1995 cp9_backward) and the posterior HMM bands
 * (cp9.rs:1517 cp9_fb2hmmbands, the new-way branch).
 *
 * Scores are the reference's scaled integer log2-odds (INTSCALE 1000) with
 * -INFTY = -987654321 and the ILogsum table (cp9.rs:21-74). Matrices are
 * [(L+1)][(M+1)], row ip = position i0 + ip - 1 (row 0 = i0 - 1). Adds are
 * 32-bit wrapping, as the reference's release build computes them.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "ixo.h"
#include "ixo_cm.h"

#define INFTY 987654321 /* cp9.rs:21 */
/* cp9O transition order (cp9.rs:906-918) */
enum { O_MM = 0, O_IM, O_DM, O_BM, O_MI, O_II, O_DI, O_MD, O_ID, O_DD, O_ME, O_MEL, O_NT };

static inline int32_t wa(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t ws(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t lsum(int32_t a, int32_t b) { return ixo_ilogsum(a, b); }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline float scorify(int32_t sc) { return (float)sc / 1000.0f; } /* cp9.rs:1257 */

#define T(k, x) h->otsc[(size_t)(k) * O_NT + (x)]
#define MSC(k, x) h->msc[(size_t)(k) * 18 + (x)]
#define ISC(k, x) h->isc[(size_t)(k) * 18 + (x)]
#define AT(mx, ip, k) mx[(size_t)(ip) * (M + 1) + (k)]
#define ELN(k) h->el_from_ct[k]
#define ELI(k, c) h->el_from_idx[h->el_from_off[k] + (c)]

/* cp9.rs:1275-1453 cp9_forward: returns best_sc; *best_pos as the reference. */
float ixo_cp9_forward(const ixo_cp9 *h, const uint8_t *dsq, int i0, int j0, int do_scan, int doing_align,
                      int32_t *mmx, int32_t *imx, int32_t *dmx, int32_t *elmx, int32_t *sca, int64_t *best_pos) {
  const int M = h->M, L = j0 - i0 + 1;
  const int32_t ninf = -INFTY;
  for (size_t x = 0; x < (size_t)(L + 1) * (M + 1); x++) mmx[x] = imx[x] = dmx[x] = elmx[x] = ninf;
  float best = -1e36f;
  int64_t bpos = -1;
  AT(mmx, 0, 0) = 0;
  AT(imx, 0, 0) = ninf;
  AT(dmx, 0, 0) = ninf;
  AT(elmx, 0, 0) = ninf;
  for (int k = 1; k <= M; k++) {
    AT(mmx, 0, k) = ninf;
    AT(imx, 0, k) = ninf;
    AT(elmx, 0, k) = ninf;
    AT(dmx, 0, k) = lsum(lsum(wa(AT(mmx, 0, k - 1), T(k - 1, O_MD)), wa(AT(imx, 0, k - 1), T(k - 1, O_ID))),
                         wa(AT(dmx, 0, k - 1), T(k - 1, O_DD)));
  }
  sca[0] = wa(AT(dmx, 0, M), T(M, O_DM));
  if (scorify(sca[0]) > best) {
    best = scorify(sca[0]);
    bpos = (int64_t)i0 - 1;
  }
  for (int j = i0; j <= j0; j++) {
    const int dj = dsq[j], cur = j - i0 + 1, prv = cur - 1;
    AT(mmx, cur, 0) = do_scan ? 0 : ninf;
    AT(dmx, cur, 0) = ninf;
    AT(elmx, cur, 0) = ninf;
    int32_t sc = lsum(lsum(wa(AT(mmx, prv, 0), T(0, O_MI)), wa(AT(imx, prv, 0), T(0, O_II))),
                      wa(AT(dmx, prv, 0), T(0, O_DI)));
    AT(imx, cur, 0) = imax(wa(sc, ISC(0, dj)), ninf);
    int32_t endsc = ninf;
    for (int k = 1; k <= M; k++) {
      sc = lsum(lsum(wa(AT(mmx, prv, k - 1), T(k - 1, O_MM)), wa(AT(imx, prv, k - 1), T(k - 1, O_IM))),
                lsum(wa(AT(dmx, prv, k - 1), T(k - 1, O_DM)), wa(AT(mmx, prv, 0), T(k, O_BM))));
      for (int c = 0; c < ELN(k); c++) sc = lsum(sc, AT(elmx, prv, ELI(k, c)));
      AT(mmx, cur, k) = imax(wa(sc, MSC(k, dj)), ninf);
      endsc = lsum(endsc, wa(AT(mmx, cur, k), T(k, O_ME)));
      sc = lsum(lsum(wa(AT(mmx, prv, k), T(k, O_MI)), wa(AT(imx, prv, k), T(k, O_II))), wa(AT(dmx, prv, k), T(k, O_DI)));
      AT(imx, cur, k) = imax(wa(sc, ISC(k, dj)), ninf);
      AT(dmx, cur, k) = lsum(lsum(wa(AT(mmx, cur, k - 1), T(k - 1, O_MD)), wa(AT(imx, cur, k - 1), T(k - 1, O_ID))),
                             wa(AT(dmx, cur, k - 1), T(k - 1, O_DD)));
      sc = ninf;
      if (h->el && h->has_el[k])
        sc = lsum(wa(AT(mmx, cur, k), T(k, O_MEL)), wa(AT(elmx, prv, k), h->el_selfsc));
      AT(elmx, cur, k) = sc;
    }
    endsc = lsum(lsum(endsc, wa(AT(dmx, cur, M), T(M, O_DM))), wa(AT(imx, cur, M), T(M, O_IM)));
    for (int c = 0; c < ELN(M + 1); c++) endsc = lsum(endsc, AT(elmx, cur, ELI(M + 1, c)));
    sca[cur] = endsc;
    if (scorify(endsc) > best) {
      best = scorify(endsc);
      bpos = j;
    }
  }
  if (doing_align) {
    best = scorify(sca[L]);
    bpos = i0;
  }
  if (best_pos) *best_pos = bpos;
  return best;
}

/* cp9.rs:1995-2208 cp9_backward. */
float ixo_cp9_backward(const ixo_cp9 *h, const uint8_t *dsq, int i0, int j0, int do_scan, int doing_align,
                       int32_t *mmx, int32_t *imx, int32_t *dmx, int32_t *elmx, int32_t *sca, int64_t *best_pos) {
  const int M = h->M, L = j0 - i0 + 1;
  const int32_t ninf = -INFTY;
  const int el = h->el;
  for (size_t x = 0; x < (size_t)(L + 1) * (M + 1); x++) mmx[x] = imx[x] = dmx[x] = elmx[x] = ninf;
  float best = -1e36f;
  int64_t bpos = -1;
  {
    const int cur = L, dj0 = dsq[j0];
    for (int k = 1; k <= M; k++) AT(elmx, cur, k) = ninf;
    if (el)
      for (int c = 0; c < ELN(M + 1); c++) AT(elmx, cur, ELI(M + 1, c)) = 0;
    AT(mmx, cur, M) = wa(lsum(wa(AT(elmx, cur, M), T(M, O_MEL)), T(M, O_ME)), MSC(M, dj0));
    AT(imx, cur, M) = wa(T(M, O_IM), ISC(M, dj0));
    AT(dmx, cur, M) = T(M, O_DM);
    for (int k = M - 1; k >= 1; k--) {
      int32_t v = T(k, O_ME);
      v = lsum(v, wa(AT(dmx, cur, k + 1), T(k, O_MD)));
      if (el) v = lsum(v, wa(AT(elmx, cur, k), T(k, O_MEL)));
      AT(mmx, cur, k) = wa(v, MSC(k, dj0));
      AT(imx, cur, k) = wa(wa(AT(dmx, cur, k + 1), T(k, O_ID)), ISC(k, dj0));
      AT(dmx, cur, k) = wa(AT(dmx, cur, k + 1), T(k, O_DD));
    }
    AT(mmx, cur, 0) = wa(AT(dmx, cur, 1), T(0, O_MD));
    AT(imx, cur, 0) = wa(wa(AT(dmx, cur, 1), T(0, O_ID)), ISC(0, dj0));
    AT(dmx, cur, 0) = ninf;
    AT(elmx, cur, 0) = ninf;
    sca[L] = AT(mmx, cur, 0);
  }
  for (int i = j0 - 1; i >= i0; i--) {
    const int cur = i - i0 + 1, prv = cur + 1, di = dsq[i];
    for (int k = 0; k <= M; k++) AT(elmx, cur, k) = ninf;
    if (el && h->has_el[M]) AT(elmx, cur, M) = wa(AT(elmx, cur, M), h->el_selfsc);
    AT(mmx, cur, M) = wa(wa(AT(imx, prv, M), T(M, O_MI)), MSC(M, di));
    if (el && h->has_el[M]) AT(mmx, cur, M) = lsum(AT(mmx, cur, M), wa(AT(elmx, cur, M), T(M, O_MEL)));
    AT(imx, cur, M) = wa(wa(AT(imx, prv, M), T(M, O_II)), ISC(M, di));
    AT(dmx, cur, M) = wa(AT(imx, prv, M), T(M, O_DI));
    if (el)
      for (int c = 0; c < ELN(M); c++) {
        const int src = ELI(M, c);
        AT(elmx, cur, src) = lsum(AT(elmx, cur, src), AT(mmx, prv, M));
      }
    if (do_scan) {
      if (el)
        for (int c = 0; c < ELN(M + 1); c++) AT(elmx, cur, ELI(M + 1, c)) = 0;
      AT(mmx, cur, M) = lsum(AT(mmx, cur, M), lsum(wa(AT(elmx, cur, M), T(M, O_MEL)), T(M, O_ME)));
      AT(imx, cur, M) = lsum(AT(imx, cur, M), T(M, O_IM));
      AT(dmx, cur, M) = lsum(AT(dmx, cur, M), T(M, O_DM));
    }
    for (int k = M - 1; k >= 1; k--) {
      if (el)
        for (int c = 0; c < ELN(k); c++) {
          const int src = ELI(k, c);
          AT(elmx, cur, src) = lsum(AT(elmx, cur, src), AT(mmx, prv, k));
        }
      if (el && h->has_el[k]) AT(elmx, cur, k) = lsum(AT(elmx, cur, k), wa(AT(elmx, prv, k), h->el_selfsc));
      int32_t mv = lsum(lsum(wa(AT(mmx, prv, k + 1), T(k, O_MM)), wa(AT(imx, prv, k), T(k, O_MI))),
                        wa(AT(dmx, cur, k + 1), T(k, O_MD)));
      if (el && h->has_el[k]) mv = lsum(mv, wa(AT(elmx, cur, k), T(k, O_MEL)));
      AT(mmx, cur, k) = wa(mv, MSC(k, di));
      AT(imx, cur, k) = wa(lsum(lsum(wa(AT(mmx, prv, k + 1), T(k, O_IM)), wa(AT(imx, prv, k), T(k, O_II))),
                                wa(AT(dmx, cur, k + 1), T(k, O_ID))),
                           ISC(k, di));
      if (do_scan) AT(mmx, cur, k) = lsum(AT(mmx, cur, k), T(k, O_ME));
      AT(dmx, cur, k) = lsum(lsum(wa(AT(mmx, prv, k + 1), T(k, O_DM)), wa(AT(imx, prv, k), T(k, O_DI))),
                             wa(AT(dmx, cur, k + 1), T(k, O_DD)));
    }
    AT(imx, cur, 0) = wa(lsum(lsum(wa(AT(mmx, prv, 1), T(0, O_IM)), wa(AT(imx, prv, 0), T(0, O_II))),
                              wa(AT(dmx, cur, 1), T(0, O_ID))),
                         ISC(0, di));
    AT(dmx, cur, 0) = ninf;
    AT(elmx, cur, 0) = ninf;
    int32_t b = ninf;
    for (int k = M; k >= 1; k--) b = lsum(b, wa(AT(mmx, prv, k), T(k, O_BM)));
    b = lsum(b, wa(AT(imx, prv, 0), T(0, O_MI)));
    b = lsum(b, wa(AT(dmx, cur, 1), T(0, O_MD)));
    AT(mmx, cur, 0) = b;
    sca[cur] = b;
    if (scorify(b) > best) {
      best = scorify(b);
      bpos = (int64_t)i + 1;
    }
  }
  {
    const int cur = 0, prv = 1;
    for (int k = 1; k <= M; k++) AT(elmx, cur, k) = ninf;
    AT(mmx, cur, M) = ninf;
    AT(imx, cur, M) = ninf;
    AT(elmx, cur, M) = ninf;
    AT(dmx, cur, M) = wa(AT(imx, prv, M), T(M, O_DI));
    if (do_scan) AT(dmx, cur, M) = lsum(AT(dmx, cur, M), T(M, O_DM));
    for (int k = M - 1; k >= 1; k--) {
      AT(mmx, cur, k) = ninf;
      AT(imx, cur, k) = ninf;
      AT(elmx, cur, k) = ninf;
      AT(dmx, cur, k) = lsum(lsum(wa(AT(mmx, prv, k + 1), T(k, O_DM)), wa(AT(imx, prv, k), T(k, O_DI))),
                             wa(AT(dmx, cur, k + 1), T(k, O_DD)));
    }
    AT(imx, cur, 0) = ninf;
    AT(dmx, cur, 0) = ninf;
    AT(elmx, cur, 0) = ninf;
    int32_t b = ninf;
    for (int k = M; k >= 1; k--) b = lsum(b, wa(AT(mmx, prv, k), T(k, O_BM)));
    b = lsum(b, wa(AT(imx, prv, 0), T(0, O_MI)));
    b = lsum(b, wa(AT(dmx, cur, 1), T(0, O_MD)));
    AT(mmx, cur, 0) = b;
    sca[0] = b;
    if (scorify(b) > best) {
      best = scorify(b);
      bpos = i0;
    }
  }
  if (doing_align) {
    best = scorify(sca[0]);
    bpos = i0;
  }
  if (best_pos) *best_pos = bpos;
  return best;
}

/* cp9.rs:977-998 Prob2Score */
static int32_t prob2score(float p, float null) {
  if (p == 0.0f) return -INFTY;
  const float ratio = p / null;
  const double lg = (double)ratio > 0.0 ? log((double)ratio) * 1.44269504 : -1e36;
  const double v = floor(0.5 + 1000.0 * lg);
  if (!isfinite(v)) return -INFTY;
  const int32_t r = (int32_t)v;
  return r > -INFTY ? r : -INFTY;
}

int32_t ixo_cp9_band_thresh(double p_thresh) { return prob2score((float)((1.0 - p_thresh) / 2.0), 1.0f); }

/* cp9.rs:1517-1737 cp9_fb2hmmbands (new way, use_sums off): posterior mass
 * crept in from both ends per state. pn: [6][M+1] = min_m, max_m, min_i,
 * max_i, min_d, max_d; -1 where unset. */
void ixo_cp9_fb2hmmbands(const ixo_cp9 *h, const uint8_t *dsq, const int32_t *fm, const int32_t *fi,
                         const int32_t *fd, const int32_t *bm, const int32_t *bi, const int32_t *bd, int i0, int j0,
                         double p_thresh, int did_fwd_scan, int32_t *pn) {
  const int M = h->M, L = j0 - i0 + 1;
  const int32_t ninf = -INFTY, thresh = ixo_cp9_band_thresh(p_thresh);
  int32_t *pmin_m = pn, *pmax_m = pn + (M + 1), *pmin_i = pn + 2 * (M + 1), *pmax_i = pn + 3 * (M + 1);
  int32_t *pmin_d = pn + 4 * (M + 1), *pmax_d = pn + 5 * (M + 1);
  memset(pn, 0, sizeof(int32_t) * 6 * (size_t)(M + 1));
  const size_t n = (size_t)(L + 1) * (M + 1);
  int32_t *pm = malloc(n * 4), *pi = malloc(n * 4), *pd = malloc(n * 4);
  int32_t *mm = malloc((M + 1) * 4), *mi = malloc((M + 1) * 4), *md = malloc((M + 1) * 4);
  char *nm = calloc(M + 1, 1), *ni = calloc(M + 1, 1), *nd = calloc(M + 1, 1);
  char *xm = calloc(M + 1, 1), *xi = calloc(M + 1, 1), *xd = calloc(M + 1, 1);
  for (size_t x = 0; x < n; x++) pm[x] = pi[x] = pd[x] = ninf;
  for (int k = 0; k <= M; k++) mm[k] = mi[k] = md[k] = ninf;
  int32_t sc;
  if (did_fwd_scan) {
    sc = ninf;
    for (int ip = 0; ip <= L; ip++) sc = lsum(sc, AT(bm, ip, 0));
  } else {
    sc = AT(bm, 0, 0);
  }
  const int32_t first = i0 - 1 < 0 ? 0 : i0 - 1;
  AT(pm, 0, 0) = ws(wa(AT(fm, 0, 0), AT(bm, 0, 0)), sc);
  AT(pi, 0, 0) = ninf;
  AT(pd, 0, 0) = ninf;
  mm[0] = AT(pm, 0, 0);
  if (mm[0] > thresh) {
    pmin_m[0] = first;
    nm[0] = 1;
  }
  mi[0] = ninf;
  md[0] = ninf;
  for (int k = 1; k <= M; k++) {
    AT(pm, 0, k) = ninf;
    AT(pi, 0, k) = ninf;
    AT(pd, 0, k) = ws(wa(AT(fd, 0, k), AT(bd, 0, k)), sc);
    md[k] = AT(pd, 0, k);
    if (md[k] > thresh) {
      pmin_d[k] = first;
      nd[k] = 1;
    }
  }
  for (int ip = 1; ip <= L; ip++) {
    const int i = i0 + ip - 1, x = dsq[i];
    AT(pm, ip, 0) = imax(ws(wa(AT(fm, ip, 0), AT(bm, ip, 0)), sc), ninf);
    if (!nm[0]) {
      mm[0] = lsum(mm[0], AT(pm, ip, 0));
      if (mm[0] > thresh) {
        pmin_m[0] = i;
        nm[0] = 1;
      }
    }
    AT(pi, ip, 0) = imax(ws(ws(wa(AT(fi, ip, 0), AT(bi, ip, 0)), ISC(0, x)), sc), ninf);
    if (!ni[0]) {
      mi[0] = lsum(mi[0], AT(pi, ip, 0));
      if (mi[0] > thresh) {
        pmin_i[0] = i;
        ni[0] = 1;
      }
    }
    AT(pd, ip, 0) = ninf;
    for (int k = 1; k <= M; k++) {
      AT(pm, ip, k) = imax(ws(ws(wa(AT(fm, ip, k), AT(bm, ip, k)), MSC(k, x)), sc), ninf);
      AT(pi, ip, k) = imax(ws(ws(wa(AT(fi, ip, k), AT(bi, ip, k)), ISC(k, x)), sc), ninf);
      AT(pd, ip, k) = imax(ws(wa(AT(fd, ip, k), AT(bd, ip, k)), sc), ninf);
      if (!nm[k]) {
        mm[k] = lsum(mm[k], AT(pm, ip, k));
        if (mm[k] > thresh) {
          pmin_m[k] = i;
          nm[k] = 1;
        }
      }
      if (!ni[k]) {
        mi[k] = lsum(mi[k], AT(pi, ip, k));
        if (mi[k] > thresh) {
          pmin_i[k] = i;
          ni[k] = 1;
        }
      }
      if (!nd[k]) {
        md[k] = lsum(md[k], AT(pd, ip, k));
        if (md[k] > thresh) {
          pmin_d[k] = i;
          nd[k] = 1;
        }
      }
    }
  }
  for (int k = 0; k <= M; k++) mm[k] = mi[k] = md[k] = ninf;
  for (int ip = L; ip >= 1; ip--) {
    const int i = i0 + ip - 1;
    for (int k = 0; k <= M; k++) {
      if (!xm[k]) {
        mm[k] = lsum(mm[k], AT(pm, ip, k));
        if (mm[k] > thresh) {
          pmax_m[k] = i;
          xm[k] = 1;
        }
      }
      if (!xi[k]) {
        mi[k] = lsum(mi[k], AT(pi, ip, k));
        if (mi[k] > thresh) {
          pmax_i[k] = i;
          xi[k] = 1;
        }
      }
      if (!xd[k]) {
        md[k] = lsum(md[k], AT(pd, ip, k));
        if (md[k] > thresh) {
          pmax_d[k] = i;
          xd[k] = 1;
        }
      }
    }
  }
  if (!xm[0]) {
    mm[0] = lsum(mm[0], AT(pm, 0, 0));
    if (mm[0] > thresh) {
      pmax_m[0] = first;
      xm[0] = 1;
    }
  }
  for (int k = 1; k <= M; k++)
    if (!xd[k]) {
      md[k] = lsum(md[k], AT(pd, 0, k));
      if (md[k] > thresh) {
        pmax_d[k] = first;
        xd[k] = 1;
      }
    }
  for (int k = 0; k <= M; k++) {
    if (!nm[k] || !xm[k] || pmax_m[k] < pmin_m[k]) pmin_m[k] = pmax_m[k] = -1;
    if (!ni[k] || !xi[k] || pmax_i[k] < pmin_i[k]) pmin_i[k] = pmax_i[k] = -1;
    if (!nd[k] || !xd[k] || pmax_d[k] < pmin_d[k]) pmin_d[k] = pmax_d[k] = -1;
  }
  pmin_d[0] = pmax_d[0] = -1;
  free(pm);
  free(pi);
  free(pd);
  free(mm);
  free(mi);
  free(md);
  free(nm);
  free(ni);
  free(nd);
  free(xm);
  free(xi);
  free(xd);
}

/* cp9.rs:1765-1776 / :1805-1814 (inside cp9_predict_start_and_end_positions):
 * the posterior occupancy of every node, occ[k] = ILogsum over rows ip =
 * 0..L of ILogsum(pmx_m[ip][k], pmx_d[ip][k]), the posterior cells as
 * cp9_fb2hmmbands (cp9.rs:1583-1640) defines them; occ[0] = -INFTY. */
void ixo_cp9_occupancy(const ixo_cp9 *h, const uint8_t *dsq, const int32_t *fm, const int32_t *fd,
                       const int32_t *bm, const int32_t *bd, int i0, int j0, int did_fwd_scan, int32_t *occ) {
  const int M = h->M, L = j0 - i0 + 1;
  const int32_t ninf = -INFTY;
  int32_t sc;
  if (did_fwd_scan) {
    sc = ninf;
    for (int ip = 0; ip <= L; ip++) sc = lsum(sc, AT(bm, ip, 0));
  } else {
    sc = AT(bm, 0, 0);
  }
  occ[0] = ninf;
  for (int k = 1; k <= M; k++) {
    int32_t o = lsum(ninf, lsum(ninf, ws(wa(AT(fd, 0, k), AT(bd, 0, k)), sc)));
    for (int ip = 1; ip <= L; ip++) {
      const int x = dsq[i0 + ip - 1];
      const int32_t m = imax(ws(ws(wa(AT(fm, ip, k), AT(bm, ip, k)), MSC(k, x)), sc), ninf);
      const int32_t d = imax(ws(wa(AT(fd, ip, k), AT(bd, ip, k)), sc), ninf);
      o = lsum(o, lsum(m, d));
    }
    occ[k] = o;
  }
}
