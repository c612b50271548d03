/*
 * envdef.c -- CPU oracle (test infrastructure only) for the F4/F5 envelope
 * definition of the STD pass: the glocal p7 profile, the generic glocal
 * Forward/Backward in nats with HMMER's FLogsum table, posterior domain
 * decoding and glocal domain definition (region finding, stochastic-trace
 * clustering of multidomain regions, isolated-domain rescoring), and the
 * per-window F4 / F4b / F5 / F5b gates of trunc_env_def. Restates
 * p7_generic.rs and cm_search.rs; not linked into the product.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ixo.h"
#include "ixo_env.h"

#define NEGINF (-INFINITY)
#define KP IXO_KP

/* p7P transition indices (p7_generic.rs:76-84) */
enum { P_MM = 0, P_IM = 1, P_DM = 2, P_BM = 3, P_MD = 4, P_DD = 5, P_MI = 6, P_II = 7, P_NT = 8 };
/* P7Profile trans order (p7_generic.rs:87-93) */
enum { H_MM = 0, H_MI = 1, H_MD = 2, H_IM = 3, H_II = 4, H_DM = 5, H_DD = 6 };
/* specials (p7_generic.rs:96-101, :400-405) */
enum { X_E = 0, X_N = 1, X_J = 2, X_C = 3 };
enum { LOOP = 0, MOVE = 1 };
enum { G_E = 0, G_N = 1, G_J = 2, G_B = 3, G_C = 4, G_NX = 5 };

/* ---- p7_generic.rs:27-69 nats FLogsum (16,000 entries, 15.7 gate) ---- */
static float p7tbl[16000];
static int p7tbl_ready = 0;
static void p7tbl_init(void) {
  if (p7tbl_ready) return;
  for (int i = 0; i < 16000; i++) p7tbl[i] = (float)log(1.0 + exp(-(double)i / 1000.0));
  p7tbl_ready = 1;
}
float ixo_p7_flogsum(float a, float b) {
  p7tbl_init();
  const float max = a > b ? a : b, min = a > b ? b : a;
  if (min == NEGINF || (max - min) >= 15.7f) return max;
  return max + p7tbl[(int)((max - min) * 1000.0f)];
}
float ixo_p7_flogsum_table(int i) {
  p7tbl_init();
  return p7tbl[i];
}
#define LS ixo_p7_flogsum

/* RNA degeneracy (p7_generic.rs:110-129) */
static const unsigned char DEGEN[KP][4] = {
    {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {0, 0, 0, 0}, {1, 0, 1, 0},
    {0, 1, 0, 1}, {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 1, 1, 0}, {1, 0, 0, 1}, {1, 1, 0, 1},
    {0, 1, 1, 1}, {1, 1, 1, 0}, {1, 0, 1, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}};

#define TSC(gm, s, k) ((gm)->tsc[(size_t)(k) * P_NT + (s)])
#define MSC(gm, x, k) ((gm)->rsc[((size_t)(x) * ((gm)->M + 1) + (k)) * 2 + 0])
#define ISC(gm, x, k) ((gm)->rsc[((size_t)(x) * ((gm)->M + 1) + (k)) * 2 + 1])

/* p7_generic.rs:379-392 reconfig_length */
void ixo_gm_reconfig_length(ixo_gm *gm, int L) {
  const float pmove = (2.0f + gm->nj) / ((float)L + 2.0f + gm->nj);
  const float ploop = 1.0f - pmove;
  const float ll = (float)log((double)ploop), lm = (float)log((double)pmove);
  const int s3[3] = {X_N, X_C, X_J};
  for (int t = 0; t < 3; t++) {
    gm->xsc[s3[t]][LOOP] = ll;
    gm->xsc[s3[t]][MOVE] = lm;
  }
  gm->L = L;
}
/* p7_generic.rs:1075-1082 */
void ixo_gm_reconfig_unihit(ixo_gm *gm, int L) {
  gm->xsc[X_E][MOVE] = 0.0f;
  gm->xsc[X_E][LOOP] = NEGINF;
  gm->nj = 0.0f;
  ixo_gm_reconfig_length(gm, L);
}
/* p7_generic.rs:1084-1094 */
void ixo_gm_reconfig_multihit(ixo_gm *gm, int L) {
  const float log2 = (float)0.69314718055994530942;
  gm->xsc[X_E][MOVE] = -log2;
  gm->xsc[X_E][LOOP] = -log2;
  gm->nj = 1.0f;
  ixo_gm_reconfig_length(gm, L);
}

/* p7_generic.rs:350-376 the 5' and 3' truncated length models */
void ixo_gm_reconfig_length_5p(ixo_gm *gm, int L) {
  const float pmove = (1.0f + gm->nj) / ((float)L + 1.0f + gm->nj);
  const float ploop = 1.0f - pmove;
  const float ll = (float)log((double)ploop), lm = (float)log((double)pmove);
  gm->xsc[X_C][LOOP] = gm->xsc[X_J][LOOP] = ll;
  gm->xsc[X_C][MOVE] = gm->xsc[X_J][MOVE] = lm;
  gm->L = L;
}
void ixo_gm_reconfig_length_3p(ixo_gm *gm, int L) {
  const float pmove = (1.0f + gm->nj) / ((float)L + 1.0f + gm->nj);
  const float ploop = 1.0f - pmove;
  const float ll = (float)log((double)ploop), lm = (float)log((double)pmove);
  gm->xsc[X_N][LOOP] = gm->xsc[X_J][LOOP] = ll;
  gm->xsc[X_N][MOVE] = gm->xsc[X_J][MOVE] = lm;
  gm->L = L;
}

/* p7_generic.rs:156-257 build_glocal_profile (multihit glocal, bg 0.25) */
int ixo_gm_build(const ixo_p7 *p7, int L, ixo_gm *gm) {
  const int M = p7->M;
  const double bgf = 0.25;
  memset(gm, 0, sizeof(*gm));
  gm->M = M;
  gm->tsc = (float *)calloc((size_t)(M + 1) * P_NT, sizeof(float));
  gm->rsc = (float *)calloc((size_t)KP * (M + 1) * 2, sizeof(float));
  if (!gm->tsc || !gm->rsc) return -1;
  const float *t0 = p7->trans;
  float z = (float)log((double)t0[H_MD]);
  TSC(gm, P_BM, 0) = (float)log(1.0 - (double)t0[H_MD]);
  for (int k = 1; k < M; k++) {
    const float *tk = p7->trans + (size_t)k * 7;
    TSC(gm, P_BM, k) = (float)((double)z + log((double)tk[H_DM]));
    z = (float)((double)z + log((double)tk[H_DD]));
  }
  for (int k = 1; k < M; k++) {
    const float *tk = p7->trans + (size_t)k * 7;
    TSC(gm, P_MM, k) = (float)log((double)tk[H_MM]);
    TSC(gm, P_MI, k) = (float)log((double)tk[H_MI]);
    TSC(gm, P_MD, k) = (float)log((double)tk[H_MD]);
    TSC(gm, P_IM, k) = (float)log((double)tk[H_IM]);
    TSC(gm, P_II, k) = (float)log((double)tk[H_II]);
    TSC(gm, P_DM, k) = (float)log((double)tk[H_DM]);
    TSC(gm, P_DD, k) = (float)log((double)tk[H_DD]);
  }
  for (int k = 1; k <= M; k++) {
    float sc[KP];
    for (int x = 0; x < 4; x++) sc[x] = (float)log((double)p7->mat[(size_t)k * 4 + x] / bgf);
    sc[4] = NEGINF;
    sc[KP - 2] = NEGINF;
    sc[KP - 1] = NEGINF;
    for (int x = 5; x <= KP - 3; x++) {
      float result = 0.0f, denom = 0.0f;
      for (int i = 0; i < 4; i++)
        if (DEGEN[x][i]) {
          result += sc[i] * (float)bgf;
          denom += (float)bgf;
        }
      sc[x] = result / denom;
    }
    for (int x = 0; x < KP; x++) MSC(gm, x, k) = sc[x];
  }
  for (int x = 0; x < KP; x++) {
    for (int k = 1; k < M; k++) ISC(gm, x, k) = 0.0f;
    ISC(gm, x, M) = NEGINF;
  }
  for (int k = 1; k <= M; k++) {
    ISC(gm, 4, k) = NEGINF;
    ISC(gm, KP - 2, k) = NEGINF;
    ISC(gm, KP - 1, k) = NEGINF;
  }
  const float log2 = (float)0.69314718055994530942;
  gm->xsc[X_E][MOVE] = -log2;
  gm->xsc[X_E][LOOP] = -log2;
  gm->nj = 1.0f;
  gm->local_end = 0;
  ixo_gm_reconfig_length(gm, L);
  return 0;
}

/* The truncated passes' profiles as cm_search.rs:1011-1016 builds them at
 * L = 100: which 0 = Rgm (p7_generic.rs:289, 5'), 1 = Lgm (:311, 3'),
 * 2 = Tgm (build_local_profile :259 + :336, 5'&3'). */
int ixo_gm_build_trunc(const ixo_p7 *p7, int which, ixo_gm *gm) {
  if (ixo_gm_build(p7, 100, gm)) return -1;
  const int M = gm->M;
  if (which == 0) {
    const float inv = (float)log(1.0 / (double)M);
    for (int k = 1; k <= M; k++) TSC(gm, P_BM, k - 1) = inv;
    gm->local_end = 0;
    gm->xsc[X_N][MOVE] = 0.0f, gm->xsc[X_E][MOVE] = 0.0f;
    gm->xsc[X_N][LOOP] = NEGINF, gm->xsc[X_E][LOOP] = NEGINF;
    gm->L = 0;
    ixo_gm_reconfig_length_5p(gm, 100);
  } else if (which == 1) {
    /* the glocal entry, recomputed as the reference does: identical */
    gm->local_end = 1;
    gm->xsc[X_C][MOVE] = 0.0f, gm->xsc[X_E][MOVE] = 0.0f;
    gm->xsc[X_C][LOOP] = NEGINF, gm->xsc[X_E][LOOP] = NEGINF;
    gm->L = 0;
    ixo_gm_reconfig_length_3p(gm, 100);
  } else {
    const float *t = p7->trans;
    float *mocc = (float *)calloc((size_t)M + 1, sizeof(float));
    mocc[1] = t[H_MI] + t[H_MM];
    for (int k = 2; k <= M; k++) {
      const float *tk = t + (size_t)(k - 1) * 7;
      mocc[k] = mocc[k - 1] * (tk[H_MM] + tk[H_MI]) + (1.0f - mocc[k - 1]) * tk[H_DM];
    }
    float z = 0.0f;
    for (int k = 1; k <= M; k++) z += mocc[k] * (float)(M - k + 1);
    for (int k = 1; k <= M; k++) TSC(gm, P_BM, k - 1) = (float)log((double)(mocc[k] / z));
    free(mocc);
    gm->local_end = 1;
    gm->xsc[X_N][MOVE] = gm->xsc[X_C][MOVE] = gm->xsc[X_E][MOVE] = 0.0f;
    gm->xsc[X_N][LOOP] = gm->xsc[X_C][LOOP] = gm->xsc[X_E][LOOP] = NEGINF;
    gm->nj = 0.0f;
    gm->L = 100;
  }
  return 0;
}

void ixo_gm_free(ixo_gm *gm) {
  free(gm->tsc);
  free(gm->rsc);
  gm->tsc = gm->rsc = NULL;
}

void ixo_gm_copy(const ixo_gm *src, ixo_gm *dst) {
  *dst = *src;
  dst->tsc = (float *)malloc((size_t)(src->M + 1) * P_NT * sizeof(float));
  dst->rsc = (float *)malloc((size_t)KP * (src->M + 1) * 2 * sizeof(float));
  memcpy(dst->tsc, src->tsc, (size_t)(src->M + 1) * P_NT * sizeof(float));
  memcpy(dst->rsc, src->rsc, (size_t)KP * (src->M + 1) * 2 * sizeof(float));
}

/* matrix access: row i = [M 0..M | I 0..M | D 0..M] (p7_generic.rs:420-437) */
#define MMX(dp, M, i, k) (dp)[((size_t)(i) * ((M) + 1) * 3) + (k)]
#define IMX(dp, M, i, k) (dp)[((size_t)(i) * ((M) + 1) * 3) + ((M) + 1) + (k)]
#define DMX(dp, M, i, k) (dp)[((size_t)(i) * ((M) + 1) * 3) + 2 * ((M) + 1) + (k)]
#define XMX(xm, i, s) (xm)[(size_t)(i) * G_NX + (s)]

/* p7_generic.rs:559-676 p7_gforward_scalar */
float ixo_gforward(const uint8_t *dsq, int L, const ixo_gm *gm, float *dp, float *xm) {
  const int M = gm->M;
  const float esc = gm->local_end ? 0.0f : NEGINF;
  XMX(xm, 0, G_N) = 0.0f;
  XMX(xm, 0, G_B) = gm->xsc[X_N][MOVE];
  XMX(xm, 0, G_E) = XMX(xm, 0, G_C) = XMX(xm, 0, G_J) = NEGINF;
  for (int k = 0; k <= M; k++) MMX(dp, M, 0, k) = IMX(dp, M, 0, k) = DMX(dp, M, 0, k) = NEGINF;
  for (int i = 1; i <= L; i++) {
    const int x = dsq[i];
    MMX(dp, M, i, 0) = IMX(dp, M, i, 0) = DMX(dp, M, i, 0) = NEGINF;
    XMX(xm, i, G_E) = NEGINF;
    for (int k = 1; k < M; k++) {
      float sc = LS(LS(MMX(dp, M, i - 1, k - 1) + TSC(gm, P_MM, k - 1), IMX(dp, M, i - 1, k - 1) + TSC(gm, P_IM, k - 1)),
                    LS(XMX(xm, i - 1, G_B) + TSC(gm, P_BM, k - 1), DMX(dp, M, i - 1, k - 1) + TSC(gm, P_DM, k - 1)));
      MMX(dp, M, i, k) = sc + MSC(gm, x, k);
      sc = LS(MMX(dp, M, i - 1, k) + TSC(gm, P_MI, k), IMX(dp, M, i - 1, k) + TSC(gm, P_II, k));
      IMX(dp, M, i, k) = sc + ISC(gm, x, k);
      DMX(dp, M, i, k) = LS(MMX(dp, M, i, k - 1) + TSC(gm, P_MD, k - 1), DMX(dp, M, i, k - 1) + TSC(gm, P_DD, k - 1));
      XMX(xm, i, G_E) = LS(LS(MMX(dp, M, i, k) + esc, DMX(dp, M, i, k) + esc), XMX(xm, i, G_E));
    }
    float sc = LS(LS(MMX(dp, M, i - 1, M - 1) + TSC(gm, P_MM, M - 1), IMX(dp, M, i - 1, M - 1) + TSC(gm, P_IM, M - 1)),
                  LS(XMX(xm, i - 1, G_B) + TSC(gm, P_BM, M - 1), DMX(dp, M, i - 1, M - 1) + TSC(gm, P_DM, M - 1)));
    MMX(dp, M, i, M) = sc + MSC(gm, x, M);
    IMX(dp, M, i, M) = NEGINF;
    DMX(dp, M, i, M) = LS(MMX(dp, M, i, M - 1) + TSC(gm, P_MD, M - 1), DMX(dp, M, i, M - 1) + TSC(gm, P_DD, M - 1));
    XMX(xm, i, G_E) = LS(LS(MMX(dp, M, i, M), DMX(dp, M, i, M)), XMX(xm, i, G_E));
    XMX(xm, i, G_J) = LS(XMX(xm, i - 1, G_J) + gm->xsc[X_J][LOOP], XMX(xm, i, G_E) + gm->xsc[X_E][LOOP]);
    XMX(xm, i, G_C) = LS(XMX(xm, i - 1, G_C) + gm->xsc[X_C][LOOP], XMX(xm, i, G_E) + gm->xsc[X_E][MOVE]);
    XMX(xm, i, G_N) = XMX(xm, i - 1, G_N) + gm->xsc[X_N][LOOP];
    XMX(xm, i, G_B) = LS(XMX(xm, i, G_N) + gm->xsc[X_N][MOVE], XMX(xm, i, G_J) + gm->xsc[X_J][MOVE]);
  }
  return XMX(xm, L, G_C) + gm->xsc[X_C][MOVE];
}

/* p7_generic.rs:939-1073 p7_gbackward */
float ixo_gbackward(const uint8_t *dsq, int L, const ixo_gm *gm, float *dp, float *xm) {
  const int M = gm->M;
  const float esc = gm->local_end ? 0.0f : NEGINF;
  XMX(xm, L, G_J) = XMX(xm, L, G_B) = XMX(xm, L, G_N) = NEGINF;
  XMX(xm, L, G_C) = gm->xsc[X_C][MOVE];
  XMX(xm, L, G_E) = XMX(xm, L, G_C) + gm->xsc[X_E][MOVE];
  MMX(dp, M, L, M) = XMX(xm, L, G_E);
  DMX(dp, M, L, M) = XMX(xm, L, G_E);
  IMX(dp, M, L, M) = NEGINF;
  for (int k = M - 1; k >= 1; k--) {
    MMX(dp, M, L, k) = LS(XMX(xm, L, G_E) + esc, DMX(dp, M, L, k + 1) + TSC(gm, P_MD, k));
    DMX(dp, M, L, k) = LS(XMX(xm, L, G_E) + esc, DMX(dp, M, L, k + 1) + TSC(gm, P_DD, k));
    IMX(dp, M, L, k) = NEGINF;
  }
  for (int i = L - 1; i >= 1; i--) {
    const int x = dsq[i + 1];
    float b = MMX(dp, M, i + 1, 1) + TSC(gm, P_BM, 0) + MSC(gm, x, 1);
    for (int k = 2; k <= M; k++) b = LS(b, MMX(dp, M, i + 1, k) + TSC(gm, P_BM, k - 1) + MSC(gm, x, k));
    XMX(xm, i, G_B) = b;
    XMX(xm, i, G_J) = LS(XMX(xm, i + 1, G_J) + gm->xsc[X_J][LOOP], XMX(xm, i, G_B) + gm->xsc[X_J][MOVE]);
    XMX(xm, i, G_C) = XMX(xm, i + 1, G_C) + gm->xsc[X_C][LOOP];
    XMX(xm, i, G_E) = LS(XMX(xm, i, G_J) + gm->xsc[X_E][LOOP], XMX(xm, i, G_C) + gm->xsc[X_E][MOVE]);
    XMX(xm, i, G_N) = LS(XMX(xm, i + 1, G_N) + gm->xsc[X_N][LOOP], XMX(xm, i, G_B) + gm->xsc[X_N][MOVE]);
    MMX(dp, M, i, M) = XMX(xm, i, G_E);
    DMX(dp, M, i, M) = XMX(xm, i, G_E);
    IMX(dp, M, i, M) = NEGINF;
    for (int k = M - 1; k >= 1; k--) {
      MMX(dp, M, i, k) = LS(LS(MMX(dp, M, i + 1, k + 1) + TSC(gm, P_MM, k) + MSC(gm, x, k + 1),
                               IMX(dp, M, i + 1, k) + TSC(gm, P_MI, k) + ISC(gm, x, k)),
                            LS(XMX(xm, i, G_E) + esc, DMX(dp, M, i, k + 1) + TSC(gm, P_MD, k)));
      IMX(dp, M, i, k) = LS(MMX(dp, M, i + 1, k + 1) + TSC(gm, P_IM, k) + MSC(gm, x, k + 1),
                            IMX(dp, M, i + 1, k) + TSC(gm, P_II, k) + ISC(gm, x, k));
      DMX(dp, M, i, k) = LS(MMX(dp, M, i + 1, k + 1) + TSC(gm, P_DM, k) + MSC(gm, x, k + 1),
                            LS(DMX(dp, M, i, k + 1) + TSC(gm, P_DD, k), XMX(xm, i, G_E) + esc));
    }
  }
  const int x1 = dsq[1];
  float b0 = MMX(dp, M, 1, 1) + TSC(gm, P_BM, 0) + MSC(gm, x1, 1);
  for (int k = 2; k <= M; k++) b0 = LS(b0, MMX(dp, M, 1, k) + TSC(gm, P_BM, k - 1) + MSC(gm, x1, k));
  XMX(xm, 0, G_B) = b0;
  XMX(xm, 0, G_J) = XMX(xm, 0, G_C) = XMX(xm, 0, G_E) = NEGINF;
  XMX(xm, 0, G_N) = LS(XMX(xm, 1, G_N) + gm->xsc[X_N][LOOP], XMX(xm, 0, G_B) + gm->xsc[X_N][MOVE]);
  for (int k = 1; k <= M; k++) MMX(dp, M, 0, k) = IMX(dp, M, 0, k) = DMX(dp, M, 0, k) = NEGINF;
  return XMX(xm, 0, G_N);
}

/* p7_generic.rs:1096-1121 p7_gdomain_decoding (btot/etot double exp, njcp expf) */
void ixo_gdomain_decoding(const ixo_gm *gm, int L, const float *fx, const float *bx, float *btot, float *etot,
                          float *mocc) {
  const float overall = XMX(fx, L, G_C) + gm->xsc[X_C][MOVE];
  btot[0] = etot[0] = mocc[0] = 0.0f;
  for (int i = 1; i <= L; i++) {
    btot[i] = btot[i - 1] + (float)exp((double)(XMX(fx, i - 1, G_B) + XMX(bx, i - 1, G_B) - overall));
    etot[i] = etot[i - 1] + (float)exp((double)(XMX(fx, i, G_E) + XMX(bx, i, G_E) - overall));
    float njcp = expf(XMX(fx, i - 1, G_N) + XMX(bx, i, G_N) + gm->xsc[X_N][LOOP] - overall);
    njcp += expf(XMX(fx, i - 1, G_J) + XMX(bx, i, G_J) + gm->xsc[X_J][LOOP] - overall);
    njcp += expf(XMX(fx, i - 1, G_C) + XMX(bx, i, G_C) + gm->xsc[X_C][LOOP] - overall);
    mocc[i] = 1.0f - njcp;
  }
}

/* ---- multidomain regions: Easel fast RNG, stochastic traceback, clustering ---- */

/* p7_generic.rs:1356-1367 esl_mix3 */
static uint32_t mix3(uint32_t a, uint32_t b, uint32_t c) {
  a -= b; a -= c; a ^= c >> 13;
  b -= c; b -= a; b ^= a << 8;
  c -= a; c -= b; c ^= b >> 13;
  a -= b; a -= c; a ^= c >> 12;
  b -= c; b -= a; b ^= a << 16;
  c -= a; c -= b; c ^= b >> 5;
  a -= b; a -= c; a ^= c >> 3;
  b -= c; b -= a; b ^= a << 10;
  c -= a; c -= b; c ^= b >> 15;
  return c;
}
typedef struct { uint32_t x; } rng_t;
/* p7_generic.rs:1368-1397 */
static void rng_init(rng_t *r, uint32_t seed) {
  r->x = mix3(seed, 87654321u, 12345678u);
  if (r->x == 0) r->x = 42;
}
static double rng_random(rng_t *r) {
  r->x = r->x * 69069u + 1u;
  return (double)r->x / 4294967296.0;
}
static int rng_fchoose(rng_t *r, const float *p, int n) {
  double norm = 0.0, sum = 0.0;
  const double roll = rng_random(r);
  for (int i = 0; i < n; i++) norm += (double)p[i];
  for (int i = 0; i < n; i++) {
    sum += (double)p[i];
    if (roll < sum / norm) return i;
  }
  return n - 1;
}
/* p7_generic.rs:1399-1433 esl_vec_FSum (Kahan), FLogSum, FLogNorm */
static float vec_fsum(const float *v, int n) {
  float sum = 0.0f, c = 0.0f;
  for (int i = 0; i < n; i++) {
    const float y = v[i] - c;
    const float t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}
static float vec_flogsum(const float *v, int n) {
  float max = NEGINF;
  for (int i = 0; i < n; i++) max = fmaxf(max, v[i]);
  if (max == INFINITY) return INFINITY;
  float sum = 0.0f;
  for (int i = 0; i < n; i++)
    if (v[i] > max - 50.0f) sum += expf(v[i] - max);
  return logf(sum) + max;
}
static void vec_flognorm(float *v, int n) {
  const float denom = vec_flogsum(v, n);
  for (int i = 0; i < n; i++) v[i] += -denom;
  for (int i = 0; i < n; i++) v[i] = expf(v[i]);
  const float sum = vec_fsum(v, n);
  if (sum != 0.0f)
    for (int i = 0; i < n; i++) v[i] /= sum;
  else
    for (int i = 0; i < n; i++) v[i] = 1.0f / (float)n;
}

enum { ST_T = 0, ST_C = 1, ST_E = 2, ST_M = 3, ST_D = 4, ST_I = 5, ST_N = 6, ST_B = 7, ST_J = 8, ST_S = 9 };
typedef struct { uint8_t st; int k, i; } tstep;
typedef struct { int idx, i, j, k, m; } seg_t;

/* p7_generic.rs:1445-1562 p7_gstochastic_trace: appends the per-domain
 * (sqfrom, sqto, hmmfrom, hmmto) of one sampled trace (trace_index,
 * :1565-1584) to segs; returns the new count. */
static int stochastic_trace_segs(rng_t *r, int lr, const ixo_gm *gm, const float *dp, const float *xm, int t,
                                 int ri, seg_t **segs, int *nseg, int *cap) {
  const int M = gm->M;
  int n = 0, tcap = 2 * lr + 2 * M + 16;
  tstep *tr = (tstep *)malloc(sizeof(tstep) * (size_t)tcap);
  int k = 0, i = lr;
  tr[n++] = (tstep){ST_T, k, i};
  tr[n++] = (tstep){ST_C, k, i};
  uint8_t sprv = ST_C;
  while (sprv != ST_S) {
    const uint8_t last = tr[n - 1].st;
    uint8_t scur = ST_S;
    float sc[4];
    switch (last) {
      case ST_C:
        sc[0] = XMX(xm, i - 1, G_C) + gm->xsc[X_C][LOOP];
        sc[1] = XMX(xm, i, G_E) + gm->xsc[X_E][MOVE];
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_C : ST_E;
        break;
      case ST_E:
        k = M;
        sc[0] = MMX(dp, M, i, M);
        sc[1] = DMX(dp, M, i, M);
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_M : ST_D;
        break;
      case ST_M: {
        sc[0] = XMX(xm, i - 1, G_B) + TSC(gm, P_BM, k - 1);
        sc[1] = MMX(dp, M, i - 1, k - 1) + TSC(gm, P_MM, k - 1);
        sc[2] = IMX(dp, M, i - 1, k - 1) + TSC(gm, P_IM, k - 1);
        sc[3] = DMX(dp, M, i - 1, k - 1) + TSC(gm, P_DM, k - 1);
        vec_flognorm(sc, 4);
        const int c = rng_fchoose(r, sc, 4);
        scur = c == 0 ? ST_B : c == 1 ? ST_M : c == 2 ? ST_I : ST_D;
        k -= 1;
        i -= 1;
        break;
      }
      case ST_D:
        sc[0] = MMX(dp, M, i, k - 1) + TSC(gm, P_MD, k - 1);
        sc[1] = DMX(dp, M, i, k - 1) + TSC(gm, P_DD, k - 1);
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_M : ST_D;
        k -= 1;
        break;
      case ST_I:
        sc[0] = MMX(dp, M, i - 1, k) + TSC(gm, P_MI, k);
        sc[1] = IMX(dp, M, i - 1, k) + TSC(gm, P_II, k);
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_M : ST_I;
        i -= 1;
        break;
      case ST_N:
        scur = i == 0 ? ST_S : ST_N;
        break;
      case ST_B:
        sc[0] = XMX(xm, i, G_N) + gm->xsc[X_N][MOVE];
        sc[1] = XMX(xm, i, G_J) + gm->xsc[X_J][MOVE];
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_N : ST_J;
        break;
      case ST_J:
        sc[0] = XMX(xm, i - 1, G_J) + gm->xsc[X_J][LOOP];
        sc[1] = XMX(xm, i, G_E) + gm->xsc[X_E][LOOP];
        vec_flognorm(sc, 2);
        scur = rng_fchoose(r, sc, 2) == 0 ? ST_J : ST_E;
        break;
      default:
        free(tr);
        return -1;
    }
    if (n == tcap) {
      tcap *= 2;
      tr = (tstep *)realloc(tr, sizeof(tstep) * (size_t)tcap);
    }
    tr[n++] = (tstep){scur, k, i};
    if ((scur == ST_N || scur == ST_J || scur == ST_C) && scur == sprv) i -= 1;
    sprv = scur;
  }
  /* trace_index over the forward-ordered trace (reverse of tr) */
  int have = 0;
  seg_t cur = {0, 0, 0, 0, 0};
  for (int q = n - 1; q >= 0; q--) {
    const tstep s = tr[q];
    if (s.st == ST_B) {
      if (have) {
        if (*nseg == *cap) {
          *cap = *cap ? 2 * *cap : 64;
          *segs = (seg_t *)realloc(*segs, sizeof(seg_t) * (size_t)*cap);
        }
        (*segs)[(*nseg)++] = cur;
      }
      have = 1;
      cur = (seg_t){t, 0, 0, 0, 0};
    } else if (s.st == ST_M && have) {
      if (cur.i == 0) cur.i = s.i;
      if (cur.k == 0) cur.k = s.k;
      cur.j = s.i;
      cur.m = s.k;
    }
  }
  if (have) {
    if (*nseg == *cap) {
      *cap = *cap ? 2 * *cap : 64;
      *segs = (seg_t *)realloc(*segs, sizeof(seg_t) * (size_t)*cap);
    }
    (*segs)[(*nseg)++] = cur;
  }
  /* region-local -> dsq-local coordinates (glocal_region_trace_ensemble) */
  free(tr);
  return 0;
  (void)ri;
}

/* p7_generic.rs:1594-1616 link_spsamples */
static int link_segs(const seg_t *h1, const seg_t *h2) {
  int nov = (h1->j < h2->j ? h1->j : h2->j) - (h1->i > h2->i ? h1->i : h2->i) + 1;
  int n = h1->j - h1->i + 1 < h2->j - h2->i + 1 ? h1->j - h1->i + 1 : h2->j - h2->i + 1;
  if ((float)nov / (float)n < 0.8f) return 0;
  nov = (h1->m < h2->m ? h1->m : h2->m) - (h1->k > h2->k ? h1->k : h2->k);
  n = h1->m - h1->k + 1 < h2->m - h2->k + 1 ? h1->m - h1->k + 1 : h2->m - h2->k + 1;
  if ((float)nov / (float)n < 0.8f) return 0;
  if (abs((h1->i - h1->k) - (h2->i - h2->k)) <= 4) return 1;
  if (abs((h1->j - h1->m) - (h2->j - h2->m)) <= 4) return 1;
  return 0;
}

/* p7_generic.rs:1619-1648 single linkage (Easel's stack order) */
static int single_linkage(const seg_t *s, int n, int *c) {
  int *a = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *b = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  for (int v = 0; v < n; v++) a[v] = n - v - 1;
  int na = n, nb = 0, nc = 0;
  while (na > 0) {
    int v = a[--na];
    b[nb++] = v;
    while (nb > 0) {
      v = b[--nb];
      c[v] = nc;
      for (int i = na - 1; i >= 0; i--)
        if (link_segs(&s[v], &s[a[i]])) {
          const int w = a[i];
          a[i] = a[na - 1];
          na--;
          b[nb++] = w;
        }
    }
    nc++;
  }
  free(a);
  free(b);
  return nc;
}

static int iargmax(const int *v, int n) {
  int best = 0;
  for (int i = 1; i < n; i++)
    if (v[i] > v[best]) best = i;
  return best;
}

typedef struct { int i, j, k, m; float p; } sigc_t;
static int sigc_cmp(const void *a, const void *b) {
  const int x = ((const sigc_t *)a)->i, y = ((const sigc_t *)b)->i;
  return (x > y) - (x < y);
}

/* p7_generic.rs:1660-1752 spensemble_cluster + dominance removal; writes
 * (i, j) pairs; returns their count. */
static int cluster_segs(const seg_t *segs, int n, int nsamples, int *oi, int *oj) {
  if (n == 0) return 0;
  int *asg = (int *)malloc(sizeof(int) * (size_t)n);
  const int nc = single_linkage(segs, n, asg);
  sigc_t *sig = (sigc_t *)malloc(sizeof(sigc_t) * (size_t)nc);
  int nsig = 0;
  for (int c = 0; c < nc; c++) {
    int ninc = 0, last = -1;
    for (int h = 0; h < n; h++)
      if (asg[h] == c) {
        if (segs[h].idx != last) ninc++;
        last = segs[h].idx;
      }
    if ((float)ninc / (float)nsamples < 0.25f) continue;
    int imin = 0, imax = 0, jmin = 0, jmax = 0, kmin = 0, kmax = 0, mmin = 0, mmax = 0, started = 0;
    for (int h = 0; h < n; h++) {
      if (asg[h] != c) continue;
      const seg_t *s = &segs[h];
      if (!started) {
        imin = imax = s->i; jmin = jmax = s->j; kmin = kmax = s->k; mmin = mmax = s->m;
        started = 1;
      } else {
        if (s->i < imin) imin = s->i;
        if (s->i > imax) imax = s->i;
        if (s->j < jmin) jmin = s->j;
        if (s->j > jmax) jmax = s->j;
        if (s->k < kmin) kmin = s->k;
        if (s->k > kmax) kmax = s->k;
        if (s->m < mmin) mmin = s->m;
        if (s->m > mmax) mmax = s->m;
      }
    }
    const int thr = (int)ceilf((float)ninc * 0.02f);
    int w = imax - imin + 1 > jmax - jmin + 1 ? imax - imin + 1 : jmax - jmin + 1;
    if (kmax - kmin + 1 > w) w = kmax - kmin + 1;
    if (mmax - mmin + 1 > w) w = mmax - mmin + 1;
    int *epc = (int *)malloc(sizeof(int) * (size_t)w);
    /* leftmost i */
    memset(epc, 0, sizeof(int) * (size_t)(imax - imin + 1));
    for (int h = 0; h < n; h++) if (asg[h] == c) epc[segs[h].i - imin]++;
    int bi = imin;
    while (bi <= imax && epc[bi - imin] < thr) bi++;
    if (bi > imax) bi = imin + iargmax(epc, imax - imin + 1);
    /* leftmost k */
    memset(epc, 0, sizeof(int) * (size_t)(kmax - kmin + 1));
    for (int h = 0; h < n; h++) if (asg[h] == c) epc[segs[h].k - kmin]++;
    int bk = kmin;
    while (bk <= kmax && epc[bk - kmin] < thr) bk++;
    if (bk > kmax) bk = kmin + iargmax(epc, kmax - kmin + 1);
    /* rightmost j */
    memset(epc, 0, sizeof(int) * (size_t)(jmax - jmin + 1));
    for (int h = 0; h < n; h++) if (asg[h] == c) epc[segs[h].j - jmin]++;
    int bj = jmax;
    while (bj >= jmin && epc[bj - jmin] < thr) bj--;
    if (bj < jmin) bj = jmin + iargmax(epc, jmax - jmin + 1);
    /* rightmost m */
    memset(epc, 0, sizeof(int) * (size_t)(mmax - mmin + 1));
    for (int h = 0; h < n; h++) if (asg[h] == c) epc[segs[h].m - mmin]++;
    int bm = mmax;
    while (bm >= mmin && epc[bm - mmin] < thr) bm--;
    if (bm < mmin) bm = mmin + iargmax(epc, mmax - mmin + 1);
    free(epc);
    if (bi > bj || bk > bm) continue;
    sig[nsig++] = (sigc_t){bi, bj, bk, bm, (float)ninc / (float)nsamples};
  }
  /* stable sort by i (Rust sort_by is stable) */
  for (int a = 1; a < nsig; a++) {
    sigc_t t = sig[a];
    int b = a - 1;
    while (b >= 0 && sig[b].i > t.i) {
      sig[b + 1] = sig[b];
      b--;
    }
    sig[b + 1] = t;
  }
  (void)sigc_cmp;
  int *dom = (int *)calloc((size_t)(nsig ? nsig : 1), sizeof(int));
  for (int d = 0; d < nsig; d++)
    for (int d2 = d + 1; d2 < nsig; d2++) {
      const int nov = (sig[d].j < sig[d2].j ? sig[d].j : sig[d2].j) - (sig[d].i > sig[d2].i ? sig[d].i : sig[d2].i) + 1;
      if (nov == 0) break;
      const int nn = sig[d].j - sig[d].i + 1 < sig[d2].j - sig[d2].i + 1 ? sig[d].j - sig[d].i + 1
                                                                       : sig[d2].j - sig[d2].i + 1;
      if ((float)nov / (float)nn >= 0.8f) {
        if (sig[d].p > sig[d2].p) dom[d2] = 1;
        else dom[d] = 1;
      }
    }
  int no = 0;
  for (int d = 0; d < nsig; d++)
    if (!dom[d]) {
      oi[no] = sig[d].i;
      oj[no] = sig[d].j;
      no++;
    }
  free(dom);
  free(sig);
  free(asg);
  return no;
}

/* p7_generic.rs:1268-1280 */
static int is_multidomain_region(const float *btot, const float *etot, int i, int j, float rt3) {
  float max = -1.0f;
  for (int z = i; z <= j; z++) {
    const float e = etot[z] - etot[i - 1];
    const float b = btot[j] - btot[z - 1];
    const float x = e < b ? e : b;
    if (x > max) max = x;
  }
  return max >= rt3;
}

/* sentinel-padded copy of dsq[i..j] */
static uint8_t *subseq(const uint8_t *dsq, int i, int j) {
  const int n = j - i + 1;
  uint8_t *s = (uint8_t *)malloc((size_t)n + 2);
  s[0] = 255;
  memcpy(s + 1, dsq + i, (size_t)n);
  s[n + 1] = 255;
  return s;
}

/* p7_generic.rs:1285-1328 glocal_rescore_isolated_domain (do_null2 FALSE) */
static float rescore(const ixo_gm *gm, const uint8_t *dsq, int i, int j) {
  const int ld = j - i + 1;
  uint8_t *s = subseq(dsq, i, j);
  float *dp = (float *)malloc(sizeof(float) * (size_t)(ld + 1) * (gm->M + 1) * 3);
  float *xm = (float *)malloc(sizeof(float) * (size_t)(ld + 1) * G_NX);
  const float sc = ixo_gforward(s, ld, gm, dp, xm);
  free(dp);
  free(xm);
  free(s);
  return sc;
}

/* p7_generic.rs:1717-1760 glocal_region_trace_ensemble; returns the number
 * of (i, j) clusters written (dsq-local coordinates). */
static int region_trace_ensemble(ixo_gm *gm, const uint8_t *dsq, int ri, int j, int save_l, int ns, int *oi,
                                 int *oj) {
  const int lr = j - ri + 1;
  ixo_gm_reconfig_multihit(gm, save_l);
  uint8_t *reg = subseq(dsq, ri, j);
  float *dp = (float *)malloc(sizeof(float) * (size_t)(lr + 1) * (gm->M + 1) * 3);
  float *xm = (float *)malloc(sizeof(float) * (size_t)(lr + 1) * G_NX);
  ixo_gforward(reg, lr, gm, dp, xm);
  rng_t r;
  rng_init(&r, 181u);
  seg_t *segs = NULL;
  int nseg = 0, cap = 0;
  for (int t = 0; t < ns; t++) {
    const int n0 = nseg;
    stochastic_trace_segs(&r, lr, gm, dp, xm, t, ri, &segs, &nseg, &cap);
    for (int q = n0; q < nseg; q++) {
      segs[q].i += ri - 1;
      segs[q].j += ri - 1;
    }
  }
  ixo_gm_reconfig_unihit(gm, save_l);
  const int nc = cluster_segs(segs, nseg, ns, oi, oj);
  free(segs);
  free(dp);
  free(xm);
  free(reg);
  return nc;
}

/* p7_generic.rs:1765-1836 p7_domaindef_glocal (do_null2 FALSE) */
int ixo_domaindef_glocal(ixo_gm *gm, const uint8_t *dsq, int L, const float *fx, const float *bx, float rt1,
                         float rt2, float rt3, int ns, int is_unihit, ixo_domain *out, int cap) {
  const int save_l = gm->L;
  float *btot = (float *)malloc(sizeof(float) * (size_t)(L + 1));
  float *etot = (float *)malloc(sizeof(float) * (size_t)(L + 1));
  float *mocc = (float *)malloc(sizeof(float) * (size_t)(L + 1));
  ixo_gdomain_decoding(gm, L, fx, bx, btot, etot, mocc);
  if (!is_unihit) ixo_gm_reconfig_unihit(gm, save_l);
  int nd = 0, i = -1, triggered = 0;
  int *ci = (int *)malloc(sizeof(int) * (size_t)(ns * 4 + 16)), *cj = (int *)malloc(sizeof(int) * (size_t)(ns * 4 + 16));
  for (int j = 1; j <= L; j++) {
    if (!triggered) {
      if (mocc[j] - (btot[j] - btot[j - 1]) < rt2) i = j;
      else if (i == -1) i = j;
      if (mocc[j] >= rt1) triggered = 1;
    } else if (mocc[j] - (etot[j] - etot[j - 1]) < rt2) {
      if (is_multidomain_region(btot, etot, i, j, rt3)) {
        const int nc = region_trace_ensemble(gm, dsq, i, j, save_l, ns, ci, cj);
        for (int c = 0; c < nc; c++) {
          if (nd < cap) out[nd] = (ixo_domain){ci[c], cj[c], rescore(gm, dsq, ci[c], cj[c]), 0.0f};
          nd++;
        }
      } else {
        if (nd < cap) out[nd] = (ixo_domain){i, j, rescore(gm, dsq, i, j), 0.0f};
        nd++;
      }
      i = -1;
      triggered = 0;
    }
  }
  if (!is_unihit) ixo_gm_reconfig_multihit(gm, save_l);
  free(ci);
  free(cj);
  free(btot);
  free(etot);
  free(mocc);
  return nd;
}

/* cm_search.rs:2449-2600 trunc_env_def for pass `pass` (1 STD with gm, 2
 * 5' with Rgm, 3 3' with Lgm, 4 5'&3' with Tgm: gm_proto is that profile):
 * glocal statistics gfmu/gflambda in the STD pass, the local tail
 * lftau/lflambda otherwise; the score correction 0 (5') or ln(1/clen) (3');
 * the first / final residue of dsq[1..L] enforced by the 5' / 3' passes.
 * The profile is one copy carried through the windows in order, as there.
 * windows are (start, end) in dsq coordinates; returns the envelope count. */
int64_t ixo_env_def_pass(const ixo_gm *gm_proto, int pass, const ixo_bias *bf, const uint8_t *dsq, int64_t L,
                         const ixo_window *wins, int64_t nwin, const ixo_envopts *o, int32_t *es, int32_t *ee,
                         int64_t cap, ixo_envacct *acct) {
  const double ln2 = 0.69314718055994530942;
  memset(acct, 0, sizeof(*acct));
  ixo_gm gm;
  ixo_gm_copy(gm_proto, &gm);
  int64_t n = 0;
  const int is_std = pass == 1, safe = pass == 2 || pass == 3;
  const float corr = pass == 3 ? logf(1.0f / (float)o->clen) : 0.0f;
  const double mu = is_std ? o->gfmu : o->lftau, lam = is_std ? o->gflambda : o->lflambda;
  const int enf_first = pass == 2 || pass == 4, enf_final = pass == 3 || pass == 4;
  for (int64_t w = 0; w < nwin; w++) {
    const int ws = (int)wins[w].start, we = (int)wins[w].end;
    if (enf_first && ws != 1) continue;
    if (enf_final && we != L) continue;
    const int wlen = we - ws + 1;
    uint8_t *sub = subseq(dsq, ws, we);
    const double nullsc = ixo_bg_null_one(wlen);
    if (pass == 2) ixo_gm_reconfig_length_5p(&gm, wlen);
    else if (pass == 3) ixo_gm_reconfig_length_3p(&gm, wlen);
    else if (pass == 1) ixo_gm_reconfig_length(&gm, wlen);
    float *fdp = (float *)malloc(sizeof(float) * (size_t)(wlen + 1) * (gm.M + 1) * 3);
    float *fxm = (float *)malloc(sizeof(float) * (size_t)(wlen + 1) * G_NX);
    const float fwdsc = ixo_gforward(sub, wlen, &gm, fdp, fxm);
    const float safe_lfwdsc = fwdsc + corr;
    const double sc_f4 = safe ? ((double)safe_lfwdsc - nullsc) / ln2 : ((double)fwdsc - nullsc) / ln2;
    if (ixo_exp_surv(sc_f4, mu, lam) > o->f4) goto next;
    acct->n_past_gfwd++;
    acct->pos_past_gfwd += (uint64_t)wlen;
    if (o->do_gfwdbias) {
      const float filtersc = ixo_bias_score(bf, sub, wlen);
      const double sc_f4b = safe ? ((double)safe_lfwdsc - nullsc) / ln2 : ((double)fwdsc - (double)filtersc) / ln2;
      if (ixo_exp_surv(sc_f4b, mu, lam) > o->f4b) goto next;
      acct->n_past_gfwdbias++;
      acct->pos_past_gfwdbias += (uint64_t)wlen;
    }
    {
      float *bdp = (float *)malloc(sizeof(float) * (size_t)(wlen + 1) * (gm.M + 1) * 3);
      float *bxm = (float *)malloc(sizeof(float) * (size_t)(wlen + 1) * G_NX);
      ixo_gbackward(sub, wlen, &gm, bdp, bxm);
      ixo_domain doms[256];
      int nd = ixo_domaindef_glocal(&gm, sub, wlen, fxm, bxm, o->rt1, o->rt2, o->rt3, o->ns, !is_std, doms, 256);
      if (nd > 256) nd = 256;
      for (int d = 0; d < nd; d++) {
        const int env_len = doms[d].jenv - doms[d].ienv + 1;
        const float env_sc =
            doms[d].envsc + ((float)wlen - (float)env_len) * logf((float)wlen / ((float)wlen + 3.0f));
        double pv = ((double)env_sc - (nullsc + 0.0)) / ln2;
        if (safe) pv += (double)corr / ln2;
        if (ixo_exp_surv(pv, mu, lam) <= o->f5) {
          acct->n_past_edef++;
          acct->pos_past_edef += (uint64_t)env_len;
          int keep = 1;
          if (o->do_edefbias) {
            const float filtersc = ixo_bias_score(bf, sub, wlen);
            double scb = ((double)env_sc - (double)filtersc) / ln2;
            if (safe) scb += (double)corr / ln2;
            if (ixo_exp_surv(scb, mu, lam) > o->f5b) keep = 0;
            else {
              acct->n_past_edefbias++;
              acct->pos_past_edefbias += (uint64_t)env_len;
            }
          }
          if (keep) {
            if (n < cap) {
              es[n] = doms[d].ienv + ws - 1;
              ee[n] = doms[d].jenv + ws - 1;
            }
            n++;
          }
        }
      }
      free(bdp);
      free(bxm);
    }
  next:
    free(fdp);
    free(fxm);
    free(sub);
  }
  ixo_gm_free(&gm);
  return n;
}

int64_t ixo_env_def_std(const ixo_gm *gm_proto, const ixo_bias *bf, const uint8_t *dsq, int64_t L,
                        const ixo_window *wins, int64_t nwin, const ixo_envopts *o, int32_t *es, int32_t *ee,
                        int64_t cap, ixo_envacct *acct) {
  return ixo_env_def_pass(gm_proto, 1, bf, dsq, L, wins, nwin, o, es, ee, cap, acct);
}
