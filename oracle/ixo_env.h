/*
 * ixo_env.h -- CPU oracle (test infrastructure only): F4/F5 envelope
 * definition (glocal p7 profile, generic glocal Forward/Backward in nats,
 * domain decoding and definition), restating p7_generic.rs and
 * cm_search.rs:2449-2600. Not linked into the product.
 */
#ifndef IXO_ENV_H
#define IXO_ENV_H
#include <stdint.h>
#include "ixo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* p7_generic.rs:134-147 GlocalProfile: tsc[(M+1)*8] (k*8 + MM IM DM BM MD DD
 * MI II), rsc[KP][(M+1)*2] (MSC, ISC), xsc[E N J C][LOOP MOVE], nats. */
typedef struct {
  int M;
  float *tsc;
  float *rsc;
  float xsc[4][2];
  float nj;
  int L;
  int local_end;
} ixo_gm;

/* p7_generic.rs:1260-1265 Domain (window-local 1-based envelope, nats) */
typedef struct {
  int ienv, jenv;
  float envsc, domcorrection;
} ixo_domain;

typedef struct {
  double f4, f4b, f5, f5b;   /* P-value thresholds */
  double gfmu, gflambda;     /* glocal Forward exponential tail */
  float rt1, rt2, rt3;       /* domain definition thresholds */
  int ns;                    /* stochastic traces per multidomain region */
  int do_gfwdbias, do_edefbias;
  double lftau, lflambda;    /* local Forward tail (truncated passes) */
  int clen, pad;             /* consensus length (the 3' pass's correction) */
} ixo_envopts;

typedef struct {
  uint64_t n_past_gfwd, pos_past_gfwd, n_past_gfwdbias, pos_past_gfwdbias;
  uint64_t n_past_edef, pos_past_edef, n_past_edefbias, pos_past_edefbias;
} ixo_envacct;

float ixo_p7_flogsum(float a, float b);
float ixo_p7_flogsum_table(int i);
int   ixo_gm_build(const ixo_p7 *p7, int L, ixo_gm *gm);
void  ixo_gm_free(ixo_gm *gm);
void  ixo_gm_copy(const ixo_gm *src, ixo_gm *dst);
void  ixo_gm_reconfig_length(ixo_gm *gm, int L);
void  ixo_gm_reconfig_unihit(ixo_gm *gm, int L);
void  ixo_gm_reconfig_multihit(ixo_gm *gm, int L);
void  ixo_gm_reconfig_length_5p(ixo_gm *gm, int L);
void  ixo_gm_reconfig_length_3p(ixo_gm *gm, int L);
/* the truncated passes' profiles: which 0 = 5' (Rgm), 1 = 3' (Lgm), 2 = 5'&3' (Tgm) */
int   ixo_gm_build_trunc(const ixo_p7 *p7, int which, ixo_gm *gm);
/* dp [(L+1)*(M+1)*3] (row i: M[0..M] I[0..M] D[0..M]), xm [(L+1)*5] (E N J B C) */
float ixo_gforward(const uint8_t *dsq, int L, const ixo_gm *gm, float *dp, float *xm);
float ixo_gbackward(const uint8_t *dsq, int L, const ixo_gm *gm, float *dp, float *xm);
void  ixo_gdomain_decoding(const ixo_gm *gm, int L, const float *fx, const float *bx, float *btot, float *etot,
                           float *mocc);
int   ixo_domaindef_glocal(ixo_gm *gm, const uint8_t *dsq, int L, const float *fx, const float *bx, float rt1,
                           float rt2, float rt3, int ns, int is_unihit, ixo_domain *out, int cap);
/* trunc_env_def for pass 1 (STD), 2 (5'), 3 (3') or 4 (5'&3') with that
 * pass's profile */
int64_t ixo_env_def_pass(const ixo_gm *gm_proto, int pass, const ixo_bias *bf, const uint8_t *dsq, int64_t L,
                         const ixo_window *wins, int64_t nwin, const ixo_envopts *o, int32_t *es, int32_t *ee,
                         int64_t cap, ixo_envacct *acct);
int64_t ixo_env_def_std(const ixo_gm *gm_proto, const ixo_bias *bf, const uint8_t *dsq, int64_t L,
                        const ixo_window *wins, int64_t nwin, const ixo_envopts *o, int32_t *es, int32_t *ee,
                        int64_t cap, ixo_envacct *acct);

#ifdef __cplusplus
}
#endif
#endif
