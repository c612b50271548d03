/*
 * ixo.h -- CPU oracle (test infrastructure only) for the infernox per-window
 * search DP. A plain-C restatement of the reference algorithms, used by the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. Nothing in the product library links or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the infernox source tree). Compile with -O2 -ffp-contract=off and no
 * fast-math: float evaluation order is part of the contract.
 */
#ifndef IXO_H
#define IXO_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IXO_KP 18          /* RNA extended alphabet (easel/alphabet.rs:46-157) */
#define IXO_NEGINF_W (-32768)

/* P7 filter HMM, natural order. trans layout [MM,MI,MD,IM,II,DM,DD]
 * (p7_hmm.rs:62-90). mat: (M+1)*4, trans: (M+1)*7. */
typedef struct {
  int M;
  const float *mat;
  const float *trans;
  float compo[4];
  double lmmu, lmlambda, lvmu, lvlambda, lftau, lflambda;
} ixo_p7;

/* cm_pipeline.rs:182-201 MsvFilter */
typedef struct {
  int m;
  float scale_b;
  uint8_t base_b, bias_b, tbm_b, tec_b, tjb_b;
  uint8_t *rbv;       /* (m+1)*KP, rbv[k*KP+x] */
  double mmu, mlambda;
  float *prefix, *suffix; /* m+1 */
  int max_length;
} ixo_msv;

/* p7_vitfilter.rs:80-118 VitFilter (natural-order and striped) */
typedef struct {
  int m;
  float scale_w;
  int16_t base_w;
  int16_t *rwv;       /* (m+1)*KP */
  int16_t *tmm, *tim, *tdm, *vbm, *tmi, *tii, *tmd, *tdd; /* m+1 each */
  int16_t xw_e_move, xw_e_loop;
  int q;
  int16_t *tw;        /* (8q)*8 */
  int16_t *rw;        /* KP*q*8 */
  int16_t ddbound_w;
} ixo_vit;

/* cm_pipeline.rs:931-963 ForwardFilter */
typedef struct {
  int m;
  float *rfv;         /* (m+1)*KP */
  float *tbm, *amm, *aim, *adm, *tmi, *tii, *tmd, *tdd; /* m+1 each */
  double ftau, flambda;
  int nq;
  float *tmain;       /* nq*28 */
  float *tddv;        /* nq*4 */
  float *rfvv;        /* KP*nq*4 */
} ixo_fwd;

/* cm_pipeline.rs:1671-1680 BiasFilter */
typedef struct {
  float eo[IXO_KP][2];
  float t1[3];
  float pi[2];
} ixo_bias;

typedef struct { int64_t n, length, k; float score; } ixo_rawwin;   /* cm_pipeline.rs:455-461 */
typedef struct { int64_t start, end; float score; } ixo_window;     /* cm_pipeline.rs:465-470 */

typedef struct {                                                     /* cm_pipeline.rs:1814-1831 */
  uint64_t n_past_msv, pos_past_msv, n_past_msvbias, pos_past_msvbias;
  uint64_t n_past_vit, pos_past_vit, n_past_vitbias, pos_past_vitbias;
  uint64_t n_past_fwd, pos_past_fwd, n_past_fwdbias, pos_past_fwdbias;
} ixo_acct;

/* builders */
int  ixo_build_msv(const ixo_p7 *p7, int max_length, ixo_msv *out);
void ixo_free_msv(ixo_msv *f);
int  ixo_build_vit(const ixo_p7 *p7, ixo_vit *out);
void ixo_free_vit(ixo_vit *f);
int  ixo_build_fwd(const ixo_p7 *p7, ixo_fwd *out);
void ixo_free_fwd(ixo_fwd *f);
void ixo_build_bias(const float compo[4], int m, ixo_bias *out);

/* statistics */
double ixo_gumbel_invsurv(double p, double mu, double lambda);
double ixo_gumbel_surv(double x, double mu, double lambda);
double ixo_exp_surv(double x, double mu, double lambda);
double ixo_bg_null_one(int64_t L);

/* kernels (dsq 1-indexed, sentinels at [0] and [L+1]) */
int64_t ixo_ssv_diagonals(const ixo_msv *f, const uint8_t *dsq, int64_t L, double f1,
                          ixo_rawwin *out, int64_t cap);
int     ixo_ssv_thresh(const ixo_msv *f, double f1);
void    ixo_ssv_rowmax(const ixo_msv *f, const uint8_t *dsq, int64_t L, uint8_t *xe_out);
int64_t ixo_f1_msv_windows(const ixo_msv *f, const uint8_t *dsq, int64_t L, double f1,
                           int64_t cmw, ixo_window *out, int64_t cap);
int64_t ixo_extend_and_merge(const ixo_msv *f, ixo_rawwin *raws, int64_t n, int64_t target_len,
                             float pct_overlap, ixo_window *out);
int64_t ixo_split_windows(const ixo_window *in, int64_t n, int64_t cmw, ixo_window *out, int64_t cap);
int     ixo_msv_score(const ixo_msv *f, const uint8_t *dsq, int64_t L, int wlen_reconfig, float *sc);
int     ixo_vit_score_scalar(const ixo_vit *f, const uint8_t *dsq, int64_t L, float *sc);
int     ixo_vit_score_striped(const ixo_vit *f, const uint8_t *dsq, int64_t L, float *sc);
float   ixo_fwd_score_striped(const ixo_fwd *f, const uint8_t *dsq, int64_t L);
float   ixo_fwd_score_scalar(const ixo_fwd *f, const uint8_t *dsq, int64_t L);
/* p7_fwdback.rs:301 p7_forward + :529 forward_score: full odds-space
 * matrices mmx/imx/dmx [(L+1)*(m+1)], specials xmx [(L+1)*6] (E N J B C
 * SCALE), totscale; returns the score in nats. unihit: the unihit length model. */
float   ixo_p7_forward(const ixo_fwd *f, const uint8_t *dsq, int64_t L, int unihit, float *mmx, float *imx,
                       float *dmx, float *xmx, float *totscale);
float   ixo_bias_score(const ixo_bias *b, const uint8_t *dsq, int64_t L);

/* filter cascade (cm_pipeline.rs:1863-2164). opts bit0 do_msv, bit1 do_fwd,
 * bit2 do_fwdbias, bit3 do_vit, bit4 do_vitbias, bit5 do_msvbias. */
typedef struct {
  double f1, f1b, f2, f2b, f3, f3b;
  int64_t cmw, maxw;
  int opts;
} ixo_f3opts;
int64_t ixo_f3_filter_sequence(const ixo_msv *mf, const ixo_vit *vf, const ixo_fwd *ff,
                               const ixo_bias *bf, const ixo_p7 *p7,
                               const uint8_t *dsq, int64_t L, const ixo_f3opts *o,
                               ixo_window *out, int64_t cap, ixo_acct *acct);

/* log-sum tables (cp9.rs:21-74, :5388-5413) */
int   ixo_ilogsum(int s1, int s2);
int   ixo_ilogsum_table(int i);
float ixo_flogsum(float a, float b);
float ixo_flogsum_table(int i);

#ifdef __cplusplus
}
#endif
#endif
