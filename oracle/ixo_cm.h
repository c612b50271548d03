/*
 * ixo_cm.h -- CPU oracle (test infrastructure only): covariance-model score
 * configuration and the HMM-banded CYK / Inside scanners of infernox.
 */
#ifndef IXO_CM_H
#define IXO_CM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IXO_MAXCONNECT 6
/* state types (constants.rs:9-18) */
enum { IXO_D_ST = 0, IXO_MP_ST = 1, IXO_ML_ST = 2, IXO_MR_ST = 3, IXO_IL_ST = 4, IXO_IR_ST = 5,
       IXO_S_ST = 6, IXO_E_ST = 7, IXO_B_ST = 8, IXO_EL_ST = 9 };
/* node types (constants.rs:24-31) */
enum { IXO_BIF_ND = 0, IXO_MATP_ND = 1, IXO_MATL_ND = 2, IXO_MATR_ND = 3, IXO_BEGL_ND = 4,
       IXO_BEGR_ND = 5, IXO_ROOT_ND = 6, IXO_END_ND = 7 };

/* A probability-parameterized CM (cm.rs:315-563, the fields the scanners'
 * score configuration reads). For B states cfirst is the left child S and
 * cnum the right child S (cm_modelmaker.rs:1001, :1083). */
typedef struct {
  int M, nodes, clen, W;
  float null[4];
  float el_selfsc;       /* log2 of the EL self-transition */
  float pbegin, pend;
  double n3_omega;
  const int8_t *sttype;  /* M */
  const int8_t *ndtype;  /* nodes */
  const int32_t *nodemap;/* nodes */
  const int32_t *cfirst, *cnum; /* M */
  const float *t;        /* M*6 transition probabilities */
  const float *e;        /* M*16 emission probabilities (first 4 for singlets) */
} ixo_cm;

/* Configured scores (cm_configure_scores output, cp9.rs:4538). */
typedef struct {
  int M;
  int local_begin;
  float el_selfsc;
  float *tsc;     /* M*6 */
  float *oesc;    /* M*18*18; singlet states use the first 18 */
  float *beginsc, *endsc; /* M */
} ixo_cmsc;

/* HMM bands (CP9Bands, cp9.rs:1453-1470): per state j band, per (v, j) d band
 * at hd_off[v] + (j - jmin[v]); hd_off has M+1 entries. */
typedef struct {
  const int32_t *jmin, *jmax;
  const int32_t *hd_off, *hdmin, *hdmax;
} ixo_bands;

typedef struct { int32_t i, j; float score, bias; } ixo_hit;

int  ixo_cm_configure(const ixo_cm *cm, ixo_cmsc *out);
void ixo_cm_free_scores(ixo_cmsc *s);
void ixo_ij2d_bands(const int8_t *sttype, int M, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                    const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax);
int64_t ixo_hb_ncells(const ixo_bands *b, int M);
/* Scans dsq[i0..j0] (1-indexed) with CYK (inside=0) or Inside (inside=1).
 * Writes vsc_root, envi, envj; optionally the root deck (state 0 rows) and
 * the whole matrix. Returns the number of root cells. */
int64_t ixo_hb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_bands *b, const uint8_t *dsq, int i0, int j0,
                    int inside, float env_cutoff, float *vsc, int64_t *envi, int64_t *envj, float *root_out,
                    float *alpha_out);
int64_t ixo_hb_report_hits(const ixo_cm *cm, const ixo_bands *b, const uint8_t *dsq, int i0, int j0,
                           const float *root, float cutoff, int do_null3, ixo_hit *out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
