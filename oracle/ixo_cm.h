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
/* Integer scores of the QDB Inside scanner (cm_nohmm.rs:279 build_integer_scores). */
typedef struct {
  int32_t *itsc;     /* M*6 */
  int32_t *ioesc;    /* M*18*18 */
  int32_t *ibeginsc, *iendsc; /* M */
  int32_t iel_selfsc;
} ixo_cmisc;
/* local != 0: the local configuration of ixo_cm_configure; local == 0: the
 * global one (cm_nohmm.rs:264). is (may be NULL) receives integer scores. */
int  ixo_cm_configure_full(const ixo_cm *cm, int local, ixo_cmsc *out, ixo_cmisc *is);
void ixo_cm_free_iscores(ixo_cmisc *is);
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

/* ---- truncated scanners (trcm.c) ---- */

/* pipeline pass indices (cm_trunc.rs:49-55) and marginal modes (:34-38) */
enum { IXO_PASS_STD_ANY = 1, IXO_PASS_5P_ONLY_FORCE = 2, IXO_PASS_3P_ONLY_FORCE = 3,
       IXO_PASS_5P_AND_3P_FORCE = 4, IXO_PASS_5P_AND_3P_ANY = 5 };
enum { IXO_TRMODE_T = 0, IXO_TRMODE_R = 1, IXO_TRMODE_L = 2, IXO_TRMODE_J = 3, IXO_TRMODE_UNKNOWN = 4 };

/* Per-state marginal-mode validity of one envelope (CP9Bands jvalid..tvalid). */
typedef struct { const uint8_t *jvalid, *lvalid, *rvalid, *tvalid; } ixo_trvalid;

typedef struct { int32_t i, j; float score, bias; int32_t mode, root; } ixo_trhit;

/* Marginal left/right emission scores lmesc/rmesc [M*18] (cp9.rs:4441-4527). */
void ixo_cm_marginals(const ixo_cm *cm, float *lmesc, float *rmesc);
/* Consensus emit map lpos/rpos per node (cp9.rs:144-204). */
void ixo_emit_map(const ixo_cm *cm, int32_t *lpos, int32_t *rpos);
/* Marginal-mode validity per state from the envelope's consensus span
 * sp1..ep1 / sp2..ep2 (cp9.rs:1925-1990); arrays of M entries. */
/* cp9.rs:752 cm_expected_state_occupancy of the un-localized CM (psi[M]) */
void ixo_cm_psi(const ixo_cm *cm, double *psi);
/* cm_trunc.rs:192 TrPenalties::new on the global CM: g, l [3][M] log2
 * (0 = 5'&3', 1 = 5' only, 2 = 3' only; -1e36 = not allowed), ig, il the
 * Prob2Score integers (-987654321 = not allowed) */
void ixo_trpenalties(const ixo_cm *cm, float *g, float *l, int32_t *ig, int32_t *il);
void ixo_tr_validity(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int pass_idx, int sp1, int sp2,
                     int ep1, int ep2, uint8_t *jv, uint8_t *lv, uint8_t *rv, uint8_t *tv);
/* HMM-banded TrCYK (inside = 0, cm_dpsearch_trunc.rs:264-1029) and FTrInside
 * (inside = 1, :1055-1794) with the truncated greedy reporting and overlap
 * removal (:2322-2495). pty: the pass's truncated-begin penalties [M].
 * planes_out (optional): J|L|R|T planes, 4 * ncells floats. Returns the number
 * of hits (hits beyond cap are counted, not written). */
int64_t ixo_trhb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const float *lmesc, const float *rmesc, const float *pty,
                      const ixo_bands *b, const ixo_trvalid *val, const uint8_t *dsq, int i0, int j0, int pass_idx,
                      int local, int inside, float cutoff, int do_null3, float env_cutoff, float *vsc, int *vmode,
                      int64_t *envi, int64_t *envj, ixo_trhit *hits, int64_t cap, float *planes_out);

/* Integer marginal emission scores ilmesc/irmesc [M*18] (cp9.rs:4441-4527). */
void ixo_cm_imarginals(const ixo_cm *cm, int32_t *ilmesc, int32_t *irmesc);
/* cm_hit_AllowTruncation (cm_dpsearch_trunc.rs:2376-2422) for a hit start..stop
 * of mode `mode` entered at state b. */
int ixo_tr_allow(const ixo_cm *cm, const int32_t *lpos, const int32_t *rpos, int pass_idx, int local, int start,
                 int stop, int i0, int j0, int mode, int b);
/* remove_overlaps_greedy_tr (cm_dpsearch_trunc.rs:2470-2495). */
int64_t ixo_tr_remove_overlaps(ixo_trhit *tmp, int64_t ntmp, ixo_trhit *hits, int64_t cap);

/* cm_dpsearch_trunc.rs:1800 generic_tr_scan over dsq[i0..j0] (the QDB /
 * non-banded truncated scanners ref_tr_cyk_scan :207 and ref_itr_inside_scan
 * :226): truncated CYK on the float tables (inside = 0: sc, lmesc, rmesc, pty)
 * or truncated Inside on the integer ones (inside = 1: is, ilmesc, irmesc,
 * ipty), with per-state dmax (NULL: non-banded, W). Writes vsc_root, its mode
 * and the full-window best (bsc_full); returns the number of hits after
 * greedy reporting and overlap removal. best_out (optional): the per-(j, d)
 * best root score in bits, [(j-i0)*(W+1) + d]. */
int64_t ixo_trqdb_scan(const ixo_cm *cm, const ixo_cmsc *sc, const ixo_cmisc *is, const float *lmesc,
                       const float *rmesc, const int32_t *ilmesc, const int32_t *irmesc, const float *pty,
                       const int32_t *ipty, const int32_t *dmax, const uint8_t *dsq, int i0, int j0, int inside,
                       int pass_idx, int local, float cutoff, int do_null3, float *vsc, int *vmode, float *bsc_full,
                       ixo_trhit *hits, int64_t cap, float *best_out);

/* ---- QDB / non-banded scanners (qdb.c) ---- */

/* cm_nohmm.rs:406 generic_scan over dsq[i0..j0]: CYK on the float scores
 * (inside = 0) or Inside on the integer scores (inside = 1), with per-state
 * d bands dmin/dmax (NULL: non-banded, 0..W). Hits after greedy reporting and
 * overlap removal; returns their number (beyond cap: counted only).
 * root_out (optional): root scores in bits per (j, d), [(j-i0)*(W+1) + d]. */
int64_t ixo_qdb_scan(const ixo_cm *cm, const ixo_cmsc *fs, const ixo_cmisc *is, const int32_t *dmin,
                     const int32_t *dmax, const uint8_t *dsq, int i0, int j0, int inside, float cutoff, int do_null3,
                     ixo_hit *out, int64_t cap, float *root_out);
/* cm_nohmm.rs:855 cyk_seq_filter over dsq[1..n]: envelopes (es, ee). */
int64_t ixo_cyk_seq_filter(const ixo_cm *cm, const ixo_cmsc *fs, const int32_t *dmin1, const int32_t *dmax1,
                           const uint8_t *dsq, int n, float cutoff, int do_null3, int32_t *es, int32_t *ee,
                           int64_t cap);

/* A configured CP9 HMM (cp9.rs:871-896): integer scores in the DP's order.
 * otsc [(M+1)][12] in cp9O order MM IM DM BM MI II DI MD ID DD ME MEL;
 * msc/isc [(M+1)][18]; EL: has_el [M+1], and per node k = 0..M+1 the
 * el_from_ct[k] source nodes el_from_idx[el_from_off[k] + c]. */
typedef struct {
  int M;
  int el; /* CPLAN9_EL set */
  int32_t el_selfsc;
  const int32_t *otsc, *msc, *isc;
  const uint8_t *has_el;
  const int32_t *el_from_ct, *el_from_off, *el_from_idx;
} ixo_cp9;

float ixo_cp9_forward(const ixo_cp9 *h, const uint8_t *dsq, int i0, int j0, int do_scan, int doing_align,
                      int32_t *mmx, int32_t *imx, int32_t *dmx, int32_t *elmx, int32_t *sca, int64_t *best_pos);
float ixo_cp9_backward(const ixo_cp9 *h, const uint8_t *dsq, int i0, int j0, int do_scan, int doing_align,
                       int32_t *mmx, int32_t *imx, int32_t *dmx, int32_t *elmx, int32_t *sca, int64_t *best_pos);
int32_t ixo_cp9_band_thresh(double p_thresh);
void ixo_cp9_fb2hmmbands(const ixo_cp9 *h, const uint8_t *dsq, const int32_t *fm, const int32_t *fi,
                         const int32_t *fd, const int32_t *bm, const int32_t *bi, const int32_t *bd, int i0, int j0,
                         double p_thresh, int did_fwd_scan, int32_t *pn);
/* per-node posterior occupancy (M + D posterior summed over the window's
 * rows, the quantity cp9_predict_start_and_end_positions thresholds) */
void ixo_cp9_occupancy(const ixo_cp9 *h, const uint8_t *dsq, const int32_t *fm, const int32_t *fd,
                       const int32_t *bm, const int32_t *bd, int i0, int j0, int did_fwd_scan, int32_t *occ);

#ifdef __cplusplus
}
#endif
#endif
