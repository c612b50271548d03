/*
 * ixo_band.h -- CPU oracle (test infrastructure only): HMM->CM band
 * conversion (cp9.rs:144, :2893-3850). Not linked into the product.
 */
#ifndef IXO_BAND_H
#define IXO_BAND_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the CM structure the conversion walks (cm.rs:315-563) */
typedef struct {
  int M, nodes;
  const int8_t *sttype, *ndtype;
  const int32_t *nodemap, *cfirst, *cnum;
  int local_begin, local_end; /* CMH_LOCAL_BEGIN / CMH_LOCAL_END */
} ixo_hbcm;

/* the CP9 terms it reads: flags, M->E (esc), D->M and I->M scores, the EL
 * source lists (cp9.rs:871-896) */
typedef struct {
  int hmm_m;
  int local_begin, local_end, el;
  const int32_t *esc, *tdm, *tim; /* [hmm_m+1] */
  const int32_t *el_from_ct, *el_from_off, *el_from_idx;
} ixo_hbcp9;

void ixo_hb_emit_map(const ixo_hbcm *cm, int32_t *lpos, int32_t *rpos);

/* The truncated-band fields of CP9Bands (cp9.rs:1471-1491): the marginal
 * start / end nodes, the R and L marginal i / j bands, and per state (M
 * entries) whether J, L, R and T alignments through it are candidates. */
typedef struct {
  int32_t sp1, sp2, ep1, ep2;
  int32_t rmarg_imin, rmarg_imax, lmarg_jmin, lmarg_jmax;
  uint8_t *jvalid, *lvalid, *rvalid, *tvalid;
} ixo_trband;
/* cp9.rs:1755 cp9_predict_start_and_end_positions from the posterior bands
 * pn [6][hmm_m+1] (before hmm2ij doctors them) and the per-node occupancy
 * occ [hmm_m+1] (ixo_cp9_occupancy); thresholds thresh1 / thresh2. */
void ixo_tr_predict_start_end(const int32_t *pn, const int32_t *occ, int hmm_m, int i0, int j0, float thresh1,
                              float thresh2, ixo_trband *tb);
/* cp9.rs:1918 cp9_marginal_candidates_from_start_end_positions over the M
 * states (the EL override concerns the EL state, which is not one of them) */
void ixo_tr_candidates(const ixo_hbcm *cm, int pass_idx, ixo_trband *tb);
/* cp9.rs:3380 cp9_hmm2ijbands with do_trunc: as ixo_hmm2ij_bands, plus the
 * L/R/T band expansion and the truncated special cases; tb's validity
 * arrays are updated as the reference updates them. */
void ixo_hmm2ij_bands_tr(const ixo_hbcm *cm, const ixo_hbcp9 *h, int32_t *pn, int i0, int j0, int doing_search,
                         ixo_trband *tb, int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax);
/* pn [6][hmm_m+1] (pn_min_m, pn_max_m, pn_min_i, pn_max_i, pn_min_d,
 * pn_max_d) is doctored in place as the reference does */
void ixo_hmm2ij_bands(const ixo_hbcm *cm, const ixo_hbcp9 *h, int32_t *pn, int i0, int j0, int doing_search,
                      int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax);
int64_t ixo_hb_ij2d(const ixo_hbcm *cm, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                       const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax,
                       int64_t cap);
double ixo_hb_mx_mb(const ixo_hbcm *cm, const int32_t *jmin, const int32_t *jmax, const int32_t *hd_off,
                    const int32_t *hdmin, const int32_t *hdmax, int L, int64_t *ncells_out);

#ifdef __cplusplus
}
#endif
#endif
