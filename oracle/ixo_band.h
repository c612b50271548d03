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
