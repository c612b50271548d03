/*
 * ixo_pipe.h -- CPU oracle (test infrastructure only): the STD pass of a
 * search over one sequence, F1 through F7 (cm_search.rs:1300-1957,
 * :2214-2860). Not linked into the product.
 */
#ifndef IXO_PIPE_H
#define IXO_PIPE_H
#include <stdint.h>
#include "ixo.h"
#include "ixo_band.h"
#include "ixo_cm.h"
#include "ixo_env.h"

#ifdef __cplusplus
extern "C" {
#endif

/* one model: the p7 filters, the glocal profile, the local CM with its
 * scores, the CP9 and the band conversion's view of both */
typedef struct {
  const ixo_msv *msv;
  const ixo_vit *vit;
  const ixo_fwd *fwd;
  const ixo_bias *bias;
  const ixo_p7 *p7;
  const ixo_gm *gm;
  const ixo_cm *cm;
  const ixo_cmsc *sc;
  const ixo_cp9 *cp9;
  const ixo_hbcm *hbcm;
  const ixo_hbcp9 *hbcp9;
  /* the truncated passes ([0] 5', [1] 3', [2] 5'&3'): their p7 profiles
   * (Rgm, Lgm, Tgm), CP9s and band-conversion views; the CM's marginal
   * emission scores and local truncated-begin penalties l [3][M] */
  const ixo_gm *gmt[3];
  const ixo_cp9 *trcp9[3];
  const ixo_hbcp9 *trhbcp9[3];
  const float *lmesc, *rmesc, *trp_l;
} ixo_pipe;

typedef struct {
  ixo_f3opts f3;
  ixo_envopts env;
  double f6, cyk_mu, cyk_lambda;
  float cyk_env_cutoff, pli_t;
  double fin_mu, fin_lambda, ez, e_report;
  double fcyk_tau, final_tau, maxtau;
  float size_limit;
  int do_fcyk, do_fcykenv, do_null3, do_cyk, do_top, do_bot;
  int trunc_mode, pad; /* 1: the terminal truncated passes as well */
} ixo_pipeopts;

typedef struct {
  int64_t start, stop;
  float score, bias;
  double evalue;
  int32_t strand, pass;
  int64_t seq; /* index of the target sequence */
} ixo_pipehit;

typedef struct {
  uint64_t nres;
  ixo_acct f;
  ixo_envacct e;
  uint64_t n_past_cyk, pos_past_cyk, n_hits;
  uint64_t tr_nres[3];
  ixo_acct tr_f[3];
  ixo_envacct tr_e[3];
  uint64_t tr_n_past_cyk[3], tr_pos_past_cyk[3], tr_hits[3];
} ixo_pipeacct;

/* seq: 1-indexed residues seq[1..L]; returns the number of reported hits
 * (beyond cap: counted, not written) */
int64_t ixo_search_std(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *seq, int64_t L, ixo_pipehit *out,
                       int64_t cap, ixo_pipeacct *A);
/* the same with the tasks on nthreads threads (identical result) */
int64_t ixo_search_std_mt(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *seq, int64_t L,
                          ixo_pipehit *out, int64_t cap, ixo_pipeacct *A, int nthreads);
/* cm_search.rs:1300 search over nseq sequences seqs[s][1..L[s]]: the flat
 * (sequence x chunk x strand) task list, one finalize over all hits
 * (overlap removal per sequence and strand; the E-value sort's third key is
 * the sequence index, cm_search.rs:1922-1931) */
int64_t ixo_search_multi_mt(const ixo_pipe *P, const ixo_pipeopts *o, const uint8_t *const *seqs, const int64_t *L,
                            int64_t nseq, ixo_pipehit *out, int64_t cap, ixo_pipeacct *A, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
