/*
 * ixg.h -- C ABI of the MI355X (gfx950) per-window search DP for infernox.
 *
 * Plain C types only: pointers, sizes, status codes. Every entry point
 * returns IXG_OK (0) or a negative IXG_E* status; nothing throws. Host
 * pointers are read before the call returns (inputs) or filled before it
 * returns (outputs). Coordinates are 1-based and inclusive, chunk-local,
 * exactly as the reference functions they replace.
 *
 * Each declaration names the reference function (file:line in infernox
 * v0.2.0) whose signature and semantics it replaces.
 */
#ifndef IXG_H
#define IXG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IXG_OK 0
#define IXG_EINVAL (-1)   /* bad argument (shape, size, null pointer) */
#define IXG_ENOMEM (-2)   /* device allocation or memory budget exceeded */
#define IXG_EDEVICE (-3)  /* HIP runtime error (message via ixg_last_error) */
#define IXG_ERANGE (-4)   /* an output list exceeded its capacity */

#define IXG_KP 18         /* digitized RNA alphabet size (easel/alphabet.rs:46-157) */

typedef struct ixg_ctx ixg_ctx;     /* one device + one stream + scratch */
typedef struct ixg_p7 ixg_p7;       /* device-resident p7 filter profile */

/* ---- context ------------------------------------------------------------ */
/* Replaces nothing in the reference (which has no device); owns the HIP
 * stream and scratch used by every call made with it. mem_budget = 0 means
 * no limit. One ctx per calling thread. */
int ixg_init(int device, size_t mem_budget, ixg_ctx **out);
int ixg_shutdown(ixg_ctx *ctx);
const char *ixg_last_error(void);
/* Device bytes currently held by ctx scratch plus uploaded profiles. */
size_t ixg_mem_in_use(const ixg_ctx *ctx);

/* ---- p7 filter HMM (P7Profile, p7_hmm.rs:62-90) -------------------------- */
typedef struct {
  int32_t M;
  const float *mat;     /* (M+1)*4, match emissions A,C,G,U; row 0 unused */
  const float *trans;   /* (M+1)*7, [MM,MI,MD,IM,II,DM,DD] */
  float compo[4];       /* mean residue composition (COMPO) */
  double lmmu, lmlambda;  /* local MSV Gumbel */
  double lvmu, lvlambda;  /* local Viterbi Gumbel */
  double lftau, lflambda; /* local Forward exponential tail */
} ixg_p7_hmm;

/* Builds the MSV (cm_pipeline.rs:204 build_msv_filter), Viterbi
 * (p7_vitfilter.rs:121 build_vit_filter), Forward (cm_pipeline.rs:969
 * build_forward_filter) and bias (cm_pipeline.rs:1683 build_bias_filter)
 * filter profiles on the host and uploads them; they stay resident in HBM
 * until ixg_p7_free. max_length is the reference's maxW (om->max_length).
 * With ctx == NULL only the host tables are built (for inspection through
 * ixg_p7_*_tables); such a handle is freed with ixg_p7_free(NULL, p). */
int ixg_p7_create(ixg_ctx *ctx, const ixg_p7_hmm *hmm, int32_t max_length, ixg_p7 **out);
int ixg_p7_free(ixg_ctx *ctx, ixg_p7 *p7);

/* Copies of the host-built tables, for parity checks against a CPU build.
 * rbv: (M+1)*18 u8. bytes: {scale_b(float bits)x4, base_b, bias_b, tbm_b,
 * tec_b, tjb_b}. prefix/suffix: (M+1) f32. Any pointer may be NULL. */
int ixg_p7_msv_tables(const ixg_p7 *p7, uint8_t *rbv, uint8_t *bytes5, float *prefix, float *suffix);
/* Viterbi: rwv (M+1)*18 i16, trans8 = 8 arrays of (M+1) i16 in the order
 * tmm,tim,tdm,vbm,tmi,tii,tmd,tdd; ddbound (1 i16). */
int ixg_p7_vit_tables(const ixg_p7 *p7, int16_t *rwv, int16_t *trans8, int16_t *ddbound);
/* Forward: tmain nq*28, tddv nq*4, rfvv 18*nq*4 (f32); nq returned. */
int ixg_p7_fwd_tables(const ixg_p7 *p7, int32_t *nq, float *tmain, float *tddv, float *rfvv);

/* ---- records ------------------------------------------------------------- */
typedef struct { int64_t n, length, k; float score; int32_t task; } ixg_rawwin; /* RawWin, cm_pipeline.rs:455 */
typedef struct { int64_t start, end; float score; int32_t task; } ixg_window;   /* Window, cm_pipeline.rs:465 */
typedef struct {                                                              /* F1F3Acct, cm_pipeline.rs:1814 */
  uint64_t n_past_msv, pos_past_msv, n_past_msvbias, pos_past_msvbias;
  uint64_t n_past_vit, pos_past_vit, n_past_vitbias, pos_past_vitbias;
  uint64_t n_past_fwd, pos_past_fwd, n_past_fwdbias, pos_past_fwdbias;
} ixg_acct;

/* ---- F1: SSV longtarget diagonals ------------------------------------------
 * Replaces cm_pipeline.rs:521 ssv_diagonals(&MsvFilter, &[u8], usize, f64)
 * -> Vec<RawWin>. dsq is 1-indexed (dsq[1..=L]; dsq[0], dsq[L+1] are not
 * read). Writes up to cap diagonals in the reference's order; *n gets the
 * full count (IXG_ERANGE if it exceeds cap). */
int ixg_ssv_diagonals(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, double f1,
                      ixg_rawwin *out, int64_t cap, int64_t *n);

/* ---- per-window scores (F1b, F2, F3, F3b) ------------------------------------
 * A batch of nwin windows, window w = seq[off[w]+1 .. off[w]+len[w]] (the
 * same 1-indexed convention: seq + off[w] plays the role of the window's
 * sentinel-padded dsq), so off[w] + len[w] < seq_len. Residue codes must be
 * below 18. ovf[w] = 1 where the reference returns None. */
/* p7_vitfilter.rs:292 vit_score(&VitFilter, &[u8], usize) -> Option<f32> */
int ixg_vit_scores(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len,
                   const int64_t *off, const int64_t *len, int64_t nwin, float *sc, uint8_t *ovf);
/* cm_pipeline.rs:1420 forward_filter_score(&ForwardFilter, &[u8], usize) -> f32 (nats) */
int ixg_fwd_scores(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len,
                   const int64_t *off, const int64_t *len, int64_t nwin, float *sc);
/* No counterpart in the reference. ixg_fwd_scores without any device
 * allocation, for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The sequence, the
 * window list and the scores are staged in a fixed device array of
 * IXG_NOALLOC_BYTES that each GPU holds until the process ends; its address
 * is looked up once per device by ixg_init. Limits: seq_len <=
 * IXG_NOALLOC_MAX_SEQ and nwin <= IXG_NOALLOC_MAX_WINDOWS (else IXG_ERANGE).
 * Calls on one GPU share the array: calls that overlap on one GPU (from
 * several threads or contexts) race. Scores equal ixg_fwd_scores'. */
#define IXG_NOALLOC_BYTES (24 << 20)
#define IXG_NOALLOC_MAX_SEQ (16 << 20)
#define IXG_NOALLOC_MAX_WINDOWS (1 << 18)
int ixg_fwd_scores_noalloc(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len,
                           const int64_t *off, const int64_t *len, int64_t nwin, float *sc);
/* cm_pipeline.rs:1722 bias_filter_score(&BiasFilter, &[u8], usize) -> f32 (nats) */
int ixg_bias_scores(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len,
                    const int64_t *off, const int64_t *len, int64_t nwin, float *sc);
/* cm_pipeline.rs:426 msv_score_reconfig(f, dsq, l, wlen=l) -> Option<f32> (F1b) */
int ixg_msv_scores(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len,
                   const int64_t *off, const int64_t *len, int64_t nwin, float *sc, uint8_t *ovf);

/* p7_fwdback.rs:301 p7_forward(&ForwardFilter, &XFactors, &[u8], l) -> Omx and
 * :529 forward_score: the full-matrix odds-space Forward over dsq[1..L]
 * (1-indexed, codes < 18) in the multihit (unihit = 0, XFactors::multihit) or
 * unihit (XFactors::unihit) length model. Writes the retained rows 0..L in
 * natural order, (L+1)*(M+1) floats each for mmx, imx, dmx, the specials
 * E N J B C SCALE per row ((L+1)*6 floats), totscale and the score (nats). */
int ixg_p7_forward(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, int unihit, float *mmx,
                   float *imx, float *dmx, float *xmx, float *totscale, float *sc);
/* Batched form: n regions of seq (region r = seq[off[r]+1 .. off[r]+len[r]],
 * off[r] + len[r] < seq_len); the matrices of region r follow those of the
 * regions before it in mat (M | I | D, 3*(len+1)*(M+1) floats per region)
 * and its specials in xmx (6*(len+1) floats per region). */
int ixg_p7_forward_batch(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                         const int64_t *len, int64_t n, int unihit, float *mat, float *xmx, float *sc,
                         float *totscale);

/* ---- the filter cascade -------------------------------------------------------- */
#define IXG_DO_MSV      (1 << 0)
#define IXG_DO_FWD      (1 << 1)
#define IXG_DO_FWDBIAS  (1 << 2)
#define IXG_DO_VIT      (1 << 3)
#define IXG_DO_VITBIAS  (1 << 4)
#define IXG_DO_MSVBIAS  (1 << 5)
typedef struct {
  double f1, f1b, f2, f2b, f3, f3b; /* P-value thresholds */
  int64_t cmw;                       /* the CM's W (window split length) */
  int64_t maxw;                      /* maxW (tiling when MSV is off) */
  int32_t opts;                      /* IXG_DO_* */
} ixg_f3opts;

/* Replaces cm_pipeline.rs:1863 f3_filter_sequence(...) -> (Vec<Window>,
 * F1F3Acct) for one chunk (dsq 1-indexed, L residues). */
int ixg_f3_filter_sequence(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *dsq, int64_t L,
                           const ixg_f3opts *o, ixg_window *out, int64_t cap, int64_t *n,
                           ixg_acct *acct);

/* ---- batched form over a device-resident target ---------------------------------
 * A target is a digitized sequence kept in HBM together with its reverse
 * complement (cm_search.rs:1822-1830: 3-d for d<4). Tasks are the
 * reference's (chunk x strand) windows (cm_search.rs:1774 window_tasks).  */
typedef struct ixg_target ixg_target;
typedef struct {
  int64_t start;   /* 1-based global start of the chunk (wgs) */
  int64_t end;     /* 1-based global end (wge) */
  int32_t strand;  /* 0 top, 1 bottom (reverse complement of [start,end]) */
  int32_t ctx;     /* residues of maxW context shared with the previous chunk */
} ixg_task;
int ixg_target_create(ixg_ctx *ctx, const uint8_t *dsq, int64_t L, ixg_target **out);
/* A target of nseq sequences (a database: cmsearch over several sequences,
 * cm_search.rs:1774-1808 window_tasks over `fulls`), each held with its
 * reverse complement; the batched scanners' envelopes and windows name the
 * sequence in their `seq` field. dsq[s] is 1-indexed with L[s] residues. */
int ixg_target_create_multi(ixg_ctx *ctx, const uint8_t *const *dsq, const int64_t *L, int64_t nseq,
                            ixg_target **out);
int ixg_target_free(ixg_ctx *ctx, ixg_target *t);
/* Builds the reference's task list for one sequence of length L
 * (cm_search.rs:1774-1808, CM_MAX_RESIDUE_COUNT = 100000). Returns the count;
 * writes at most cap tasks. */
int64_t ixg_window_tasks(int64_t L, int64_t maxw, int do_top, int do_bot, ixg_task *out, int64_t cap);

/* Runs f3_filter_sequence for every task in one pass over the device:
 * one F1 sweep over all tasks, then one F2/F3/F3b launch over every surviving
 * window. out[] receives each task's merged windows (chunk-local coordinates,
 * tagged with their task index) grouped by task in task order; acct[t] per
 * task (may be NULL). *n = total windows (IXG_ERANGE if above cap). */
int ixg_f3_filter_tasks(ixg_ctx *ctx, const ixg_p7 *p7, const ixg_target *tgt,
                        const ixg_task *tasks, int64_t ntasks, const ixg_f3opts *o,
                        ixg_window *out, int64_t cap, int64_t *n, ixg_acct *acct);

/* ---- covariance models: HMM-banded CYK / Inside scanners (F6/F7) --------------- */
typedef struct ixg_cm ixg_cm;   /* device-resident configured CM */
/* A probability-parameterized CM (cm.rs:315-563): states 0..M-1 with types
 * (constants.rs:9-18), children cfirst[v]..cfirst[v]+cnum[v]-1 (for a B state
 * cfirst is the left S and cnum the right S), node types and first states,
 * transition probabilities t[M][6] and emission probabilities e[M][16]
 * (pair states use all 16, singlets the first 4). */
typedef struct {
  int32_t M, nodes, clen, W;
  float null[4];
  float el_selfsc;      /* log2 of the EL self-transition */
  float pbegin, pend;   /* local begin/end totals */
  double n3_omega;
  const int8_t *sttype; /* M */
  const int8_t *ndtype; /* nodes */
  const int32_t *nodemap; /* nodes */
  const int32_t *cfirst, *cnum; /* M */
  const float *t;       /* M*6 */
  const float *e;       /* M*16 */
} ixg_cm_params;
/* Configures the CM for local search as cp9.rs:4538 cm_configure_scores does
 * (cm_config_local + cm_logoddsify + the EL self-score bound) and uploads the
 * score tables; ctx == NULL builds the host tables only. */
int ixg_cm_create(ixg_ctx *ctx, const ixg_cm_params *p, ixg_cm **out);
int ixg_cm_free(ixg_ctx *ctx, ixg_cm *cm);
/* tsc M*6, oesc M*324, beginsc M, endsc M, el_selfsc 1 (any may be NULL). */
int ixg_cm_tables(const ixg_cm *cm, float *tsc, float *oesc, float *beginsc, float *endsc, float *el_selfsc);
/* As ixg_cm_create with a choice of configuration: local != 0 is
 * cm_configure_scores (cp9.rs:4538), local == 0 the global configuration
 * of cm_nohmm.rs:264 cm_configure_scores_global (no local begins or ends).
 * Both also build the integer scores of cm_nohmm.rs:279
 * build_integer_scores for the QDB Inside scanner. */
int ixg_cm_create_cfg(ixg_ctx *ctx, const ixg_cm_params *p, int local, ixg_cm **out);
/* Integer scores: itsc M*6, ioesc M*324, ibeginsc M, iendsc M, iel_selfsc 1. */
int ixg_cm_itables(const ixg_cm *cm, int32_t *itsc, int32_t *ioesc, int32_t *ibeginsc, int32_t *iendsc,
                   int32_t *iel_selfsc);

/* HMM bands of one envelope (CP9Bands, cp9.rs:1453-1470): j band per state,
 * d band per (v, j) at hd_off[v] + (j - jmin[v]); hd_off has M+1 entries. */
typedef struct {
  const int32_t *jmin, *jmax, *hd_off, *hdmin, *hdmax;
} ixg_bands;
typedef struct { int32_t i, j; float score, bias; int32_t env; } ixg_hit;

/* cp9.rs:4682 fast_cyk_scan_hb(&CM, &CP9Bands, &[u8], i0, j0, env_cutoff)
 * -> (vsc_root, envi, envj); envi/envj = -1 when no cell reaches env_cutoff.
 * dsq is 1-indexed with L residues. */
int ixg_cyk_scan_hb(ixg_ctx *ctx, const ixg_cm *cm, const ixg_bands *b, const uint8_t *dsq, int64_t L,
                    int32_t i0, int32_t j0, float env_cutoff, float *vsc, int64_t *envi, int64_t *envj);
/* cp9.rs:5465 fast_finside_scan_hb (inside = 1) and cp9.rs:4704
 * fast_fcyk_scan_hb (inside = 0): the scan plus ReportHitsGreedily with the
 * null3 correction (cp9.rs:5950-5997). */
int ixg_inside_scan_hb(ixg_ctx *ctx, const ixg_cm *cm, const ixg_bands *b, const uint8_t *dsq, int64_t L,
                       int32_t i0, int32_t j0, int inside, float env_cutoff, float cutoff, int do_null3,
                       float *vsc, int64_t *envi, int64_t *envj, ixg_hit *hits, int64_t cap, int64_t *nhits);

/* Batched form: n envelopes of one CM over a device-resident target, all in
 * one launch. Envelope k scans [i0, j0] of strand `strand` (1-based
 * coordinates of that strand's sequence; strand 1 is the reverse complement).
 * Per-envelope results go to vsc/envi/envj[k]; with report != 0 the greedy
 * hits of every envelope are appended to hits (tagged with env = k, grouped
 * by envelope in order). cells (may be NULL) receives the number of banded
 * cells filled. */
typedef struct {
  int32_t i0, j0, strand;
  int32_t seq;  /* sequence of a multi-sequence target (ixg_target_create_multi), else 0 */
  ixg_bands bands;
} ixg_envelope;
int ixg_hb_scan_batch(ixg_ctx *ctx, const ixg_cm *cm, const ixg_target *tgt, const ixg_envelope *env, int64_t n,
                      int inside, float env_cutoff, int report, float cutoff, int do_null3, float *vsc,
                      int64_t *envi, int64_t *envj, ixg_hit *hits, int64_t cap, int64_t *nhits, int64_t *cells);
/* The same over envelopes of several CMs in one persistent launch (cmscan's
 * per-query loop over models, cm_search.rs:1357-1366, batched): envelope k
 * belongs to cms[env_cm[k]] (ncm models, all created on ctx's device).
 * Workgroups claim envelopes, largest first, from a device work queue.
 * Results are per envelope exactly as ixg_hb_scan_batch gives for that
 * envelope's model alone. */
int ixg_hb_scan_multi(ixg_ctx *ctx, const ixg_cm *const *cms, int32_t ncm, const int32_t *env_cm,
                      const ixg_target *tgt, const ixg_envelope *env, int64_t n, int inside, float env_cutoff,
                      int report, float cutoff, int do_null3, float *vsc, int64_t *envi, int64_t *envj,
                      ixg_hit *hits, int64_t cap, int64_t *nhits, int64_t *cells);

/* ---- truncated HMM-banded scanners: TrCYK / FTrInside (F6/F7, truncated passes) -- */
/* Pipeline pass indices (cm_trunc.rs:49-55) and marginal modes (:34-38). */
enum { IXG_PASS_STD_ANY = 1, IXG_PASS_5P_ONLY_FORCE = 2, IXG_PASS_3P_ONLY_FORCE = 3,
       IXG_PASS_5P_AND_3P_FORCE = 4, IXG_PASS_5P_AND_3P_ANY = 5 };
enum { IXG_TRMODE_T = 0, IXG_TRMODE_R = 1, IXG_TRMODE_L = 2, IXG_TRMODE_J = 3, IXG_TRMODE_UNKNOWN = 4 };

/* Truncated-begin penalties of the CM (TrPenalties, cm_trunc.rs:138-150): g
 * (global) and l (local) are [3][M] log2 scores, first index 0 = 5'&3',
 * 1 = 5' only, 2 = 3' only; <= -1e36 where a truncated entry is not allowed.
 * Must be set once before the truncated scanners run. The marginal emission
 * scores (cp9.rs:4441 fill_marginal_emissions) are built by ixg_cm_create. */
int ixg_cm_set_trpenalties(ixg_ctx *ctx, ixg_cm *cm, const float *g, const float *l);
/* Host builder of those penalties from the probability CM, replacing
 * cm_trunc.rs:192 TrPenalties::new(&cm_global, &emap, &psi) as
 * cm_search.rs:927-930 calls it (psi: cm_expected_state_occupancy,
 * cp9.rs:752, of the un-localized CM): g, l [3][M] for
 * ixg_cm_set_trpenalties and ig, il [3][M] (Prob2Score, -987654321 where
 * not allowed) for ixg_cm_set_itrpenalties. */
int ixg_cm_trpenalties(const ixg_cm_params *p, float *g, float *l, int32_t *ig, int32_t *il);
/* Host copies: lmesc/rmesc M*18; lpos/rpos per state: the consensus span of
 * the state's subtree (the emit map, cp9.rs:144-204, with the per-node-type
 * offsets of cm_dpsearch_trunc.rs:2394-2403). Any may be NULL. */
int ixg_cm_trtables(const ixg_cm *cm, float *lmesc, float *rmesc, int32_t *lpos, int32_t *rpos);

/* Marginal-mode validity of one envelope's states (CP9Bands jvalid, lvalid,
 * rvalid, tvalid, cp9.rs:1488-1491): M bytes each, nonzero = valid. */
typedef struct {
  const uint8_t *jvalid, *lvalid, *rvalid, *tvalid;
} ixg_trvalid;
/* A truncated hit (cm_dpsearch_trunc.rs:110-118 TrHit) + its envelope. */
typedef struct { int32_t i, j; float score, bias; int32_t mode, root, env; } ixg_trhit;

/* cm_dpsearch_trunc.rs:264 tr_cyk_scan_hb(cm, trp, cp9b, dsq, i0, j0, cutoff,
 * do_null3, pass_idx, env_cutoff, local) -> (hits, vsc, vmode, envi, envj):
 * the four marginal planes J/L/R/T over the HMM bands, truncated begins into
 * the root, greedy reporting with allow_truncation and null3, and greedy
 * overlap removal. pass_idx is one of IXG_PASS_5P_ONLY_FORCE ..
 * IXG_PASS_5P_AND_3P_ANY. d bands may start at max(left, right) emission
 * count (cp9.rs:3832-3836). dsq is 1-indexed with L residues. Hits beyond
 * cap are counted in *nhits and IXG_ERANGE is returned. */
int ixg_trcyk_scan_hb(ixg_ctx *ctx, const ixg_cm *cm, const ixg_bands *b, const ixg_trvalid *valid,
                      const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0, float cutoff, int do_null3,
                      int pass_idx, float env_cutoff, int local, float *vsc, int32_t *vmode, int64_t *envi,
                      int64_t *envj, ixg_trhit *hits, int64_t cap, int64_t *nhits);
/* cm_dpsearch_trunc.rs:1055 ftr_inside_scan_hb: the same with FLogsum child
 * and split combines. */
int ixg_trinside_scan_hb(ixg_ctx *ctx, const ixg_cm *cm, const ixg_bands *b, const ixg_trvalid *valid,
                         const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0, float cutoff, int do_null3,
                         int pass_idx, float env_cutoff, int local, float *vsc, int32_t *vmode, int64_t *envi,
                         int64_t *envj, ixg_trhit *hits, int64_t cap, int64_t *nhits);
/* Batched form over a device-resident target (as ixg_hb_scan_batch):
 * inside selects FTrInside; per-envelope vsc/vmode/envi/envj; hits of every
 * envelope appended in envelope order (env = k). */
typedef struct {
  int32_t i0, j0, strand, seq;  /* seq as in ixg_envelope */
  ixg_bands bands;
  ixg_trvalid valid;
} ixg_trenvelope;
int ixg_trhb_scan_batch(ixg_ctx *ctx, const ixg_cm *cm, const ixg_target *tgt, const ixg_trenvelope *env,
                        int64_t n, int inside, int pass_idx, int local, float env_cutoff, float cutoff,
                        int do_null3, float *vsc, int32_t *vmode, int64_t *envi, int64_t *envj, ixg_trhit *hits,
                        int64_t cap, int64_t *nhits, int64_t *cells);

/* ---- query-dependent-banded / non-banded scanners (cm_nohmm.rs) ------------------ */
/* cm_nohmm.rs:406 generic_scan over dsq[i0..j0]: CYK on the float scores
 * (inside = 0; cyk_seq_filter's scan, :855) or Inside on the integer scores
 * (inside = 1; final_stage_inside_opt, :919), with per-state d bands
 * dmin/dmax (the CM's dmin1/dmax1 or dmin2/dmax2; both NULL = non-banded,
 * 0..W). Hits after greedy reporting with null3 and overlap removal, in
 * window coordinates; IXG_ERANGE if more than cap. */
int ixg_qdb_scan(ixg_ctx *ctx, const ixg_cm *cm, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0,
                 const int32_t *dmin, const int32_t *dmax, int inside, float cutoff, int do_null3, ixg_hit *hits,
                 int64_t cap, int64_t *nhits);
/* Batched form: n windows of a device-resident target (strand 1 = reverse
 * complement coordinates); hits tagged with env = window index, grouped by
 * window in order (the per-envelope loop of final_stage_inside_opt). */
typedef struct { int32_t i0, j0, strand, seq; } ixg_qdbwin;  /* seq as in ixg_envelope */
int ixg_qdb_scan_batch(ixg_ctx *ctx, const ixg_cm *cm, const ixg_target *tgt, const ixg_qdbwin *win, int64_t n,
                       const int32_t *dmin, const int32_t *dmax, int inside, float cutoff, int do_null3,
                       ixg_hit *hits, int64_t cap, int64_t *nhits);
/* cm_nohmm.rs:855 cyk_seq_filter(cm, dsq, n, cutoff, do_null3): CYK over
 * dsq[1..n] with dmin1/dmax1, hits padded by W-1 and merged into envelopes
 * (es[k], ee[k]); *nenv = count (IXG_ERANGE above cap). */
int ixg_cyk_seq_filter(ixg_ctx *ctx, const ixg_cm *cm, const uint8_t *dsq, int64_t n, const int32_t *dmin1,
                       const int32_t *dmax1, float cutoff, int do_null3, int32_t *es, int32_t *ee, int64_t cap,
                       int64_t *nenv);

/* ---- QDB / non-banded truncated scanners (cm_dpsearch_trunc.rs generic_tr_scan) -- */
/* Integer truncated-begin penalties (TrPenalties ig/il, cm_trunc.rs:145-150,
 * Prob2Score of the same probabilities as g/l): [3][M] each, same row order
 * as ixg_cm_set_trpenalties, -987654321 where a truncated begin is not
 * allowed. Required by the integer Inside scanner. */
int ixg_cm_set_itrpenalties(ixg_ctx *ctx, ixg_cm *cm, const int32_t *ig, const int32_t *il);
/* Host copies of the integer marginal emission scores ilmesc/irmesc (M*18
 * each, cp9.rs:4441 fill_marginal_emissions). Either may be NULL. */
int ixg_cm_imtables(const ixg_cm *cm, int32_t *ilmesc, int32_t *irmesc);
/* cm_dpsearch_trunc.rs:207 ref_tr_cyk_scan (inside = 0: float scores and
 * penalties) and :226 ref_itr_inside_scan (inside = 1: integer scores,
 * integer marginals and penalties, ILogsum) over dsq[i0..j0], i.e.
 * generic_tr_scan (:1800): the J/L/R/T planes with per-state dmax (the CM's
 * dmax1 or dmax2; NULL = non-banded), truncated begins into root rows kept
 * across j, greedy reporting with cm_hit_AllowTruncation and null3, overlap
 * removal. Writes vsc_root and its mode, and bsc_full (the best root score
 * of the full window, reported by the reference as the fourth result;
 * may be NULL). pass_idx: a truncated pass (IXG_PASS_5P_ONLY_FORCE ..
 * IXG_PASS_5P_AND_3P_ANY); local selects the local penalties and truncation
 * rules. Hits beyond cap are counted in *nhits and IXG_ERANGE is returned. */
int ixg_trqdb_scan(ixg_ctx *ctx, const ixg_cm *cm, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0,
                   const int32_t *dmax, int inside, int pass_idx, int local, float cutoff, int do_null3, float *vsc,
                   int32_t *vmode, float *bsc_full, ixg_trhit *hits, int64_t cap, int64_t *nhits);
/* Batched form over a device-resident target: per-window vsc, vmode,
 * bsc_full (may be NULL); hits tagged with env = window index, grouped by
 * window in order. */
int ixg_trqdb_scan_batch(ixg_ctx *ctx, const ixg_cm *cm, const ixg_target *tgt, const ixg_qdbwin *win, int64_t n,
                         const int32_t *dmax, int inside, int pass_idx, int local, float cutoff, int do_null3,
                         float *vsc, int32_t *vmode, float *bsc_full, ixg_trhit *hits, int64_t cap,
                         int64_t *nhits);

/* ---- CP9 HMM banding (cp9.rs:1275 cp9_forward, :1995 cp9_backward, :1517
 * cp9_fb2hmmbands): the band-producing HMM's integer Forward and Backward
 * over a window and its posterior HMM bands ---- */
typedef struct ixg_cp9 ixg_cp9;
/* A configured CP9 HMM (cp9.rs:871-896 CP9 after cp9_logoddsify): M nodes;
 * otsc [(M+1)][12] in the reference's cp9O order MM IM DM BM MI II DI MD ID
 * DD ME MEL (cp9.rs:906-918); msc / isc [(M+1)][18] over the digitized
 * alphabet; el != 0 when local ends (CPLAN9_EL) are on, with el_selfsc,
 * has_el [M+1] and, per node k = 0..M+1, the el_from_ct[k] source nodes
 * el_from_idx[el_from_off[k] + c] (cp9.rs:1042-1076). Scores are scaled
 * log2-odds x 1000, -987654321 for impossible. */
typedef struct {
  int32_t M, el, el_selfsc, pad;
  const int32_t *otsc, *msc, *isc;
  const uint8_t *has_el;
  const int32_t *el_from_ct, *el_from_off, *el_from_idx;
} ixg_cp9_params;
/* Uploads the HMM (device-resident until ixg_cp9_free). IXG_EINVAL when a
 * local-end source node is not below the node it ends at (a CP9 built from
 * a CM never has one) or M exceeds 2100 (the kernels' per-window LDS). */
int ixg_cp9_create(ixg_ctx *ctx, const ixg_cp9_params *p, ixg_cp9 **out);
int ixg_cp9_free(ixg_ctx *ctx, ixg_cp9 *h);
/* mode bits: the Forward scans (a parse may start at any position,
 * do_fwd_scan), the Backward scans (do_bck_scan), the reported scores are the
 * whole window's (doing_align) rather than the best over positions.
 * cp9_seq2bands passes FWD_SCAN | BCK_SCAN in search and ALIGN in alignment
 * (cp9.rs:4063-4066); cp9_seq2bands_trunc passes each scan unless the pass
 * forces the first (Forward) or final (Backward) residue (cp9.rs:4147-4152).
 * The posterior pass sums the Backward's begin cells over every row when the
 * Forward scanned (did_fwd_scan). */
#define IXG_CP9_FWD_SCAN 1
#define IXG_CP9_BCK_SCAN 2
#define IXG_CP9_ALIGN 4
/* One window dsq[i0..j0] (1-based, sentinel-padded, length L): Forward and
 * Backward in the given mode, their matrices (each
 * [(j0-i0+2)][(M+1)], row ip = position i0+ip-1; any may be NULL), the two
 * best scores (bits) and the posterior HMM bands at tau: pn [6][M+1] =
 * pn_min_m, pn_max_m, pn_min_i, pn_max_i, pn_min_d, pn_max_d, -1 where the
 * state has no band (the CP9Bands fields cp9_fb2hmmbands fills). */
int ixg_cp9_bands(ixg_ctx *ctx, const ixg_cp9 *h, const uint8_t *dsq, int64_t L, int32_t i0, int32_t j0,
                  double tau, int mode, int32_t *fmx, int32_t *fix, int32_t *fdx, int32_t *bmx, int32_t *bix,
                  int32_t *bdx, float *fwd_sc, float *bwd_sc, int32_t *pn);
/* Batched form over a device-resident target (the per-envelope loop that
 * calls cp9_seq2bands, cm_search.rs:2740 / :2816): pn [n][6][M+1];
 * fwd_sc / bwd_sc [n] may be NULL. */
typedef struct { int32_t i0, j0, strand, seq; } ixg_cp9win;  /* seq as in ixg_envelope */
int ixg_cp9_bands_batch(ixg_ctx *ctx, const ixg_cp9 *h, const ixg_target *tgt, const ixg_cp9win *win, int64_t n,
                        double tau, int mode, float *fwd_sc, float *bwd_sc, int32_t *pn);

/* ---- F4/F5: glocal envelope definition (STD pass) -----------------------------
 * The glocal profile (p7_generic.rs:156 build_glocal_profile, bg 0.25) is
 * built from the same ixg_p7_hmm by ixg_p7_create and kept resident.
 *
 * ixg_gfb_scores replaces p7_generic.rs:519 p7_gforward and :939
 * p7_gbackward (nats, the 16,000-entry FLogsum table) over a batch of items
 * seq[off[i]+1 .. off[i]+len[i]]: the profile is configured multihit (or
 * unihit with IXG_GFB_UNIHIT) for length cfg_len[i] (len[i] when cfg_len is
 * NULL; p7_generic.rs:379 reconfig_length). sc[i]: the score; xmx (optional):
 * the special states [(len[i]+1)][E N J B C] of every item, concatenated;
 * mat (optional, Forward only, IXG_GFB_MATRIX): [(len[i]+1)][M | I | D][M+1]
 * per item, concatenated. */
#define IXG_GFB_BACKWARD 1
#define IXG_GFB_UNIHIT 2
#define IXG_GFB_MATRIX 4
int ixg_gfb_scores(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *seq, int64_t seq_len, const int64_t *off,
                   const int64_t *len, const int32_t *cfg_len, int64_t n, int mode, float *sc, float *xmx,
                   float *mat);
/* Thresholds of the envelope definition (cm_search.rs:2449 trunc_env_def):
 * F4 / F4b / F5 / F5b P-values against the glocal Forward tail (gfmu,
 * gflambda) in the STD pass and the local Forward tail (lftau, lflambda) in
 * the truncated passes; rt1 rt2 rt3 and ns of p7_generic.rs:1765
 * p7_domaindef_glocal (the multidomain traces use Easel's fast RNG, seed
 * 181); clen, the CM's consensus length, sets the 3' pass's score
 * correction ln(1/clen) (cm_search.rs:2476-2478). */
typedef struct {
  double f4, f4b, f5, f5b;
  double gfmu, gflambda;
  float rt1, rt2, rt3;
  int32_t ns, do_gfwdbias, do_edefbias;
  double lftau, lflambda;
  int32_t clen, pad;
} ixg_envopts;
/* Per-pass accounting of the stage (cm_search.rs add_gfwd / add_gfwdbias /
 * add_edef / add_edefbias). */
typedef struct {
  uint64_t n_past_gfwd, pos_past_gfwd, n_past_gfwdbias, pos_past_gfwdbias;
  uint64_t n_past_edef, pos_past_edef, n_past_edefbias, pos_past_edefbias;
} ixg_envacct;
/* Replaces the STD-pass call of cm_search.rs:2449 trunc_env_def(dsq, L,
 * merged windows, PLI_PASS_STD_ANY, ...) -> Vec<(es, ee)>: the F3b-surviving
 * windows of one chunk (dsq 1-indexed, L residues) to envelopes in chunk
 * coordinates, in the reference's order. */
int ixg_env_def(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, const ixg_window *wins,
                int64_t nwin, const ixg_envopts *o, int32_t *es, int32_t *ee, int64_t cap, int64_t *nenv,
                ixg_envacct *acct);
/* The same for any pass of trunc_env_def (cm_search.rs:2449-2600):
 * IXG_PASS_STD_ANY (gm), IXG_PASS_5P_ONLY_FORCE (Rgm, windows must start at
 * residue 1), IXG_PASS_3P_ONLY_FORCE (Lgm, must end at L, score correction
 * ln(1/o->clen)), IXG_PASS_5P_AND_3P_FORCE (Tgm, both); the truncated
 * passes' statistics are o->lftau / o->lflambda. The windows are one
 * chunk's in the reference's order: a truncated pass's profile carries
 * over from window to window as there. */
int ixg_env_def_pass(ixg_ctx *ctx, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, const ixg_window *wins,
                     int64_t nwin, int pass_idx, const ixg_envopts *o, int32_t *es, int32_t *ee, int64_t cap,
                     int64_t *nenv, ixg_envacct *acct);

/* ---- HMM -> CM band conversion (cp9.rs:2893-3879) ---------------------------
 * A host map of one CM and its CP9 (ixg_cm_params / ixg_cp9_params, the CP9
 * having one node per consensus column): the guide-tree walk and the CP9
 * terms the conversion reads. cm_local / cp9_local: local begins and ends on
 * (cmsearch's default configuration of both). */
typedef struct ixg_hbmap ixg_hbmap;
int ixg_hbmap_create(const ixg_cm_params *cm, int cm_local, const ixg_cp9_params *cp9, int cp9_local,
                     ixg_hbmap **out);
int ixg_hbmap_free(ixg_hbmap *m);
/* cp9.rs:3380 cp9_hmm2ijbands (non-truncated) with
 * hmm_bands_enforce_valid_parse: pn [6][hmm_m+1] (as ixg_cp9_bands returns
 * it; doctored in place when local begins/ends are off) -> per-state i and j
 * bands imin/imax/jmin/jmax [M] (-1/-2 empty). */
int ixg_hmm2ij_bands(const ixg_hbmap *m, int32_t *pn, int32_t i0, int32_t j0, int doing_search, int32_t *imin,
                     int32_t *imax, int32_t *jmin, int32_t *jmax);
/* cp9.rs:3815 ij2d_bands into the ixg_bands layout: hd_off [M+1], rows
 * hdmin/hdmax; *nrows = rows needed (IXG_ERANGE above cap). */
/* The truncated form (cp9_seq2bands_trunc's steps 3-4, cp9.rs:4130-4175):
 * from pn and the per-node posterior occupancy occ [hmm_m+1], the predicted
 * start / end nodes and marginal bands (cp9.rs:1755), the J/L/R/T
 * candidacy of every state for the pass (:1918), then cp9_hmm2ijbands with
 * do_trunc; jvalid..tvalid [M]; spep (may be NULL) receives sp1 sp2 ep1 ep2
 * rmarg_imin rmarg_imax lmarg_jmin lmarg_jmax. */
int ixg_hmm2ij_bands_trunc(const ixg_hbmap *m, int32_t *pn, const int32_t *occ, int32_t i0, int32_t j0,
                           int pass_idx, int32_t *imin, int32_t *imax, int32_t *jmin, int32_t *jmax,
                           uint8_t *jvalid, uint8_t *lvalid, uint8_t *rvalid, uint8_t *tvalid, int32_t *spep);
int ixg_ij2d_bands(const ixg_hbmap *m, const int32_t *imin, const int32_t *imax, const int32_t *jmin,
                   const int32_t *jmax, int do_trunc, int32_t *hd_off, int32_t *hdmin, int32_t *hdmax, int64_t cap,
                   int64_t *nrows);

/* ---- a search, F1 through F7 ------------------------------------------------
 * Replaces FaithfulSearcher::search (cm_search.rs:1300) for one resident
 * sequence: the (chunk x strand) tasks of cm_search.rs:1774, per task
 * trunc_pipeline_one_strand (:2214) with the STD pass (the filter cascade,
 * cm_pipeline.rs:1863; the envelope definition, cm_search.rs:2449;
 * cm_std_stage, :2707: CP9 bands iterated to the size limit (cp9.rs:4075),
 * HB CYK filter with its P-value gate and envelope redefinition, HB Inside
 * with greedy reporting, null3 and per-envelope overlap removal) and, with
 * trunc_mode 1 (cmsearch's default), the terminal truncated passes: 5' on
 * the first maxW residues of a chunk that holds the sequence's 5' end, 3'
 * on the last maxW residues of one that holds its 3' end, 5'&3' on a chunk
 * of at most maxW residues holding both (cm_search.rs:2265-2280, :2380-2420),
 * each with its own filter cascade, envelope definition (Rgm / Lgm / Tgm)
 * and cm_trunc_stage (:2863: cp9_seq2bands_trunc, TrCYK filter, FTrInside);
 * then finalize (cm_search.rs:1915: genome coordinates, E-values, global
 * overlap removal, the E-value sort, the reporting threshold). The per-hit
 * alignment (mdl from/to, trunc) is not computed. */
typedef struct {
  ixg_f3opts f3;                   /* F1..F3b thresholds, cmw, maxw, IXG_DO_* */
  ixg_envopts env;                 /* F4..F5b */
  double f6, cyk_mu, cyk_lambda;   /* F6 P-value against the local CYK tail */
  float cyk_env_cutoff;            /* CYK envelope redefinition cutoff (bits) */
  float pli_t;                     /* final reporting threshold (bits) */
  double fin_mu, fin_lambda, ez;   /* E-value = exp_surv(score, fin_mu, fin_lambda) * ez */
  double e_report;                 /* hits with E > e_report are dropped */
  double fcyk_tau, final_tau, maxtau;
  float size_limit;                /* HB matrix limit in Mb (cm_search.rs:890-896) */
  int32_t do_fcyk, do_fcykenv, do_null3, do_cyk, do_top, do_bot;
  int32_t trunc_mode;              /* 0: STD pass only (--notrunc); 1: + the terminal truncated passes */
  int32_t pad;
} ixg_searchopts;
typedef struct {
  int64_t start, stop;  /* genome coordinates (start > stop on the bottom strand) */
  float score, bias;
  double evalue;
  int32_t strand, pass; /* pass: IXG_PASS_* that found the hit */
  int64_t seq;          /* the target's sequence (0 for a one-sequence target) */
} ixg_searchhit;
typedef struct {
  uint64_t nres;        /* residues searched: both strands, overlap context excluded */
  ixg_acct f;           /* F1..F3b, summed over tasks */
  ixg_envacct e;        /* F4..F5b */
  uint64_t n_past_cyk, pos_past_cyk, n_hits;
  /* the truncated passes, [0] 5', [1] 3', [2] 5'&3': residues charged
   * (min(maxW, chunk) per chunk that runs the pass), filter, envelope and
   * CYK accounting (cm_search.rs:2283-2300) and the hits they contribute
   * before the global overlap removal */
  uint64_t tr_nres[3];
  ixg_acct tr_f[3];
  ixg_envacct tr_e[3];
  uint64_t tr_n_past_cyk[3], tr_pos_past_cyk[3], tr_hits[3];
} ixg_searchacct;
int ixg_search_std(ixg_ctx *ctx, const ixg_p7 *p7, const ixg_cm *cm, const ixg_cp9 *cp9, const ixg_hbmap *map,
                   const ixg_target *tgt, const ixg_searchopts *o, ixg_searchhit *out, int64_t cap, int64_t *n,
                   ixg_searchacct *acct);
/* The same with the truncated passes when o->trunc_mode is 1: trcp9[k] and
 * trmap[k] (k = 0 5', 1 3', 2 5'&3') are the pass's CP9 (Rcp9 / Lcp9 /
 * Tcp9 of cm_search.rs:953-960, local ends as the search's configuration)
 * and its band map; the CM's truncation penalties must be set
 * (ixg_cm_set_trpenalties). ixg_search_std is this with trunc_mode 0.
 * Over a multi-sequence target (ixg_target_create_multi) it is the
 * reference's search over several sequences (cm_search.rs:1300): one flat
 * task list, overlap removal per sequence and strand, the E-value sort
 * keyed (E, score desc, sequence, start, pass desc); hits name their
 * sequence. o->ez is the caller's, from the whole database's size. */
int ixg_search(ixg_ctx *ctx, const ixg_p7 *p7, const ixg_cm *cm, const ixg_cp9 *cp9, const ixg_hbmap *map,
               const ixg_cp9 *const *trcp9, const ixg_hbmap *const *trmap, const ixg_target *tgt,
               const ixg_searchopts *o, ixg_searchhit *out, int64_t cap, int64_t *n, ixg_searchacct *acct);
/* Host wall time (ms) of each stage of the last ixg_search call on the
 * context (IXG_SS_* index; a stage's time includes its launches, waits and
 * host work) and the HB cells its STD CYK (cells[0]) and Inside (cells[1])
 * launches computed. The truncated passes' CM stages run concurrently with
 * the STD pass's F6 / F7 (each pass on an auxiliary context's stream):
 * IXG_SS_TRUNC is the longest pass's wall time on its thread (0 when no
 * truncated pass had envelopes), and IXG_SS_FINALIZE includes what of it
 * was left after F7. A measurement hook with no
 * counterpart in the reference. */
enum { IXG_SS_F1_F3B = 0, IXG_SS_F4_F5 = 1, IXG_SS_F6_BANDS = 2, IXG_SS_F6_CYK = 3, IXG_SS_F7_BANDS = 4,
       IXG_SS_F7_INSIDE = 5, IXG_SS_FINALIZE = 6, IXG_SS_TOTAL = 7, IXG_SS_TRUNC = 8, IXG_SS_N = 9 };
int ixg_search_stats(ixg_ctx *ctx, double *ms, int64_t *cells);

/* Per-kernel device time, cumulative since the last reset: HIP events around
 * every launch of each kernel (IXG_K_* index), total milliseconds and launch
 * counts; reset != 0 clears them after reading. A measurement hook with no
 * counterpart in the reference. The events are recorded without waiting and
 * read when the timings are queried (or a pool of 256 launches fills). */
enum { IXG_K_SSV_SWEEP = 0, IXG_K_SSV_RESOLVE = 1, IXG_K_FWD = 2, IXG_K_BIAS = 3, IXG_K_VIT = 4, IXG_K_MSV = 5,
       IXG_K_HB_SCAN = 6, IXG_K_TRHB_SCAN = 7, IXG_K_COPY = 8, IXG_K_QDB_SCAN = 9, IXG_K_FWD_FULL = 10,
       IXG_K_TRQDB_SCAN = 11, IXG_K_CP9 = 12, IXG_K_GFWD = 13, IXG_K_GBWD = 14, IXG_K_GFB = 15,
       IXG_NKERNELS = 16 };
int ixg_kernel_timings(ixg_ctx *ctx, double *ms, int64_t *launches, int reset);
/* Launch-shape autotuning (no counterpart in the reference): the F3 Forward
 * has three launch shapes with identical scores (stripe per wave with 64
 * windows per wave, 4 lanes per window, the generic LDS form). mode 1 (the
 * default) times every applicable shape (a warm run, then a timed one) on
 * the first batch of each (stripe count, log2 window count) class of at
 * least 256 windows and keeps the fastest for the context; mode 0 takes the
 * fixed order; mode 2 + v forces shape v. */
int ixg_set_autotune(ixg_ctx *ctx, int mode);
/* The context's choices so far: (stripe count, log2 window count class, shape). */
int ixg_autotune_choices(ixg_ctx *ctx, int32_t *nq, int32_t *cls, int32_t *shape, int64_t cap, int64_t *n);

/* Measurement hook (no counterpart in the reference): copies `bytes`
 * device-to-device with a plain kernel reps times after one warm-up copy;
 * *ms_per_copy = mean kernel time. Its counters calibrate HBM traffic
 * readings (FETCH_SIZE / WRITE_SIZE) and its rate is the achievable copy
 * bandwidth. */
int ixg_copy_probe(ixg_ctx *ctx, size_t bytes, int reps, double *ms_per_copy);

/* Per-stage device timings (milliseconds, HIP events) of the most recent
 * ixg_f3_filter_tasks call on ctx: [0] F1 sweep, [1] F1 resolve, [2] F2/F3/F3b
 * window kernels, [3] host glue, [4] total. Also the launch counts. */
int ixg_last_timings(const ixg_ctx *ctx, double *ms5, int64_t *cells5);

/* ---- serving: continuous batching of search requests --------------------------
 * No counterpart in the reference, which runs each (model, sequence) search to
 * completion. A request is the filter stage of one search: the sequence's
 * (chunk x strand) tasks on both strands (cm_search.rs:1774-1808) filtered
 * with f3_filter_sequence (cm_pipeline.rs:1863) against a resident p7 model;
 * its result equals ixg_window_tasks(L, o->maxw, 1, 1) + ixg_f3_filter_tasks
 * on the sequence alone. The server owns nstreams worker contexts on owner's
 * device (mem_budget/nstreams bytes each, 0 = no cap); each idle worker takes
 * the oldest queued request and every later queued one for the same model and
 * options, up to max_batch residues, as one batch. The head of the queue is
 * admitted once the residues in flight plus its own fit max_inflight (alone
 * when larger). Models stay owned by the caller and must outlive the server. */
typedef struct ixg_server ixg_server;
int ixg_server_create(ixg_ctx *owner, int nstreams, size_t mem_budget, int64_t max_inflight, int64_t max_batch,
                      ixg_server **out);
/* Copies dsq[1..L] (1-indexed) and queues the request; returns its ticket. */
int ixg_server_submit(ixg_server *s, const ixg_p7 *p7, const uint8_t *dsq, int64_t L, const ixg_f3opts *o,
                      uint64_t *ticket);
/* *done = 1 once the request has finished; waits up to timeout_us (0 polls,
 * negative waits until done). */
int ixg_server_wait(ixg_server *s, uint64_t ticket, int64_t timeout_us, int *done);
/* Waits for the request, then writes its tasks, windows (task index local to
 * the request, grouped by task), per-task acct and times_ms[4] = {queue wait,
 * run, submit-to-done latency, requests in its batch}, and retires the
 * ticket. IXG_ERANGE (ticket kept) if tcap or wcap is too small; a failed
 * batch's status is returned for each of its requests. */
int ixg_server_result(ixg_server *s, uint64_t ticket, ixg_task *tasks, int64_t tcap, int64_t *ntasks,
                      ixg_window *wins, int64_t wcap, int64_t *nwins, ixg_acct *acct, double *times_ms);
/* A search request: the whole search (ixg_search) of dsq[1..L] against one
 * resident model bundle with the caller's options; o->ez and the thresholds
 * are the caller's (as cmsearch -Z fixes the E-value scale), so a request's
 * result does not depend on what it is batched with. Queued requests for
 * the same bundle and options (compared bytewise) run as one ixg_search over
 * a multi-sequence target of their sequences. A served search of cmsearch
 * (cm_search.rs:1300) with no counterpart in the reference's control flow. */
typedef struct {
  const ixg_p7 *p7;
  const ixg_cm *cm;
  const ixg_cp9 *cp9;
  const ixg_hbmap *map;
  const ixg_cp9 *trcp9[3];   /* the truncated passes' (used when o->trunc_mode is 1) */
  const ixg_hbmap *trmap[3];
} ixg_search_model;
int ixg_server_submit_search(ixg_server *s, const ixg_search_model *m, const uint8_t *dsq, int64_t L,
                             const ixg_searchopts *o, uint64_t *ticket);
/* Waits for a search request, then writes its hits (ixg_search's for the
 * sequence alone, seq = 0) and times_ms[4] as ixg_server_result, and retires
 * the ticket; IXG_ERANGE (ticket kept) if cap is too small. */
int ixg_server_search_result(ixg_server *s, uint64_t ticket, ixg_searchhit *hits, int64_t cap, int64_t *n,
                             double *times_ms);
/* stats4 = {batches, requests, largest batch in requests, residues searched}. */
int ixg_server_stats(ixg_server *s, int64_t *stats4);
/* Runs every queued request to completion, then stops the workers. */
int ixg_server_free(ixg_server *s);

#ifdef __cplusplus
}
#endif
#endif
