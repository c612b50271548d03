// Internal declarations of the covariance-model scanners.
#pragma once
#include "ixg_internal.h"

namespace ixg {

constexpr int MAXC = 6;
constexpr float IMPOSSIBLE = -1e36f;  // cp9.rs:4189
constexpr int FLOGSUM_TBL = 23000;     // cp9.rs:5388
constexpr int INFTY_I = 987654321;     // cp9.rs:21, the integer scores' -inf
enum { D_ST = 0, MP_ST = 1, ML_ST = 2, MR_ST = 3, IL_ST = 4, IR_ST = 5, S_ST = 6, E_ST = 7, B_ST = 8 };
enum { BIF_ND = 0, MATP_ND = 1, MATL_ND = 2, MATR_ND = 3, BEGL_ND = 4, BEGR_ND = 5, ROOT_ND = 6, END_ND = 7 };

// Host copy of a configured CM (the tables the scanners read).
struct CmHost {
  int M = 0;
  int local_begin = 1;
  float el_selfsc = 0.f;
  float null[4];
  double n3_omega = 0;
  std::vector<int8_t> sttype;
  std::vector<int32_t> cfirst, cnum;
  std::vector<float> tsc, oesc, beginsc, endsc;  // M*6, M*324, M, M
  std::vector<int32_t> itsc, ioesc, ibeginsc, iendsc;  // integer scores (same shapes)
  int32_t iel_selfsc = 0;
  // QDB scanner inputs (cm_nohmm.rs:406-532): W, BEGL_S deck index per state
  // (-1 otherwise), per semiring the d = 0 base value of each state and deck
  // and el_sca[0..W], and the states y >= 1 with a local begin
  int W = 0;
  std::vector<int32_t> qbegl, qibase0, qiel, qibbase0, qblist, ilog;
  std::vector<float> qfbase0, qfel, qfbbase0;
  std::vector<int32_t> begin_list;               // states y >= 1 with a local begin
  // dependency levels: a state's level is one more than its children's
  // highest (self transitions excluded), so the states of one level are
  // mutually independent; lvl_states lists states level by level, level l
  // at lvl_off[l] .. lvl_off[l+1]-1
  std::vector<int32_t> lvl_states, lvl_off, lvl;  // lvl: each state's level
  // truncated scanning
  int clen = 0;
  std::vector<float> lmesc, rmesc;  // M*18 marginal emission scores
  std::vector<int32_t> lpos, rpos;  // per state: consensus span of its subtree
  std::vector<float> trp_g, trp_l;  // 3*M truncated-begin penalties (empty until set)
  std::vector<int32_t> ilmesc, irmesc;  // M*18 integer marginal emission scores
  std::vector<int32_t> trp_ig, trp_il;  // 3*M integer truncated-begin penalties (empty until set)
};
void marginal_emissions(const ixg_cm_params &p, CmHost &h);
int state_spans(const ixg_cm_params &p, CmHost &h);
int configure_cm(const ixg_cm_params &p, int local, CmHost &h);
void dependency_levels(const ixg_cm_params &p, CmHost &h);
int qdb_tables(const ixg_cm_params &p, CmHost &h);
float null3_score(const CmHost &h, const std::vector<float> &ftbl, const float *comp, int len);
void flogsum_table(std::vector<float> &t);
float host_flogsum(const std::vector<float> &t, float a, float b);

struct CmDev {
  const int8_t *sttype;
  const int32_t *cfirst, *cnum;
  const float *tsc, *oesc, *beginsc, *endsc;
  const float *lmesc, *rmesc;
  const int32_t *begin_list;
  int nbegin;
  const int32_t *lvl_states, *lvl_off;
  int nlev;
  int M;
  int local_begin;
  float el_selfsc;
};

// One envelope of a batched banded scan. Band and layout arrays are indexed
// like the reference's CP9Bands: jmin/jmax per state, hd_off per state (M+1),
// and per (v, j - jmin[v]) at hd_off[v] + jp: hdmin, hdmax and roff, the
// offset of the row's first cell in alpha.
// The truncated scanners add four planes (J, L, R, T) of nc cells each at
// alpha + p * nc, a per-state mode mask (bit 0 J, 1 L, 2 R, 3 T, with the
// pass's fill flags applied) and, per root cell, the entry state and mode of
// the best truncated begin (state * 8 + mode, int32 bits in rout).
struct EnvDev {
  const uint8_t *dsq;
  const int32_t *jmin, *jmax, *hd_off, *hdmin, *hdmax, *roff;
  float *alpha;
  int32_t i0, j0;
  int32_t nc;
  const int32_t *mask;
  float *rout;  // root rows out: standard nroot floats; truncated J|L|R|T|aux, nroot each
  // per state: work units of one level step (cells; rows of an IL state;
  // i diagonals of an IR state; 0 to skip) and the first i of an IR state
  const int32_t *units, *ibase;
  int32_t cmi;  // model of the envelope in a several-model launch (index into its CmDev array)
  // standard scans (wavefront over d), packed contiguously from rect: per
  // state HB_RECT ints (jlo, nj, ilo, nx, deck base, level * 4 + 2 when the
  // band is its (i, j) box + 1 when x counts j (0: x counts i), dlo, nd,
  // first band row, state type | child count << 4 | first child << 8, i band
  // width, 0); the nbegin local-begin states and
  // their begin scores (float bits); per root row the offset of its first
  // cell in rout; per band row its d band (hdmin, hdmax). nsteps: the
  // wavefront's steps.
  const int32_t *rect, *rootoff, *rowx;
  int32_t nbegin, nsteps;
};

constexpr int HB_THREADS = 1024;     // lanes per envelope workgroup (standard scans)
constexpr int TRHB_THREADS = 512;    // lanes per envelope workgroup (truncated scans)
constexpr int HB_LDS_STATES = 3200;  // states whose per-state band scalars fit in LDS

constexpr int HB_RECT = 12;  // ints per state of a standard scan's rectangle table
// LDS ints of a standard scan's per-step prefix sum (pre and jst, M each, and
// 32 wave sums), rounded up to 4 so that the rectangle table after it is
// 16-byte aligned for its int4 reads
__host__ __device__ constexpr int hb_scan_ints(int M) { return (2 * M + 32 + 3) & ~3; }
// LDS ints of an envelope's residues i0..j0 (bytes), rounded up to 16 bytes
__host__ __device__ constexpr int hb_seq_ints(int len) { return ((len + 15) / 16) * 4; }
// lds_ints: LDS ints per workgroup: the per-step scan (hb_scan_ints), the
// residues (hb_seq_ints) and, when they fit, the Inside FLogsum table,
// rectangle table and begin list.
hipError_t launch_hb_scan_queue(const CmDev *cms, const EnvDev *envs, const int32_t *order, int64_t n,
                                uint32_t *counter, int grid, const float *ftbl, int inside, size_t lds_ints,
                                hipStream_t s);
hipError_t launch_hb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, int inside,
                          size_t lds_ints, hipStream_t s);
// pty: the pass's truncated-begin penalty per state; ylist: the ny states
// y >= 1 with an allowed truncated begin, ascending.
hipError_t launch_trhb_scan(const CmDev &cm, const EnvDev *envs, int64_t n, const float *ftbl, const float *pty,
                            const int32_t *ylist, int ny, int inside, hipStream_t s);

// ---- QDB / non-banded scanners (cm_nohmm.rs generic_scan) ----

// Score tables of one semiring (float CYK or integer Inside): T is float or
// int32; base0 holds each state's d = 0 value, el_sca[d] the EL score of d
// residues (repeated additions), bbase0 the d = 0 value of each BEGL_S deck.
template <typename T>
struct QdbTabs {
  const T *tsc, *oesc, *endsc, *beginsc, *base0, *el_sca, *bbase0;
};
enum QdbUnitKind { QU_NONE = 0, QU_CELL = 1, QU_IL = 2, QU_ROOT = 3 };
// One lane's work at a step: a cell (state, d), an IL row, or a root cell,
// with everything the lane needs to address it (no dependent metadata loads).
// A ring is packed as first row | log2(depth) << 27.
struct QdbUnit {
  int32_t code;      // kind << 28 | state
  int32_t d;         // the cell's d (IL: the row's first d)
  int32_t lvl;       // the state's level: it works on j = i0 + t - lvl at step t
  int32_t info;      // sttype | child count << 4 | (local-begin slot + 1) << 8
  int32_t y0;        // first child (bifurcation: the left child, a BEGL_S)
  int32_t y1;        // bifurcation: the right child; IL: the row's last d
  int32_t ring;      // the state's own ring
  int32_t deck;      // BEGL_S and bifurcation: the deck index; -1 otherwise
  int32_t cring[6];  // the children's rings; bifurcation: right child's ring, deck offset, deck gap
  int32_t k0, k1;    // bifurcation: split range; BEGL_S: deck offset, deck gap
};
static_assert(sizeof(QdbUnit) == 64, "QdbUnit is four 16-byte loads");
constexpr int RING_SHIFT = 27;

// The wavefront schedule of one scan call: at step t every state v works on
// the end position j = i0 + t - lvl[v] (lvl: one more than its highest
// child). All of a step's cells read only cells written at earlier steps, so
// one barrier per step suffices. Per state the rows of the j's still to be
// read live in a ring (a power of two deep); each BEGL_S deck is a ring of
// w+1+gap rows; local begins are folded into per-(j, slot, d) maxima that
// the root reads and clears.
struct QdbDev {
  const int8_t *sttype;
  const int32_t *cfirst, *cnum;
  const int32_t *begl;  // per state: BEGL_S deck index, -1 otherwise
  const int32_t *dn, *dxcap;  // per state, this call's bands: max(dmin, emissions), min(dmax, W)
  const int32_t *dminw;       // per state, min(dmin, W): a bifurcation's split range
  const int32_t *ilog;        // ILogsum table
  const int32_t *lvl;         // per state: dependency level
  const int32_t *acc;         // per state: local-begin slot, -1 without a local begin
  const int2 *ring;           // per state: {first ring row, ring depth - 1}
  const int32_t *row_state;   // per ring row: its state
  const int2 *deck;           // per BEGL_S deck: {extra rows before it, extra depth beyond w+1}
  const int4 *item;           // per item of 64 units: {lowest level, highest level, lowest d, 0}
  const QdbUnit *unit;        // [item][64]
  const int32_t *wave_item;   // per wave: its items wave_item[w] .. wave_item[w+1]-1
  int M, W, nbegl, local, nw;
  int nrows;       // ring rows over all states
  int deck_extra;  // extra deck rows over all decks
  int nslot, amask;  // local-begin slots per j, ring depth - 1 of those rows
  int lroot;         // the root's level
};
constexpr int QDB_MAX_WAVES = 16, QDB_MAX_THREADS = 64 * QDB_MAX_WAVES;
struct QdbWinDev {
  const uint8_t *dsq;
  int32_t i0, j0, w, win;
  void *rows;  // standard scanners: state rings, decks, local-begin rows; truncated: plane rows
  void *begl;  // truncated scanners: the BEGL_S decks
};
struct QdbCand {  // a root cell at or above the reporting cutoff
  int32_t win, j, d;
  float sc;
};
// maxlen: the longest window of the launch (windows up to 32768 residues are
// staged in LDS)
hipError_t launch_qdb_scan(const QdbDev &q, const void *tabs, int inside, const QdbWinDev *wins, int64_t n,
                           float cutoff, QdbCand *cand, int64_t cap, unsigned long long *counter, int maxlen,
                           hipStream_t s);

// The QDB / non-banded truncated scanners (generic_tr_scan).
template <typename T>
struct TrqTabs {
  const T *tsc, *oesc, *endsc, *el_sca, *base0, *bbase0;
  const T *lmesc, *rmesc;  // marginal emission scores [M][18]
  const T *pty;            // the pass's truncated-begin penalties [M]
};
// The truncated scanners' wavefront schedule: as QdbDev, with four planes
// (J, L, R, T) per ring row and three (J, L, R) per deck row; the root keeps
// its two parity rows of each plane for the whole window.
struct TrqDev {
  const int8_t *sttype;
  const int32_t *cfirst, *cnum;
  const int32_t *begl;   // per state: BEGL_S deck index, -1 otherwise
  const int32_t *dxcap;  // per state: min(dmax, W)
  const int32_t *ilog;   // ILogsum table
  const int32_t *lvl, *row_state, *wave_item;
  const int2 *ring, *deck;
  const QdbUnit *unit;
  const int4 *item;
  const int4 *tblist;  // the root's begin states: {ring, penalty bits, dxcap, state | bifurcation << 30}
  int ntb;
  int M, W, nbegl;
  int fill;       // bit 1 L, bit 2 R, bit 3 T planes filled by the pass
  int nonbanded;  // no dmax given: bifurcation splits over every k
  int nw, nrows, deck_extra, lroot;
};
struct TrqCand {  // a root cell whose best truncated begin is at or above the cutoff
  int32_t win, j, d;
  float sc;
  int32_t root, mode;
};
struct TrqSummary {  // per window: best root score (first in (j, d) order), its mode, full-window best
  float vsc;
  int32_t vmode;
  float bsc_full;
};
hipError_t launch_trqdb_scan(const TrqDev &q, const void *tabs, int inside, const QdbWinDev *wins, int64_t n,
                             float cutoff, TrqCand *cand, int64_t cap, unsigned long long *counter, TrqSummary *summ,
                             int maxlen, hipStream_t s);

// ---- CP9 HMM banding (cp9.rs:1275 cp9_forward, :1995 cp9_backward, :1517
// cp9_fb2hmmbands) ----
struct Cp9Dev {
  int M, el, el_selfsc;
  int nz;                           // ILogsum table entries up to the last nonzero one
  const int32_t *otsc, *msc, *isc;  // [(M+1)][12] cp9O order, [(M+1)][18], [(M+1)][18]
  const uint8_t *has_el;            // [M+1]
  const int32_t *el_ct, *el_off, *el_idx;  // local-end sources per node 0..M+1
  // The same lists per source node y, for the Backward's local-end cells:
  // how often y ends at node M, whether it ends at M+1 (a scan resets the
  // cell there), and the other nodes it ends at, in decreasing order.
  const int32_t *bel_nM, *bel_reset, *bel_off, *bel_k;  // [M+1], [M+1], [M+2], list
  const int32_t *ilog;                                  // ILogsum table
  // LDS slots of the wavefront kernels: per node, the slot holding its
  // local-end cell while a later node of the next row may read it (Forward)
  // and the slot holding its match cell while a local-end cell of the next
  // row may read it (Backward); -1 for nodes that need none
  const int32_t *fslot, *bslot;  // [M+1] each
  int nsf, nsb;                  // slots in use
  // per entry of the local-end lists: the slot of el_idx's source (Forward)
  // and of bel_k's end node (Backward); list lengths
  const int32_t *el_slot, *bel_slot;
  int nel, nbel;
  // per node, two packed words staged in LDS by the wavefront kernels.
  // Forward (nodes 0..M+1): list offset << 16 | list length; slot (16 bits,
  // 0xFFFF = none) | has_el << 16. Backward (nodes 0..M): list offset << 16 |
  // list length; slot | has_el << 16 | ends at M+1 << 17 | times it ends at
  // M << 18.
  const uint32_t *finfo, *binfo;  // [2][M+2] each
  int wave;                       // bit 0: the Forward wavefront may run, bit 1: the Backward
};
struct Cp9WinDev {
  const uint8_t *dsq;
  int32_t i0, j0;
  int64_t out;  // the window's index in the caller's order
  int32_t *mx;  // Forward M, I, D then Backward M, I, D, each [(L+1)][(M+1)], then 64 words
                // the wavefront kernels' idle lanes store to
};
// Dynamic LDS of the CP9 kernels: two local-end rows, the two serial chains'
// terms, a transition column, the ILogsum table as 16-bit entries.
inline size_t cp9_lds_bytes(int M, int nz) { return (size_t)5 * (M + 1) * 4 + (((size_t)(nz + 1) * 2 + 15) & ~(size_t)15); }
// Dynamic LDS of the wavefront kernels: nslots slots of 65 lanes, two
// local-end rows and two match rows, the packed node words, the nlist entries of the local-end lists (node and slot,
// 16 bits each), the ILogsum table (nz entries and a zero).
inline size_t cp9_wave_lds_bytes(int M, int nz, int nslots, int nlist) {
  return (size_t)nslots * 65 * 4 + (size_t)4 * (M + 1) * 4 + (size_t)2 * (M + 2) * 4 + (size_t)nlist * 4 +
         (((size_t)(nz + 1) * 2 + 15) & ~(size_t)15);
}
constexpr size_t CP9_LDS_MAX = 65536 - 64;
hipError_t launch_cp9_bands(const Cp9Dev &h, const Cp9WinDev *wins, int64_t n, int mode, int32_t thresh,
                            float *fsc, float *bsc, int32_t *pn, int32_t *occ, hipStream_t s);

}  // namespace ixg
